"""GPU: the BF16 and FP8 (e4m3) matrix-core likelihood against the oracle.

The dense first pass (by default; every pass with FLAG_MMA_EVERY_PASS) is
approximated on the matrix cores and each 8-position group that the error
bound cannot prove negative is recomputed with the exact FP32 chain (later
passes run the FP32 kernel on the changed pairs), so spikes, labels, order and
the residual must equal the oracle's bit for bit, for both formats, both pass
plans and every screen path."""
import numpy as np
import pytest

from oracle import fbp_oracle
from spikesorting_fullpursuit_amd import _native, synthetic
from spikesorting_fullpursuit_amd.overlap import binary_pursuit_parallel as bpp
from tests import cases
from tests.test_gpu_parity import assert_same
from tests.test_oracle_bookkeeping import CASES

pytestmark = pytest.mark.gpu

FMTS = {"bf16": _native.FLAG_PRECISION_BF16, "fp8": _native.FLAG_PRECISION_FP8}


@pytest.fixture(scope="module")
def handle():
    h = _native.Handle(0)
    yield h
    h.close()


def run_fmt(h, V, T, bias, thr, n, fmt, screen="split", recheck=True, every=False):
    h.set_templates(T, bias, thr)
    flags = FMTS[fmt] | (_native.FLAG_OVERLAP_RECHECK if recheck else 0)
    if every:
        flags |= _native.FLAG_MMA_EVERY_PASS
    if screen == "valu":
        flags |= _native.FLAG_SCREEN_FP32
    elif screen == "literal":
        flags |= _native.FLAG_RECHECK_LITERAL
    print(f"[precision] {fmt} every={every} screen={screen} C={V.shape[0]} N={V.shape[1]} U={T.shape[0]} L={T.shape[2]} n={n}",
          flush=True)
    h.run(V, n, flags)
    idx, lab = h.result()
    return idx, lab, h.residual(V.shape[1])


@pytest.mark.parametrize("every", [False, True])
@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("seed", range(6))
def test_noisy_cases_match_oracle(handle, fmt, seed, every):
    V, T, bias, thr, on, lab = cases.noisy_case(seed)
    n = T.shape[2] - 1
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt, every=every),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("C,L,U,N", [(2, 8, 3, 600), (5, 20, 17, 1500), (9, 33, 40, 2000), (12, 48, 16, 3000),
                                     (17, 64, 33, 2500), (3, 16, 1, 400), (64, 48, 20, 1200), (4, 17, 6, 900),
                                     (1, 1, 2, 300), (3, 49, 5, 1000)])
@pytest.mark.parametrize("every", [False, True])
def test_shapes_match_oracle(handle, fmt, C, L, U, N, every):
    """Rows per window 1..4 (L up to 64), one, two and three k-blocks per channel, templates
    with more channels than one register load of fragments (dense 64-channel ones), tiles
    that end inside a window, windows at the end of the chunk; the dense first pass where L
    is a multiple of 16 (a partial last span of 16 windows), the per-template kernel elsewhere."""
    V, T, bias, thr, on, lab = cases.noisy_case(70 + C + L, C=C, L=L, U=U, N=N, nsp=max(4, N // 25))
    n = L - 1
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt, every=every),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("screen", ["split", "valu"])
@pytest.mark.parametrize("seed", [0, 1])
def test_channel_dynamic_range_matches_oracle(handle, fmt, screen, seed):
    """Channels whose amplitudes span 10^-7 .. 10^3 of each other: FP8 operands lose the
    small channels entirely (below its sub-normals) and a spike may push the voltage past
    the format's range (NaN approximations); the bound and the NaN test must send every
    such group to the exact pass."""
    rng = np.random.default_rng(40 + seed)
    U, C, L, N = 6, 8, 24, 3000
    scale = np.array([1e3, 1.0, 1e-3, 1e-7, 3e2, 1e-5, 0.5, 2e-2], np.float32)
    T = cases.dog_templates(U, C, L, seed).astype(np.float32) * scale[None, :, None]
    V = (rng.normal(size=(C, N)).astype(np.float32) * 0.05) * scale[:, None]
    for p in range(30, N - 2 * L, 41):
        V[:, p:p + L] += T[(p // 41) % U] * np.float32(0.8 + 0.4 * rng.random())
    V[0, 1500:1510] *= 40.0  # far beyond the first pass's range after scaling
    bias, thr = cases.bias_thr(T.astype(np.float64), 0.05)
    n = L - 1
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt, screen=screen),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("every", [False, True])
@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("name,dur", [("small", 3.0), ("probe16", 0.5), ("probe64", 0.3)])
def test_workload_slices_match_oracle(handle, fmt, name, dur, every):
    d = synthetic.generate(name, duration_s=dur)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt, every=every),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))
    st = handle.stats()
    # the matrix cores take the first pass, or every pass with FLAG_MMA_EVERY_PASS
    assert st["n_likelihood_mfma_launches"] == (st["passes"] if every else 1)


@pytest.mark.parametrize("fmt", sorted(FMTS))
def test_without_recheck_matches_oracle(handle, fmt):
    d = synthetic.generate("small", duration_s=3.0)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt, recheck=False),
                fbp_oracle.pursuit(V, T, bias, thr, n, do_recheck=False, return_residual=True))


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_known_answer_traces(handle, fmt, name):
    c = CASES[name]()
    idx, lab, R = run_fmt(handle, c["V"], c["T"], c["bias"], c["thr"], c["n"], fmt)
    assert [tuple(r) for r in handle.pass_log().tolist()] == c["trace"]
    assert list(zip(idx.tolist(), lab.tolist())) == c["spikes"]
    np.testing.assert_array_equal(R, c["residual"])


@pytest.mark.parametrize("fmt", sorted(FMTS))
def test_dropin_precision_option(fmt):
    d = synthetic.generate("small", duration_s=2.0)
    m1, m2 = dict(d["separability_metrics"]), dict(d["separability_metrics"])
    a = bpp.binary_pursuit(d["voltage"], np.float32, d["sort_info"], m1, n_max_shift_inds=d["n_max_shift_inds"],
                           precision=fmt)
    b = bpp.binary_pursuit(d["voltage"], np.float32, d["sort_info"], m2, n_max_shift_inds=d["n_max_shift_inds"],
                           precision="fp32")
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


_FP64 = {}


def _fp64_case(name, dur, n):
    """A workload slice and the FP64 NumPy oracle's spikes on it (computed once)."""
    key = (name, dur, n)
    if key not in _FP64:
        from oracle import fbp_fp64
        d = synthetic.generate(name, duration_s=dur)
        T, bias, thr = synthetic.kernel_inputs(d)
        nn = d["n_max_shift_inds"] if n is None else n
        _FP64[key] = (d["voltage"], T, bias, thr, nn, fbp_fp64.pursuit(d["voltage"], T, bias, thr, nn))
    return _FP64[key]


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("name,dur,n", [("small", 1.0, None), ("probe16", 0.1, 8), ("probe64", 0.05, 3)])
def test_decisions_agree_with_fp64_oracle(handle, fmt, name, dur, n):
    """Decision agreement with the independent FP64 restatement (sliding-window products,
    energy-drop overlap values) on slices it finishes in seconds: the shift range is cut
    on the 16- and 64-channel slices (its overlap items are materialised)."""
    V, T, bias, thr, nn, (i64, l64) = _fp64_case(name, dur, n)
    idx, lab, _ = run_fmt(handle, V, T, bias, thr, nn, fmt)
    assert len(i64) > 20
    np.testing.assert_array_equal(idx, i64)
    np.testing.assert_array_equal(lab, l64)


def _first_pass(h, V, T, bias, thr, flags):
    h.set_templates(T, bias, thr)
    h.run(V, T.shape[2] - 1, flags | _native.FLAG_FIRST_PASS_ONLY)
    return h.likelihood(V.shape[1])


@pytest.mark.parametrize("name,dur", [("small", 1.0), ("probe64", 0.2)])
def test_first_pass_likelihood_fp32_equals_oracle(handle, name, dur):
    d = synthetic.generate(name, duration_s=dur)
    T, bias, thr = synthetic.kernel_inputs(d)
    LIK = _first_pass(handle, d["voltage"], T, bias, thr, _native.FLAG_SCREEN_FP32)
    np.testing.assert_array_equal(LIK, fbp_oracle.full_likelihood(d["voltage"], T, bias))


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("name,dur", [("small", 1.0), ("probe16", 0.3), ("probe64", 0.2)])
def test_first_pass_positive_values_exact(handle, fmt, name, dur):
    """After the exact rescoring every positive likelihood is the FP32 value bit for bit
    and every other value is <= 0; without the rescoring the approximations stay
    within a few units of the format's rounding of sum |t v|."""
    d = synthetic.generate(name, duration_s=dur)
    T, bias, thr = synthetic.kernel_inputs(d)
    V = d["voltage"]
    exact = fbp_oracle.full_likelihood(V, T, bias)
    got = _first_pass(handle, V, T, bias, thr, FMTS[fmt])
    pos = exact > 0
    np.testing.assert_array_equal(got[pos], exact[pos])
    assert np.all(got[~pos] <= 0)
    raw = _first_pass(handle, V, T, bias, thr, FMTS[fmt] | _native.FLAG_NO_RESCORE)
    eps = {"bf16": 2.0 ** -8, "fp8": 2.0 ** -4}[fmt]
    U, C, L = T.shape
    N = V.shape[1]
    # Q(t, x) = sum_c sum_j |T[t,c,j]| |V[c,x+j]|, the scale of the rounding error
    Q = np.zeros((U, N))
    A = np.abs(V.astype(np.float64))
    for t in range(U):
        for c in range(C):
            if np.any(T[t, c]):
                Q[t, :N - L + 1] += np.correlate(A[c], np.abs(T[t, c].astype(np.float64)), "valid")
    Q[:, N - L:] = 0.0
    err = np.abs(raw.astype(np.float64) - exact)
    assert np.all(err <= 4 * eps * Q + 1e-6 * np.abs(bias.astype(np.float64))[:, None] + 1e-30)
