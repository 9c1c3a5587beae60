"""GPU: the BF16 and FP8 (e4m3) matrix-core likelihood against the oracle.

Every pass's likelihood is approximated on the matrix cores and each
8-position group that the error bound cannot prove negative is recomputed with
the exact FP32 chain, so spikes, labels, order and the residual must equal the
oracle's bit for bit, for both formats and every screen path."""
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


def run_fmt(h, V, T, bias, thr, n, fmt, screen="split", recheck=True):
    h.set_templates(T, bias, thr)
    flags = FMTS[fmt] | (_native.FLAG_OVERLAP_RECHECK if recheck else 0)
    if screen == "valu":
        flags |= _native.FLAG_SCREEN_FP32
    elif screen == "literal":
        flags |= _native.FLAG_RECHECK_LITERAL
    print(f"[precision] {fmt} screen={screen} C={V.shape[0]} N={V.shape[1]} U={T.shape[0]} L={T.shape[2]} n={n}",
          flush=True)
    h.run(V, n, flags)
    idx, lab = h.result()
    return idx, lab, h.residual(V.shape[1])


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("seed", range(6))
def test_noisy_cases_match_oracle(handle, fmt, seed):
    V, T, bias, thr, on, lab = cases.noisy_case(seed)
    n = T.shape[2] - 1
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt), fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("C,L,U,N", [(2, 8, 3, 600), (5, 20, 17, 1500), (9, 33, 40, 2000), (12, 48, 16, 3000),
                                     (17, 64, 33, 2500), (3, 16, 1, 400), (64, 48, 20, 1200), (4, 17, 6, 900),
                                     (1, 1, 2, 300), (3, 49, 5, 1000)])
def test_shapes_match_oracle(handle, fmt, C, L, U, N):
    """Rows per window 1..4 (L up to 64), one, two and three k-blocks per channel, templates
    with more channels than one register load of fragments (dense 64-channel ones), tiles
    that end inside a window, windows at the end of the chunk."""
    V, T, bias, thr, on, lab = cases.noisy_case(70 + C + L, C=C, L=L, U=U, N=N, nsp=max(4, N // 25))
    n = L - 1
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt), fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


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


@pytest.mark.parametrize("fmt", sorted(FMTS))
@pytest.mark.parametrize("name,dur", [("small", 3.0), ("probe16", 0.5), ("probe64", 0.3)])
def test_workload_slices_match_oracle(handle, fmt, name, dur):
    d = synthetic.generate(name, duration_s=dur)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    assert_same(run_fmt(handle, V, T, bias, thr, n, fmt), fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))
    st = handle.stats()
    assert st["n_likelihood_mfma_launches"] == st["passes"]


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
