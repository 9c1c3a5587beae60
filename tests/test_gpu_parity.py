"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Decisions (spike indices, labels, order) must be identical and the residual
bit-identical: both sides use the same FP32 operation order."""
import sys

import numpy as np
import pytest

from oracle import fbp_oracle
from spikesorting_fullpursuit_amd import _native, synthetic
from spikesorting_fullpursuit_amd.overlap import binary_pursuit_parallel as bpp
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def handle():
    h = _native.Handle(0)
    yield h
    h.close()


MODES = ["literal", "split", "valu", "stream"]


def gpu_run(h, V, T, bias, thr, n_max, recheck=True, clips=False, literal="literal", mixed=False):
    """literal: "literal" (every overlap item), "split" (screen with split-fp16
    matrix-core correlations, the default path), "valu" (screen with FP32
    vector correlations) or "stream" (split screen with the streaming
    selection kernel, the path for U*(2n+1) > 24576); booleans map to
    literal / split.  mixed: True takes the matrix-core first likelihood pass
    whatever the tiles' padding, "auto" lets the library choose it."""
    h.set_templates(T, bias, thr)
    flags = (_native.FLAG_OVERLAP_RECHECK if recheck else 0) | (_native.FLAG_ADJUSTED_CLIPS if clips else 0)
    mode = {True: "literal", False: "split"}.get(literal, literal)
    if mode == "literal":
        flags |= _native.FLAG_RECHECK_LITERAL
    elif mode == "valu":
        flags |= _native.FLAG_SCREEN_FP32
    elif mode == "stream":
        flags |= _native.FLAG_SELECT_STREAMING
    if mixed:
        flags |= _native.FLAG_PRECISION_MIXED
        if mixed != "auto":
            flags |= _native.FLAG_LIKELIHOOD_MFMA
    print(f"[gpu_run] C={V.shape[0]} N={V.shape[1]} U={T.shape[0]} L={T.shape[2]} n_max={n_max} "
          f"recheck={recheck}", flush=True)
    h.run(V, n_max, flags)
    idx, lab = h.result()
    R = h.residual(V.shape[1])
    out = [idx, lab, R]
    if clips:
        out.append(h.clips(len(idx)))
    return out


def assert_same(gpu, ora):
    gi, gl, gR = gpu[:3]
    oi, ol, oR = ora[:3]
    assert len(gi) == len(oi), (len(gi), len(oi))
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gl, ol)
    np.testing.assert_array_equal(gR, oR)


@pytest.mark.parametrize("literal", MODES)
@pytest.mark.parametrize("seed", range(10))
def test_noisy_cases_match_oracle(handle, seed, literal):
    V, T, bias, thr, on, lab = cases.noisy_case(seed)
    n = T.shape[2] - 1
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("literal", MODES)
@pytest.mark.parametrize("seed", [0, 1, 5, 9])
def test_correlated_pairs_match_oracle(handle, seed, literal):
    U, C, L = 4, 3, 16
    T = cases.dog_templates(U, C, L, seed).astype(np.float32)
    V, truth = cases.planted_pairs(T, 2000, 100 + seed, min_d=3)
    V = V + np.random.default_rng(seed).normal(size=V.shape).astype(np.float32) * 0.2
    bias, thr = cases.bias_thr(T)
    assert_same(gpu_run(handle, V, T, bias, thr, L - 1, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, L - 1, return_residual=True))


@pytest.mark.parametrize("seed", [2, 3])
def test_overlapping_pairs_on_channel_subsets_match_oracle(handle, seed):
    """Overlapping spikes of templates that hold different channel subsets: the
    residual subtracts each template on its own channels only, and a sample
    covered by two spikes on a shared channel is written once."""
    U, C, L = 5, 6, 16
    rng = np.random.default_rng(seed)
    T = cases.dog_templates(U, C, L, seed)
    for u in range(U):
        keep = np.zeros(C, bool)
        keep[rng.choice(C, int(rng.integers(2, 5)), replace=False)] = True
        T[u, ~keep] = 0.0
    T = T.astype(np.float32)
    V, truth = cases.planted_pairs(T, 2400, 300 + seed, min_d=2)
    V = V + rng.normal(size=V.shape).astype(np.float32) * 0.2
    bias, thr = cases.bias_thr(T)
    ora = fbp_oracle.pursuit(V, T, bias, thr, L - 1, return_residual=True)
    for mode, mixed in (("split", False), ("valu", False), ("split", True)):
        assert_same(gpu_run(handle, V, T, bias, thr, L - 1, literal=mode, mixed=mixed), ora)


@pytest.mark.parametrize("recheck,literal", [(True, "literal"), (True, "split"), (True, "valu"), (False, "split")])
def test_small_workload_3s_matches_oracle(handle, recheck, literal):
    d = synthetic.generate("small", duration_s=3.0)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    assert_same(gpu_run(handle, V, T, bias, thr, n, recheck=recheck, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, n, do_recheck=recheck, return_residual=True))


_FULL_SMALL = {}


def _small_full():
    """The 4-channel configuration at its full length (40 kHz, 30 s, 5 units) and its oracle run."""
    if not _FULL_SMALL:
        d = synthetic.generate("small")
        T, bias, thr = synthetic.kernel_inputs(d)
        n = d["n_max_shift_inds"]
        _FULL_SMALL["args"] = (d["voltage"], T, bias, thr, n)
        _FULL_SMALL["oracle"] = fbp_oracle.pursuit(d["voltage"], T, bias, thr, n, return_residual=True)
    return _FULL_SMALL["args"], _FULL_SMALL["oracle"]


@pytest.mark.parametrize("literal,mixed", [("literal", False), ("split", False), ("valu", False), ("split", "auto"),
                                           ("split", True)])
def test_small_workload_full_30s_matches_oracle(handle, literal, mixed):
    args, ora = _small_full()
    assert_same(gpu_run(handle, *args, literal=literal, mixed=mixed), ora)


@pytest.mark.parametrize("name,dur", [("probe64", 10.0), ("probe16", 30.0)])
def test_screened_recheck_matches_literal_gpu(handle, name, dur):
    """The 64-channel configuration on 10 s and the 16-channel one on 30 s, where the CPU
    oracle's literal recheck is too slow: every screen path against the literal GPU recheck
    (every overlap item evaluated in the reference's order)."""
    d = synthetic.generate(name, duration_s=dur)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    ref = gpu_run(handle, V, T, bias, thr, n, literal="literal")
    ref_log = handle.pass_log()
    assert len(ref[0]) > 1000
    for mode, mixed in (("split", "auto"), ("valu", False), ("stream", "auto")):
        assert_same(gpu_run(handle, V, T, bias, thr, n, literal=mode, mixed=mixed), ref)
        np.testing.assert_array_equal(handle.pass_log(), ref_log)


@pytest.mark.parametrize("seed", range(6))
def test_mixed_precision_noisy_cases_match_oracle(handle, seed):
    V, T, bias, thr, on, lab = cases.noisy_case(seed)
    n = T.shape[2] - 1
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal="split", mixed=True),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("mixed", [True, "auto"])
@pytest.mark.parametrize("name,dur", [("small", 3.0), ("probe16", 0.5), ("probe64", 0.1)])
def test_mixed_precision_workloads_match_oracle(handle, name, dur, mixed):
    d = synthetic.generate(name, duration_s=dur)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal="split", mixed=mixed),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("C,L,U,N", [(2, 8, 3, 600), (5, 20, 17, 1500), (9, 33, 40, 2000), (12, 48, 16, 3000),
                                     (17, 64, 33, 2500), (3, 16, 1, 400), (64, 48, 20, 1200)])
def test_mixed_precision_shapes_match_oracle(handle, C, L, U, N):
    """Tile padding (U not a multiple of 16), channel padding (C not a multiple
    of 8), positions past L within the last 16-column tile, dense templates and
    windows at the end of the chunk."""
    V, T, bias, thr, on, lab = cases.noisy_case(7 + C, C=C, L=L, U=U, N=N, nsp=max(4, N // 25))
    n = L - 1
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal="split", mixed=True),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("groups", [(12, 12), (5, 20, 9)])
def test_mixed_precision_partial_tiles_match_oracle(handle, groups):
    """Templates on disjoint 8-channel blocks, in groups that do not fill a
    16-template tile: the matrix-core tables cut the tiles where their summed
    k-blocks are fewest, so tiles with unused rows sit between full ones (first
    likelihood pass and screen correlations on the matrix cores)."""
    rng = np.random.default_rng(sum(groups))
    U, C, L, N = sum(groups), 8 * len(groups), 24, 4000
    T = cases.dog_templates(U, C, L, 3).astype(np.float32)
    u0 = 0
    for g, cnt in enumerate(groups):  # group g on channels 8g .. 8g + 7 only
        keep = np.zeros(C, bool)
        keep[8 * g:8 * g + 8] = True
        T[u0:u0 + cnt][:, ~keep, :] = 0.0
        u0 += cnt
    V = rng.normal(size=(C, N)).astype(np.float32) * 0.05
    for k, p in enumerate(range(30, N - 2 * L, 37)):
        V[:, p:p + L] += T[(7 * k) % U] * np.float32(0.8 + 0.4 * rng.random())
    bias, thr = cases.bias_thr(T.astype(np.float64), 0.05)
    n = L - 1
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal="split", mixed=True),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("literal", ["split", "valu"])
@pytest.mark.parametrize("seed", [0, 1])
def test_channel_dynamic_range_matches_oracle(handle, seed, literal):
    """Channels whose amplitudes span 10^-7 .. 10^3 of each other, with tiny
    values next to large ones in one window: the fp16 operands of the matrix
    cores lose the small channels entirely (sub-normal or flushed), which the
    bounds must absorb so that the decisions stay those of FP32."""
    rng = np.random.default_rng(40 + seed)
    U, C, L, N = 6, 8, 24, 3000
    scale = np.array([1e3, 1.0, 1e-3, 1e-7, 3e2, 1e-5, 0.5, 2e-2], np.float32)
    T = cases.dog_templates(U, C, L, seed).astype(np.float32) * scale[None, :, None]
    V = (rng.normal(size=(C, N)).astype(np.float32) * 0.05) * scale[:, None]
    for p in range(30, N - 2 * L, 41):
        V[:, p:p + L] += T[(p // 41) % U] * np.float32(0.8 + 0.4 * rng.random())
    bias, thr = cases.bias_thr(T.astype(np.float64), 0.05)
    n = L - 1
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal=literal, mixed=True),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


def test_mixed_precision_rescores_few_positions(handle):
    d = synthetic.generate("probe64", duration_s=0.5)
    T, bias, thr = synthetic.kernel_inputs(d)
    gpu_run(handle, d["voltage"], T, bias, thr, d["n_max_shift_inds"], literal="split", mixed=True)
    st = handle.stats()
    assert 0 < st["likelihood_rescored"] < 0.2 * st["likelihood_positions"]


def test_adjusted_clips_match_oracle(handle):
    V, T, bias, thr, on, lab = cases.noisy_case(3)
    n = T.shape[2] - 1
    g = gpu_run(handle, V, T, bias, thr, n, clips=True)
    oi, ol, oR = fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True)
    np.testing.assert_array_equal(g[3], fbp_oracle.adjusted_clips(oR, T, oi, ol))


@pytest.mark.parametrize("literal", MODES)
@pytest.mark.parametrize("n_max", [0, 1, 3, 20, 70])
def test_shift_range_variants(handle, n_max, literal):
    V, T, bias, thr, on, lab = cases.noisy_case(20 + n_max, U=3, L=8)
    assert_same(gpu_run(handle, V, T, bias, thr, n_max, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, n_max, return_residual=True))


@pytest.mark.parametrize("literal", MODES)
def test_zero_channel_templates_and_odd_length(handle, literal):
    rng = np.random.default_rng(8)
    U, C, L, N = 4, 5, 11, 900
    T = (rng.normal(size=(U, C, L)) * 4).astype(np.float32)
    T[0, 1:] = 0
    T[1, [0, 4]] = 0
    T[3, 2] = 0
    V = (rng.normal(size=(C, N)) * 0.3).astype(np.float32)
    for p in range(20, N - 2 * L, 37):
        V[:, p:p + L] += T[(p // 37) % U]
    bias, thr = cases.bias_thr(T.astype(np.float64), 0.05)
    assert_same(gpu_run(handle, V, T, bias, thr, L - 1, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, L - 1, return_residual=True))


def test_short_chunk_edges(handle):
    # N just above 4 template widths, a spike at the very end and one at 0
    U, C, L = 2, 2, 8
    T = cases.random_templates(U, C, L, 4).astype(np.float32)
    N = 4 * L + 3
    V = np.zeros((C, N), np.float32)
    V[:, 0:L] += T[0]
    V[:, N - L:] += T[1]
    V[:, 13:13 + L] += T[1]
    bias, thr = cases.bias_thr(T.astype(np.float64))
    assert_same(gpu_run(handle, V, T, bias, thr, L - 1),
                fbp_oracle.pursuit(V, T, bias, thr, L - 1, return_residual=True))


def test_empty_result(handle):
    T = cases.random_templates(3, 2, 8, 0).astype(np.float32)
    bias, thr = cases.bias_thr(T.astype(np.float64))
    g = gpu_run(handle, np.zeros((2, 500), np.float32), T, bias, thr, 7)
    assert len(g[0]) == 0


@pytest.mark.parametrize("literal", MODES)
def test_probe16_slice_matches_oracle(handle, literal):
    d = synthetic.generate("probe16", duration_s=0.5)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("literal", ["split", "valu"])
def test_probe64_slice_screened_matches_oracle(handle, literal):
    d = synthetic.generate("probe64", duration_s=0.1)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = d["voltage"]
    assert_same(gpu_run(handle, V, T, bias, thr, n, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True))


@pytest.mark.parametrize("literal", MODES)
def test_exact_ties_identical_templates(handle, literal):
    # two identical units: every overlap value ties between the labels, the
    # winner must be the lowest item id in both paths
    T1 = cases.random_templates(2, 3, 12, 21)
    T = np.concatenate([T1, T1[:1]]).astype(np.float32)
    V, truth = cases.planted_pairs(T, 1800, 21)
    V = V + np.random.default_rng(21).normal(size=V.shape).astype(np.float32) * 0.05
    bias, thr = cases.bias_thr(T.astype(np.float64))
    assert_same(gpu_run(handle, V, T, bias, thr, 11, literal=literal),
                fbp_oracle.pursuit(V, T, bias, thr, 11, return_residual=True))


def _ref_output(d, chunks=None):
    """Oracle through the reference's output assembly (single chunk)."""
    T, bias, thr = synthetic.kernel_inputs(d)
    idx, lab = fbp_oracle.pursuit(d["voltage"], T, bias, thr, d["n_max_shift_inds"])
    return idx.astype(np.int64) + d["clip_init_samples"], lab.astype(np.int64)


def test_dropin_binary_pursuit_matches_oracle():
    d = synthetic.generate("small", duration_s=5.0)
    m = dict(d["separability_metrics"])
    ev, lab, bpbool, clips = bpp.binary_pursuit(d["voltage"], np.float32, d["sort_info"], m,
                                                n_max_shift_inds=d["n_max_shift_inds"])
    rev, rlab = _ref_output(d)
    assert ev.dtype == np.int64 and lab.dtype == np.int64 and bpbool.dtype == bool
    np.testing.assert_array_equal(ev, rev)
    np.testing.assert_array_equal(lab, rlab)
    assert bpbool.all()
    assert clips.shape == (1, 0)  # np.vstack([[]]) as in binary_pursuit_parallel.py:1509
    np.testing.assert_array_equal(m["n_spikes"], np.bincount(rlab, minlength=m["templates"].shape[0]))


def test_dropin_chunked_matches_oracle_per_chunk():
    d = synthetic.generate("small", duration_s=2.0)
    si = dict(d["sort_info"])
    m = dict(d["separability_metrics"])
    cap = 1 * 2**20
    ev, lab, _, _ = bpp.binary_pursuit(d["voltage"], np.float32, si, m, n_max_shift_inds=d["n_max_shift_inds"],
                                       max_gpu_memory=cap)
    T, bias, thr = synthetic.kernel_inputs(d)
    V = d["voltage"]
    L = T.shape[2]
    n_per, onsets = bpp.chunk_plan(V.shape[0], V.shape[1], si["sampling_rate"], T.shape[0], L,
                                   d["n_max_shift_inds"], 288 * 2**30, max_gpu_memory=cap)
    assert len(onsets) > 1
    rev, rlab = [], []
    for s0 in onsets:
        s1 = min(V.shape[1], s0 + n_per)
        i, l = fbp_oracle.pursuit(V[:, s0:s1], T, bias, thr, d["n_max_shift_inds"])
        rev.append(i.astype(np.int64) + s0)
        rlab.append(l.astype(np.int64))
    np.testing.assert_array_equal(ev, np.concatenate(rev) + d["clip_init_samples"])
    np.testing.assert_array_equal(lab, np.concatenate(rlab))


def test_dropin_adjusted_clips_dtype_and_values():
    d = synthetic.generate("small", duration_s=1.0)
    si = dict(d["sort_info"], get_adjusted_clips=True)
    m = dict(d["separability_metrics"])
    ev, lab, _, clips = bpp.binary_pursuit(d["voltage"], np.float64, si, m, n_max_shift_inds=d["n_max_shift_inds"])
    T, bias, thr = synthetic.kernel_inputs(d)
    i, l, R = fbp_oracle.pursuit(d["voltage"], T, bias, thr, d["n_max_shift_inds"], return_residual=True)
    assert clips.dtype == np.float64
    np.testing.assert_array_equal(clips, fbp_oracle.adjusted_clips(R, T, i, l).astype(np.float64))


def test_dropin_mixed_precision_identical_to_fp32():
    d = synthetic.generate("probe16", duration_s=1.0)
    m1, m2 = dict(d["separability_metrics"]), dict(d["separability_metrics"])
    a = bpp.binary_pursuit(d["voltage"], np.float32, d["sort_info"], m1, n_max_shift_inds=d["n_max_shift_inds"],
                           precision="fp32", return_residual=True)
    b = bpp.binary_pursuit(d["voltage"], np.float32, d["sort_info"], m2, n_max_shift_inds=d["n_max_shift_inds"],
                           precision="mixed", return_residual=True)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[4][0], b[4][0])


def test_dropin_device_resident_voltage_matches_host():
    import torch
    d = synthetic.generate("small", duration_s=1.0)
    m1, m2 = dict(d["separability_metrics"]), dict(d["separability_metrics"])
    ev1, lab1, _, _ = bpp.binary_pursuit(d["voltage"], np.float32, d["sort_info"], m1,
                                         n_max_shift_inds=d["n_max_shift_inds"])
    vt = torch.from_numpy(d["voltage"]).cuda()
    ev2, lab2, _, _ = bpp.binary_pursuit(vt, np.float32, d["sort_info"], m2,
                                         n_max_shift_inds=d["n_max_shift_inds"])
    np.testing.assert_array_equal(ev1, ev2)
    np.testing.assert_array_equal(lab1, lab2)


def test_dropin_no_spikes_returns_empty_lists():
    d = synthetic.generate("small", duration_s=0.5)
    m = dict(d["separability_metrics"])
    out = bpp.binary_pursuit(np.zeros_like(d["voltage"]), np.float32, d["sort_info"], m,
                             n_max_shift_inds=d["n_max_shift_inds"])
    assert out == ([], [], [], [])
    assert np.all(m["n_spikes"] == 0)


@pytest.mark.parametrize("name,dur,offset", [("probe64", 10.5, 7), ("small", 150.0, 3)])
def test_host_input_staged_in_pieces_matches_device_input(handle, name, dur, offset):
    """Host voltage moves to the device in 32 MiB pinned pieces and the residual comes back the
    same way: a float32 row view with row stride above N (no host copy) spanning several pieces,
    the last one partial, must give the spikes and residual of the same voltage passed as a
    device tensor."""
    import torch

    d = synthetic.generate(name, duration_s=dur)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    C, N = d["voltage"].shape
    wide = np.zeros((C, N + offset + 5), np.float32)
    wide[:, offset:offset + N] = d["voltage"]
    view = wide[:, offset:offset + N]
    assert view.strides[0] > 4 * N and 4 * C * N > 2 * (32 << 20)
    flags = _native.FLAG_OVERLAP_RECHECK | _native.FLAG_PRECISION_MIXED
    handle.set_templates(T, bias, thr)
    print(f"[staging] {name} C={C} N={N} host view, then device tensor", flush=True)
    handle.run(view, n, flags)
    hi, hl = handle.result()
    hR = handle.residual(N)
    Vd = torch.tensor(d["voltage"], device="cuda")
    torch.cuda.synchronize()
    handle.run(None, n, flags, ld=Vd.stride(0), n_samples=N, device_ptr=Vd.data_ptr())
    di, dl = handle.result()
    dR = handle.residual(N)
    assert len(hi) > 1000
    np.testing.assert_array_equal(hi, di)
    np.testing.assert_array_equal(hl, dl)
    np.testing.assert_array_equal(hR, dR)
    np.testing.assert_array_equal(wide[:, offset:offset + N], d["voltage"])  # the caller's array is untouched
    # the residual is the voltage minus every found template
    ref = d["voltage"].astype(np.float64).copy()
    L = T.shape[2]
    for p, u in zip(hi.tolist(), hl.tolist()):
        m = min(L, N - p)
        ref[:, p:p + m] -= T[u, :, :m]
    assert np.max(np.abs(hR - ref)) <= 1e-3 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("literal", ["split", "valu"])
def test_host_input_first_pass_regions_equal_device_input(handle, literal):
    """Host input rechecks the first pass region by region (16 384 windows each)
    as the voltage lands; device input in one launch (both with the vector-unit
    first likelihood pass, which the regions need).  The 16-channel
    configuration on 60 s (37 500 windows, three regions): same spikes, labels,
    order, residual, per-pass trace and recheck statistics."""
    import torch
    d = synthetic.generate("probe16", duration_s=60.0)
    T, bias, thr = synthetic.kernel_inputs(d)
    n = d["n_max_shift_inds"]
    V = np.ascontiguousarray(d["voltage"], dtype=np.float32)
    host = gpu_run(handle, V, T, bias, thr, n, literal=literal)
    host_log, host_st = handle.pass_log(), handle.stats()
    assert (V.shape[1] + T.shape[2] - 1) // T.shape[2] > 2 * 16384
    Vd = torch.tensor(V, device="cuda")
    flags = _native.FLAG_OVERLAP_RECHECK
    if literal == "valu":
        flags |= _native.FLAG_SCREEN_FP32
    torch.cuda.synchronize()
    handle.run(None, n, flags, ld=Vd.stride(0), n_samples=Vd.shape[1], device_ptr=Vd.data_ptr())
    idx, lab = handle.result()
    np.testing.assert_array_equal(host[0], idx)
    np.testing.assert_array_equal(host[1], lab)
    np.testing.assert_array_equal(host[2], handle.residual(V.shape[1]))
    np.testing.assert_array_equal(host_log, handle.pass_log())
    dev_st = handle.stats()
    for k in ("windows_flagged", "recheck_items", "screen_shared_pairs", "screen_separable_pairs"):
        assert host_st[k] == dev_st[k], k
    assert len(idx) > 1000
