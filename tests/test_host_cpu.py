"""Host logic of the drop-in and the C ABI surface, without a GPU."""
import re
import os

import numpy as np
import pytest

from spikesorting_fullpursuit_amd import _native, synthetic
from spikesorting_fullpursuit_amd.overlap import binary_pursuit_parallel as bpp
from spikesorting_fullpursuit_amd.overlap import neuron_separability
from spikesorting_fullpursuit_amd.processing.conversions import time_window_to_samples

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_time_window_to_samples_closed_form():
    # SF/processing/conversions.py:32-37
    w, tw = time_window_to_samples([0.8e-3, 0.775e-3], 40000)
    assert w == [-32, 32] and tw[0] == -0.8e-3
    w, _ = time_window_to_samples([-0.5e-3, 0.0], 30000)
    assert w == [-15, 2]  # second element at least 1, then +1
    w, _ = time_window_to_samples([0.0, 1e-3], 30000)
    assert w == [0, 31]


def test_add_n_spikes_counts_per_label():
    m = {"templates": np.zeros((4, 10))}
    neuron_separability.add_n_spikes(m, np.array([0, 2, 2, 3, 3, 3], np.int64))
    assert m["n_spikes"].dtype == np.float64
    np.testing.assert_array_equal(m["n_spikes"], [1, 0, 2, 3])
    neuron_separability.add_n_spikes(m, [])
    np.testing.assert_array_equal(m["n_spikes"], [0, 0, 0, 0])


def test_set_bp_threshold_takes_max_of_thr_zero_and_ci():
    m = {"neuron_lower_thresholds": np.array([-1.0, 2.0, 3.0]), "neuron_lower_CI": np.array([0.5, -np.inf, 5.0])}
    neuron_separability.set_bp_threshold(m)
    np.testing.assert_array_equal(m["neuron_lower_thresholds"], [0.5, 2.0, 5.0])


def test_template_thresholds_formula():
    rng = np.random.default_rng(0)
    U, C, L = 3, 2, 5
    T = rng.normal(size=(U, C * L))
    covs = [np.eye(L) * (c + 1) for c in range(C)]
    SS, thr, ci = neuron_separability.template_thresholds(T, covs, L, 3.719, 2.0)
    var = np.array([sum((c + 1) * np.sum(T[u, c * L:(c + 1) * L] ** 2) for c in range(C)) for u in range(U)])
    np.testing.assert_allclose(SS, np.sum(T ** 2, 1))
    np.testing.assert_allclose(thr, -0.5 * SS + 3.719 * np.sqrt(var))
    np.testing.assert_allclose(ci, 0.5 * SS - 2.0 * np.sqrt(var))


def test_chunk_plan_single_chunk_on_large_memory():
    n, onsets = bpp.chunk_plan(64, 18_000_000, 30000.0, 120, 48, 47, 288 * 2**30)
    assert onsets == [0] and n >= 18_000_000


def test_chunk_plan_formula_and_overlap():
    # binary_pursuit_parallel.py:398-507, by hand for a small memory cap
    C, fs, U, L, nmax = 4, 40000.0, 5, 64, 63
    cap = 64 * 2**20
    n, onsets = bpp.chunk_plan(C, 1_200_000, fs, U, L, nmax, 288 * 2**30, max_gpu_memory=cap)
    items = (2 * nmax + 1) ** 2 * U
    nloc = np.int64((items - 1) / 256 + 1)
    per_s = C * fs * 4 + U * fs * 4 + (fs / L) * (20 + (nloc / 3) * 8)
    expect = int(np.floor((cap - (4 * U * C * L + 4 * U)) / per_s * fs))
    assert n == expect
    assert onsets[0] == 0 and all(b - a == n - 3 * L for a, b in zip(onsets, onsets[1:]))
    assert onsets[-1] < 1_200_000 - L <= onsets[-1] + n - 3 * L


def _fbp_run_device_bytes(C, N, U, L, nmax):
    """Every device buffer fbp_run holds for an N-sample chunk with every template pair
    sharing a channel, listed from csrc/fbp_api.hip: working voltage, likelihood,
    per-template window maxima (f32 + u16), 5 per-window u32/f32 arrays, 5 flag
    arrays, 4 window lists, the dirty masks, spikes with twice the windows'
    capacity after growth, BF16-likelihood buffers, screen tables and buffer, staging, templates."""
    S = 2 * nmax + 1
    nwin = -(-N // L)
    pairs = U * U
    b = 4 * C * N + 4 * U * N + 6 * U * nwin + 20 * nwin + 5 * nwin + 16 * nwin + 8 * nwin + 2 * 2 * 4 * nwin
    ldq = (N + L + 96 + 63) // 64 * 64
    b += 2 * C * ldq + 4 * C * (ldq // 16) + U * nwin  # BF16 likelihood: quantized voltage, block maxima, groups
    b += pairs * S * ((S + 3) // 4 * 4) * 4 + pairs * S * 4 + pairs * S * (-(-S // 8)) * 4 + U * S * 4
    b += min(65536, (2 << 30) // (U * S * 4)) * U * S * 4
    b += 3 * (32 << 20) + 3 * U * C * L * 4
    return b


@pytest.mark.parametrize("hours,on_device", [(1 / 6, False), (3, False), (10, False), (40, False), (1 / 6, True),
                                             (3, True), (6, True)])
def test_chunk_plan_fits_library_buffers_probe64(hours, on_device):
    """64 channels, 30 kHz, 120 units, L = 48 on a 288 GB device: every chunk the plan makes
    fits fbp_run's buffers beside a voltage tensor passed in place; short recordings keep
    the reference's single chunk."""
    C, fs, U, L, nmax = 64, 30000.0, 120, 48, 47
    total = 288 * 10**9
    N = int(hours * 3600 * fs)
    resident = 4 * C * N if on_device else 0
    n, onsets = bpp.chunk_plan(C, N, fs, U, L, nmax, total, resident_bytes=resident)
    assert _fbp_run_device_bytes(C, n, U, L, nmax) + resident <= total - (1 << 30)
    assert onsets[0] == 0 and all(b - a == n - 3 * L for a, b in zip(onsets, onsets[1:]))
    assert onsets[-1] + n >= N - L
    if hours <= 1:
        assert onsets == [0]


def test_chunk_plan_too_small_raises():
    with pytest.raises(ValueError):
        bpp.chunk_plan(4, 100000, 40000.0, 5, 64, 63, 288 * 2**30, max_gpu_memory=4096)


def test_chunk_plan_short_voltage_has_no_chunks():
    _, onsets = bpp.chunk_plan(4, 64, 40000.0, 5, 64, 63, 2**34)
    assert onsets == []


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "fbp.h")) as fh:
        header = fh.read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(fbp_\w+)\s*\(", header, re.M))
    assert declared == set(_native.EXPORTS)
    lib = _native.load()
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU error path")
def test_no_gpu_raises_like_reference():
    # binary_pursuit_parallel.py:292-293
    with pytest.raises(RuntimeError, match="No GPU available"):
        _native.Handle(0)
    V = np.zeros((2, 1000), np.float32)
    m = {"templates": np.ones((1, 16)), "template_SS": np.array([16.0]),
         "template_SS_by_chan": np.ones((1, 2)), "neuron_lower_thresholds": np.array([1.0])}
    si = {"sampling_rate": 40000.0, "clip_width": [-4 / 40000, 3 / 40000], "n_samples_per_chan": 8,
          "do_overlap_recheck": True, "get_adjusted_clips": False}
    with pytest.raises(RuntimeError, match="No GPU available"):
        bpp.binary_pursuit(V, np.float32, si, m)


def test_missing_sort_info_key_raises_keyerror():
    V = np.zeros((2, 1000), np.float32)
    m = {"templates": np.ones((1, 16)), "template_SS": np.array([16.0]),
         "template_SS_by_chan": np.ones((1, 2)), "neuron_lower_thresholds": np.array([1.0])}
    si = {"sampling_rate": 40000.0, "clip_width": [-4 / 40000, 3 / 40000], "n_samples_per_chan": 8,
          "get_adjusted_clips": False}
    with pytest.raises(KeyError):
        bpp.binary_pursuit(V, np.float32, si, m)


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU error path")
def test_batch_and_refine_have_no_cpu_fallback():
    from spikesorting_fullpursuit_amd import batch, refine
    V = np.zeros((2, 1000), np.float32)
    m = {"templates": np.ones((1, 16)), "template_SS": np.array([16.0]),
         "template_SS_by_chan": np.ones((1, 2)), "neuron_lower_thresholds": np.array([1.0])}
    si = {"sampling_rate": 40000.0, "clip_width": [-4 / 40000, 3 / 40000], "n_samples_per_chan": 8,
          "do_overlap_recheck": True, "get_adjusted_clips": False}
    with pytest.raises(RuntimeError, match="No GPU available"):
        batch.binary_pursuit_batch([{"voltage": V, "sort_info": si, "separability_metrics": m}])
    import torch
    table = type("T", (), {"N": 1000, "U": 1, "n": 0})()
    with pytest.raises(ValueError, match="device tensors"):
        refine.forward_residual(torch.zeros(2, 1000), torch.zeros(1, 2, 8), torch.zeros(0), table)


def test_batch_graph_mode_runs_one_lane():
    """Graph mode keeps its graph calls on one host thread: more lanes, or the
    streams mode, are refused before any device work."""
    from spikesorting_fullpursuit_amd import batch, _native
    assert _native.FLAG_BATCH_GRAPH == 8192
    for kw in ({"lanes": 2}, {"lanes": 1, "mode": "streams"}):
        with pytest.raises(ValueError, match="one lane"):
            batch.binary_pursuit_batch([], graph=True, **kw)
    # graph replays are switched off in this build
    with pytest.raises(NotImplementedError, match="disabled"):
        batch.binary_pursuit_batch([], graph=True, lanes=1)


def test_batch_workloads_cover_the_configured_mix():
    from spikesorting_fullpursuit_amd import synthetic
    ws = synthetic.batch_workloads()
    assert len(ws) == 256
    assert {w.n_chans for w in ws} == set(synthetic.BATCH_CHANNELS)
    assert all(5.0 <= w.duration_s <= 120.0 and 4 <= w.n_units <= 60 and w.L == 64 for w in ws)
    total = sum(w.n_samples * w.n_chans for w in ws) / 1e6
    assert 4500 < total < 7200  # expected 256 * 62.5 s * 12.2 ch * 30 kHz = 5856 Msamples*ch
    assert len({w.seed for w in ws}) == 256


def test_refine_abi_checks_arguments_without_a_gpu():
    import ctypes
    lib = _native.load()
    nb = ctypes.c_int64(0)
    assert lib.fbp_refine_workspace_bytes(1000, 10, 3, 4, 32, ctypes.byref(nb)) == _native.FBP_OK
    # forward: 4 + 1 (256-B aligned) partials; backward: chunk table + (ceil(10/256) + 3) chunks of C*L floats
    assert nb.value == max(256 + 256, 256 + 4 * 4 * 32 * 4)
    assert lib.fbp_refine_workspace_bytes(0, 10, 3, 4, 32, ctypes.byref(nb)) == _native.FBP_E_INVALID
    assert b"positive" in lib.fbp_refine_last_error()
    assert lib.fbp_refine_workspace_bytes(1000, 10, 3, 1 << 13, 1 << 12, ctypes.byref(nb)) == _native.FBP_E_UNSUPPORTED
    vp = ctypes.c_void_p
    rc = lib.fbp_refine_forward(None, 1000, 1000, 4, None, 3, 32, None, None, None, 10, None, 1000, None, None, 0,
                                None)
    assert rc == _native.FBP_E_INVALID and b"null" in lib.fbp_refine_last_error()
    rc = lib.fbp_refine_backward(vp(8), 1000, 1000, 4, vp(8), 3, 32, vp(8), vp(8), vp(8), 10, vp(8), vp(8), None,
                                 None, vp(8), 1 << 20, None)
    assert rc == _native.FBP_E_INVALID and b"no gradient" in lib.fbp_refine_last_error()


def test_probe16_overlap_fraction():
    """The 16-channel configuration asks that >= 15 % of spikes lie within 1 ms of a
    spike of another unit (40 units, 2-80 Hz, 300 s)."""
    w = synthetic.workload("probe16")
    rng = np.random.default_rng(w.seed)
    synthetic.make_templates(w, rng)
    on, lab, _ = synthetic.spike_trains(w, rng, w.n_samples)
    o = np.argsort(on, kind="stable")
    on, lab = on[o], lab[o]
    win = int(round(0.001 * w.fs))
    close = np.zeros(len(on), bool)
    d = 1
    while d < len(on) and np.any(on[d:] - on[:-d] <= win):
        hit = (on[d:] - on[:-d] <= win) & (lab[d:] != lab[:-d])
        close[d:] |= hit
        close[:-d] |= hit
        d += 1
    assert close.mean() >= 0.15
