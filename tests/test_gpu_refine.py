"""GPU: residual energy and its gradients (template refinement) against the
FP64 oracle, central differences and run-to-run determinism.

Tolerances are the FP32 rounding bounds of the kernels' own summation orders
(u = 2^-24), evaluated per element on the FP64 data:
  residual   |R_gpu - R_64| <= dR = 2 (m + 1) u S,  S = |V| + sum_cover |a_k||T|,
             m = spikes covering the sample (one fmaf per covering spike)
  energy     |E_gpu - E_64| <= sum |R_64| dR + 1e-12 E_64
  dE/da_k    <= sum |T||dR| + 2 (C ceil(L/64) + 7) u sum |T||R_64|   (lane chains + wave tree)
  dE/dT      <= sum |a||dR| + 2 (256 + chunks + 1) u sum |a||R_64|   (256-spike chunks, then chunk sum)
On the realistic cases these are ~1e-6 relative.  Central differences of the
GPU energy (exact for a quadratic in exact arithmetic) agree to 1e-3 max|g|
with step 1.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import refine_fp64
from spikesorting_fullpursuit_amd import _native, refine, synthetic

pytestmark = pytest.mark.gpu


def _case(seed, C=4, N=3000, U=5, L=32, n=120, n_empty_unit=True):
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(C, N)).astype(np.float32)
    T = (rng.normal(size=(U, C, L)) * 3).astype(np.float32)
    idx = np.sort(rng.integers(0, N, size=n))
    idx[-1] = N - 5  # truncated at the end of the recording
    lab = rng.integers(0, U - 1 if n_empty_unit else U, size=n)  # the last unit has no spikes
    amp = rng.normal(1.0, 0.1, size=n).astype(np.float32)
    return V, T, idx, lab, amp


def _gpu(V, T, idx, lab, amp):
    print(f"[refine] C={V.shape[0]} N={V.shape[1]} U={T.shape[0]} L={T.shape[2]} n={len(idx)}", flush=True)
    Vd = torch.tensor(V, device="cuda")
    Td = torch.tensor(T, device="cuda")
    table = refine.SpikeTable(idx, lab, T.shape[0], V.shape[1])
    a = table.to_sorted(torch.tensor(amp, device="cuda"))
    R, E = refine.forward_residual(Vd, Td, a, table)
    gT, ga = refine.backward_grads(R, Td, a, table)
    torch.cuda.synchronize()
    return (R.cpu().numpy(), float(E), gT.cpu().numpy(), table.from_sorted(ga).cpu().numpy(), table)


U_RND = 2.0 ** -24


def _bounds(V, T, idx, lab, amp, R64):
    C, N = V.shape
    U, _, L = T.shape
    S = refine_fp64.residual(np.abs(V), -np.abs(T), idx, lab, np.abs(amp))
    m = refine_fp64.residual(np.zeros((1, N)), -np.ones((U, 1, L)), idx, lab, np.ones(len(idx)))[0]
    dR = 2 * (m + 1) * U_RND * S
    dE = float(np.sum(np.abs(R64) * dR)) + 1e-12 * 0.5 * float(np.sum(R64 ** 2))
    T64 = np.abs(T.astype(np.float64))
    dga = np.zeros(len(idx))
    dgT_abs = np.zeros_like(T64)
    dgT_mag = np.zeros_like(T64)
    n_u = np.bincount(lab, minlength=U) if len(idx) else np.zeros(U, np.int64)
    for k, (p, u, a) in enumerate(zip(idx, lab, amp)):
        n = min(L, N - p)
        dga[k] = np.sum(T64[u, :, :n] * dR[:, p:p + n]) + \
            2 * (C * -(-L // 64) + 7) * U_RND * np.sum(T64[u, :, :n] * np.abs(R64[:, p:p + n]))
        dgT_abs[u, :, :n] += abs(a) * dR[:, p:p + n]
        dgT_mag[u, :, :n] += abs(a) * np.abs(R64[:, p:p + n])
    chunks = -(-n_u // 256)
    dgT = dgT_abs + 2 * (256 + chunks[:, None, None] + 1) * U_RND * dgT_mag
    return dR, dE, dgT, dga


def _check(V, T, idx, lab, amp):
    R, E, gT, ga, _ = _gpu(V, T, idx, lab, amp)
    E64, gT64, ga64, R64 = refine_fp64.gradients(V, T, idx, lab, amp)
    dR, dE, dgT, dga = _bounds(V, T, idx, lab, amp, R64)
    assert np.all(np.abs(R - R64) <= dR + 1e-30)
    assert abs(E - E64) <= dE
    assert np.all(np.abs(gT - gT64) <= dgT + 1e-30)
    assert np.all(np.abs(ga - ga64) <= dga + 1e-30)
    # and the bounds are tight enough to mean something on these cases
    assert np.max(dgT) <= 1e-2 * max(1.0, np.max(np.abs(gT64)))
    return R, E, gT, ga


def test_matches_fp64_oracle():
    _check(*_case(1))


@pytest.mark.parametrize("shape", [dict(C=1, L=1), dict(C=7, L=100), dict(C=3, L=64, U=2), dict(C=16, L=48, N=700)])
def test_matches_fp64_oracle_shapes(shape):
    _check(*_case(2, **shape))


def test_dense_spikes_take_several_staging_rounds():
    # ~1000 onsets can cover one 256-sample tile: more than one round of 512 staged spikes,
    # and units with more than 256 spikes (several gradient chunks)
    V, T, idx, lab, amp = _case(3, C=2, N=2000, U=3, L=64, n=6000)
    _check(V, T, idx, lab, amp)


def test_no_spikes():
    V, T, _, _, _ = _case(4)
    idx = np.zeros(0, np.int64)
    R, E, gT, ga, _ = _gpu(V, T, idx, idx, np.zeros(0, np.float32))
    np.testing.assert_array_equal(R, V)
    assert abs(E - 0.5 * np.sum(V.astype(np.float64) ** 2)) <= 1e-9 * E
    assert not np.any(gT)


def test_central_differences_of_gpu_energy():
    V, T, idx, lab, amp = _case(5)
    _, _, gT, ga, _ = _gpu(V, T, idx, lab, amp)
    rng = np.random.default_rng(6)
    gmax = np.max(np.abs(gT))
    for _ in range(8):
        u = int(rng.integers(T.shape[0] - 1))
        c, i = int(rng.integers(T.shape[1])), int(rng.integers(T.shape[2]))
        Tp, Tm = T.copy(), T.copy()
        Tp[u, c, i] += 1.0
        Tm[u, c, i] -= 1.0
        Ep = _gpu(V, Tp, idx, lab, amp)[1]
        Em = _gpu(V, Tm, idx, lab, amp)[1]
        assert abs((Ep - Em) / 2 - gT[u, c, i]) <= 1e-3 * gmax
    amax = np.max(np.abs(ga))
    for k in (0, 7, len(idx) - 1):
        ap, am = amp.copy(), amp.copy()
        ap[k] += 1.0
        am[k] -= 1.0
        Ep = _gpu(V, T, idx, lab, ap)[1]
        Em = _gpu(V, T, idx, lab, am)[1]
        assert abs((Ep - Em) / 2 - ga[k]) <= 1e-3 * amax


def test_run_to_run_identical():
    case = _case(7, n=400)
    a = _gpu(*case)
    b = _gpu(*case)
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(np.asarray(x), np.asarray(y))


def test_autograd_function_scales_by_incoming_gradient():
    V, T, idx, lab, amp = _case(8)
    Vd = torch.tensor(V, device="cuda", requires_grad=True)
    Td = torch.tensor(T, device="cuda", requires_grad=True)
    table = refine.SpikeTable(idx, lab, T.shape[0], V.shape[1])
    a = table.to_sorted(torch.tensor(amp, device="cuda")).requires_grad_(True)
    E = refine.residual_energy(Vd, Td, a, table)
    (3.0 * E).backward()
    E64, gT64, ga64, R64 = refine_fp64.gradients(V, T, idx, lab, amp)
    assert E.dtype == torch.float64
    np.testing.assert_allclose(Vd.grad.cpu().numpy(), 3.0 * R64, rtol=0, atol=1e-4)
    assert np.max(np.abs(Td.grad.cpu().numpy() - 3.0 * gT64)) <= 3e-5 * np.max(np.abs(gT64))
    ga = table.from_sorted(a.grad).cpu().numpy()
    assert np.max(np.abs(ga - 3.0 * ga64)) <= 3e-5 * np.max(np.abs(ga64))


def test_unsorted_spikes_give_nan_energy():
    V, T, idx, lab, amp = _case(9)
    lib = _native.load()
    Vd = torch.tensor(V, device="cuda")
    Td = torch.tensor(T, device="cuda")
    bad = torch.tensor(idx[::-1].copy().astype(np.int32), device="cuda")
    labd = torch.tensor(lab.astype(np.int32), device="cuda")
    ad = torch.tensor(amp, device="cuda")
    C, N = V.shape
    U, _, L = T.shape
    R = torch.empty_like(Vd)
    E = torch.zeros((), dtype=torch.float64, device="cuda")
    nb = ctypes.c_int64(0)
    _native.refine_check(lib.fbp_refine_workspace_bytes(N, len(idx), U, C, L, ctypes.byref(nb)))
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    vp = ctypes.c_void_p
    rc = lib.fbp_refine_forward(vp(Vd.data_ptr()), N, N, C, vp(Td.data_ptr()), U, L, vp(bad.data_ptr()),
                                vp(labd.data_ptr()), vp(ad.data_ptr()), len(idx), vp(R.data_ptr()), N,
                                vp(E.data_ptr()), vp(work.data_ptr()), work.numel(), None)
    assert rc == _native.FBP_OK
    torch.cuda.synchronize()
    assert np.isnan(float(E))
    rc = lib.fbp_refine_forward(vp(Vd.data_ptr()), N, N, C, vp(Td.data_ptr()), U, L, vp(bad.data_ptr()),
                                vp(labd.data_ptr()), vp(ad.data_ptr()), len(idx), vp(R.data_ptr()), N,
                                vp(E.data_ptr()), vp(work.data_ptr()), work.numel() - 1, None)
    assert rc == _native.FBP_E_INVALID


def test_refinement_reduces_template_error():
    d = synthetic.generate("probe32", duration_s=4.0)
    truth = d["templates"].astype(np.float32)
    rng = np.random.default_rng(10)
    T0 = (truth * (1 + 0.1 * rng.standard_normal(truth.shape))).astype(np.float32)
    T, a, hist = refine.refine_templates(d["voltage"], T0, d["truth_onsets"], d["truth_labels"], steps=30,
                                         truth=truth)
    print(hist[0], hist[-1], flush=True)
    assert hist[-1]["energy"] < hist[0]["energy"]
    assert hist[-1]["template_error"] < 0.7 * hist[0]["template_error"]
    assert a.shape == (len(d["truth_onsets"]),)


def _bounds_chunked(V, T, idx, lab, amp, R64, chunk=4096):
    """_bounds for many spikes: the same per-element terms, windows gathered in chunks
    (samples at or past N contribute nothing, as in the kernels)."""
    C, N = V.shape
    U, _, L = T.shape
    S = refine_fp64.residual(np.abs(V), -np.abs(T), idx, lab, np.abs(amp))
    m = refine_fp64.residual(np.zeros((1, N)), -np.ones((U, 1, L)), idx, lab, np.ones(len(idx)))[0]
    dR = 2 * (m + 1) * U_RND * S
    dE = float(np.sum(np.abs(R64) * dR)) + 1e-12 * 0.5 * float(np.sum(R64 ** 2))
    T64 = np.abs(T.astype(np.float64))
    dRp = np.concatenate([dR, np.zeros((C, L))], 1)
    aRp = np.concatenate([np.abs(R64), np.zeros((C, L))], 1)
    idx = np.asarray(idx, np.int64)
    lab = np.asarray(lab, np.int64)
    a = np.abs(np.asarray(amp, np.float64))
    cl = 2 * (C * -(-L // 64) + 7) * U_RND
    dga = np.zeros(len(idx))
    dgT_abs = np.zeros_like(T64)
    dgT_mag = np.zeros_like(T64)
    for k0 in range(0, len(idx), chunk):
        k1 = min(len(idx), k0 + chunk)
        pos = idx[k0:k1, None] + np.arange(L)[None, :]
        wd = dRp[:, pos].transpose(1, 0, 2)
        wa = aRp[:, pos].transpose(1, 0, 2)
        Tk = T64[lab[k0:k1]]
        dga[k0:k1] = np.sum(Tk * wd, axis=(1, 2)) + cl * np.sum(Tk * wa, axis=(1, 2))
        for u in np.unique(lab[k0:k1]):
            s = lab[k0:k1] == u
            dgT_abs[u] += np.einsum("k,kcl->cl", a[k0:k1][s], wd[s])
            dgT_mag[u] += np.einsum("k,kcl->cl", a[k0:k1][s], wa[s])
    n_u = np.bincount(lab, minlength=U)
    chunks = -(-n_u // 256)
    dgT = dgT_abs + 2 * (256 + chunks[:, None, None] + 1) * U_RND * dgT_mag
    return dR, dE, dgT, dga


def test_probe32_full_recording_matches_fp64():
    """The differentiable configuration at its full size: 32 channels, 30 kHz, 120 s,
    60 units, 64-sample templates perturbed from the truth by 10 % Gaussian noise, spike
    times from one pursuit with the perturbed templates.  Every residual sample, the energy
    and every gradient element within its FP32 rounding bound of the FP64 oracle."""
    from spikesorting_fullpursuit_amd.overlap import binary_pursuit_parallel as bpp

    d = synthetic.generate("probe32")
    V = d["voltage"]
    C, N = V.shape
    w = d["workload"]
    truth = d["templates"].astype(np.float32)
    U, _, L = truth.shape
    rng = np.random.default_rng(w.seed + 1)
    Tp = (truth * (1.0 + 0.1 * rng.standard_normal(truth.shape))).astype(np.float32)
    SS, SS_by_chan, thr = synthetic.separability_inputs(Tp.astype(np.float64), w.fs)
    m = {"templates": Tp.reshape(U, C * L).astype(np.float64), "template_SS": SS,
         "template_SS_by_chan": SS_by_chan, "neuron_lower_thresholds": thr}
    print(f"[refine] probe32 full: C={C} N={N} U={U} L={L}", flush=True)
    ev, lab, _, _ = bpp.binary_pursuit(V, np.float32, d["sort_info"], m, n_max_shift_inds=d["n_max_shift_inds"])
    idx = np.asarray(ev, np.int64) - d["clip_init_samples"]
    lab = np.asarray(lab, np.int64)
    assert len(idx) > 100_000
    amp = (1.0 + 0.05 * rng.standard_normal(len(idx))).astype(np.float32)
    R, E, gT, ga, _ = _gpu(V, Tp, idx, lab, amp)
    E64, gT64, ga64, R64 = refine_fp64.gradients(V, Tp, idx, lab, amp)
    dR, dE, dgT, dga = _bounds_chunked(V, Tp, idx, lab, amp, R64)
    assert np.all(np.abs(R - R64) <= dR + 1e-30)
    assert abs(E - E64) <= dE
    assert np.all(np.abs(gT - gT64) <= dgT + 1e-30)
    assert np.all(np.abs(ga - ga64) <= dga + 1e-30)
    assert np.max(np.abs(gT - gT64)) <= 1e-5 * np.max(np.abs(gT64))
    assert np.max(np.abs(ga - ga64)) <= 1e-5 * np.max(np.abs(ga64))
