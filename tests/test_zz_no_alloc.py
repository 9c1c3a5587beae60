"""The allocation-free refinement forward pass (fbp_refine_forward_noalloc,
static device scratch) against fbp_refine_forward with a caller workspace:
residual and energy bit-identical, on calls in a row whose sizes shrink, and
after a call that flagged bad input."""
import ctypes

import numpy as np
import pytest
import torch

from spikesorting_fullpursuit_amd import _native, refine

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _inputs(seed, C, N, U, L, n):
    rng = np.random.default_rng(seed)
    V = torch.tensor(rng.normal(size=(C, N)).astype(np.float32), device="cuda")
    T = torch.tensor((rng.normal(size=(U, C, L)) * 3).astype(np.float32), device="cuda")
    idx = np.sort(rng.integers(0, N, size=n)).astype(np.int32)
    lab = rng.integers(0, U, size=n).astype(np.int32)
    amp = torch.tensor(rng.normal(1.0, 0.1, size=n).astype(np.float32), device="cuda")
    return V, T, torch.tensor(idx, device="cuda"), torch.tensor(lab, device="cuda"), amp


def _both(V, T, idx, lab, amp):
    """(R, E) of the workspace entry point and of the allocation-free one."""
    C, N = V.shape
    U, _, L = T.shape
    n = idx.numel()
    lib = _native.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    nbytes = ctypes.c_int64(0)
    _native.refine_check(lib.fbp_refine_workspace_bytes(N, n, U, C, L, ctypes.byref(nbytes)))
    work = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda")
    R1 = torch.empty((C, N), dtype=torch.float32, device="cuda")
    E1 = torch.empty((), dtype=torch.float64, device="cuda")
    _native.refine_check(lib.fbp_refine_forward(_p(V), N, N, C, _p(T), U, L, _p(idx), _p(lab), _p(amp), n,
                                                _p(R1), N, _p(E1), _p(work), work.numel(), stream))
    R2 = torch.full((C, N), 7.0, dtype=torch.float32, device="cuda")
    E2 = torch.full((), 7.0, dtype=torch.float64, device="cuda")
    print(f"[noalloc] C={C} N={N} U={U} L={L} n={n}", flush=True)
    _native.refine_check(lib.fbp_refine_forward_noalloc(_p(V), N, N, C, _p(T), U, L, _p(idx), _p(lab), _p(amp), n,
                                                        _p(R2), N, _p(E2), stream))
    torch.cuda.synchronize()
    return R1.cpu(), E1.cpu(), R2.cpu(), E2.cpu()


def _same(R1, E1, R2, E2):
    assert torch.equal(R1.view(torch.int32), R2.view(torch.int32))
    assert E1.view(torch.int64).item() == E2.view(torch.int64).item()


def test_noalloc_matches_workspace_entry_point_shrinking_sizes():
    for seed, (C, N, U, L, n) in enumerate([(8, 400_000, 6, 48, 3000), (4, 30_000, 3, 32, 200)]):
        R1, E1, R2, E2 = _both(*_inputs(seed, C, N, U, L, n))
        assert np.isfinite(E1.item())
        _same(R1, E1, R2, E2)


def test_noalloc_flags_do_not_leak_into_the_next_call():
    V, T, idx, lab, amp = _inputs(5, 3, 200_000, 4, 40, 1500)
    bad = idx.clone()
    bad[100], bad[101] = idx[101] + 1, idx[100]  # unsorted: energy is NaN
    R1, E1, R2, E2 = _both(V, T, bad, lab, amp)
    assert np.isnan(E1.item()) and np.isnan(E2.item())
    R1, E1, R2, E2 = _both(*_inputs(6, 2, 50_000, 3, 40, 300))
    assert np.isfinite(E2.item())
    _same(R1, E1, R2, E2)


def test_noalloc_rejects_recordings_beyond_its_scratch():
    lib = _native.load()
    rc = lib.fbp_refine_forward_noalloc(None, (1 << 26) + 1, (1 << 26) + 1, 1, None, 1, 8, None, None, None, 0,
                                        None, (1 << 26) + 1, None, None)
    assert rc == _native.FBP_E_UNSUPPORTED
    assert b"67108864" in lib.fbp_refine_last_error()
