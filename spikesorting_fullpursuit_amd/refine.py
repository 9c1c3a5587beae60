"""Template refinement at fixed spike times on the GPU.

The reference has no counterpart: binary pursuit subtracts each found template
once (SF/kernels/binary_pursuit.cl:57-102, SF = spikesorting_fullpursuit/) and
never differentiates.  Here the residual energy

    R[c, x] = V[c, x] - sum_k a_k T[lab_k, c, x - p_k]     (0 <= x - p_k < L, x < N)
    E       = 0.5 * sum R^2

is a differentiable function of the templates T [U][C][L] and of per-spike
amplitudes a_k, at spike onsets p_k and labels fixed by a pursuit pass.  The
forward pass and both gradients are hand-written HIP kernels behind the C ABI
(fbp_refine_forward / fbp_refine_backward in include/fbp.h); this module wraps
them in a torch.autograd.Function and runs an Adam refinement loop.

There is no CPU fallback: without the HIP library or a GPU the calls raise.
"""
import ctypes
import time

import numpy as np
import torch

from . import _native


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class SpikeTable:
    """Spike onsets, labels and the by-unit CSR index on the device.

    Onsets are sorted ascending (labels, then input order, break ties) as the
    forward kernel requires; `order` maps the sorted positions back to the
    caller's spike order.  The table is built once per spike set.
    """

    def __init__(self, indices, labels, n_units, n_samples, device="cuda"):
        idx = torch.as_tensor(np.asarray(indices, dtype=np.int64) if not torch.is_tensor(indices) else indices)
        lab = torch.as_tensor(np.asarray(labels, dtype=np.int64) if not torch.is_tensor(labels) else labels)
        idx = idx.to(device=device, dtype=torch.int64).reshape(-1)
        lab = lab.to(device=device, dtype=torch.int64).reshape(-1)
        if idx.shape != lab.shape:
            raise ValueError("indices and labels must have the same length")
        n = idx.numel()
        if n_units <= 0:
            raise ValueError("n_units must be positive")
        if n_samples >= 2 ** 31:
            raise ValueError("recordings of 2^31 samples or more are not supported")
        if n:
            if int(idx.min()) < 0 or int(idx.max()) >= n_samples:
                raise ValueError("spike onsets must lie in [0, n_samples)")
            if int(lab.min()) < 0 or int(lab.max()) >= n_units:
                raise ValueError("labels must lie in [0, n_units)")
        key = idx * n_units + lab
        order = torch.sort(key, stable=True).indices
        self.order = order
        self.idx = idx[order].to(torch.int32).contiguous()
        self.lab = lab[order].to(torch.int32).contiguous()
        # spikes grouped by unit, ascending onset within a unit
        by_unit = torch.sort(self.lab.to(torch.int64) * (n_samples + 1) + self.idx.to(torch.int64), stable=True).indices
        self.by_unit = by_unit.to(torch.int32).contiguous()
        counts = torch.bincount(self.lab.to(torch.int64), minlength=n_units)
        off = torch.zeros(n_units + 1, dtype=torch.int64, device=idx.device)
        off[1:] = torch.cumsum(counts, 0)
        self.offsets = off.to(torch.int32).contiguous()
        self.n = n
        self.U = n_units
        self.N = n_samples

    def to_sorted(self, per_spike):
        """Reorder a per-spike tensor from the caller's order into the table's."""
        return per_spike[self.order]

    def from_sorted(self, per_spike):
        out = torch.empty_like(per_spike)
        out[self.order] = per_spike
        return out


def _workspace(N, n, U, C, L, device):
    b = ctypes.c_int64(0)
    _native.refine_check(_native.load().fbp_refine_workspace_bytes(int(N), int(n), int(U), int(C), int(L),
                                                                   ctypes.byref(b)))
    return torch.empty(int(b.value), dtype=torch.uint8, device=device)


def _check_inputs(V, T, amp, table):
    if not (V.is_cuda and T.is_cuda and amp.is_cuda):
        raise ValueError("voltage, templates and amplitudes must be device tensors")
    if V.dtype != torch.float32 or T.dtype != torch.float32 or amp.dtype != torch.float32:
        raise ValueError("voltage, templates and amplitudes must be float32")
    if V.dim() != 2 or T.dim() != 3:
        raise ValueError("voltage must be [C][N] and templates [U][C][L]")
    U, C, L = T.shape
    if V.shape[0] != C or V.shape[1] != table.N or U != table.U:
        raise ValueError("voltage, templates and spike table disagree in shape")
    if amp.shape != (table.n,):
        raise ValueError("one amplitude per spike")


def forward_residual(V, T, amp_sorted, table):
    """(R [C][N] float32, E float64 scalar tensor) on the current stream."""
    _check_inputs(V, T, amp_sorted, table)
    V = V if V.stride(1) == 1 else V.contiguous()
    T = T.contiguous()
    amp_sorted = amp_sorted.contiguous()
    U, C, L = T.shape
    N = V.shape[1]
    R = torch.empty((C, N), dtype=torch.float32, device=V.device)
    E = torch.empty((), dtype=torch.float64, device=V.device)
    work = _workspace(N, table.n, U, C, L, V.device)
    lib = _native.load()
    _native.refine_check(lib.fbp_refine_forward(
        _ptr(V), V.stride(0), N, C, _ptr(T), U, L, _ptr(table.idx), _ptr(table.lab), _ptr(amp_sorted), table.n,
        _ptr(R), R.stride(0), _ptr(E), _ptr(work), work.numel(), _stream(V.device)))
    return R, E


def backward_grads(R, T, amp_sorted, table, need_T=True, need_amp=True):
    """(dE/dT, dE/da in table order) from the residual of forward_residual."""
    U, C, L = T.shape
    N = R.shape[1]
    T = T.contiguous()
    gT = torch.empty_like(T) if need_T else None
    ga = torch.empty_like(amp_sorted) if need_amp else None
    if not need_T and not need_amp:
        return None, None
    work = _workspace(N, table.n, U, C, L, R.device)
    lib = _native.load()
    _native.refine_check(lib.fbp_refine_backward(
        _ptr(R), R.stride(0), N, C, _ptr(T), U, L, _ptr(table.idx), _ptr(table.lab), _ptr(amp_sorted), table.n,
        _ptr(table.by_unit), _ptr(table.offsets), _ptr(gT), _ptr(ga), _ptr(work), work.numel(), _stream(R.device)))
    return gT, ga


class ResidualEnergy(torch.autograd.Function):
    """E = 0.5 * ||V - sum_k a_k T[lab_k] placed at p_k||^2 (float64 scalar).

    Inputs: voltage [C][N], templates [U][C][L], amplitudes [n] (table order),
    table (SpikeTable).  Gradients flow to voltage (R), templates and
    amplitudes; the backward kernels give dE/dT and dE/da, scaled here by the
    incoming gradient.
    """

    @staticmethod
    def forward(ctx, V, T, amp_sorted, table):
        R, E = forward_residual(V, T, amp_sorted, table)
        ctx.table = table
        ctx.save_for_backward(R, T, amp_sorted)
        return E

    @staticmethod
    def backward(ctx, gE):
        R, T, amp_sorted = ctx.saved_tensors
        need_V, need_T, need_a = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        gT, ga = backward_grads(R, T, amp_sorted, ctx.table, need_T, need_a)
        g = gE.to(torch.float32)
        return (R * g if need_V else None, gT * g if need_T else None, ga * g if need_a else None, None)


def residual_energy(V, T, amp_sorted, table):
    return ResidualEnergy.apply(V, T, amp_sorted, table)


def refine_templates(voltage, templates, indices, labels, steps=50, lr=0.1, lr_amplitudes=0.005, amplitudes=None,
                     refine_amplitudes=True, channel_mask=None, truth=None, device="cuda"):
    """Adam on the residual energy at fixed spike times (step sizes lr for the
    templates, lr_amplitudes for the amplitudes, which start at 1).

    voltage [C][N], templates [U][C][L] (array or tensor), indices/labels of the
    spikes (onsets within the recording, labels < U).  channel_mask [U][C]
    (default: the template's non-zero channels) limits which template rows move.
    With `truth` [U][C][L], the history records the relative template error
    ||T_eff - truth|| / ||truth|| where T_eff[u] = T[u] * mean amplitude of u.

    Returns (templates [U][C][L] tensor, amplitudes in the caller's spike
    order, history list of dicts).
    """
    V = torch.as_tensor(voltage, device=device, dtype=torch.float32)
    T0 = torch.as_tensor(templates, device=device, dtype=torch.float32)
    U, C, L = T0.shape
    N = V.shape[1]
    table = SpikeTable(indices, labels, U, N, device=device)
    if channel_mask is None:
        channel_mask = (T0.abs().amax(dim=2) > 0)
    mask = torch.as_tensor(channel_mask, device=device).to(torch.float32)[:, :, None].expand(U, C, L).contiguous()
    T = (T0 * mask).clone().requires_grad_(True)
    if amplitudes is None:
        a = torch.ones(table.n, device=device, dtype=torch.float32)
    else:
        a = table.to_sorted(torch.as_tensor(amplitudes, device=device, dtype=torch.float32))
    a = a.clone().requires_grad_(bool(refine_amplitudes))
    groups = [{"params": [T], "lr": lr}]
    if refine_amplitudes:
        groups.append({"params": [a], "lr": lr_amplitudes})
    opt = torch.optim.Adam(groups)
    truth_t = torch.as_tensor(truth, device=device, dtype=torch.float32) if truth is not None else None
    lab_long = table.lab.to(torch.int64)
    counts = torch.bincount(lab_long, minlength=U).clamp(min=1).to(torch.float32)

    def template_error():
        if truth_t is None:
            return None
        with torch.no_grad():
            mean_a = torch.zeros(U, device=device).index_add_(0, lab_long, a.detach()) / counts
            eff = T.detach() * mean_a[:, None, None]
            return float(torch.linalg.vector_norm(eff - truth_t) / torch.linalg.vector_norm(truth_t))

    history = []
    err0 = template_error()
    for it in range(steps + 1):
        opt.zero_grad(set_to_none=True)
        E = residual_energy(V, T, a, table)
        if it == steps:
            history.append({"step": it, "energy": float(E.detach()), "template_error": template_error()})
            break
        E.backward()
        with torch.no_grad():
            T.grad.mul_(mask)
        opt.step()
        if it == 0:
            history.append({"step": 0, "energy": float(E.detach()), "template_error": err0})
    a_out = table.from_sorted(a.detach())
    return T.detach(), a_out, history


def time_gradient(V, T, amp_sorted, table, reps=10):
    """Seconds per forward + backward (both gradients) on the current stream."""
    T = T.detach().clone().requires_grad_(True)
    a = amp_sorted.detach().clone().requires_grad_(True)
    E = residual_energy(V, T, a, table)
    E.backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        T.grad = None
        a.grad = None
        E = residual_energy(V, T, a, table)
        E.backward()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps
