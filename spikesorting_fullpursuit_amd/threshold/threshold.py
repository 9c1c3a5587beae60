"""Per-channel spike-detection thresholds on the GPU.

Restates median_threshold, single_thresholds and single_thresholds_and_samples
of SF/threshold/threshold.py:6-70 (SF = spikesorting_fullpursuit/) for float32
voltage, with numpy's float32 arithmetic: the nanmedian of |V_c| (exact order
statistics, fbp_abs_medians), then sigma and the 0.6745 MAD factor applied in
the reference's order and types (sigma a Python float).  There is no CPU
fallback.
"""
import ctypes

import numpy as np

from .. import _native


def _device_voltage(voltage, device):
    import torch

    if torch.is_tensor(voltage):
        V = voltage if voltage.is_cuda else voltage.to(torch.device("cuda", device))
        if V.dtype != torch.float32:
            raise TypeError("GPU thresholds take float32 voltage")
    else:
        Vh = np.asarray(voltage)
        if Vh.dtype != np.float32:
            raise TypeError("GPU thresholds take float32 voltage")
        if not torch.cuda.is_available():
            raise RuntimeError("No GPU available")
        V = torch.from_numpy(np.ascontiguousarray(Vh)).to(torch.device("cuda", device))
    if V.dim() == 1:
        V = V[None, :]
    if V.stride(1) != 1:
        V = V.contiguous()
    return V


def _abs_medians(V):
    import torch

    C, N = V.shape
    lib = _native.load()
    nb = ctypes.c_int64(0)
    _native.preproc_check(lib.fbp_preproc_workspace_bytes(N, C, ctypes.byref(nb)))
    work = torch.empty(nb.value, dtype=torch.uint8, device=V.device)
    med = torch.empty(C, dtype=torch.float32, device=V.device)
    vp = ctypes.c_void_p
    _native.preproc_check(lib.fbp_abs_medians(vp(V.data_ptr()), V.stride(0), N, C, vp(med.data_ptr()),
                                              vp(work.data_ptr()), work.numel(),
                                              vp(torch.cuda.current_stream(V.device).cuda_stream)))
    return med.cpu().numpy()


def median_threshold(voltage, sigma, *, device=0):
    """nanmedian(|V_c|) / 0.6745 per channel (float32), then * sigma (float64)
    (threshold.py:6-21)."""
    med = _abs_medians(_device_voltage(voltage, device))
    thresholds = np.empty((med.shape[0],))
    for chan in range(med.shape[0]):
        thresholds[chan] = med[chan] / 0.6745
    thresholds *= sigma
    return thresholds


def single_thresholds(voltage, sigma, *, device=0):
    """sigma * nanmedian(|V_c|) / 0.6745 per channel, in float32 (threshold.py:24-41)."""
    med = _abs_medians(_device_voltage(voltage, device))
    thresholds = np.empty((med.shape[0],))
    for chan in range(med.shape[0]):
        thresholds[chan] = sigma * med[chan] / 0.6745
    return thresholds


def single_thresholds_and_samples(voltage, sigma, *, device=0):
    """(thresholds, samples_over_thresh): single_thresholds and the count of
    samples with |V_c| above the channel's threshold (threshold.py:44-70)."""
    import torch

    V = _device_voltage(voltage, device)
    thresholds = single_thresholds(V, sigma, device=device)
    C, N = V.shape
    thr = torch.from_numpy(thresholds).to(V.device)
    cnt = torch.zeros(C, dtype=torch.int64, device=V.device)
    vp = ctypes.c_void_p
    _native.preproc_check(_native.load().fbp_count_above(
        vp(V.data_ptr()), V.stride(0), N, C, vp(thr.data_ptr()), vp(cnt.data_ptr()),
        vp(torch.cuda.current_stream(V.device).cuda_stream)))
    return thresholds, [int(k) for k in cnt.cpu().numpy()]
