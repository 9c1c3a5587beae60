"""Multi-channel spike clips on the GPU.

Restates get_clips and its edge-case rules, SF/processing/clip_utils.py:402-632
(SF = spikesorting_fullpursuit/), with the gather done by the HIP kernel behind
fbp_get_clips (include/fbp.h).  Full binary pursuit calls it over every found
spike and every channel right after the pursuit
(SF/overlap/full_binary_pursuit.py:495-502); here the voltage may be a float32
device tensor already in HBM (for example the recording the pursuit just read),
so the clips are gathered without a host round trip.

There is no CPU fallback: without the HIP library or a GPU the call raises.
"""
import ctypes
import os

import numpy as np

from .. import _native
from .conversions import time_window_to_samples


def check_edge_cases(event_indices, n_samples, valid_event_indices, window):
    """Marks the events get_clips drops and returns (start_ind, stop_ind), the
    first and last kept event (clip_utils.py:402-498).

    The rule only looks at the ends of the (ordered) list: the leading run of
    events whose clip would start before sample 0 and the trailing run whose
    clip would end after n_samples are dropped.  start_ind is len(events) when
    the leading run covers the whole list, stop_ind is -1 when the trailing run
    does; either way every event is then marked invalid.
    """
    ev = np.asarray(event_indices, dtype=np.int64)
    n = ev.shape[0]
    early = ev + window[0] < 0
    late = ev + window[1] > n_samples
    start_ind = int(np.argmin(early)) if not early.all() else n
    stop_ind = n - 1 - int(np.argmin(late[::-1])) if not late.all() else -1
    if start_ind == n or stop_ind < 0:
        valid_event_indices[:] = False
    else:
        valid_event_indices[:start_ind] = False
        valid_event_indices[stop_ind + 1:] = False
    return start_ind, stop_ind


def get_clips(probe_dict, voltage, neighbors, event_indices, clip_width_s, use_memmap=False, check_valid=True, *,
              device=0, return_device=False):
    """Clips of every channel in `neighbors` (in that order) around each event.

    Same arguments, returns and edge rules as clip_utils.py:501-632:
    returns (spike_clips [n_valid][W * len(neighbors)] of probe_dict["v_dtype"],
    valid_event_indices bool[n]); events must be ordered; leading events whose
    clip starts before sample 0 and trailing events whose clip ends after
    probe_dict["n_samples"] are dropped; (None, all-False) when none is kept;
    an empty event list gives a (0, width) array.  `voltage` is [C][N], a host
    array or a torch tensor on the GPU, of any 16-, 32- or 64-bit sample type;
    the clips are moved bit for bit and cast to v_dtype on the host, as the
    reference's slice assignment does.  Keyword-only: `return_device` returns
    the clips as a device tensor of the recording's type instead.
    """
    import torch

    event_indices = np.asarray(event_indices)
    if event_indices.ndim > 1:
        raise ValueError("Event_indices must be one dimensional array of indices")
    sampling_rate = probe_dict["sampling_rate"]
    n_samples = probe_dict["n_samples"]
    v_dtype = probe_dict["v_dtype"]
    window, clip_width_s = time_window_to_samples(clip_width_s, sampling_rate)
    neighbors = np.atleast_1d(np.asarray(neighbors, dtype=np.int64))
    W = window[1] - window[0]
    width = W * len(neighbors)
    if len(event_indices) == 0:
        return np.zeros((0, width), dtype=v_dtype), np.ones(0, dtype="bool")

    valid_event_indices = np.ones(event_indices.shape[0], dtype="bool")
    start_ind = 0
    if check_valid:
        start_ind, stop_ind = check_edge_cases(event_indices, n_samples, valid_event_indices, window)
        if not np.any(valid_event_indices):
            return None, valid_event_indices
    else:
        stop_ind = len(event_indices) - 1
    ev = np.ascontiguousarray(event_indices[start_ind:stop_ind + 1], dtype=np.int64)

    C = int(voltage.shape[0])
    N = int(voltage.shape[1])
    if np.any(neighbors < 0) or np.any(neighbors >= C):
        raise IndexError("neighbors must be channel indices of voltage")
    # the reference slices voltage[chan, n + w0 : n + w1]; a clip outside the
    # recording here (unordered events) would not be a slice of it
    if np.any(ev + window[0] < 0) or np.any(ev + window[1] > N):
        raise ValueError("event clips outside the recording: event_indices must be ordered")

    dev = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
    if torch.is_tensor(voltage):
        Vd = voltage if voltage.is_cuda else voltage.to(dev)
        dev = Vd.device
    else:
        if not torch.cuda.is_available():
            raise RuntimeError("No GPU available")
        Vh = np.asarray(voltage)
        if Vh.dtype.itemsize not in (2, 4, 8):
            Vh = Vh.astype(np.float32)  # exact for 8-bit samples
        Vd = torch.from_numpy(np.ascontiguousarray(Vh)).to(dev)
    if Vd.element_size() not in (2, 4, 8):
        Vd = Vd.to(torch.float32)
    if Vd.stride(1) != 1:
        Vd = Vd.contiguous()
    chans = torch.from_numpy(neighbors.astype(np.int32)).to(dev)
    ev_d = torch.from_numpy(ev).to(dev)
    # the clips keep the recording's element type; the cast to v_dtype follows
    out = torch.empty((len(ev), width), dtype=Vd.dtype, device=dev)
    lib = _native.load()
    vp = ctypes.c_void_p
    _native.clips_check(lib.fbp_get_clips(
        vp(Vd.data_ptr()), Vd.element_size(), Vd.stride(0), N, C, vp(chans.data_ptr()), len(neighbors),
        vp(ev_d.data_ptr()), len(ev), window[0], window[1], vp(out.data_ptr()),
        vp(torch.cuda.current_stream(dev).cuda_stream)))
    if return_device:
        return out, valid_event_indices
    spike_clips = out.cpu().numpy().astype(v_dtype, copy=False)
    if use_memmap:
        clip_fname = os.path.join(probe_dict["memmap_dir"],
                                  "{0}clips_{1}.bin".format(probe_dict["memmap_fID"], str(probe_dict["ID"])))
        mm = np.memmap(clip_fname, dtype=v_dtype, mode="w+", shape=spike_clips.shape)
        mm[:] = spike_clips
        mm.flush()
        del mm
        spike_clips = np.memmap(clip_fname, dtype=v_dtype, mode="r", shape=(len(ev), width))
    return spike_clips, valid_event_indices


def calculate_robust_template(all_channel_clips, *, device=0):
    """Robust template of one unit's clips [n][W] (clip_utils.py:202-247).

    Per column: the median of the values within one median absolute deviation
    (scaled by 1 / 0.6745) of the column median, or the median when none is;
    with numpy's float32 semantics (the mean of the two middle values for an
    even count, rounded to float32).  One clip (or a 1-D input) is returned
    squeezed, as the reference does.  float32 clips, a host array or a device
    tensor; returns a host float32 array.
    """
    import torch

    if all_channel_clips.shape[0] == 1 or all_channel_clips.ndim == 1:
        out = all_channel_clips.squeeze()
        return out.cpu().numpy() if torch.is_tensor(out) else np.squeeze(out)
    dev = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
    if torch.is_tensor(all_channel_clips):
        X = all_channel_clips if all_channel_clips.is_cuda else all_channel_clips.to(dev)
        if X.dtype != torch.float32:
            raise TypeError("the GPU robust template takes float32 clips")
        dev = X.device
    else:
        Xh = np.asarray(all_channel_clips)
        if Xh.dtype != np.float32:
            raise TypeError("the GPU robust template takes float32 clips")
        if not torch.cuda.is_available():
            raise RuntimeError("No GPU available")
        X = torch.from_numpy(np.ascontiguousarray(Xh)).to(dev)
    if X.stride(1) != 1:
        X = X.contiguous()
    n, W = X.shape
    out = torch.empty(W, dtype=torch.float32, device=dev)
    _native.inputs_check(_native.load().fbp_robust_template(
        ctypes.c_void_p(X.data_ptr()), n, X.stride(0), W, ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
    return out.cpu().numpy()
