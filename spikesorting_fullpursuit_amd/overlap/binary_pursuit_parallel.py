"""Drop-in replacement of the reference's binary_pursuit() on MI355X.

Same call signature, inputs read and return arrays as
SF/overlap/binary_pursuit_parallel.py:175-1532 (SF = spikesorting_fullpursuit/).
The device work runs in libfbp_hip.so (HIP kernels for gfx950, C ABI in
include/fbp.h); this module keeps the reference's host logic: argument
handling, the GPU-memory chunk plan and the output assembly.
"""
import numpy as np

from .. import _native
from ..processing.conversions import time_window_to_samples
from . import neuron_separability

# Work-group size the reference's memory model divides the overlap items by
# (CL_KERNEL_WORK_GROUP_SIZE of overlap_recheck_indices,
# binary_pursuit_parallel.py:336-340); 256 is the AMD OpenCL value.
OVERLAP_LOCAL_WORK_SIZE = 256
DEVICE_ADDRESS_BITS = 64

_handles = {}


def _handle(device):
    h = _handles.get(device)
    if h is None:
        h = _native.Handle(device)
        _handles[device] = h
    return h


def release_device_buffers():
    """Free the cached device handles and their buffers."""
    for h in list(_handles.values()):
        h.close()
    _handles.clear()


# Limits of libfbp_hip.so's buffers (csrc/fbp_api.hip) used by library_device_bytes.
_MAX_GRAM_BYTES = 8 << 30
_MAX_SCREEN_BUFFER_BYTES = 2 << 30
_SCREEN_BATCH_BYTES = 8 << 30
_SCREEN_BATCH_WINDOWS = 262144
_STAGING_BYTES = 3 * (32 << 20)
# Bytes per window besides the per-template maxima: best index / label / value and
# the overlap copies (5 x 4), five flag bytes, four window lists (4 x 4), the
# dirty-channel mask (8) and the spike list with room to grow (4 x 8).
_PER_WINDOW_BYTES = 20 + 5 + 16 + 8 + 32


def library_device_bytes(n_chans, n_samples, n_templates, template_samples_per_chan, n_max_shift_inds,
                         shared_pairs=None):
    """Device memory libfbp_hip.so allocates for one chunk of n_samples:
    (constant bytes, bytes per sample).  shared_pairs is the number of ordered
    template pairs that share a channel (U^2 when None: every pair)."""
    C, U, L = int(n_chans), int(n_templates), int(template_samples_per_chan)
    S = 2 * int(n_max_shift_inds) + 1
    pairs = U * U if shared_pairs is None else max(1, int(shared_pairs))
    gram = pairs * S * (((S + 3) // 4) * 4) * 4 + pairs * S * 4 + pairs * S * (-(-S // 8)) * 4
    if pairs * S * S * 4 > _MAX_GRAM_BYTES:
        gram = 0  # the literal recheck runs instead
    screen = 0
    if U * S * 4 <= _MAX_SCREEN_BUFFER_BYTES:
        n_windows = -(-max(int(n_samples), 1) // L)  # the library holds no more rows than windows
        batch = max(1, min(_SCREEN_BATCH_WINDOWS, n_windows, _SCREEN_BATCH_BYTES // (U * S * 4)))
        screen = batch * U * S * 4 + U * S * 4 + batch * 65 * 4  # B rows, Ntr, deferred-rescoring records
    templates = 3 * U * C * L * 4
    constant = gram + screen + templates + _STAGING_BYTES
    # BF16 likelihood (the largest of the precisions): the quantized voltage (2 C), its
    # 16-sample block maxima (C / 4) and one rescore byte per (window, template)
    per_sample = 4.0 * C + 4.0 * U + (6.0 * U + _PER_WINDOW_BYTES) / L + 2.25 * C + float(U) / L
    return constant, per_sample


def chunk_plan(n_chans, n_samples, sampling_rate, n_templates, template_samples_per_chan,
               n_max_shift_inds, gpu_total_memory, max_gpu_memory=None, template_bytes=None,
               resident_bytes=0, shared_pairs=None):
    """Chunk length and onsets, binary_pursuit_parallel.py:387-507.

    Returns (num_indices_per_chunk, chunk_onsets).  Raises ValueError when a
    chunk would hold fewer than 4 template widths (:472-475).

    The reference's formula gives the chunk length wherever that chunk fits
    this library's own buffers (library_device_bytes) beside `resident_bytes`
    of caller data already on the device (a voltage tensor passed in place)
    with 1 GiB to spare; otherwise the chunk is shortened to the longest that
    fits.  Its cost model counts buffers this library does not have (the
    overlap kernels' per-work-group partials) and misses ones it has (the
    screen tables and buffer), so on long recordings the two differ.
    """
    gpu_memory_size = float(gpu_total_memory)
    gpu_memory_size = max(gpu_memory_size * 0.5, gpu_memory_size - (1024 * 1024 * 1024))
    if max_gpu_memory is not None:
        gpu_memory_size = min(gpu_memory_size, max_gpu_memory)
    assert gpu_memory_size > 0
    L = int(template_samples_per_chan)
    items_per_index = np.int64((2 * int(n_max_shift_inds) + 1) ** 2 * int(n_templates))
    n_local_ID = np.int64(((items_per_index - 1) / OVERLAP_LOCAL_WORK_SIZE) + 1)
    if template_bytes is None:
        template_bytes = 4 * int(n_templates) * int(n_chans) * L
    constant_memory_usage = template_bytes + 4 * int(n_templates)
    memory_usage_per_second = n_chans * sampling_rate * 4.0
    memory_usage_per_second += n_templates * sampling_rate * 4.0
    memory_usage_per_second += (sampling_rate / L) * (
        4 + 4 + 4 + 4 + 4 + (n_local_ID / 3) * 4 + (n_local_ID / 3) * 4)
    num_seconds_per_chunk = (gpu_memory_size - constant_memory_usage) / memory_usage_per_second
    max_addressable_seconds = ((1 << DEVICE_ADDRESS_BITS) - 1) / (4.0 * sampling_rate * n_chans)
    num_seconds_per_chunk = min(num_seconds_per_chunk, np.floor(max_addressable_seconds))
    num_indices_per_chunk = int(np.floor(num_seconds_per_chunk * sampling_rate))
    constant, per_sample = library_device_bytes(n_chans, n_samples, n_templates, L, n_max_shift_inds,
                                                shared_pairs)
    fit = (float(gpu_total_memory) - (1 << 30) - float(resident_bytes) - constant) / per_sample
    num_indices_per_chunk = min(num_indices_per_chunk, int(np.floor(max(fit, 0.0))))
    if num_indices_per_chunk < 4 * L:
        raise ValueError(
            "Cannot fit enough data on GPU to run binary pursuit. Decrease template samples or get a new GPU.")
    onsets = []
    curr = 0
    while curr < n_samples - L:
        onsets.append(curr)
        curr += num_indices_per_chunk - 3 * L
    return num_indices_per_chunk, onsets


def _is_device_tensor(x):
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def prepare_inputs(voltage, sort_info, separability_metrics, n_max_shift_inds=None, *, recheck_literal=False,
                   precision="mixed"):
    """The drop-in's reading of its arguments (binary_pursuit_parallel.py:175-300,
    :480-507): kernel inputs (templates [U][C][L], bias, thresholds), the
    library flags, the shift range and what the chunk plan needs."""
    n_chans = int(voltage.shape[0])
    n_samples = int(voltage.shape[1])
    sampling_rate = sort_info["sampling_rate"]
    chan_win, _ = time_window_to_samples(sort_info["clip_width"], sampling_rate)
    template_samples_per_chan = int(sort_info["n_samples_per_chan"])

    templates = separability_metrics["templates"]
    if templates.ndim == 1:
        templates = np.reshape(templates, (1, -1))
    templates = np.float32(templates)
    n_templates = templates.shape[0]
    if templates.shape[1] != n_chans * template_samples_per_chan:
        raise ValueError("templates must have n_channels * n_samples_per_chan columns")
    clip_init_samples = int(np.abs(chan_win[0]))
    template_sum_squared = np.float32(-0.5 * separability_metrics["template_SS"])
    _ = np.float32(separability_metrics["template_SS_by_chan"].flatten(order="C"))
    neuron_lower_thresholds = np.float32(separability_metrics["neuron_lower_thresholds"])
    do_overlap_recheck = sort_info["do_overlap_recheck"]
    get_adjusted_clips = sort_info["get_adjusted_clips"]
    if n_max_shift_inds is None or n_max_shift_inds < 0:
        n_max_shift_inds = template_samples_per_chan // 2
    n_max_shift_inds = np.uint32(n_max_shift_inds)

    on_device = _is_device_tensor(voltage)
    resident = 0
    if on_device:
        import torch
        resident = voltage.numel() * voltage.element_size()
        if voltage.dtype != torch.float32 or voltage.stride(1) != 1:
            resident += voltage.numel() * 4  # the float32 copy made below
    nz = np.any(templates.reshape(n_templates, n_chans, template_samples_per_chan) != 0, axis=2)
    shared_pairs = int(np.count_nonzero((nz.astype(np.int64) @ nz.T.astype(np.int64)) > 0))
    flags = 0
    if do_overlap_recheck:
        flags |= _native.FLAG_OVERLAP_RECHECK
    if get_adjusted_clips:
        flags |= _native.FLAG_ADJUSTED_CLIPS
    if recheck_literal:
        flags |= _native.FLAG_RECHECK_LITERAL
    if precision == "mixed":
        flags |= _native.FLAG_PRECISION_MIXED
    elif precision == "bf16":
        flags |= _native.FLAG_PRECISION_BF16
    elif precision == "fp8":
        flags |= _native.FLAG_PRECISION_FP8
    else:
        flags |= _native.FLAG_SCREEN_FP32
    return {
        "n_chans": n_chans, "n_samples": n_samples, "sampling_rate": sampling_rate,
        "template_samples_per_chan": template_samples_per_chan, "n_templates": n_templates,
        "T3": templates.reshape(n_templates, n_chans, template_samples_per_chan),
        "bias": np.atleast_1d(template_sum_squared), "thr": np.atleast_1d(neuron_lower_thresholds),
        "template_bytes": templates.nbytes, "clip_init_samples": clip_init_samples,
        "n_max_shift_inds": n_max_shift_inds, "flags": flags, "get_adjusted_clips": get_adjusted_clips,
        "on_device": on_device, "resident": resident, "shared_pairs": shared_pairs,
    }


def assemble_outputs(secret_spike_indices, secret_spike_labels, secret_spike_bool, adjusted_spike_clips, v_dtype,
                     clip_init_samples):
    """The drop-in's return arrays from its per-chunk lists
    (binary_pursuit_parallel.py:1506-1532)."""
    if len(secret_spike_indices) > 0:
        # one pass per output array: a single chunk is converted directly (new arrays,
        # the inputs are not modified), several are joined first
        if len(secret_spike_indices) == 1:
            event_indices = np.add(secret_spike_indices[0], clip_init_samples, dtype=np.int64)
            neuron_labels = np.asarray(secret_spike_labels[0]).astype(np.int64)
        else:
            event_indices = np.hstack(secret_spike_indices).astype(np.int64, copy=False)
            event_indices += clip_init_samples
            neuron_labels = np.hstack(secret_spike_labels).astype(np.int64, copy=False)
        adjusted_clips = (np.vstack(adjusted_spike_clips)).astype(v_dtype)
        binary_pursuit_spike_bool = np.hstack(secret_spike_bool)
        return event_indices, neuron_labels, binary_pursuit_spike_bool, adjusted_clips
    return [], [], [], []



def binary_pursuit(voltage, v_dtype, sort_info, separability_metrics, n_max_shift_inds=None,
                   kernels_path=None, max_gpu_memory=None, *, device=0, return_residual=False,
                   recheck_literal=False, precision="mixed", stats=None, handle=None):
    """Binary pursuit over every chunk of `voltage` ([C][N], any real dtype;
    a float32 torch tensor already on the GPU is used in place).

    Returns (event_indices int64, neuron_labels int64, binary_pursuit_spike_bool,
    adjusted_clips) exactly as binary_pursuit_parallel.py:1506-1532, or four
    empty lists when nothing is found.  `kernels_path` is accepted and ignored
    (there is no OpenCL source).  Keyword-only extras: `return_residual` appends
    the list of per-chunk residual voltages; `stats` (a list) receives the
    device statistics of every chunk; `precision` "mixed" (default) computes the
    likelihood and the overlap screen's correlations with fp16 matrix-core
    products and recomputes in exact FP32 every value that can influence a
    decision, "fp32" computes both on the FP32 vector units -- the two give the
    same spikes and residual; `recheck_literal` evaluates every overlap item
    instead of screening them; `handle` runs on that library handle (its own
    stream and buffers) instead of the cached one of `device`, so that several
    threads can run pursuits concurrently.
    """
    if precision not in ("mixed", "fp32", "bf16", "fp8"):
        raise ValueError("precision must be 'mixed', 'fp32', 'bf16' or 'fp8'")
    p = prepare_inputs(voltage, sort_info, separability_metrics, n_max_shift_inds, recheck_literal=recheck_literal,
                       precision=precision)
    h = handle if handle is not None else _handle(device)
    on_device = p["on_device"]
    n_samples = p["n_samples"]
    n_max_shift_inds = p["n_max_shift_inds"]
    flags = p["flags"]
    get_adjusted_clips = p["get_adjusted_clips"]
    num_indices_per_chunk, chunk_onsets = chunk_plan(
        p["n_chans"], n_samples, p["sampling_rate"], p["n_templates"], p["template_samples_per_chan"],
        n_max_shift_inds, h.device_memory(), max_gpu_memory, template_bytes=p["template_bytes"],
        resident_bytes=p["resident"], shared_pairs=p["shared_pairs"])
    h.set_templates(p["T3"], p["bias"], p["thr"])
    if on_device:
        import torch
        if voltage.dtype != torch.float32 or voltage.stride(1) != 1:
            voltage = voltage.to(torch.float32).contiguous()
        # the library works on its own stream: the producer's writes must be done
        torch.cuda.current_stream(voltage.device).synchronize()

    secret_spike_indices, secret_spike_labels, secret_spike_bool = [], [], []
    adjusted_spike_clips, residuals = [], []
    for start_index in chunk_onsets:
        stop_index = min(n_samples, start_index + num_indices_per_chunk)
        if on_device:
            h.run(None, n_max_shift_inds, flags, ld=voltage.stride(0), n_samples=stop_index - start_index,
                  device_ptr=voltage.data_ptr() + 4 * start_index)
        else:
            # float32 rows are read in place (the library stages them through pinned
            # pieces); other types are cast per chunk as binary_pursuit_parallel.py:524-527
            chunk_voltage = voltage[:, start_index:stop_index]
            if (not isinstance(chunk_voltage, np.ndarray) or chunk_voltage.dtype != np.float32
                    or chunk_voltage.strides[1] != 4 or chunk_voltage.strides[0] % 4 != 0):
                chunk_voltage = np.float32(chunk_voltage)
            h.run(chunk_voltage, n_max_shift_inds, flags)
        idx, lab = h.result()
        if stats is not None:
            stats.append(h.stats())
        if return_residual:
            residuals.append(h.residual(stop_index - start_index))
        if idx.shape[0] > 0:
            clips = h.clips(idx.shape[0]) if get_adjusted_clips else []
            chunk_indices = idx.astype(np.int64)
            if start_index:
                chunk_indices += start_index
            secret_spike_indices.append(chunk_indices)
            secret_spike_labels.append(lab)
            adjusted_spike_clips.append(clips)
            secret_spike_bool.append(np.ones(idx.shape[0], dtype="bool"))

    event_indices, neuron_labels, binary_pursuit_spike_bool, adjusted_clips = assemble_outputs(
        secret_spike_indices, secret_spike_labels, secret_spike_bool, adjusted_spike_clips, v_dtype,
        p["clip_init_samples"])
    neuron_separability.add_n_spikes(separability_metrics, neuron_labels)
    out = (event_indices, neuron_labels, binary_pursuit_spike_bool, adjusted_clips)
    if return_residual:
        out = out + (residuals,)
    return out
