"""ctypes binding of libfbp_hip.so (the C ABI declared in include/fbp.h).

The library is built in-tree by spikesorting_fullpursuit_amd.build.  There is
no CPU fallback: if the library or a GPU is missing, calls raise.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libfbp_hip.so")

FBP_OK = 0
FBP_E_INVALID = -1
FBP_E_NODEVICE = -2
FBP_E_HIP = -3
FBP_E_NOMEM = -4
FBP_E_UNSUPPORTED = -5
FBP_E_STATE = -6

FLAG_OVERLAP_RECHECK = 1
FLAG_ADJUSTED_CLIPS = 2
FLAG_RECHECK_LITERAL = 4
FLAG_TIMING = 8
FLAG_SCREEN_FP32 = 16
FLAG_PRECISION_MIXED = 32
FLAG_SELECT_STREAMING = 64
FLAG_LIKELIHOOD_MFMA = 128
FLAG_PRECISION_BF16 = 256
FLAG_PRECISION_FP8 = 512
FLAG_FIRST_PASS_ONLY = 1024
FLAG_NO_RESCORE = 2048
FLAG_MMA_EVERY_PASS = 4096
FLAG_BATCH_GRAPH = 8192

# Every symbol include/fbp.h declares.
EXPORTS = (
    "fbp_create", "fbp_destroy", "fbp_last_error", "fbp_device_memory", "fbp_set_templates",
    "fbp_run", "fbp_run_batch", "fbp_result_count", "fbp_copy_result", "fbp_copy_clips", "fbp_copy_residual",
    "fbp_get_stats", "fbp_copy_likelihood", "fbp_get_pass_log", "fbp_refine_workspace_bytes", "fbp_refine_forward", "fbp_refine_forward_noalloc", "fbp_refine_backward",
    "fbp_refine_last_error", "fbp_get_clips", "fbp_clips_last_error", "fbp_covariance_workspace_bytes",
    "fbp_channel_covariance", "fbp_robust_template", "fbp_template_variances", "fbp_inputs_last_error",
    "fbp_preproc_workspace_bytes", "fbp_abs_medians", "fbp_count_above", "fbp_channel_cov_rows", "fbp_zca_apply",
    "fbp_preproc_last_error", "fbp_overlap_residuals", "fbp_qa_last_error",
)


class FbpStats(ctypes.Structure):
    _fields_ = [
        ("passes", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("spikes", ctypes.c_int64),
        ("windows_listed", ctypes.c_int64),
        ("windows_flagged", ctypes.c_int64),
        ("likelihood_positions", ctypes.c_int64),
        ("likelihood_flops", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("ms_likelihood", ctypes.c_double),
        ("ms_window_max", ctypes.c_double),
        ("ms_recheck", ctypes.c_double),
        ("ms_pursuit", ctypes.c_double),
        ("ms_residual", ctypes.c_double),
        ("ms_other", ctypes.c_double),
        ("n_likelihood_launches", ctypes.c_int64),
        ("n_recheck_launches", ctypes.c_int64),
        ("recheck_items", ctypes.c_int64),
        ("recheck_overflow", ctypes.c_int64),
        ("likelihood_rescored", ctypes.c_int64),
        ("screen_shared_pairs", ctypes.c_int64),
        ("screen_separable_pairs", ctypes.c_int64),
        ("ms_select", ctypes.c_double),
        ("n_select_launches", ctypes.c_int64),
        ("likelihood_pairs_computed", ctypes.c_int64),
        ("n_likelihood_mfma_launches", ctypes.c_int64),
        ("ms_rescore", ctypes.c_double),
        ("n_rescore_launches", ctypes.c_int64),
        ("n_graph_builds", ctypes.c_int64),
        ("n_graph_replays", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved0"}


_lib = None


def load():
    """Load libfbp_hip.so and declare its signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m spikesorting_fullpursuit_amd.build`")
    # One HIP runtime per process: torch ships its own libamdhip64 with the same
    # SONAME (libamdhip64.so.7) as /opt/rocm's, so whichever is loaded first serves
    # every later user, and torch cannot initialise its devices on a runtime other
    # than its own.  Load torch's first when torch is installed.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32
    L.fbp_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.fbp_destroy.argtypes = [vp]
    L.fbp_last_error.argtypes = [vp]
    L.fbp_last_error.restype = ctypes.c_char_p
    L.fbp_device_memory.argtypes = [vp, ctypes.POINTER(i64)]
    L.fbp_set_templates.argtypes = [vp, vp, i32, i32, i32, vp, vp]
    L.fbp_run.argtypes = [vp, vp, i64, i64, i32, u32, u32]
    L.fbp_run_batch.argtypes = [vp, i32, vp, vp, vp, i32, vp, u32, i32, i32]
    L.fbp_result_count.argtypes = [vp, ctypes.POINTER(i64)]
    L.fbp_copy_result.argtypes = [vp, vp, vp]
    L.fbp_copy_clips.argtypes = [vp, vp]
    L.fbp_copy_residual.argtypes = [vp, vp]
    L.fbp_get_stats.argtypes = [vp, ctypes.POINTER(FbpStats)]
    L.fbp_get_pass_log.argtypes = [vp, vp, i32, ctypes.POINTER(i32)]
    L.fbp_copy_likelihood.argtypes = [vp, vp]
    L.fbp_refine_workspace_bytes.argtypes = [i64, i64, i32, i32, i32, ctypes.POINTER(i64)]
    L.fbp_refine_forward.argtypes = [vp, i64, i64, i32, vp, i32, i32, vp, vp, vp, i64, vp, i64, vp, vp, i64, vp]
    L.fbp_refine_forward_noalloc.argtypes = [vp, i64, i64, i32, vp, i32, i32, vp, vp, vp, i64, vp, i64, vp, vp]
    L.fbp_refine_backward.argtypes = [vp, i64, i64, i32, vp, i32, i32, vp, vp, vp, i64, vp, vp, vp, vp, vp, i64,
                                      vp]
    L.fbp_refine_last_error.argtypes = []
    L.fbp_refine_last_error.restype = ctypes.c_char_p
    L.fbp_get_clips.argtypes = [vp, i32, i64, i64, i32, vp, i32, vp, i64, i32, i32, vp, vp]
    L.fbp_clips_last_error.argtypes = []
    L.fbp_clips_last_error.restype = ctypes.c_char_p
    L.fbp_covariance_workspace_bytes.argtypes = [i64, i32, i32, ctypes.POINTER(i64)]
    L.fbp_channel_covariance.argtypes = [vp, i64, i64, i32, i32, vp, i64, vp, vp, i64, vp]
    L.fbp_robust_template.argtypes = [vp, i64, i64, i64, vp, vp]
    L.fbp_template_variances.argtypes = [vp, i32, i32, i32, vp, vp, vp]
    L.fbp_inputs_last_error.argtypes = []
    L.fbp_inputs_last_error.restype = ctypes.c_char_p
    L.fbp_preproc_workspace_bytes.argtypes = [i64, i32, ctypes.POINTER(i64)]
    L.fbp_abs_medians.argtypes = [vp, i64, i64, i32, vp, vp, i64, vp]
    L.fbp_count_above.argtypes = [vp, i64, i64, i32, vp, vp, vp]
    L.fbp_channel_cov_rows.argtypes = [vp, i64, i64, i32, i32, vp, vp, i64, vp]
    L.fbp_zca_apply.argtypes = [vp, vp, i64, i64, i32, vp, i64, vp]
    L.fbp_preproc_last_error.argtypes = []
    L.fbp_preproc_last_error.restype = ctypes.c_char_p
    L.fbp_overlap_residuals.argtypes = [vp, i32, i32, i32, i32, i32, vp, i64, vp, vp, vp]
    L.fbp_qa_last_error.argtypes = []
    L.fbp_qa_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if not name.endswith("last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def refine_check(rc):
    """Raise for a failed handle-free refinement call."""
    if rc != FBP_OK:
        msg = load().fbp_refine_last_error()
        _raise(rc, msg.decode() if msg else f"error {rc}")


def inputs_check(rc):
    """Raise for a failed covariance / robust template / variance call."""
    if rc != FBP_OK:
        msg = load().fbp_inputs_last_error()
        _raise(rc, msg.decode() if msg else f"error {rc}")


def preproc_check(rc):
    """Raise for a failed preprocessing call."""
    if rc != FBP_OK:
        msg = load().fbp_preproc_last_error()
        _raise(rc, msg.decode() if msg else f"error {rc}")


def qa_check(rc):
    """Raise for a failed template-QA call."""
    if rc != FBP_OK:
        msg = load().fbp_qa_last_error()
        _raise(rc, msg.decode() if msg else f"error {rc}")


def clips_check(rc):
    """Raise for a failed fbp_get_clips call."""
    if rc != FBP_OK:
        msg = load().fbp_clips_last_error()
        _raise(rc, msg.decode() if msg else f"error {rc}")


def _raise(code, msg):
    if code == FBP_E_NODEVICE:
        raise RuntimeError("No GPU available")
    if code in (FBP_E_INVALID, FBP_E_UNSUPPORTED):
        raise ValueError(msg)
    if code == FBP_E_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


class Handle:
    """One device + stream + buffers (fbp_handle)."""

    def __init__(self, device=0):
        self._lib = load()
        self._h = ctypes.c_void_p()
        rc = self._lib.fbp_create(int(device), ctypes.byref(self._h))
        if rc != FBP_OK:
            _raise(rc, f"fbp_create failed ({rc})")
        self.C = self.L = self.U = None

    def close(self):
        if self._h:
            self._lib.fbp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != FBP_OK:
            msg = self._lib.fbp_last_error(self._h)
            _raise(rc, msg.decode() if msg else f"error {rc}")

    def device_memory(self):
        n = ctypes.c_int64(0)
        self._check(self._lib.fbp_device_memory(self._h, ctypes.byref(n)))
        return int(n.value)

    def set_templates(self, T, bias, thr):
        T = np.ascontiguousarray(T, dtype=np.float32)
        if T.ndim != 3:
            raise ValueError("templates must be [U][C][L]")
        bias = np.ascontiguousarray(bias, dtype=np.float32)
        thr = np.ascontiguousarray(thr, dtype=np.float32)
        U, C, L = T.shape
        if bias.shape != (U,) or thr.shape != (U,):
            raise ValueError("bias and thresholds must have one entry per template")
        self._check(self._lib.fbp_set_templates(self._h, T.ctypes.data, U, C, L, bias.ctypes.data,
                                                thr.ctypes.data))
        self.U, self.C, self.L = U, C, L

    def run(self, V, n_max_shift, flags, ld=None, n_samples=None, device_ptr=None):
        """Run one chunk.  V: host float32 [C][>=N] array, or pass device_ptr
        (int address of a device float32 [C][ld] buffer) with ld and n_samples."""
        if device_ptr is not None:
            self._check(self._lib.fbp_run(self._h, ctypes.c_void_p(int(device_ptr)), int(ld), int(n_samples), 1,
                                          int(n_max_shift), int(flags)))
            return
        V = np.asarray(V)
        if V.dtype != np.float32 or V.strides[1] != 4:
            V = np.ascontiguousarray(V, dtype=np.float32)
        if V.shape[0] != self.C:
            raise ValueError("voltage channels do not match templates")
        ld = V.strides[0] // 4
        self._check(self._lib.fbp_run(self._h, V.ctypes.data, int(ld), int(V.shape[1]), 0, int(n_max_shift),
                                      int(flags)))

    def result(self):
        n = ctypes.c_int64(0)
        self._check(self._lib.fbp_result_count(self._h, ctypes.byref(n)))
        idx = np.empty(n.value, np.uint32)  # every entry is written by the copy
        lab = np.empty(n.value, np.uint32)
        self._check(self._lib.fbp_copy_result(self._h, idx.ctypes.data, lab.ctypes.data))
        return idx, lab

    def clips(self, n):
        out = np.zeros((n, self.C * self.L), np.float32)
        self._check(self._lib.fbp_copy_clips(self._h, out.ctypes.data))
        return out

    def residual(self, n_samples):
        out = np.zeros((self.C, n_samples), np.float32)
        self._check(self._lib.fbp_copy_residual(self._h, out.ctypes.data))
        return out

    def stats(self):
        s = FbpStats()
        self._check(self._lib.fbp_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def likelihood(self, n_samples):
        """[U][N] float32 likelihood the last run left (diagnostics)."""
        out = np.zeros((self.U, n_samples), np.float32)
        self._check(self._lib.fbp_copy_likelihood(self._h, out.ctypes.data))
        return out

    def pass_log(self):
        """[passes][3] int64: spikes added, windows listed, windows flagged."""
        n = ctypes.c_int32(0)
        self._check(self._lib.fbp_get_pass_log(self._h, None, 0, ctypes.byref(n)))
        rows = np.zeros((max(n.value, 1), 3), np.int64)
        self._check(self._lib.fbp_get_pass_log(self._h, rows.ctypes.data, n.value, ctypes.byref(n)))
        return rows[: n.value]


def run_batch(handles, voltages, n_max_shift, flags, block_cap=0, lik_windows=0):
    """fbp_run_batch: pursue segment g (voltages[g], host float32 [C][N] arrays,
    or (device_ptr, ld, n_samples) tuples) with handles[g]'s templates; read the
    results back through each handle.  block_cap / lik_windows: tile shape (0:
    the library's default)."""
    n = len(handles)
    if n == 0:
        return
    if len(voltages) != n:
        raise ValueError("one voltage per handle")
    lib = handles[0]._lib
    hs = (ctypes.c_void_p * n)(*[h._h.value for h in handles])
    ptrs = (ctypes.c_void_p * n)()
    lds = (ctypes.c_int64 * n)()
    ns = (ctypes.c_int64 * n)()
    nms = np.ascontiguousarray(np.broadcast_to(np.asarray(n_max_shift, np.uint32), (n,)))
    keep = []
    on_device = isinstance(voltages[0], tuple)
    for g, (h, V) in enumerate(zip(handles, voltages)):
        if isinstance(V, tuple) != on_device:
            raise ValueError("voltages must be all host arrays or all device pointers")
        if on_device:
            ptr, ld, nsmp = V
            ptrs[g], lds[g], ns[g] = int(ptr), int(ld), int(nsmp)
        else:
            V = np.asarray(V)
            if V.dtype != np.float32 or V.strides[1] != 4:
                V = np.ascontiguousarray(V, dtype=np.float32)
            if V.shape[0] != h.C:
                raise ValueError(f"segment {g}: voltage channels do not match templates")
            keep.append(V)
            ptrs[g], lds[g], ns[g] = V.ctypes.data, V.strides[0] // 4, V.shape[1]
    handles[0]._check(lib.fbp_run_batch(hs, n, ptrs, lds, ns, 1 if on_device else 0, nms.ctypes.data, int(flags),
                                        int(block_cap), int(lik_windows)))
