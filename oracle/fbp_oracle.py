"""ctypes wrapper of the C restatement of the binary-pursuit pass loop.

TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped).

Restates SF/overlap/binary_pursuit_parallel.py:907-1355 and
SF/kernels/binary_pursuit.cl:1-1130 for one chunk (SF = spikesorting_fullpursuit/).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfbp_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle library (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def bind(path):
    """Load an oracle library at `path` and declare its signatures."""
    L = ctypes.CDLL(path)
    i64, i32, u32 = ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32
    L.fbpo_pursuit.restype = ctypes.c_int
    L.fbpo_pursuit.argtypes = [
        _f32p, i64, i64, _f32p, i32, i32, _f32p, _f32p, i64, i32, i32, i64,
        _u32p, _u32p, ctypes.POINTER(i64), ctypes.c_void_p, i32, ctypes.c_void_p,
        ctypes.POINTER(i32),
    ]
    L.fbpo_adjusted_clips.restype = ctypes.c_int
    L.fbpo_adjusted_clips.argtypes = [
        _f32p, i64, i64, _f32p, i32, i32, _u32p, _u32p, i64, _f32p]
    L.fbpo_full_likelihood.restype = ctypes.c_int
    L.fbpo_full_likelihood.argtypes = [_f32p, i64, i64, _f32p, i32, i32, _f32p, _f32p]
    L.fbpo_likelihood_at.restype = ctypes.c_float
    L.fbpo_likelihood_at.argtypes = [_f32p, i64, i64, _f32p, i32, i32, _f32p, i32, i64]
    L.fbpo_combo_value.restype = ctypes.c_float
    L.fbpo_combo_value.argtypes = [_f32p, i64, i64, _f32p, i32, i32, u32, u32, i64, i64, i64, i64]
    L.fbpo_window_max_step.restype = ctypes.c_int
    L.fbpo_window_max_step.argtypes = [_f32p, i64, i32, i32, _u32p, i64, _u32p, _u32p, _f32p, _u32p, _u32p,
                                       _u8p, _u8p]
    L.fbpo_reassign_step.restype = ctypes.c_int
    L.fbpo_reassign_step.argtypes = [i64, i32, _u32p, i64, _u32p, _u32p, _u32p, _u32p, _u8p, _u8p]
    L.fbpo_pursuit_step.restype = ctypes.c_int64
    L.fbpo_pursuit_step.argtypes = [i64, i32, _f32p, _u32p, i64, _u32p, _u32p, _f32p, _u8p, _u8p, _u32p, _u32p]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = bind(_LIB_PATH)
    return _lib


def _prep(V, T, bias, thr):
    V = np.ascontiguousarray(V, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    if T.ndim != 3:
        raise ValueError("templates must be [U][C][L]")
    bias = np.ascontiguousarray(bias, dtype=np.float32)
    thr = np.ascontiguousarray(thr, dtype=np.float32)
    return V, T, bias, thr


def pursuit(V, T, bias, thr, n_max, do_recheck=True, nthreads=0, return_residual=False,
            return_log=False, cap=None, lib_=None):
    """One chunk of binary pursuit.  V [C][N], T [U][C][L] float32.

    Returns (idx uint32, lab uint32[, residual][, pass_log]) with spikes in
    pass order, ascending window order within a pass.  Indices refer to the
    template onset within the chunk.  pass_log rows are
    (spikes added, windows listed, windows flagged, recheck items evaluated).
    """
    V, T, bias, thr = _prep(V, T, bias, thr)
    C, N = V.shape
    U, C2, L = T.shape
    if C2 != C:
        raise ValueError("channel mismatch")
    nwin = (N + L - 1) // L
    cap = int(cap if cap is not None else max(64, 4 * nwin))
    while True:
        idx = np.zeros(cap, np.uint32)
        lab = np.zeros(cap, np.uint32)
        n = ctypes.c_int64(0)
        npass = ctypes.c_int32(0)
        resid = np.zeros((C, N), np.float32) if return_residual else None
        maxlog = 4096
        log = np.zeros((maxlog, 4), np.int64)
        rc = (lib_ or lib()).fbpo_pursuit(
            V, C, N, T, U, L, bias, thr, int(n_max), int(bool(do_recheck)), int(nthreads), cap,
            idx, lab, ctypes.byref(n),
            resid.ctypes.data if resid is not None else None, maxlog, log.ctypes.data,
            ctypes.byref(npass))
        if rc == -3:
            cap *= 4
            continue
        if rc != 0:
            raise RuntimeError(f"oracle failed rc={rc}")
        break
    out = [idx[: n.value].copy(), lab[: n.value].copy()]
    if return_residual:
        out.append(resid)
    if return_log:
        out.append(log[: min(npass.value, maxlog)].copy())
    return tuple(out)


def adjusted_clips(R, T, idx, lab):
    R = np.ascontiguousarray(R, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    lab = np.ascontiguousarray(lab, dtype=np.uint32)
    C, N = R.shape
    U, _, L = T.shape
    clips = np.zeros((idx.shape[0], C * L), np.float32)
    rc = lib().fbpo_adjusted_clips(R, C, N, T, U, L, idx, lab, idx.shape[0], clips)
    if rc != 0:
        raise RuntimeError("oracle clips failed")
    return clips


def full_likelihood(V, T, bias):
    V, T, bias, _ = _prep(V, T, bias, np.zeros(1, np.float32))
    C, N = V.shape
    U, _, L = T.shape
    LIK = np.zeros((U, N), np.float32)
    lib().fbpo_full_likelihood(V, C, N, T, U, L, bias, LIK)
    return LIK


def likelihood_at(V, T, bias, t, x):
    V, T, bias, _ = _prep(V, T, bias, np.zeros(1, np.float32))
    C, N = V.shape
    U, _, L = T.shape
    return float(lib().fbpo_likelihood_at(V, C, N, T, U, L, bias, int(t), int(x)))


def combo_value(V, T, lab, u, b, f, s, n):
    V, T, _, _ = _prep(V, T, np.zeros(1, np.float32), np.zeros(1, np.float32))
    C, N = V.shape
    U, _, L = T.shape
    return float(lib().fbpo_combo_value(V, C, N, T, U, L, int(lab), int(u), int(b), int(f),
                                        int(s), int(n)))


def window_max_step(LIK, L, win, state, lib_=None):
    """compute_template_maximum_likelihood over the listed windows on `state`
    (a dict of per-window arrays: best_idx, best_lab, best_lik, obest_idx,
    obest_lab, check, orc), updated in place.  LIK is [U][N] float32."""
    LIK = np.ascontiguousarray(LIK, dtype=np.float32)
    U, N = LIK.shape
    win = np.ascontiguousarray(win, dtype=np.uint32)
    rc = (lib_ or lib()).fbpo_window_max_step(
        LIK, N, U, int(L), win, win.shape[0], state["best_idx"], state["best_lab"], state["best_lik"],
        state["obest_idx"], state["obest_lab"], state["check"], state["orc"])
    if rc != 0:
        raise RuntimeError(f"oracle failed rc={rc}")


def reassign_step(N, L, flist, state, lib_=None):
    """check_overlap_reassignments over the flagged windows `flist`, in place."""
    flist = np.ascontiguousarray(flist, dtype=np.uint32)
    rc = (lib_ or lib()).fbpo_reassign_step(
        int(N), int(L), flist, flist.shape[0], state["best_idx"], state["best_lab"], state["obest_idx"],
        state["obest_lab"], state["check"], state["orc"])
    if rc != 0:
        raise RuntimeError(f"oracle failed rc={rc}")


def pursuit_step(N, L, thr, win, state, lib_=None):
    """binary_pursuit over the listed windows, in place on state["check"].
    Returns the spikes (indices, labels) in ascending window order."""
    thr = np.ascontiguousarray(thr, dtype=np.float32)
    win = np.ascontiguousarray(win, dtype=np.uint32)
    out_idx = np.zeros(max(1, win.shape[0]), np.uint32)
    out_lab = np.zeros(max(1, win.shape[0]), np.uint32)
    n = (lib_ or lib()).fbpo_pursuit_step(
        int(N), int(L), thr, win, win.shape[0], state["best_idx"], state["best_lab"], state["best_lik"],
        state["check"], state["orc"], out_idx, out_lab)
    if n < 0:
        raise RuntimeError(f"oracle failed rc={n}")
    return out_idx[:n].copy(), out_lab[:n].copy()


def new_state(nwin):
    """Zeroed per-window state arrays (binary_pursuit_parallel.py:528-655)."""
    return {
        "best_idx": np.zeros(nwin, np.uint32), "best_lab": np.zeros(nwin, np.uint32),
        "best_lik": np.zeros(nwin, np.float32), "obest_idx": np.zeros(nwin, np.uint32),
        "obest_lab": np.zeros(nwin, np.uint32), "check": np.zeros(nwin, np.uint8),
        "orc": np.zeros(nwin, np.uint8),
    }
