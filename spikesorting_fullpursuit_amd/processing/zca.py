"""ZCA whitening of the voltage on the GPU.

Restates get_full_zca_matrix of SF/processing/zca.py:5-31 (SF =
spikesorting_fullpursuit/) and the whitening step its callers run,
`(zca_matrix @ voltage).astype(v_dtype)` (SF/parallel/spikesorting_parallel.py:295):
the channel covariance np.cov(V) is an FP64 kernel (fbp_channel_cov_rows), the
C x C SVD and scaling stay on the host (scipy, as the reference), and the
product runs in FP64 on the GPU (fbp_zca_apply), rounded to float32.
There is no CPU fallback for the N-long passes.
"""
import ctypes

import numpy as np
from scipy import linalg

from .. import _native


def _dev(x, device):
    import torch

    if torch.is_tensor(x):
        X = x if x.is_cuda else x.to(torch.device("cuda", device))
    else:
        if not torch.cuda.is_available():
            raise RuntimeError("No GPU available")
        X = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(torch.device("cuda", device))
    if X.dtype != torch.float32:
        X = X.to(torch.float32)
    return X if X.stride(-1) == 1 else X.contiguous()


def get_full_zca_matrix(data, rowvar=True, *, device=0):
    """ZCA matrix of `data` (channels x samples when rowvar, zca.py:5-31):
    U diag(1 / sqrt(S + 1e-9)) U^T of sigma = np.cov(data, rowvar), each row
    divided by its diagonal element.  1.0 for one channel."""
    import torch

    if data.ndim == 1 or data.shape[0] == 1:
        return 1.0
    X = _dev(data, device)
    if not rowvar:
        X = X.t().contiguous()
    C, N = X.shape
    lib = _native.load()
    nb = ctypes.c_int64(0)
    _native.preproc_check(lib.fbp_preproc_workspace_bytes(N, C, ctypes.byref(nb)))
    work = torch.empty(nb.value, dtype=torch.uint8, device=X.device)
    cov = torch.empty((C, C), dtype=torch.float64, device=X.device)
    vp = ctypes.c_void_p
    _native.preproc_check(lib.fbp_channel_cov_rows(vp(X.data_ptr()), X.stride(0), N, C, 1, vp(cov.data_ptr()),
                                                   vp(work.data_ptr()), work.numel(),
                                                   vp(torch.cuda.current_stream(X.device).cuda_stream)))
    sigma = cov.cpu().numpy()
    U, S, _ = linalg.svd(sigma)
    zca_matrix = U @ np.diag(1.0 / np.sqrt(S + 1e-9)) @ U.T
    return zca_matrix / np.diag(zca_matrix)[:, None]


def zca_whiten(zca_matrix, voltage, v_dtype=np.float32, *, device=0, return_device=False):
    """(zca_matrix @ voltage).astype(v_dtype) with FP64 products and sums on the
    GPU (float32 output; v_dtype float32).  voltage [C][N] host or device."""
    import torch

    if np.dtype(v_dtype) != np.float32:
        raise TypeError("GPU whitening writes float32")
    V = _dev(voltage, device)
    C, N = V.shape
    Z = torch.from_numpy(np.ascontiguousarray(zca_matrix, dtype=np.float64)).to(V.device)
    if Z.shape != (C, C):
        raise ValueError("zca_matrix must be C x C")
    out = torch.empty((C, N), dtype=torch.float32, device=V.device)
    vp = ctypes.c_void_p
    _native.preproc_check(_native.load().fbp_zca_apply(vp(Z.data_ptr()), vp(V.data_ptr()), V.stride(0), N, C,
                                                       vp(out.data_ptr()), out.stride(0),
                                                       vp(torch.cuda.current_stream(V.device).cuda_stream)))
    return out if return_device else out.cpu().numpy()
