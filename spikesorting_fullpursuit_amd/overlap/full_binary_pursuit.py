"""Inputs of binary pursuit prepared on the GPU.

Restates get_channel_covariance_matrix of SF/overlap/full_binary_pursuit.py:863-923
(SF = spikesorting_fullpursuit/): the per-channel noise covariance of the
residuals between each unit's clips and its robust template, from which the
pursuit's thresholds are derived.  The covariance runs in the FP64 kernel
behind fbp_channel_covariance (include/fbp.h).  There is no CPU fallback.
"""
import ctypes

import numpy as np

from .. import _native


def get_channel_covariance_matrix(clip_template_residuals, n_channels, n_samples_per_chan, n_cov_samples, verbose,
                                  *, device=0):
    """List of n_channels [L][L] float64 covariance matrices (np.cov, ddof=0).

    clip_template_residuals: list (one per unit) of [n_u][C*L] arrays, stacked
    as the reference stacks them.  When n_cov_samples is below the number of
    rows, each channel draws its own n_cov_samples rows with
    np.random.randint, as the reference does (full_binary_pursuit.py:902-913),
    so a seeded numpy global state selects the same rows.
    """
    import torch

    X = np.vstack(clip_template_residuals)
    if verbose:
        print("Computing clip noise covariance for each channel with", n_cov_samples, "clip samples")
    C, L = int(n_channels), int(n_samples_per_chan)
    n_total = X.shape[0]
    if X.shape[1] != C * L:
        raise ValueError("residual rows must hold n_channels * n_samples_per_chan samples")
    if not torch.cuda.is_available():
        raise RuntimeError("No GPU available")
    dev = torch.device("cuda", device)
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(dev)
    lib = _native.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    vp = ctypes.c_void_p

    def run(rows, Cc, col0, n_rows):
        nb = ctypes.c_int64(0)
        _native.inputs_check(lib.fbp_covariance_workspace_bytes(n_rows, Cc, L, ctypes.byref(nb)))
        work = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        cov = torch.empty((Cc, L, L), dtype=torch.float64, device=dev)
        rows_d = None if rows is None else torch.from_numpy(np.ascontiguousarray(rows, dtype=np.int64)).to(dev)
        _native.inputs_check(lib.fbp_channel_covariance(
            vp(Xd.data_ptr() + 4 * col0), Xd.stride(0), n_total, Cc, L,
            None if rows_d is None else vp(rows_d.data_ptr()), n_rows, vp(cov.data_ptr()), vp(work.data_ptr()),
            work.numel(), stream))
        return cov.cpu().numpy()

    if n_cov_samples >= n_total:
        cov = run(None, C, 0, n_total)
        return [cov[c] for c in range(C)]
    mats = []
    for chan in range(C):
        rows = np.random.randint(0, n_total, n_cov_samples)
        mats.append(run(rows, 1, chan * L, len(rows))[0])
    return mats
