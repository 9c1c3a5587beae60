"""CPU restatements of the pursuit's inputs (TEST INFRASTRUCTURE ONLY).

* robust_template: calculate_robust_template (SF/processing/clip_utils.py:202-247)
  with numpy's float32 semantics, one column at a time;
* channel_covariance: get_channel_covariance_matrix
  (SF/overlap/full_binary_pursuit.py:863-923), np.cov(rowvar=False, ddof=0)
  per channel over the same rows;
* template_variances: sum_c t_c^T Sigma_c t_c (SF/overlap/neuron_separability.py:158-172).
"""
import numpy as np


def robust_template(clips):
    X = np.asarray(clips)
    if X.shape[0] == 1 or X.ndim == 1:
        return np.squeeze(X)
    out = np.zeros((X.shape[1],), dtype=X.dtype)
    med = np.median(X, axis=0)
    for j in range(X.shape[1]):
        col = X[:, j]
        mad = np.median(np.abs(col - med[j])) / 0.6745
        keep = (col > med[j] - mad) & (col < med[j] + mad)
        out[j] = np.median(col[keep]) if keep.any() else med[j]
    return out


def channel_covariance(residuals, n_channels, L, rows_per_channel=None):
    X = np.vstack(residuals)
    mats = []
    for c in range(n_channels):
        rows = np.arange(X.shape[0]) if rows_per_channel is None else rows_per_channel[c]
        mats.append(np.cov(X[rows, c * L:(c + 1) * L], rowvar=False, ddof=0))
    return mats


def template_variances(T, covs, L):
    T = np.atleast_2d(np.asarray(T, dtype=np.float64))
    C = T.shape[1] // L
    return np.array([sum(T[u, c * L:(c + 1) * L] @ covs[c] @ T[u, c * L:(c + 1) * L] for c in range(C))
                     for u in range(T.shape[0])])
