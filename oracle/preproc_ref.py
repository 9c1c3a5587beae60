"""CPU restatements of the voltage preprocessing (TEST INFRASTRUCTURE ONLY).

* thresholds: SF/threshold/threshold.py:6-70 with numpy 2 scalar rules
  (float32 medians, sigma a Python float);
* zca_matrix: SF/processing/zca.py:5-31;
* whiten: the callers' (zca_matrix @ voltage).astype(float32)
  (SF/parallel/spikesorting_parallel.py:295).
"""
import numpy as np
from scipy import linalg


def median_threshold(voltage, sigma):
    V = np.atleast_2d(voltage)
    thr = np.empty((V.shape[0],))
    for c in range(V.shape[0]):
        thr[c] = np.nanmedian(np.abs(V[c])) / 0.6745
    thr *= sigma
    return thr


def single_thresholds(voltage, sigma):
    V = np.atleast_2d(voltage)
    thr = np.empty((V.shape[0],))
    for c in range(V.shape[0]):
        thr[c] = sigma * np.nanmedian(np.abs(V[c])) / 0.6745
    return thr


def single_thresholds_and_samples(voltage, sigma):
    V = np.atleast_2d(voltage)
    thr = single_thresholds(V, sigma)
    return thr, [int(np.count_nonzero(np.abs(V[c]) > thr[c])) for c in range(V.shape[0])]


def zca_matrix(data):
    sigma = np.cov(data, rowvar=True)
    U, S, _ = linalg.svd(sigma)
    Z = U @ np.diag(1.0 / np.sqrt(S + 1e-9)) @ U.T
    return Z / np.diag(Z)[:, None]


def whiten(Z, V):
    return (Z @ V).astype(np.float32)
