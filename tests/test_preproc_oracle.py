"""Pins the preprocessing restatements (oracle/preproc_ref.py): the median by
brute force, the float32 operation order of the thresholds, and the
whitening property of the ZCA matrix."""
import numpy as np

from oracle import preproc_ref


def test_threshold_medians_and_float32_order():
    rng = np.random.default_rng(1)
    V = rng.normal(size=(3, 1001)).astype(np.float32)
    V[1, 5] = np.nan
    thr = preproc_ref.single_thresholds(V, 4.0)
    for c in range(3):
        a = np.sort(np.abs(V[c][~np.isnan(V[c])]))
        n = a.size
        med = a[n // 2] if n % 2 else np.float32(np.float32(a[n // 2 - 1] + a[n // 2]) / np.float32(2))
        assert thr[c] == np.float64(np.float32(np.float32(np.float32(4.0) * med) / np.float32(0.6745)))
    mt = preproc_ref.median_threshold(V, 4.0)
    assert np.array_equal(mt, np.float64(np.float32(np.nanmedian(np.abs(V), axis=1) / np.float32(0.6745))) * 4.0)
    _, counts = preproc_ref.single_thresholds_and_samples(V, 4.0)
    assert counts == [int(np.sum(np.abs(V[c]) > thr[c])) for c in range(3)]


def test_zca_matrix_whitens():
    rng = np.random.default_rng(2)
    A = rng.normal(size=(6, 6))
    V = (A @ rng.normal(size=(6, 20000))).astype(np.float32)
    Z = preproc_ref.zca_matrix(V)
    S = np.cov(V)
    W = Z @ S @ Z.T
    d = np.diag(W)
    assert np.max(np.abs(W - np.diag(d))) < 1e-6 * np.max(d)
    np.testing.assert_allclose(np.diag(Z), 1.0)
