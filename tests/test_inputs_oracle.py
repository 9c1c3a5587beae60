"""Pins the CPU restatements of the pursuit's inputs (oracle/inputs_ref.py).
The reference holds no fixtures for them; the checks below are known answers
and special cases the mathematics fixes."""
import numpy as np

from oracle import inputs_ref


def test_robust_template_known_answers():
    # one column, n = 7: median 4, |x - 4| = [3,2,1,0,1,2,96] -> MAD = 2 / 0.6745 = 2.965;
    # kept: 1.035 < x < 6.965 -> {2, 3, 4, 5, 6}; median 4 (the outlier 100 and 1 drop out)
    x = np.array([1, 2, 3, 4, 5, 6, 100], np.float32)[:, None]
    assert inputs_ref.robust_template(x)[0] == np.float32(4.0)
    # n = 6 with an outlier: median (4 + 5) / 2 = 4.5, |x - 4.5| = [2.5,1.5,.5,.5,1.5,95.5] ->
    # MAD = 1.5 / 0.6745 = 2.224; kept: 2.276 < x < 6.724 -> {3, 4, 5, 6}; median 4.5
    y = np.array([2, 3, 4, 5, 6, 100], np.float32)[:, None]
    assert inputs_ref.robust_template(y)[0] == np.float32(4.5)
    z = np.array([2, 3, 4, 5, 6, 7, 100], np.float32)[:, None]  # odd: median 5, MAD 2/0.6745 -> {3..7} -> 5
    assert inputs_ref.robust_template(z)[0] == np.float32(5.0)
    # a constant column: MAD 0, nothing strictly inside -> the median
    c = np.full((6, 1), 3.25, np.float32)
    assert inputs_ref.robust_template(c)[0] == np.float32(3.25)
    # even count: mean of the middle two in float32
    e = np.array([[1.0], [2.0], [4.0], [8.0]], np.float32)
    assert inputs_ref.robust_template(e).dtype == np.float32
    # one clip comes back squeezed
    one = np.arange(5, dtype=np.float32)[None, :]
    assert np.array_equal(inputs_ref.robust_template(one), np.arange(5, dtype=np.float32))


def test_robust_template_ignores_outliers_and_matches_brute_force():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(41, 5)).astype(np.float32)
    X[:3] += 1000.0  # outliers far beyond one MAD
    got = inputs_ref.robust_template(X)
    for j in range(5):
        col = np.sort(X[:, j])
        med = col[20]
        mad = np.float32(np.sort(np.abs(X[:, j] - med))[20] / np.float32(0.6745))
        kept = np.sort(X[:, j][(X[:, j] > med - mad) & (X[:, j] < med + mad)])
        k = len(kept)
        want = kept[k // 2] if k % 2 else np.float32(np.float32(kept[k // 2 - 1] + kept[k // 2]) / np.float32(2))
        assert got[j] == want
        assert abs(got[j]) < 10.0


def test_covariance_and_variances_closed_forms():
    rng = np.random.default_rng(3)
    L, C = 6, 3
    res = [rng.normal(size=(50, C * L)).astype(np.float32), rng.normal(size=(30, C * L)).astype(np.float32)]
    covs = inputs_ref.channel_covariance(res, C, L)
    X = np.vstack(res).astype(np.float64)
    for c in range(C):
        Xc = X[:, c * L:(c + 1) * L] - X[:, c * L:(c + 1) * L].mean(axis=0)
        np.testing.assert_allclose(covs[c], Xc.T @ Xc / X.shape[0], rtol=1e-12, atol=1e-14)
    T = rng.normal(size=(4, C * L))
    eye = [np.eye(L)] * C
    np.testing.assert_allclose(inputs_ref.template_variances(T, eye, L), np.sum(T ** 2, axis=1), rtol=1e-14)
