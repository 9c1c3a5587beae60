"""GPU voltage preprocessing against the numpy restatements: thresholds and
counts bit-exact (exact order statistics, same float32 arithmetic), the ZCA
matrix to 1e-9 relative (FP64 covariance summed in another order), whitened
voltage within 1 float32 ulp (FP64 sums in another order, then rounding)."""
import numpy as np
import pytest
import torch

from oracle import preproc_ref
from spikesorting_fullpursuit_amd.processing import zca
from spikesorting_fullpursuit_amd.threshold import threshold

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [1, 2, 1001, 100000, 123457])
def test_thresholds_match_oracle(N):
    rng = np.random.default_rng(N)
    V = rng.normal(size=(5, N)).astype(np.float32)
    V[1] = np.round(V[1])                     # ties
    if N > 4:
        V[2, :3] = np.nan                     # NaNs are ignored
    V[3] = -np.abs(V[3]) * 1e-30              # tiny values
    print(f"[thresholds] N={N}", flush=True)
    for fn in ("median_threshold", "single_thresholds"):
        got = getattr(threshold, fn)(V, 4.0)
        want = getattr(preproc_ref, fn)(V, 4.0)
        assert np.array_equal(got, want), fn
    t, c = threshold.single_thresholds_and_samples(V, 3.5)
    t2, c2 = preproc_ref.single_thresholds_and_samples(V, 3.5)
    assert np.array_equal(t, t2) and c == c2


def test_thresholds_all_nan_row_and_device_input():
    V = np.ones((2, 10), np.float32)
    V[0] = np.nan
    with np.errstate(all="ignore"), pytest.warns(RuntimeWarning):
        want = preproc_ref.single_thresholds(V, 4.0)
    got = threshold.single_thresholds(torch.tensor(V, device="cuda"), 4.0)
    assert np.isnan(got[0]) and np.isnan(want[0]) and got[1] == want[1]


def test_zca_matrix_and_whitening_match_numpy():
    rng = np.random.default_rng(3)
    C, N = 16, 200000
    A = np.eye(C) + 0.3 * rng.normal(size=(C, C))
    V = (A @ rng.normal(size=(C, N))).astype(np.float32)
    Z = zca.get_full_zca_matrix(V)
    Zr = preproc_ref.zca_matrix(V)
    assert np.max(np.abs(Z - Zr)) <= 1e-9 * np.max(np.abs(Zr))
    W = zca.zca_whiten(Zr, V)
    Wr = preproc_ref.whiten(Zr, V)
    ulp = np.spacing(np.abs(Wr))
    assert np.all(np.abs(W - Wr) <= ulp)
    assert np.mean(W == Wr) > 0.999
    assert zca.get_full_zca_matrix(V[:1]) == 1.0
