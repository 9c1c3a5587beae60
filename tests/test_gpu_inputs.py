"""GPU inputs of the pursuit against the CPU restatements: robust templates
bit-exact (order statistics and float32 arithmetic), covariance and template
variances to FP64 rounding (|diff| <= 1e-12 max|ref|), the separability
metrics dict as the host formula."""
import numpy as np
import pytest
import torch

from oracle import inputs_ref
from spikesorting_fullpursuit_amd.overlap import full_binary_pursuit, neuron_separability
from spikesorting_fullpursuit_amd.processing import clip_utils

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,W", [(2, 5), (3, 33), (40, 64), (41, 100), (1000, 96), (4097, 40)])
def test_robust_template_matches_oracle(n, W):
    rng = np.random.default_rng(n * 7 + W)
    X = rng.normal(size=(n, W)).astype(np.float32)
    X[: max(1, n // 10)] *= 50.0                     # outliers
    X[:, 0] = 1.5                                    # constant column: MAD 0
    X[:, 1] = np.round(X[:, 1])                      # many ties
    if W > 3:
        X[:, 3] = -np.abs(X[:, 3])                   # negative values only
    print(f"[robust] n={n} W={W}", flush=True)
    got = clip_utils.calculate_robust_template(X)
    want = inputs_ref.robust_template(X)
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got, want)


def test_robust_template_device_input_and_single_clip():
    rng = np.random.default_rng(9)
    X = rng.normal(size=(17, 48)).astype(np.float32)
    np.testing.assert_array_equal(clip_utils.calculate_robust_template(torch.tensor(X, device="cuda")),
                                  inputs_ref.robust_template(X))
    np.testing.assert_array_equal(clip_utils.calculate_robust_template(X[:1]), X[0])


def test_channel_covariance_matches_np_cov():
    rng = np.random.default_rng(4)
    C, L = 5, 48
    res = [rng.normal(size=(n, C * L)).astype(np.float32) * (1 + u) for u, n in enumerate([700, 1500, 30])]
    got = full_binary_pursuit.get_channel_covariance_matrix(res, C, L, 10 ** 9, False)
    want = inputs_ref.channel_covariance(res, C, L)
    for g, w in zip(got, want):
        assert g.shape == (L, L) and g.dtype == np.float64
        assert np.max(np.abs(g - w)) <= 1e-12 * np.max(np.abs(w))


def test_channel_covariance_subsampled_rows_follow_numpy_random():
    rng = np.random.default_rng(5)
    C, L = 3, 20
    res = [rng.normal(size=(500, C * L)).astype(np.float32)]
    np.random.seed(123)
    got = full_binary_pursuit.get_channel_covariance_matrix(res, C, L, 100, False)
    np.random.seed(123)
    rows = [np.random.randint(0, 500, 100) for _ in range(C)]
    want = inputs_ref.channel_covariance(res, C, L, rows)
    for g, w in zip(got, want):
        assert np.max(np.abs(g - w)) <= 1e-12 * np.max(np.abs(w))


def test_separability_metrics_match_host_formula():
    rng = np.random.default_rng(6)
    C, L, U = 4, 32, 7
    A = rng.normal(size=(C, L, L))
    covs = [a @ a.T / L + np.eye(L) for a in A]
    T = [rng.normal(size=C * L) * 3 for _ in range(U)]
    si = {"n_channels": C, "n_samples_per_chan": L, "n_samples": 10 ** 6, "n_threshold_crossings": np.arange(C) * 10,
          "sigma": 4.0, "sigma_bp_noise": 3.719, "sigma_bp_CI": 12.0}
    m = neuron_separability.compute_separability_metrics(T, covs, si)
    var = inputs_ref.template_variances(np.vstack(T), covs, L)
    assert np.max(np.abs(m["neuron_variances"] - var)) <= 1e-12 * np.max(var)
    SS = np.sum(np.vstack(T) ** 2, axis=1)
    np.testing.assert_allclose(m["neuron_lower_thresholds"], -0.5 * SS + 3.719 * np.sqrt(var), rtol=1e-12)
    np.testing.assert_allclose(m["neuron_lower_CI"], 0.5 * SS - 12.0 * np.sqrt(var), rtol=1e-12)
    assert np.array_equal(m["peak_channel"], np.argmax(np.abs(np.vstack(T)), axis=1) // L)
    assert not np.any(m["template_SS_by_chan"])
    np.testing.assert_allclose(m["channel_p_noise"], (10 ** 6 - np.arange(C) * 10) / 10 ** 6)
