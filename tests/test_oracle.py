"""Pins the CPU oracle (oracle/fbp_oracle.c) before it is trusted as the
parity reference.  The reference repository holds no golden vectors for this
path (no test calls binary_pursuit), so the oracle is pinned by closed forms,
an independent FP64 restatement, brute force on tiny inputs and planted-spike
known answers."""
import numpy as np
import pytest

from oracle import fbp_fp64, fbp_oracle
from tests import cases


def _fma32(a, b, c):
    """fmaf for float32 operands: the product is exact in float64."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def test_likelihood_closed_form():
    rng = np.random.default_rng(1)
    C, N, U, L = 3, 300, 4, 12
    V = rng.normal(size=(C, N)).astype(np.float32)
    T = rng.normal(size=(U, C, L)).astype(np.float32)
    T[1, 2] = 0.0  # an all-zero channel row
    T[3, 0] = 0.0
    bias = np.float32(-0.5 * np.sum(T.reshape(U, -1).astype(np.float64) ** 2, 1))
    LIK = fbp_oracle.full_likelihood(V, T, bias)
    ref = np.zeros((U, N))
    for t in range(U):
        acc = np.full(N - L + 1, float(bias[t]))
        for c in range(C):
            acc += np.correlate(V[c].astype(np.float64), T[t, c].astype(np.float64), "valid")
        ref[t, : N - L] = acc[: N - L]  # x + L >= N gives 0 (binary_pursuit.cl:231-234)
    assert np.all(LIK[:, N - L:] == 0)
    np.testing.assert_allclose(LIK, ref, rtol=1e-5, atol=1e-4)


def test_likelihood_operation_order_is_bias_then_channel_then_sample():
    rng = np.random.default_rng(2)
    C, N, U, L = 3, 80, 2, 10
    V = rng.normal(size=(C, N)).astype(np.float32)
    T = (rng.normal(size=(U, C, L)) * 3).astype(np.float32)
    T[0, 1] = 0.0
    bias = np.float32([-7.25, 3.5])
    for t in range(U):
        for x in (0, 7, N - L - 1):
            lik = bias[t]
            for c in range(C):
                if not np.any(T[t, c]):
                    continue
                for i in range(L):
                    lik = _fma32(T[t, c, i], V[c, x + i], lik)
            assert fbp_oracle.likelihood_at(V, T, bias, t, x) == lik


def _shifted(Tu, shift, L):
    return fbp_fp64.shifted_truncated(Tu.astype(np.float64), shift, L)


def test_combo_value_equals_energy_drop():
    rng = np.random.default_rng(3)
    C, N, U, L = 3, 200, 3, 9
    V = rng.normal(size=(C, N)).astype(np.float32)
    T = (rng.normal(size=(U, C, L)) * 2).astype(np.float32)
    T[2, 1] = 0.0
    n = L - 1
    S = 2 * n + 1
    for _ in range(200):
        lab, u = int(rng.integers(U)), int(rng.integers(U))
        b = int(rng.integers(n, N - L - n))
        f, s = int(rng.integers(S)), int(rng.integers(S))
        got = fbp_oracle.combo_value(V, T, lab, u, b, f, s, n)
        vw = V[:, b:b + L].astype(np.float64)
        a = _shifted(T[lab], f - n, L)
        bb = _shifted(T[u], s - n, L)
        ref = 0.5 * (np.sum(vw * vw) - np.sum((vw - a - bb) ** 2))
        assert abs(got - ref) <= 1e-4 * (1 + abs(ref)), (lab, u, b, f, s, got, ref)


def test_combo_value_extreme_shifts_single_sample_overlap():
    # f = 0 keeps one sample of the fixed template at the window's first
    # position; s = 2n keeps one sample of the other at the window's last.
    C, N, U, L = 1, 64, 2, 5
    V = np.zeros((C, N), np.float32)
    V[0, 20:25] = [1, 2, 3, 4, 5]
    T = np.zeros((U, C, L), np.float32)
    T[0, 0] = [10, 20, 30, 40, 50]
    T[1, 0] = [1, 1, 1, 1, 2]
    n = L - 1
    # window starts at b = 20; f = 0 puts T0 onset at b - n: only T0[4] = 50 at j = 0
    got = fbp_oracle.combo_value(V, T, 0, 1, 20, 0, 2 * n, n)
    a = np.array([50, 0, 0, 0, 0.0])
    bb = np.array([0, 0, 0, 0, 1.0])  # T1 onset at b + n: only T1[0] = 1 at j = 4
    v = V[0, 20:25].astype(np.float64)
    ref = np.dot(a + bb, v) - 0.5 * np.sum((a + bb) ** 2)
    assert got == pytest.approx(ref, abs=1e-3)


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_independent_fp64_restatement(seed):
    V, T, bias, thr, on, lab = cases.noisy_case(seed)
    n = T.shape[2] - 1
    i1, l1, r1 = fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True)
    i2, l2, r2 = fbp_fp64.pursuit(V, T, bias, thr, n, return_residual=True)
    np.testing.assert_array_equal(i1, i2)
    np.testing.assert_array_equal(l1, l2)
    np.testing.assert_allclose(r1, r2, atol=1e-5)


def test_oracle_matches_fp64_without_recheck():
    for seed in range(5):
        V, T, bias, thr, on, lab = cases.noisy_case(seed + 100, U=4, nsp=40)
        n = T.shape[2] - 1
        i1, l1 = fbp_oracle.pursuit(V, T, bias, thr, n, do_recheck=False)
        i2, l2 = fbp_fp64.pursuit(V, T, bias, thr, n, do_recheck=False)
        np.testing.assert_array_equal(i1, i2)
        np.testing.assert_array_equal(l1, l2)


def test_known_answer_isolated_spikes():
    U, C, L = 5, 4, 16
    T = cases.dog_templates(U, C, L, 7)
    N = 4000
    V = np.zeros((C, N))
    truth = []
    rng = np.random.default_rng(7)
    for k, p in enumerate(range(2 * L, N - 4 * L, 3 * L)):
        u = int(rng.integers(U))
        V[:, p:p + L] += T[u]
        truth.append((p, u))
    bias, thr = cases.bias_thr(T)
    idx, lab, res = fbp_oracle.pursuit(V.astype(np.float32), T.astype(np.float32), bias, thr, L - 1,
                                       return_residual=True)
    assert sorted(zip(idx.tolist(), lab.tolist())) == truth
    assert np.abs(res).max() < 1e-5


@pytest.mark.parametrize("seed", range(8))
def test_known_answer_overlapping_pairs(seed):
    U, C, L = 4, 3, 16
    T = cases.random_templates(U, C, L, seed)
    V, truth = cases.planted_pairs(T, 2400, seed)
    bias, thr = cases.bias_thr(T)
    idx, lab, res = fbp_oracle.pursuit(V, T.astype(np.float32), bias, thr, L - 1, return_residual=True)
    assert sorted(zip(idx.tolist(), lab.tolist())) == sorted(truth)
    assert np.abs(res).max() < 1e-5


@pytest.mark.parametrize("seed", [1, 5, 6, 7, 8])
def test_known_answer_pairs_resolved_only_by_recheck(seed):
    # Strongly correlated templates: greedy single-template pursuit explains
    # some summed pairs with the wrong unit; the overlap recheck recovers them.
    U, C, L = 4, 3, 16
    T = cases.dog_templates(U, C, L, seed)
    V, truth = cases.planted_pairs(T, 2000, 100 + seed, min_d=3)
    bias, thr = cases.bias_thr(T)
    T32 = T.astype(np.float32)
    idx, lab, res = fbp_oracle.pursuit(V, T32, bias, thr, L - 1, return_residual=True)
    assert sorted(zip(idx.tolist(), lab.tolist())) == sorted(truth)
    assert np.abs(res).max() < 1e-5
    idx2, lab2 = fbp_oracle.pursuit(V, T32, bias, thr, L - 1, do_recheck=False)
    assert sorted(zip(idx2.tolist(), lab2.tolist())) != sorted(truth)


def test_ties_resolve_to_lowest_template():
    # Two identical templates: every likelihood ties, the lexicographic arg-max
    # (binary_pursuit.cl:397-404, strict > in template order) labels all 0.
    U, C, L = 2, 2, 12
    T = cases.random_templates(1, C, L, 3)
    T = np.concatenate([T, T])
    V = np.zeros((C, 1200))
    truth = []
    for p in range(2 * L, 1200 - 4 * L, 4 * L):
        V[:, p:p + L] += T[1]
        truth.append(p)
    bias, thr = cases.bias_thr(T)
    idx, lab = fbp_oracle.pursuit(V.astype(np.float32), T.astype(np.float32), bias, thr, L - 1)
    assert sorted(idx.tolist()) == truth
    assert np.all(lab == 0)


def test_threshold_quirk_without_recheck():
    # With do_overlap_recheck False every window whose best likelihood is
    # positive inside it is added whatever the threshold, because
    # overlap_recheck keeps the value compute_template_maximum_likelihood gave
    # it (binary_pursuit_parallel.py:980-983, binary_pursuit.cl:423-433, 982).
    # With the recheck on, a below-threshold best is skipped (cl:518-522) and
    # parse clears the flag, so nothing is added.
    U, C, L = 3, 2, 10
    T = cases.random_templates(U, C, L, 11)
    V, truth = cases.planted_pairs(T, 1500, 11)
    bias, _ = cases.bias_thr(T)
    huge = np.float32(np.full(U, 1e9))
    idx_off, _ = fbp_oracle.pursuit(V, T.astype(np.float32), bias, huge, L - 1, do_recheck=False)
    idx_on, _ = fbp_oracle.pursuit(V, T.astype(np.float32), bias, huge, L - 1, do_recheck=True)
    assert len(idx_off) >= len(truth)
    assert len(idx_on) == 0


def test_no_spikes_in_zero_voltage():
    T = cases.random_templates(3, 2, 8, 0)
    bias, thr = cases.bias_thr(T)
    idx, lab = fbp_oracle.pursuit(np.zeros((2, 500), np.float32), T.astype(np.float32), bias, thr, 7)
    assert idx.size == 0 and lab.size == 0


def test_adjusted_clips_are_residual_plus_template():
    V, T, bias, thr, on, lab = cases.noisy_case(4)
    n = T.shape[2] - 1
    idx, labs, R = fbp_oracle.pursuit(V, T, bias, thr, n, return_residual=True)
    clips = fbp_oracle.adjusted_clips(R, T, idx, labs)
    C, N = V.shape
    L = T.shape[2]
    for k in range(len(idx)):
        seg = np.zeros((C, L), np.float32)
        hi = min(N, idx[k] + L)
        seg[:, : hi - idx[k]] = R[:, idx[k]:hi] + T[labs[k]][:, : hi - idx[k]]
        np.testing.assert_array_equal(clips[k], seg.ravel())
