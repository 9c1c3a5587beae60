"""Pins the FP64 refinement oracle (oracle/refine_fp64.py) before the GPU
gradients are compared with it.  The reference has no differentiable path and
no fixtures for one, so the oracle is pinned by what the mathematics fixes:

* E is quadratic in T at fixed a and quadratic in a at fixed T, so a central
  difference with ANY step equals the exact partial derivative -- a dropped
  amplitude factor, a sign, or an index off by one in the gradient formulas
  fails these checks;
* the residual equals V minus per-unit convolutions of weighted spike trains
  with the templates (a different form of the same sum);
* noise-free planted spikes give zero energy and zero gradients.
"""
import numpy as np

from oracle import refine_fp64


def _case(seed, C=3, N=200, U=3, L=9, n=25):
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(C, N))
    T = rng.normal(size=(U, C, L))
    idx = np.sort(rng.integers(0, N, size=n))
    idx[-1] = N - 3  # truncated at the end of the recording
    lab = rng.integers(0, U, size=n)
    amp = rng.normal(1.0, 0.2, size=n)
    return V, T, idx, lab, amp


def test_template_gradient_equals_central_difference():
    V, T, idx, lab, amp = _case(1)
    _, gT, _, _ = refine_fp64.gradients(V, T, idx, lab, amp)
    rng = np.random.default_rng(11)
    for _ in range(30):
        u, c, i = (int(rng.integers(s)) for s in T.shape)
        for h in (1.0, 0.37):
            Tp = T.copy()
            Tm = T.copy()
            Tp[u, c, i] += h
            Tm[u, c, i] -= h
            fd = (refine_fp64.energy(V, Tp, idx, lab, amp) - refine_fp64.energy(V, Tm, idx, lab, amp)) / (2 * h)
            assert abs(fd - gT[u, c, i]) <= 1e-9 * max(1.0, abs(fd))


def test_amplitude_gradient_equals_central_difference():
    V, T, idx, lab, amp = _case(2)
    _, _, ga, _ = refine_fp64.gradients(V, T, idx, lab, amp)
    for k in range(len(idx)):
        h = 0.5
        ap = amp.copy()
        am = amp.copy()
        ap[k] += h
        am[k] -= h
        fd = (refine_fp64.energy(V, T, idx, lab, ap) - refine_fp64.energy(V, T, idx, lab, am)) / (2 * h)
        assert abs(fd - ga[k]) <= 1e-9 * max(1.0, abs(fd))


def test_residual_equals_convolution_form():
    V, T, idx, lab, amp = _case(3)
    C, N = V.shape
    U, _, L = T.shape
    ref = V.copy()
    for u in range(U):
        train = np.zeros(N)
        np.add.at(train, idx[lab == u], amp[lab == u])
        for c in range(C):
            ref[c] -= np.convolve(train, T[u, c])[:N]
    np.testing.assert_allclose(refine_fp64.residual(V, T, idx, lab, amp), ref, rtol=0, atol=1e-12)


def test_planted_spikes_give_zero_energy_and_gradients():
    rng = np.random.default_rng(4)
    C, N, U, L = 2, 150, 2, 7
    T = rng.normal(size=(U, C, L))
    idx = np.array([3, 5, 40, 41, 100, 146])  # overlapping pairs and an end truncation
    lab = np.array([0, 1, 1, 0, 1, 0])
    amp = np.array([1.0, 0.8, 1.2, 0.9, 1.1, 1.0])
    V = np.zeros((C, N))
    for p, u, a in zip(idx, lab, amp):
        n = min(L, N - p)
        V[:, p:p + n] += a * T[u, :, :n]
    E, gT, ga, R = refine_fp64.gradients(V, T, idx, lab, amp)
    assert E < 1e-24
    assert np.max(np.abs(R)) < 1e-12
    assert np.max(np.abs(gT)) < 1e-12 and np.max(np.abs(ga)) < 1e-12
