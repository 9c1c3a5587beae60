"""FP64 NumPy restatement of residual energy and its gradients at fixed spike
times (template refinement).

TEST INFRASTRUCTURE ONLY.  The reference has no counterpart: it never
differentiates.  The quantity is the one its residual pass leaves behind
(SF/kernels/binary_pursuit.cl:57-102 subtracts each found template from the
voltage), with a per-spike amplitude a_k:

  R[c, x] = V[c, x] - sum_k a_k * T[lab_k, c, x - p_k]      (0 <= x - p_k < L, x < N)
  E       = 0.5 * sum_{c, x} R[c, x]^2

  dE/dT[u, c, i] = - sum_{k : lab_k = u, p_k + i < N} a_k * R[c, p_k + i]
  dE/da_k        = - sum_{c, i : p_k + i < N} T[lab_k, c, i] * R[c, p_k + i]
  dE/dV          = R

E is a quadratic function of T (at fixed a) and of a (at fixed T), so a central
difference (E(z + h) - E(z - h)) / 2h equals the exact partial derivative for
every step h; the tests pin these formulas that way.
"""
import numpy as np


def residual(V, T, idx, lab, amp):
    """R (float64 [C][N]) after subtracting every spike, clipped at N."""
    R = np.array(V, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    C, N = R.shape
    L = T.shape[2]
    for p, u, a in zip(np.asarray(idx, np.int64), np.asarray(lab, np.int64), np.asarray(amp, np.float64)):
        n = min(L, N - p)
        if n > 0:
            R[:, p:p + n] -= a * T[u, :, :n]
    return R


def energy(V, T, idx, lab, amp):
    R = residual(V, T, idx, lab, amp)
    return 0.5 * float(np.sum(R * R))


def gradients(V, T, idx, lab, amp):
    """(E, dE/dT [U][C][L], dE/da [n], R) in float64."""
    T = np.asarray(T, dtype=np.float64)
    U, C, L = T.shape
    R = residual(V, T, idx, lab, amp)
    N = R.shape[1]
    gT = np.zeros_like(T)
    ga = np.zeros(len(idx))
    for k, (p, u, a) in enumerate(zip(np.asarray(idx, np.int64), np.asarray(lab, np.int64),
                                      np.asarray(amp, np.float64))):
        n = min(L, N - p)
        if n <= 0:
            continue
        seg = R[:, p:p + n]
        gT[u, :, :n] -= a * seg
        ga[k] = -float(np.sum(T[u, :, :n] * seg))
    return 0.5 * float(np.sum(R * R)), gT, ga, R
