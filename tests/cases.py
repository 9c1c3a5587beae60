"""Seeded small pursuit problems shared by the CPU and GPU tests."""
import numpy as np


def random_templates(U, C, L, seed, amp=5.0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(U, C, L)) * np.hanning(L + 2)[1:-1] * amp


def dog_templates(U, C, L, seed):
    """Trough-dominated templates of similar shape (strongly correlated)."""
    rng = np.random.default_rng(seed)
    t = np.arange(L)
    T = np.zeros((U, C, L))
    for u in range(U):
        c0 = rng.uniform(0.3, 0.5) * L
        w = rng.uniform(1.5, 3)
        wave = -np.exp(-0.5 * ((t - c0) / w) ** 2) + 0.4 * np.exp(-0.5 * ((t - c0 - 2 * w) / (2 * w)) ** 2)
        gains = rng.uniform(0.2, 1.0, C)
        gains[rng.integers(C)] = 1
        T[u] = rng.uniform(5, 10) * gains[:, None] * wave[None, :]
    return T


def bias_thr(T, thr_frac=0.1):
    U = T.shape[0]
    SS = np.sum(T.reshape(U, -1) ** 2, 1)
    return np.float32(-0.5 * SS), np.float32(thr_frac * SS)


def noisy_case(seed, C=2, L=8, U=3, N=600, nsp=25, noise=0.3, thr=2.0):
    """Random templates planted at random onsets in white noise."""
    rng = np.random.default_rng(seed)
    T = rng.normal(size=(U, C, L)) * np.hanning(L + 2)[1:-1] * 4
    V = rng.normal(size=(C, N)) * noise
    on = np.sort(rng.choice(N - L - 1, nsp, replace=False))
    lab = rng.integers(0, U, nsp)
    for o, l in zip(on, lab):
        V[:, o:o + L] += T[l]
    SS = np.sum(T.reshape(U, -1) ** 2, 1)
    return (V.astype(np.float32), T.astype(np.float32), np.float32(-0.5 * SS),
            np.float32(np.full(U, thr)), on, lab)


def planted_pairs(T, N, seed, min_d=2):
    """Noise-free voltage with overlapping pairs (offset d in [min_d, L-min_d))
    separated by 5 template widths.  Returns (V float32, truth list)."""
    U, C, L = T.shape
    rng = np.random.default_rng(seed)
    V = np.zeros((C, N))
    truth = []
    p = 2 * L
    while p + 3 * L < N - 2 * L:
        a, b = rng.choice(U, 2, replace=False)
        d = int(rng.integers(min_d, L - min_d))
        V[:, p:p + L] += T[a]
        V[:, p + d:p + d + L] += T[b]
        truth += [(p, int(a)), (p + d, int(b))]
        p += 5 * L
    return V.astype(np.float32), truth
