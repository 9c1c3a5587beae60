"""CPU restatement of the template QA (TEST INFRASTRUCTURE ONLY).

* find_overlap_templates: SF/c_cython/sort_cython.pyx:333-462, the literal
  loops (descending template energy; pairs with >= 4x the test unit's spikes;
  min/max channel screens in float32; every shift pair; FP64 residual with
  np.sum; the first strict minimum).  Small inputs only.
* check_template_pair: SF/overlap/neuron_separability.py:59-110.
"""
import numpy as np
from scipy.stats import norm


def _shift(template, shift, C, L):
    out = np.zeros(C * L)
    for c in range(C):
        for j in range(L):
            if 0 <= j - shift < L:
                out[c * L + j] = template[c * L + j - shift]
    return out


def find_overlap_templates(templates, L, C, n_pre, n_post, n_template_spikes):
    T = np.asarray(templates, dtype=np.float32)
    U = T.shape[0]
    if U < 3:
        return []
    SS = np.sum(T ** 2, axis=1)
    shifts = list(range(-n_pre, n_post + 1))
    allsh = [[_shift(T[u], s, C, L) for s in shifts] for u in range(U)]
    out = []
    for t in np.argsort(SS)[-1::-1]:
        test = T[t]
        cmin, cmax = np.argmin(test) // L, np.argmax(test) // L
        best, pick = np.inf, None
        for n1 in range(U):
            if n_template_spikes[n1] < 4 * n_template_spikes[t] or n1 == t:
                continue
            for n2 in range(n1 + 1, U):
                if n_template_spikes[n2] < 4 * n_template_spikes[t] or n2 == t:
                    continue
                if np.amin(T[n1, cmin * L:(cmin + 1) * L]) + np.amin(T[n2, cmin * L:(cmin + 1) * L]) > \
                        0.75 * np.amin(test[cmin * L:(cmin + 1) * L]):
                    continue
                if np.amax(T[n1, cmax * L:(cmax + 1) * L]) + np.amax(T[n2, cmax * L:(cmax + 1) * L]) < \
                        0.75 * np.amax(test[cmax * L:(cmax + 1) * L]):
                    continue
                for i1 in range(len(shifts)):
                    for i2 in range(len(shifts)):
                        s = allsh[n1][i1] + allsh[n2][i2]
                        r = np.sum((test - s) ** 2)
                        if r < best:
                            best, pick = r, (n1, n2, s.copy())
        if pick is not None:
            out.append([int(t), pick[2], [pick[0], pick[1]]])
    return out


def check_template_pair(t1, t2, covs, C, L):
    E1 = 0.5 * np.dot(t1, t1)
    E2 = np.dot(t1, t2) - 0.5 * np.dot(t2, t2)
    var = 0.0
    for c in range(C):
        d = t1[c * L:(c + 1) * L] - t2[c * L:(c + 1) * L]
        var += d @ covs[c] @ d
    return norm.cdf(0, E1 - E2, np.sqrt(var)) if var > 0 else 1.0
