"""CPU restatement of get_clips (SF/processing/clip_utils.py:402-632).

TEST INFRASTRUCTURE ONLY.  Slices the voltage row by row, as the reference
does: for each kept event and each channel of `neighbors` in order, the
samples [n + w0, n + w1).  The kept events are the reference's: drop the
leading run of events whose clip starts before 0 and the trailing run whose
clip ends after n_samples (scanned one event at a time from each end).
"""
import numpy as np


def get_clips(voltage, neighbors, event_indices, w0, w1, n_samples, v_dtype=np.float32):
    ev = np.asarray(event_indices, dtype=np.int64)
    W = w1 - w0
    width = W * len(neighbors)
    if ev.size == 0:
        return np.zeros((0, width), dtype=v_dtype), np.ones(0, dtype=bool)
    valid = np.ones(ev.size, dtype=bool)
    s = 0
    while s < ev.size and ev[s] + w0 < 0:
        valid[s] = False
        s += 1
    e = ev.size - 1
    while e >= 0 and ev[e] + w1 > n_samples:
        valid[e] = False
        e -= 1
    if s == ev.size or e < 0:
        valid[:] = False
        return None, valid
    out = np.empty((e - s + 1, width), dtype=v_dtype)
    for row, k in enumerate(range(s, e + 1)):
        for q, c in enumerate(neighbors):
            out[row, q * W:(q + 1) * W] = voltage[c, ev[k] + w0:ev[k] + w1]
    return out, valid
