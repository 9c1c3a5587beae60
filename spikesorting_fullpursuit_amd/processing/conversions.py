"""Clip-window conversion used by the pursuit output offset.

Restates SF/processing/conversions.py:4-39 (SF = spikesorting_fullpursuit/).
"""
import copy


def time_window_to_samples(time_window_s, sampling_rate):
    """[t0, t1] seconds -> ([w0, w1], new_time_window_s).

    w0 = min(round(-|t0| * fs), 0); w1 = max(round(t1 * fs), 1) + 1 so the
    window [w0, w1) includes sample round(t1 * fs).  The first element of the
    returned time window is made non-positive.
    """
    new_time_window_s = copy.copy(time_window_s)
    if new_time_window_s[0] > 0:
        new_time_window_s[0] *= -1
    w0 = min(int(round(new_time_window_s[0] * sampling_rate)), 0)
    w1 = max(int(round(new_time_window_s[1] * sampling_rate)), 1) + 1
    return [w0, w1], new_time_window_s
