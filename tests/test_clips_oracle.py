"""Pins the clip oracle (oracle/clips_ref.py) and the host edge rule of the
GPU get_clips.  The reference holds no fixtures for get_clips; an index-coded
voltage V[c][x] = 1e6 c + x makes every clip value name its own channel and
sample, so a wrong channel order, offset or window width fails."""
import numpy as np

from oracle import clips_ref
from spikesorting_fullpursuit_amd.processing import clip_utils, conversions


def test_clip_values_name_their_channel_and_sample():
    C, N = 5, 400
    V = (1e6 * np.arange(C)[:, None] + np.arange(N)[None, :]).astype(np.float64)
    window, _ = conversions.time_window_to_samples([-3 / 1000, 4 / 1000], 1000.0)
    w0, w1 = window
    ev = np.array([1, 2, 50, 51, 300, 397, 399])
    nbrs = [3, 0, 4]
    clips, valid = clips_ref.get_clips(V, nbrs, ev, w0, w1, N, np.float64)
    kept = ev[valid]
    assert list(valid) == [False, False, True, True, True, False, False]  # w0 = -3, w1 = 5
    W = w1 - w0
    for row, n in enumerate(kept):
        for q, c in enumerate(nbrs):
            np.testing.assert_array_equal(clips[row, q * W:(q + 1) * W], 1e6 * c + n + np.arange(w0, w1))


def test_edge_rule_matches_the_oracle_scan():
    rng = np.random.default_rng(3)
    for _ in range(500):
        n = int(rng.integers(1, 9))
        N = 60
        w0, w1 = -int(rng.integers(0, 6)), int(rng.integers(1, 8))
        ev = np.sort(rng.integers(-8, 70, size=n))
        valid = np.ones(n, dtype=bool)
        s, e = clip_utils.check_edge_cases(ev, N, valid, [w0, w1])
        out, v2 = clips_ref.get_clips(np.zeros((1, N)), [0], ev, w0, w1, N)
        assert np.array_equal(valid, v2)
        if v2.any():
            assert (e - s + 1) == out.shape[0] and np.all(valid[s:e + 1])


def test_empty_and_all_invalid():
    out, valid = clips_ref.get_clips(np.zeros((2, 10)), [0, 1], np.zeros(0, np.int64), -2, 3, 10)
    assert out.shape == (0, 10) and valid.shape == (0,)
    out, valid = clips_ref.get_clips(np.zeros((2, 10)), [0, 1], np.array([0, 1]), -2, 3, 10)
    assert out is None and not valid.any()
