"""GPU: batched pursuit of many independent segments (the fused pass loop and
concurrent streams) gives, per segment, exactly the spikes of a separate
drop-in call, and those equal the CPU oracle's."""
import numpy as np
import pytest
import torch

from oracle import fbp_oracle
from spikesorting_fullpursuit_amd import batch, synthetic
from spikesorting_fullpursuit_amd.overlap import binary_pursuit_parallel as bpp

pytestmark = pytest.mark.gpu


def _segments(n, seed, on_device=False):
    ws = synthetic.batch_workloads(n, seed=seed, duration_range=(0.3, 1.5), unit_range=(4, 16))
    segs = []
    for w in ws:
        d = synthetic.generate_workload(w)
        V = torch.tensor(d["voltage"], device="cuda") if on_device else d["voltage"]
        segs.append({"voltage": V, "sort_info": d["sort_info"], "separability_metrics": d["separability_metrics"],
                     "n_max_shift_inds": d["n_max_shift_inds"], "data": d})
    print(f"[batch] {n} segments, channels {[w.n_chans for w in ws]}, units {[w.n_units for w in ws]}", flush=True)
    return segs


def _separate(seg):
    m = dict(seg["separability_metrics"])
    return bpp.binary_pursuit(seg["voltage"], np.float32, seg["sort_info"], m, n_max_shift_inds=seg["n_max_shift_inds"])


@pytest.mark.parametrize("mode,workers,lanes", [("fused", 1, 1), ("fused", 1, 3), ("streams", 1, 1), ("streams", 4, 1)])
def test_batch_equals_separate_calls(mode, workers, lanes):
    segs = _segments(10, 11)
    out = batch.binary_pursuit_batch(segs, n_workers=workers, mode=mode, lanes=lanes)
    for seg, got in zip(segs, out):
        ref = _separate(seg)
        assert len(got[0]) == len(ref[0])
        np.testing.assert_array_equal(np.asarray(got[0]), np.asarray(ref[0]))
        np.testing.assert_array_equal(np.asarray(got[1]), np.asarray(ref[1]))
        assert np.array_equal(seg["separability_metrics"]["n_spikes"],
                              np.bincount(np.asarray(ref[1], np.int64), minlength=len(seg["separability_metrics"]["n_spikes"])))


@pytest.mark.parametrize("mode", ["fused", "streams"])
def test_batch_matches_oracle(mode):
    segs = _segments(3, 12)
    out = batch.binary_pursuit_batch(segs, n_workers=3, mode=mode)
    for seg, got in zip(segs, out):
        d = seg["data"]
        T, bias, thr = synthetic.kernel_inputs(d)
        oi, ol = fbp_oracle.pursuit(d["voltage"], T, bias, thr, d["n_max_shift_inds"])
        np.testing.assert_array_equal(np.asarray(got[0], np.int64), oi.astype(np.int64) + d["clip_init_samples"])
        np.testing.assert_array_equal(np.asarray(got[1], np.int64), ol.astype(np.int64))


def test_batch_device_resident_segments():
    host = _segments(6, 13)
    dev = _segments(6, 13, on_device=True)
    a = batch.binary_pursuit_batch(host, n_workers=4, mode="streams")
    b = batch.binary_pursuit_batch(dev)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(np.asarray(x[0]), np.asarray(y[0]))
        np.testing.assert_array_equal(np.asarray(x[1]), np.asarray(y[1]))


def test_autotune_picks_a_candidate():
    segs = _segments(8, 14)
    best, rates = batch.autotune_workers(segs, candidates=(1, 2, 4), sample=6)
    assert best in (1, 2, 4) and set(rates) == {1, 2, 4}
    assert all(r > 0 for r in rates.values())


def test_lane_autotune_picks_a_candidate_and_lanes_give_the_same_spikes():
    segs = _segments(8, 16)
    best, rates = batch.autotune_lanes(segs, candidates=(1, 2, 4), sample=8)
    assert best in (1, 2, 4) and set(rates) == {1, 2, 4}
    ref = batch.binary_pursuit_batch(segs, lanes=1)
    got = batch.binary_pursuit_batch(segs, lanes=4)
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(np.asarray(a[0]), np.asarray(b[0]))
        np.testing.assert_array_equal(np.asarray(a[1]), np.asarray(b[1]))


def test_tile_autotune_and_every_shape_gives_the_same_spikes():
    segs = _segments(8, 16)
    best, rates = batch.autotune_tiles(segs, block_caps=(16, 1024), lik_windows=(0, 1, 3), sample=6)
    assert best in rates and set(rates) == {(c, w) for c in (16, 1024) for w in (0, 1, 3)}
    ref = batch.binary_pursuit_batch(segs, mode="streams", n_workers=2)
    for tiles in [(1, 1), (16, 3), (4096, 16)]:
        got = batch.binary_pursuit_batch(segs, tiles=tiles)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(np.asarray(a[0]), np.asarray(b[0]))
            np.testing.assert_array_equal(np.asarray(a[1]), np.asarray(b[1]))


def test_fused_batch_adjusted_clips():
    """The fused batch returns every segment's adjusted clips as the drop-in does."""
    segs = _segments(4, 17)
    for s in segs:
        s["sort_info"] = dict(s["sort_info"], get_adjusted_clips=True)
    got = batch.binary_pursuit_batch(segs)
    for seg, g in zip(segs, got):
        ref = _separate(seg)
        np.testing.assert_array_equal(np.asarray(g[0]), np.asarray(ref[0]))
        np.testing.assert_array_equal(np.asarray(g[3]), np.asarray(ref[3]))


def test_batch_reports_failing_segment():
    segs = _segments(3, 15)
    bad = dict(segs[1])
    m = dict(bad["separability_metrics"])
    m["templates"] = m["templates"][:, :-1]  # wrong width
    bad["separability_metrics"] = m
    segs[1] = bad
    with pytest.raises(RuntimeError, match="segment 1"):
        batch.binary_pursuit_batch(segs, n_workers=2, mode="streams")
    with pytest.raises(RuntimeError, match="segment 1"):
        batch.binary_pursuit_batch(segs)
