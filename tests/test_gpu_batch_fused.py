"""GPU: the batched pass loop (fbp_run_batch) gives, per segment, exactly the
spikes, labels, order, residual and per-pass trace of a separate fbp_run on
that segment, and those equal the CPU oracle's."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import fbp_oracle
from spikesorting_fullpursuit_amd import _native, synthetic

pytestmark = pytest.mark.gpu

CHANNELS = (1, 4, 8, 16, 32)
FLAGS = _native.FLAG_OVERLAP_RECHECK | _native.FLAG_SCREEN_FP32


_GRAPH_OFF = pytest.mark.skip(reason="hipGraph batch replays are switched off: a one-lane graph batch ended in a host "
                               "segmentation fault whose cause is not known")


def _segments(n, seed, dur=(0.3, 1.0), units=(4, 60)):
    """n segments of the batched workload's kind, channel counts cycling over
    1, 4, 8, 16 and 32, unit counts drawn from `units`."""
    ws = synthetic.batch_workloads(n, seed=seed, duration_range=dur, unit_range=units)
    out = []
    for i, w in enumerate(ws):
        C = CHANNELS[i % len(CHANNELS)]
        xy = np.stack([np.zeros(C), 25.0 * np.arange(C)], axis=1)
        w = dataclasses.replace(w, n_chans=C, chan_xy=xy, neighbourhood=0 if C <= 4 else 5)
        d = synthetic.generate_workload(w)
        T, bias, thr = synthetic.kernel_inputs(d)
        out.append({"V": d["voltage"], "T": T, "bias": bias, "thr": thr, "n": d["n_max_shift_inds"]})
    print(f"[batch-fused] {n} segments: C {[s['V'].shape[0] for s in out]} U {[s['T'].shape[0] for s in out]} "
          f"N {[s['V'].shape[1] for s in out]}", flush=True)
    return out


def _handles(segs):
    hs = []
    for s in segs:
        h = _native.Handle(0)
        h.set_templates(s["T"], s["bias"], s["thr"])
        hs.append(h)
    return hs


def _separate(s, flags):
    with _native.Handle(0) as h:
        h.set_templates(s["T"], s["bias"], s["thr"])
        h.run(s["V"], s["n"], flags)
        idx, lab = h.result()
        return idx, lab, h.residual(s["V"].shape[1]), h.pass_log()


def _check(segs, hs, flags):
    for g, (s, h) in enumerate(zip(segs, hs)):
        idx, lab = h.result()
        R = h.residual(s["V"].shape[1])
        ri, rl, rR, rlog = _separate(s, flags)
        assert len(idx) == len(ri), g
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(lab, rl)
        np.testing.assert_array_equal(R, rR)
        np.testing.assert_array_equal(h.pass_log(), rlog)


def test_batch_of_20_equals_separate_runs():
    """20 segments, C in {1, 4, 8, 16, 32}, U in [4, 60]: batch == separate run."""
    segs = _segments(20, 31)
    assert {s["V"].shape[0] for s in segs} == set(CHANNELS)
    assert max(s["T"].shape[0] for s in segs) > 40
    hs = _handles(segs)
    _native.run_batch(hs, [s["V"] for s in segs], [s["n"] for s in segs], FLAGS)
    _check(segs, hs, FLAGS)
    for h in hs:
        h.close()


def test_batch_matches_oracle():
    """16 segments with at most 16 units and a short shift range (the oracle's
    literal recheck stays fast): batch == oracle, residual included."""
    segs = _segments(16, 32, units=(4, 16))
    for s in segs:
        s["n"] = 15
    hs = _handles(segs)
    _native.run_batch(hs, [s["V"] for s in segs], [s["n"] for s in segs], FLAGS)
    for g, (s, h) in enumerate(zip(segs, hs)):
        idx, lab = h.result()
        oi, ol, oR = fbp_oracle.pursuit(s["V"], s["T"], s["bias"], s["thr"], s["n"], return_residual=True)
        np.testing.assert_array_equal(idx, oi)
        np.testing.assert_array_equal(lab, ol)
        np.testing.assert_array_equal(h.residual(s["V"].shape[1]), oR)
        h.close()


def test_batch_without_recheck_and_device_voltage():
    segs = _segments(17, 33, units=(4, 30))
    hs = _handles(segs)
    dev = [torch.tensor(s["V"], device="cuda") for s in segs]
    _native.run_batch(hs, [(v.data_ptr(), v.stride(0), v.shape[1]) for v in dev], [s["n"] for s in segs], 0)
    _check(segs, hs, 0)
    for h in hs:
        h.close()


def test_batch_mixed_shift_ranges_and_lengths():
    """Per-segment shift ranges (0, 1, 7, 63) and one segment much longer than
    the others (it keeps converging after the rest have left)."""
    segs = _segments(8, 34, units=(4, 24))
    for g, s in enumerate(segs):
        s["n"] = (0, 1, 7, 63)[g % 4]
    long = _segments(1, 35, dur=(4.0, 4.0), units=(30, 30))[0]
    segs.append(long)
    hs = _handles(segs)
    _native.run_batch(hs, [s["V"] for s in segs], [s["n"] for s in segs], FLAGS)
    _check(segs, hs, FLAGS)
    for h in hs:
        h.close()


GRAPH = _native.FLAG_BATCH_GRAPH


def _graph_stats(hs):
    st = hs[0].stats()
    print(f"[batch-graph] builds {st['n_graph_builds']} replays {st['n_graph_replays']}", flush=True)
    return st


@_GRAPH_OFF
def test_graph_batch_of_20_equals_separate_runs():
    """Captured passes (one hipGraph replay per pass after the first), 20
    segments, C in {1, 4, 8, 16, 32}, U in [4, 60]: batch == separate run,
    pass logs included."""
    segs = _segments(20, 31)
    hs = _handles(segs)
    _native.run_batch(hs, [s["V"] for s in segs], [s["n"] for s in segs], FLAGS | GRAPH)
    st = _graph_stats(hs)
    assert st["n_graph_builds"] >= 1
    assert st["n_graph_replays"] >= 1
    _check(segs, hs, FLAGS)
    for h in hs:
        h.close()


@_GRAPH_OFF
def test_graph_batch_matches_oracle():
    segs = _segments(16, 32, units=(4, 16))
    for s in segs:
        s["n"] = 15
    hs = _handles(segs)
    _native.run_batch(hs, [s["V"] for s in segs], [s["n"] for s in segs], FLAGS | GRAPH)
    assert _graph_stats(hs)["n_graph_replays"] >= 1
    for g, (s, h) in enumerate(zip(segs, hs)):
        idx, lab = h.result()
        oi, ol, oR = fbp_oracle.pursuit(s["V"], s["T"], s["bias"], s["thr"], s["n"], return_residual=True)
        np.testing.assert_array_equal(idx, oi)
        np.testing.assert_array_equal(lab, ol)
        np.testing.assert_array_equal(h.residual(s["V"].shape[1]), oR)
        h.close()


@_GRAPH_OFF
def test_graph_batch_without_recheck_mixed_lengths_repeated():
    """No recheck; shift ranges 0..63 and one long segment; the same handles
    run twice (graphs of the second call are built afresh): both == separate."""
    segs = _segments(8, 34, units=(4, 24))
    for g, s in enumerate(segs):
        s["n"] = (0, 1, 7, 63)[g % 4]
    segs.append(_segments(1, 35, dur=(4.0, 4.0), units=(30, 30))[0])
    hs = _handles(segs)
    for flags in (GRAPH, FLAGS | GRAPH):
        _native.run_batch(hs, [s["V"] for s in segs], [s["n"] for s in segs], flags)
        _graph_stats(hs)
        _check(segs, hs, flags & ~GRAPH)
    for h in hs:
        h.close()


def test_batch_rejects_mismatched_template_lengths():
    segs = _segments(2, 36, units=(4, 8))
    hs = _handles(segs)
    T = segs[1]["T"][:, :, :32].copy()
    hs[1].set_templates(T, segs[1]["bias"], segs[1]["thr"])
    with pytest.raises(Exception, match="template length"):
        _native.run_batch(hs, [s["V"] for s in segs], [s["n"] for s in segs], FLAGS)
    for h in hs:
        h.close()
