"""Batched binary pursuit over many independent recording segments.

The reference pursues one recording (or chunk) at a time: its chunk loop
(SF/overlap/binary_pursuit_parallel.py:508-1500, SF = spikesorting_fullpursuit/)
runs the pass loop of each chunk to completion before the next.  A batch of
segments of uneven length, channel count and unit count is the workload of a
lab that sorts many short sessions or tetrodes; one small segment alone leaves
most of the 256 compute units idle.

Two ways to run a batch:

- "fused" (default): the pass loops of all segments run together in the
  library (fbp_run_batch): every kernel of a pass is launched once over all
  segments still converging, with two host reads of all segments' counters per
  pass.  Segments are grouped so that each group's buffers fit the device.  The
  likelihood and the screened recheck run in FP32 (the same spikes as every
  other precision).
- "streams": a pool of worker threads, each owning one library handle (one HIP
  stream and its device buffers), takes segments from a queue ordered largest
  first and runs the drop-in binary_pursuit on each; the pass loops of
  different segments interleave on the device.

Either way every segment gets exactly the spikes of a separate drop-in call.
autotune_workers() picks the stream count of the "streams" mode;
autotune_block_cap() picks the fused mode's blocks per segment and launch.
"""
import threading
import time

import numpy as np

from . import _native
from .overlap import binary_pursuit_parallel as bpp


def segment_cost(seg):
    """Work estimate used to order the queue: samples x channels x units."""
    V = seg["voltage"]
    tm = np.asarray(seg["separability_metrics"]["templates"])
    U = tm.shape[0] if tm.ndim == 2 else 1
    return int(V.shape[0]) * int(V.shape[1]) * U


class _Pool:
    """Worker handles, created once and reused across batches."""

    def __init__(self, device):
        self.device = device
        self.handles = []

    def get(self, n):
        while len(self.handles) < n:
            self.handles.append(_native.Handle(self.device))
        return self.handles[:n]

    def close(self):
        for h in self.handles:
            h.close()
        self.handles = []


_pools = {}


def _pool(device, lane=0):
    """The handles of `device` for lane `lane` of a fused batch (lane 0 also
    serves the streams mode)."""
    p = _pools.get((device, lane))
    if p is None:
        p = _Pool(device)
        _pools[(device, lane)] = p
    return p


def release_device_buffers():
    for p in list(_pools.values()):
        p.close()
    _pools.clear()


def _batch_fused(segments, device, stats, tiles, lanes=1, graph=False):
    """Fused mode of binary_pursuit_batch: the drop-in's argument handling,
    chunk plan and output assembly per segment, one fbp_run_batch per group.
    The chunks are shared out over `lanes` lanes (largest first, each to the
    lane with the least work so far); the lanes run side by side, each its
    groups in turn on its own handles and stream.  graph: every pass after
    the first is one replay of a hipGraph of the pass."""
    t0 = time.perf_counter()
    prepared = []
    for i, seg in enumerate(segments):
        try:
            prepared.append(bpp.prepare_inputs(seg["voltage"], seg["sort_info"], seg["separability_metrics"],
                                               seg.get("n_max_shift_inds"), precision="fp32"))
        except Exception as e:
            raise RuntimeError(f"segment {i} failed: {e}") from e
    dev_mem = _pool(device).get(1)[0].device_memory()
    # work items: (segment, chunk start, chunk stop, bytes), chunks as the drop-in plans them
    items = []
    for i, p in enumerate(prepared):
        per_chunk, onsets = bpp.chunk_plan(
            p["n_chans"], p["n_samples"], p["sampling_rate"], p["n_templates"], p["template_samples_per_chan"],
            p["n_max_shift_inds"], dev_mem, None, template_bytes=p["template_bytes"], resident_bytes=p["resident"],
            shared_pairs=p["shared_pairs"])
        const, per_sample = bpp.library_device_bytes(p["n_chans"], per_chunk, p["n_templates"],
                                                     p["template_samples_per_chan"], p["n_max_shift_inds"],
                                                     shared_pairs=p["shared_pairs"])
        for a in onsets:
            b = min(p["n_samples"], a + per_chunk)
            items.append((i, int(a), int(b), const + per_sample * (b - a)))
    lanes = max(1, min(int(lanes), len(items)))
    # lanes: largest chunk first, each to the lane with the least work so far
    cost = [(b - a) * prepared[i]["n_chans"] * prepared[i]["n_templates"] for (i, a, b, _) in items]
    load = [0] * lanes
    lane_items = [[] for _ in range(lanes)]
    for j in sorted(range(len(items)), key=lambda j: -cost[j]):
        k = min(range(lanes), key=lambda k: load[k])
        lane_items[k].append(items[j])
        load[k] += cost[j]
    # groups that fit the device together (the lanes share it), one kind of input and one flag set each
    budget = (0.8 * dev_mem - sum(p["resident"] for p in prepared)) / lanes
    lane_groups = []
    for its in lane_items:
        groups, cur, used, key = [], [], 0.0, None
        for it in its:
            p = prepared[it[0]]
            k = (p["on_device"], p["flags"] & (_native.FLAG_OVERLAP_RECHECK | _native.FLAG_ADJUSTED_CLIPS))
            if cur and (k != key or used + it[3] > budget or len(cur) >= 65535):
                groups.append((key, cur))
                cur, used = [], 0.0
            cur.append(it)
            used += it[3]
            key = k
        if cur:
            groups.append((key, cur))
        lane_groups.append(groups)
    chunks = {}  # (segment, chunk start) -> (indices, labels, spike bools, clips)
    dev_views = {}
    if any(prepared[i]["on_device"] for i, _, _, _ in items):
        import torch
        for i, p in enumerate(prepared):
            if p["on_device"]:
                V = segments[i]["voltage"]
                if V.dtype != torch.float32 or V.stride(1) != 1:
                    V = V.to(torch.float32).contiguous()
                torch.cuda.current_stream(V.device).synchronize()
                dev_views[i] = V

    def run_lane(lane, groups):
        pool = _pool(device, lane)
        for (on_device, flags), group in groups:
            hs = pool.get(len(group))
            volts, nms = [], []
            for h, (i, a, b, _) in zip(hs, group):
                p = prepared[i]
                h.set_templates(p["T3"], p["bias"], p["thr"])
                if on_device:
                    V = dev_views[i]
                    volts.append((V.data_ptr() + 4 * a, V.stride(0), b - a))
                else:
                    chunk = segments[i]["voltage"][:, a:b]
                    if (not isinstance(chunk, np.ndarray) or chunk.dtype != np.float32 or chunk.strides[1] != 4
                            or chunk.strides[0] % 4 != 0):
                        chunk = np.float32(chunk)
                    volts.append(chunk)
                nms.append(int(p["n_max_shift_inds"]))
            _native.run_batch(hs, volts, nms, flags | (_native.FLAG_BATCH_GRAPH if graph else 0),
                              block_cap=tiles[0], lik_windows=tiles[1])
            for h, (i, a, b, _) in zip(hs, group):
                idx, lab = h.result()
                if idx.shape[0] > 0:
                    clips = h.clips(idx.shape[0]) if prepared[i]["get_adjusted_clips"] else []
                    chunks[(i, a)] = (idx.astype(np.int64) + a, lab, np.ones(idx.shape[0], dtype="bool"), clips)

    errors = []

    def guarded(lane, groups):
        try:
            run_lane(lane, groups)
        except BaseException as e:  # re-raised in the caller's thread
            errors.append(e)

    if lanes == 1:
        run_lane(0, lane_groups[0])
    else:
        threads = [threading.Thread(target=guarded, args=(k, g), daemon=True) for k, g in enumerate(lane_groups)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
    per_seg = [([], [], [], []) for _ in segments]
    for (i, a) in sorted(chunks):
        for dst, v in zip(per_seg[i], chunks[(i, a)]):
            dst.append(v)
    results = []
    for seg, p, (si, sl, sb, sc) in zip(segments, prepared, per_seg):
        r = bpp.assemble_outputs(si, sl, sb, sc, seg.get("v_dtype", np.float32), p["clip_init_samples"])
        bpp.neuron_separability.add_n_spikes(seg["separability_metrics"], r[1])
        results.append(r)
    if stats is not None:
        stats.update({"seconds": time.perf_counter() - t0, "mode": "fused",
                      "groups": sum(len(g) for g in lane_groups), "lanes": lanes, "chunks": len(items),
                      "tiles": tuple(tiles), "graph": bool(graph)})
    return results


def binary_pursuit_batch(segments, n_workers=4, device=0, precision="mixed", stats=None, mode="fused",
                         tiles=(0, 0), lanes=2, graph=False):
    """Run binary_pursuit on every segment: all pass loops together ("fused",
    tile shape `tiles` = (blocks per segment and launch, windows per likelihood
    block), 0 = the library's default; `lanes` fused loops side by side;
    `graph`: passes after the first as hipGraph replays) or on `n_workers`
    concurrent streams ("streams", `precision` as the drop-in's).

    segments: list of dicts with the drop-in's arguments: "voltage" ([C][N]
    array or float32 device tensor), "sort_info", "separability_metrics" and
    optionally "n_max_shift_inds", "v_dtype" (default float32).
    Returns the list of drop-in results (event_indices, neuron_labels,
    binary_pursuit_spike_bool, adjusted_clips), in input order.  Each
    separability_metrics dict gets its "n_spikes" as from the drop-in.
    `stats` (a dict) receives the wall time and the mode's parameters.
    """
    n = len(segments)
    if graph and (mode != "fused" or int(lanes) != 1):
        raise ValueError("graph=True runs the fused mode on one lane (graph calls stay on one host thread)")
    if graph:
        raise NotImplementedError("graph=True is disabled in this build (the library refuses FBP_FLAG_BATCH_GRAPH)")
    if n == 0:
        return []
    if mode == "fused":
        return _batch_fused(segments, device, stats, tiles, lanes, graph)
    if mode != "streams":
        raise ValueError("mode must be 'fused' or 'streams'")
    n_workers = max(1, min(int(n_workers), n))
    order = sorted(range(n), key=lambda i: -segment_cost(segments[i]))
    handles = _pool(device).get(n_workers)
    results = [None] * n
    errors = []
    lock = threading.Lock()
    nxt = [0]

    def worker(h):
        while True:
            with lock:
                if errors or nxt[0] >= n:
                    return
                i = order[nxt[0]]
                nxt[0] += 1
            seg = segments[i]
            try:
                results[i] = bpp.binary_pursuit(
                    seg["voltage"], seg.get("v_dtype", np.float32), seg["sort_info"], seg["separability_metrics"],
                    n_max_shift_inds=seg.get("n_max_shift_inds"), device=device, precision=precision, handle=h)
            except BaseException as e:  # re-raised in the caller's thread
                with lock:
                    errors.append((i, e))
                return

    t0 = time.perf_counter()
    if n_workers == 1:
        worker(handles[0])
    else:
        threads = [threading.Thread(target=worker, args=(h,), daemon=True) for h in handles]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    wall = time.perf_counter() - t0
    if errors:
        i, e = errors[0]
        raise RuntimeError(f"segment {i} failed: {e}") from e
    if stats is not None:
        stats.update({"seconds": wall, "n_workers": n_workers, "order": order})
    return results


def autotune_workers(segments, candidates=(1, 2, 4, 8), sample=16, device=0, precision="mixed", seed=0):
    """Pick the worker count with the highest throughput on a sample of the batch.

    Times binary_pursuit_batch on `sample` segments drawn evenly from the
    batch's cost order (so the sample has the batch's mix of sizes) once per
    candidate, after one untimed warm-up run.  Returns (best, {count: Msamples*ch/s}).
    """
    if not segments:
        return 1, {}
    order = sorted(range(len(segments)), key=lambda i: -segment_cost(segments[i]))
    k = min(sample, len(segments))
    pick = [order[int(j)] for j in np.linspace(0, len(order) - 1, k).round().astype(int)]
    sub = [segments[i] for i in pick]
    work = sum(int(s["voltage"].shape[0]) * int(s["voltage"].shape[1]) for s in sub)
    binary_pursuit_batch(sub, n_workers=max(candidates), device=device, precision=precision, mode="streams")
    rates = {}
    for c in candidates:
        st = {}
        binary_pursuit_batch(sub, n_workers=c, device=device, precision=precision, stats=st, mode="streams")
        rates[c] = work / st["seconds"] / 1e6
    best = max(rates, key=lambda c: rates[c])
    return best, rates


def _sample(segments, sample):
    order = sorted(range(len(segments)), key=lambda i: -segment_cost(segments[i]))
    k = min(sample, len(segments))
    pick = [order[int(j)] for j in np.linspace(0, len(order) - 1, k).round().astype(int)]
    return [segments[i] for i in pick]


def autotune_tiles(segments, block_caps=(256, 1024, 4096), lik_windows=(0, 1, 4), sample=32, device=0):
    """Pick the fused mode's tile shape with the highest throughput on a sample
    of the batch: every (blocks per segment and launch, windows per likelihood
    block) pair is timed on `sample` segments drawn evenly from the batch's cost
    order, after one untimed warm-up run.  Returns (best, {shape: Msamples*ch/s}).
    Results do not depend on the shape."""
    if not segments:
        return (0, 0), {}
    sub = _sample(segments, sample)
    work = sum(int(s["voltage"].shape[0]) * int(s["voltage"].shape[1]) for s in sub)
    binary_pursuit_batch(sub, device=device, mode="fused")
    rates = {}
    for cap in block_caps:
        for wb in lik_windows:
            st = {}
            binary_pursuit_batch(sub, device=device, stats=st, mode="fused", tiles=(cap, wb))
            rates[(cap, wb)] = work / st["seconds"] / 1e6
    best = max(rates, key=lambda k: rates[k])
    return best, rates


def autotune_lanes(segments, candidates=(1, 2, 3, 4), sample=64, tiles=(0, 0), device=0, graphs=(False,)):
    """Pick the fused mode's lane count (fused loops side by side) with the
    highest throughput on `sample` segments drawn evenly from the batch's cost
    order, after one untimed warm-up run.  Returns (best, {lanes: Msamples*ch/s}).
    With several `graphs` settings the (lanes, graph) pairs are timed (graph
    mode at one lane only) and the keys are such pairs.  Results do not depend
    on either."""
    if not segments:
        return (1, graphs[0]) if len(graphs) > 1 else 1, {}
    sub = _sample(segments, sample)
    work = sum(int(s["voltage"].shape[0]) * int(s["voltage"].shape[1]) for s in sub)
    binary_pursuit_batch(sub, device=device, mode="fused", tiles=tiles, lanes=max(candidates))
    rates = {}
    for k in candidates:
        for gr in graphs:
            if gr and k != 1:
                continue  # graph mode runs one lane
            st = {}
            binary_pursuit_batch(sub, device=device, stats=st, mode="fused", tiles=tiles, lanes=k, graph=gr)
            rates[(k, gr) if len(graphs) > 1 else k] = work / st["seconds"] / 1e6
    best = max(rates, key=lambda k: rates[k])
    return best, rates
