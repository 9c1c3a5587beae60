"""Batched binary pursuit over many independent recording segments.

The reference pursues one recording (or chunk) at a time: its chunk loop
(SF/overlap/binary_pursuit_parallel.py:508-1500, SF = spikesorting_fullpursuit/)
runs the pass loop of each chunk to completion before the next.  A batch of
segments of uneven length, channel count and unit count is the workload of a
lab that sorts many short sessions or tetrodes; one small segment alone leaves
most of the 256 compute units idle.

Here a pool of worker threads, each owning one library handle (one HIP stream
and its device buffers), takes segments from a queue ordered largest first (by
samples x channels x units, so the long ones start early) and runs the drop-in
binary_pursuit on each.  The library calls release the GIL, so the pass loops
of different segments interleave on the device.  Every segment's pass loop is
the same deterministic sequence of kernels on its own stream, so the results
equal separate calls bit for bit.

autotune_workers() picks the number of concurrent streams for a batch by
timing a sample of its segments at each candidate count.
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


def _pool(device):
    p = _pools.get(device)
    if p is None:
        p = _Pool(device)
        _pools[device] = p
    return p


def release_device_buffers():
    for p in list(_pools.values()):
        p.close()
    _pools.clear()


def binary_pursuit_batch(segments, n_workers=4, device=0, precision="mixed", stats=None):
    """Run binary_pursuit on every segment with `n_workers` concurrent streams.

    segments: list of dicts with the drop-in's arguments: "voltage" ([C][N]
    array or float32 device tensor), "sort_info", "separability_metrics" and
    optionally "n_max_shift_inds", "v_dtype" (default float32).
    Returns the list of drop-in results (event_indices, neuron_labels,
    binary_pursuit_spike_bool, adjusted_clips), in input order.  Each
    separability_metrics dict gets its "n_spikes" as from the drop-in.
    `stats` (a dict) receives the wall time, the worker count and the segment
    order.
    """
    n = len(segments)
    if n == 0:
        return []
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
    binary_pursuit_batch(sub, n_workers=max(candidates), device=device, precision=precision)
    rates = {}
    for c in candidates:
        st = {}
        binary_pursuit_batch(sub, n_workers=c, device=device, precision=precision, stats=st)
        rates[c] = work / st["seconds"] / 1e6
    best = max(rates, key=lambda c: rates[c])
    return best, rates
