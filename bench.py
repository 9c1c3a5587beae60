"""Benchmark: full binary pursuit on the 64-channel recording, one MI355X.

One step = one drop-in binary_pursuit() call over the whole recording
(64 ch, 30 kHz, 600 s, 120 units, 48-sample templates on 7-channel
neighbourhoods, n_max_shift_inds = L - 1, overlap recheck on) with the voltage
a host numpy array, so the step includes its host-to-device copy; the rate
with the voltage already resident in HBM is reported beside it.  Prints one
JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--duration S]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pursuit Msamples·ch/s at 64ch; mixed-prec vs FP32 GPU at identical spike set"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


TRAFFIC_PROFILE = os.path.join("profiles", "pursuit_probe64_600s_hbm_traffic.md")


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-kernel scaled); None when the summary does not cover this workload."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if workload != "probe64" or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(line.split(":")[1])
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="probe64")
    ap.add_argument("--duration", type=float, default=None, help="recording length in s (default: full)")
    ap.add_argument("--literal", action="store_true", help="evaluate every overlap item literally")
    ap.add_argument("--cpu-sample-s", type=float, default=1.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: os.cpu_count(), capped by OMP_NUM_THREADS when set)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the BF16 / FP8 likelihood report")
    ap.add_argument("--variant", default="pursuit", choices=["pursuit", "refine", "batch"],
                    help="refine: template refinement on the 32-channel 120 s recording; batch: 256 independent "
                         "segments (5-120 s, 1-32 ch, 4-60 units) on concurrent streams (each its own JSON line)")
    ap.add_argument("--segments", type=int, default=256, help="segments of the batch variant")
    ap.add_argument("--profile-run", action="store_true",
                    help="for hardware counters: three 2 GiB device-to-device copies (the copy kernel of known "
                         "size that scales FETCH_SIZE / WRITE_SIZE), then one mixed-precision drop-in call "
                         "(FP32 with --profile-fp32) on the device-resident voltage; no JSON")
    ap.add_argument("--profile-fp32", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(the path would shard along time with one all-gather of the spike lists)"}))
        sys.exit(2)

    if args.variant == "refine":
        return run_refine(args)
    if args.variant == "batch":
        return run_batch(args)

    import torch
    from spikesorting_fullpursuit_amd import synthetic
    from spikesorting_fullpursuit_amd.overlap import binary_pursuit_parallel as bpp

    t0 = time.time()
    d = synthetic.generate_on_device(args.workload, duration_s=args.duration)
    V = d["voltage"]
    C, N = V.shape
    w = d["workload"]
    U, _, L = d["templates"].shape
    # The drop-in's usual input is a host array (the reference reads the voltage from host memory and
    # copies it to the device per chunk, binary_pursuit_parallel.py:524-545): the headline times that.
    Vh = V.cpu().numpy()
    log(f"generated {args.workload}: C={C} N={N} U={U} L={L} spikes={len(d['truth_onsets'])} "
        f"in {time.time() - t0:.1f}s (host copy {Vh.nbytes / 1e9:.2f} GB)")

    def step(precision, voltage, literal=args.literal):
        m = dict(d["separability_metrics"])
        return bpp.binary_pursuit(voltage, np.float32, d["sort_info"], m, n_max_shift_inds=d["n_max_shift_inds"],
                                  recheck_literal=literal, precision=precision)

    if args.profile_run:
        nbytes = 2 << 30
        src = torch.ones(nbytes // 4, device=V.device, dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        out = step("fp32" if args.profile_fp32 else "mixed", V)
        torch.cuda.synchronize()
        log(f"profile run: {nbytes} B copied 3 times; one {'FP32' if args.profile_fp32 else 'mixed'} call, "
            f"spikes={len(out[0])}")
        return

    def timed(precision, voltage, k):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(k):
            o = step(precision, voltage)
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / k, o

    for i in range(args.warmup):
        t1 = time.time()
        out = step("mixed", Vh)
        out32 = step("fp32", Vh)
        step("mixed", V)
        torch.cuda.synchronize()
        log(f"warmup {i}: {time.time() - t1:.3f}s spikes={len(out[0])}")
    k32 = max(1, args.steps // 2)
    sec_mixed, out = timed("mixed", Vh, args.steps)
    sec_fp32, out32 = timed("fp32", Vh, k32)
    sec_mixed_dev, out_dev = timed("mixed", V, args.steps)
    sec_fp32_dev, out32_dev = timed("fp32", V, k32)

    def same(a, b):
        return bool(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]))

    identical = same(out, out32)
    identical_all = identical and same(out, out_dev) and same(out, out32_dev)
    ms_per_step = 1000.0 * sec_mixed
    value = N * C / sec_mixed / 1e6
    value32 = N * C / sec_fp32 / 1e6
    value_dev = N * C / sec_mixed_dev / 1e6
    value32_dev = N * C / sec_fp32_dev / 1e6
    log(f"host input: mixed {ms_per_step:.1f} ms/step = {value:.1f} Msamples*ch/s, fp32 {1000 * sec_fp32:.1f} ms = "
        f"{value32:.1f}; device-resident: mixed {1000 * sec_mixed_dev:.1f} ms = {value_dev:.1f}, fp32 "
        f"{1000 * sec_fp32_dev:.1f} ms = {value32_dev:.1f}; identical spike sets: {identical_all}; "
        f"spikes={len(out[0])}")

    # Untimed runs with per-kernel HIP-event timing for the roofline.
    from spikesorting_fullpursuit_amd import _native
    h = bpp._handle(0)
    T3, bias, thr = synthetic.kernel_inputs(d)
    h.set_templates(T3, bias, thr)
    base = _native.FLAG_OVERLAP_RECHECK | _native.FLAG_TIMING | (_native.FLAG_RECHECK_LITERAL if args.literal else 0)
    h.run(None, d["n_max_shift_inds"], base | _native.FLAG_SCREEN_FP32, ld=V.stride(0), n_samples=N,
          device_ptr=V.data_ptr())
    st = h.stats()
    log("kernel stats fp32:", json.dumps(st))
    h.run(None, d["n_max_shift_inds"], base | _native.FLAG_PRECISION_MIXED, ld=V.stride(0), n_samples=N,
          device_ptr=V.data_ptr())
    stm = h.stats()
    log("kernel stats mixed:", json.dumps(stm))
    classes = {"likelihood": stm["ms_likelihood"], "recheck": stm["ms_recheck"], "window_max": stm["ms_window_max"],
               "pursuit": stm["ms_pursuit"], "residual": stm["ms_residual"], "other": stm["ms_other"]}
    classes32 = {"likelihood": st["ms_likelihood"], "recheck": st["ms_recheck"], "window_max": st["ms_window_max"],
                 "pursuit": st["ms_pursuit"], "residual": st["ms_residual"], "other": st["ms_other"]}
    S = 2 * int(d["n_max_shift_inds"]) + 1
    nz = np.count_nonzero(np.any(np.asarray(T3) != 0, axis=2), axis=1)
    # dominant kernel of the mixed path: the overlap screen selection.  Its
    # algorithmic traffic per flagged window is B(u, s) and the sign of
    # LIK[u][b+s-n] for every (unit, shift) pair plus the window's voltage.
    sel_bytes_per_window = U * S * 8 + C * L * 4
    sel_launch_ms = stm["ms_select"] / max(1, stm["n_select_launches"])
    sel_bytes_per_launch = stm["windows_flagged"] * sel_bytes_per_window / max(1, stm["n_select_launches"])
    sel_gbs = sel_bytes_per_launch / (sel_launch_ms * 1e-3) / 1e9 if sel_launch_ms > 0 else 0.0
    traffic = profiled_traffic("k_screen_select_reg", args.workload) if args.duration is None else None
    # FP32 path: its dominant kernel is k_screen_corr, the window-truncated correlations B(u, s) of every
    # unit and shift of each flagged window, 2 L nz(u) flops per (unit, shift) on the vector units.
    corr_flops_per_window = 2.0 * L * S * float(np.sum(nz))
    corr_ms = st["ms_recheck"] - st["ms_select"] - st.get("ms_rescore", 0.0)
    corr_launch_ms = corr_ms / max(1, st["n_select_launches"])
    corr_flops_per_launch = st["windows_flagged"] * corr_flops_per_window / max(1, st["n_select_launches"])
    corr_tflops = corr_flops_per_launch / (corr_launch_ms * 1e-3) / 1e12 if corr_launch_ms > 0 else 0.0
    lik_avg_ms = st["ms_likelihood"] / max(1, st["n_likelihood_launches"])
    # flops the FP32 likelihood kernel performs: after the first pass it recomputes
    # only the (window, template) pairs whose channels changed, each 2 L^2 nz flops
    lik_done = st["likelihood_pairs_computed"] * 2.0 * L * L * float(np.mean(nz))
    lik_done_tflops = lik_done / (st["ms_likelihood"] * 1e-3) / 1e12 if st["ms_likelihood"] > 0 else 0.0
    roofline = {
        "kernel": "k_screen_select_reg (overlap screen selection, dominant kernel of the mixed path)",
        "bound": "hbm",
        "achieved": round(sel_gbs, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(sel_gbs / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "traffic_source": (f"{TRAFFIC_PROFILE}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --profile-run, "
                           "copy-kernel scaled (not measured in this run)") if traffic is not None else None,
        "avg_launch_ms": round(sel_launch_ms, 4),
        "algorithmic_bytes_per_window": sel_bytes_per_window,
        "bytes_per_launch": round(sel_bytes_per_launch),
        "ms_by_class_mixed": {k: round(v, 3) for k, v in classes.items()},
        "ms_by_class_fp32": {k: round(v, 3) for k, v in classes32.items()},
        "ms_select_mixed": round(stm["ms_select"], 3),
        "ms_rescore_mixed": round(stm.get("ms_rescore", 0.0), 3),
        "fp32_path": {
            "kernel": "k_screen_corr (FP32 overlap-screen correlations, dominant kernel of the FP32 path)",
            "bound": "fp32 vector FMA", "achieved": round(corr_tflops, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(corr_tflops / FP32_PEAK_TFLOPS, 4),
            "peak_unpacked": FP32_PEAK_TFLOPS / 2, "frac_unpacked": round(corr_tflops / (FP32_PEAK_TFLOPS / 2), 4),
            "avg_launch_ms": round(corr_launch_ms, 4), "flops_per_window": corr_flops_per_window,
            "note": "algorithmic flops 2 L S sum_u nz(u) per flagged window / HIP-event time of the recheck class "
                    "minus the selection; the build issues no packed FP32 instructions, so one FMA per lane per "
                    "cycle (peak_unpacked) bounds it"},
        "likelihood_fp32": {"kernel": "k_likelihood (FP32 likelihood)", "achieved_tflops": round(lik_done_tflops, 3),
                            "peak": FP32_PEAK_TFLOPS, "frac": round(lik_done_tflops / FP32_PEAK_TFLOPS, 4),
                            "peak_unpacked": FP32_PEAK_TFLOPS / 2,
                            "frac_unpacked": round(lik_done_tflops / (FP32_PEAK_TFLOPS / 2), 4),
                            "avg_launch_ms": round(lik_avg_ms, 4),
                            "pairs_computed": st["likelihood_pairs_computed"],
                            "note": "recomputed (window, template) pairs x 2 L^2 x mean non-zero channels / kernel "
                                    "time"},
        "passes": stm["passes"],
    }
    if stm["n_likelihood_mfma_launches"] > 0:
        mf_ms = stm["ms_likelihood"]
        roofline["likelihood_mixed"] = {
            "kernel": "likelihood class of the mixed run: k_likmfma (matrix-core pass 1) + k_likelihood after",
            "launches_matrix_core": stm["n_likelihood_mfma_launches"], "ms": round(mf_ms, 3)}

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import fbp_oracle
        ns = int(round(args.cpu_sample_s * w.fs))
        Vs = np.ascontiguousarray(Vh[:, :ns])
        omp = os.environ.get("OMP_NUM_THREADS")
        nthr = args.cpu_threads or min(os.cpu_count() or 1, int(omp) if omp and omp.isdigit() else 1 << 30)
        t2 = time.perf_counter()
        oi, ol = fbp_oracle.pursuit(Vs, T3, bias, thr, d["n_max_shift_inds"], nthreads=nthr)
        tc = time.perf_counter() - t2
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": round(ns * C / tc / 1e6, 6), "unit": "Msamples·ch/s", "cores": nthr, "kind": "port",
               "sample": f"first {args.cpu_sample_s} s of the same recording ({ns} samples x {C} ch, "
                         f"{len(oi)} spikes), C restatement with OpenMP on {nthr} threads, literal overlap recheck, "
                         f"{tc:.1f} s",
               "cpu_model": model, "os_cpu_count": os.cpu_count(), "omp_num_threads": omp}
        h.run(None, d["n_max_shift_inds"], _native.FLAG_OVERLAP_RECHECK | _native.FLAG_PRECISION_MIXED |
              (_native.FLAG_RECHECK_LITERAL if args.literal else 0), ld=V.stride(0), n_samples=ns,
              device_ptr=V.data_ptr())
        gi, gl = h.result()
        parity = bool(np.array_equal(gi, oi) and np.array_equal(gl, ol))
        log(f"cpu baseline {cpu['value']} Msamples*ch/s on {nthr} threads; slice parity {parity} ({len(oi)} spikes)")

    precision = None if args.no_precision else precision_report(args, d, V, h, T3, bias, thr, out32_dev, oi_ol=(
        (oi, ol, ns) if cpu is not None else None))

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples·ch/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 MFMA + f32 (decisions in exact f32)",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {C} ch, {w.fs / 1000:g} kHz, {N / w.fs:g} s, {U} units, L={L}, "
                        f"n_max_shift_inds={d['n_max_shift_inds']}, overlap recheck on; voltage a host numpy "
                        f"float32 array, so each step includes its H2D (pinned staging) and the D2H of the spikes",
            "spikes_found": int(len(out[0])),
            "spikes_planted": int(len(d["truth_onsets"])),
            "recheck": "literal" if args.literal else "screened",
            "slice_parity_vs_oracle": parity,
            "fp32_value": round(value32, 3),
            "fp32_ms_per_step": round(1000 * sec_fp32, 3),
            "mixed_over_fp32": round(value / value32, 3),
            "identical_spike_set_mixed_vs_fp32": identical,
            "identical_spike_set_all_runs": identical_all,
            "device_resident_value": round(value_dev, 3),
            "device_resident_ms_per_step": round(1000 * sec_mixed_dev, 3),
            "device_resident_fp32_value": round(value32_dev, 3),
            "device_resident_mixed_over_fp32": round(value_dev / value32_dev, 3),
            "h2d_ms_per_step": round(1000 * (sec_mixed - sec_mixed_dev), 3),
            "likelihood_rescored_frac": round(stm["likelihood_rescored"] / max(1, stm["likelihood_positions"]), 5),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "precision_variants": precision,
    }
    print(json.dumps(line), flush=True)


def precision_report(args, d, V, h, T3, bias, thr, ref, oi_ol=None):
    """BF16 and FP8 (e4m3) matrix-core likelihood on the same recording: step time (device-resident
    voltage), likelihood class time, rescored fraction, agreement of the spike set with the FP32 path
    (whole recording) and with the C oracle (the CPU-baseline slice), and the error of the first-pass
    approximations against the exact FP32 likelihood (first 60 s) and the FP64 one (first 1 s)."""
    import torch
    from oracle import fbp_fp64
    from spikesorting_fullpursuit_amd import _native

    C, N = V.shape
    w = d["workload"]
    n = d["n_max_shift_inds"]
    fp_flags = {"bf16": _native.FLAG_PRECISION_BF16, "fp8": _native.FLAG_PRECISION_FP8}
    h.set_templates(T3, bias, thr)
    rec = _native.FLAG_OVERLAP_RECHECK

    def run(flags, ns=N):
        h.run(None, n, flags, ld=V.stride(0), n_samples=ns, device_ptr=V.data_ptr())

    n60 = min(N, int(round(60 * w.fs)))
    n1 = min(N, int(round(1 * w.fs)))
    run(_native.FLAG_SCREEN_FP32 | _native.FLAG_FIRST_PASS_ONLY, n60)
    lik32 = h.likelihood(n60)
    # FP64 likelihood of the first second (sliding-window products in float64)
    V1 = V[:, :n1].double().cpu().numpy()
    lik64 = fbp_fp64.likelihood_rows(V1, T3.astype(np.float64), bias.astype(np.float64), np.arange(n1))
    out = {}
    for name, f in fp_flags.items():
        run(rec | f)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k = max(1, args.steps)
        for _ in range(k):
            run(rec | f)
        sec = (time.perf_counter() - t0) / k
        gi, gl = h.result()
        run(rec | f | _native.FLAG_TIMING)
        st = h.stats()
        run(f | _native.FLAG_FIRST_PASS_ONLY | _native.FLAG_NO_RESCORE, n60)
        approx = h.likelihood(n60)
        e32 = np.abs(approx.astype(np.float64) - lik32.astype(np.float64))
        live = np.zeros(n60, bool)
        live[: n60 - T3.shape[2]] = True  # positions with x + L < N
        e64 = np.abs(approx[:, :n1].astype(np.float64) - lik64)
        e32_64 = np.abs(lik32[:, :n1].astype(np.float64) - lik64)
        slice_ok = None
        if oi_ol is not None:
            oi, ol, ns = oi_ol
            run(rec | f, ns)
            si, sl = h.result()
            slice_ok = bool(np.array_equal(si, oi) and np.array_equal(sl, ol))
        out[name] = {
            "value": round(N * C / sec / 1e6, 3), "unit": "Msamples·ch/s (device-resident voltage)",
            "ms_per_step": round(1000 * sec, 3),
            "ms_likelihood": round(st["ms_likelihood"], 3), "ms_total_device": round(st["ms_total"], 3),
            "likelihood_rescored_frac_of_listed": round(st["likelihood_rescored"] / max(1, st["likelihood_positions"]), 5),
            "identical_spike_set_vs_fp32_whole_recording": bool(
                np.array_equal(gi.astype(np.int64) + d["clip_init_samples"], ref[0]) and
                np.array_equal(gl.astype(np.int64), ref[1])),
            "identical_spike_set_vs_c_oracle_cpu_slice": slice_ok,
            "first_pass_error_vs_fp32": {"max": float(e32[:, live].max()), "mean": float(e32[:, live].mean()),
                                         "max_rel_to_max_abs_lik": float(e32[:, live].max() /
                                                                         np.abs(lik32[:, live]).max()),
                                         "positions": int(live.sum()) * T3.shape[0], "span_s": n60 / w.fs},
            "first_pass_error_vs_fp64": {"max": float(e64[:, :n1 - T3.shape[2]].max()),
                                         "mean": float(e64[:, :n1 - T3.shape[2]].mean()), "span_s": n1 / w.fs,
                                         "fp32_chain_max": float(e32_64[:, :n1 - T3.shape[2]].max())},
        }
        log(f"{name}: {out[name]}")
    out["note"] = ("decisions: every likelihood the approximation cannot prove <= 0 is recomputed with the exact FP32 "
                   "chain, so the spike sets equal the FP32 path's; agreement with the FP64 oracle is checked in "
                   "tests/test_gpu_precision.py (its overlap items are materialised, too slow for this recording)")
    return out


def run_batch(args):
    """Batched variant: independent segments with lengths 5-120 s, channel counts from {1, 4, 8, 16, 32}
    and 4-60 units (64-sample templates, 30 kHz), voltage resident in HBM.  The fused batch (every pass
    kernel launched once over all converging segments, tile shape from the autotuner) against the batch
    on concurrent streams and per-segment plain calls; all must give the identical spike set."""
    import torch
    from spikesorting_fullpursuit_amd import batch, synthetic

    t0 = time.time()
    ws = synthetic.batch_workloads(args.segments)
    segs = []
    dense_flops = 0.0
    for i, w in enumerate(ws):
        d = synthetic.generate_workload_on_device(w, pow2_fft=True)
        segs.append({"voltage": d["voltage"], "sort_info": d["sort_info"],
                     "separability_metrics": d["separability_metrics"], "n_max_shift_inds": d["n_max_shift_inds"]})
        T = np.asarray(d["separability_metrics"]["templates"]).reshape(w.n_units, w.n_chans, w.L)
        nz = int(np.count_nonzero(np.any(T != 0, axis=2)))
        dense_flops += 2.0 * w.n_samples * nz * w.L  # first-pass likelihood: every position of every template
        if i % 32 == 31:
            log(f"generated {i + 1} segments in {time.time() - t0:.1f}s")
    work = sum(w.n_samples * w.n_chans for w in ws)
    log(f"generated {len(segs)} segments, {work / 1e6:.1f} Msamples*ch, in {time.time() - t0:.1f}s")

    tiles, trates = batch.autotune_tiles(segs, sample=32)
    log(f"autotune: tiles {tiles}, Msamples*ch/s by (blocks per segment, windows per block) {trates}")
    lanes, lrates = batch.autotune_lanes(segs, candidates=(2, 4, 6, 8), sample=64, tiles=tiles)
    log(f"autotune: lanes {lanes}, Msamples*ch/s by lanes {lrates}")
    workers, wrates = batch.autotune_workers(segs, candidates=(2, 4, 6, 8), sample=12)
    log(f"autotune: workers {workers}, Msamples*ch/s by workers {wrates}")
    for _ in range(args.warmup):
        batch.binary_pursuit_batch(segs[:16], tiles=tiles, lanes=lanes)
    stf, sts, st1 = {}, {}, {}
    torch.cuda.synchronize()
    fused = batch.binary_pursuit_batch(segs, tiles=tiles, stats=stf, lanes=lanes)
    torch.cuda.synchronize()
    log(f"fused batch: {stf['seconds']:.2f}s in {stf['groups']} group(s)")
    streams = batch.binary_pursuit_batch(segs, n_workers=workers, stats=sts, mode="streams")
    torch.cuda.synchronize()
    log(f"batch on {workers} streams: {sts['seconds']:.2f}s")
    one = batch.binary_pursuit_batch(segs, n_workers=1, stats=st1, mode="streams")
    torch.cuda.synchronize()
    log(f"per-segment calls on one stream: {st1['seconds']:.2f}s")
    # one lane, passes planned on the host (hipGraph replays are switched off in this build)
    sth = {}
    host1 = batch.binary_pursuit_batch(segs, tiles=tiles, stats=sth, lanes=1)
    torch.cuda.synchronize()
    log(f"one lane: host-planned passes {sth['seconds']:.2f}s")

    def same(x, y):
        return all(np.array_equal(np.asarray(a[0]), np.asarray(b[0])) and np.array_equal(np.asarray(a[1]),
                                                                                         np.asarray(b[1]))
                   for a, b in zip(x, y))
    identical = same(fused, one) and same(streams, one) and same(host1, one)
    n_spk = int(sum(len(r[0]) for r in fused))
    value = work / stf["seconds"] / 1e6
    achieved = dense_flops / stf["seconds"] / 1e12
    line = {
        "metric": "batched pursuit Msamples·ch/s over 256 independent segments (fused pass loop) vs "
                  "per-segment plain calls at identical spike set",
        "value": round(value, 3),
        "unit": "Msamples·ch/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * stf["seconds"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(segs)} segments, 5-120 s, channels {{1,4,8,16,32}}, 4-60 units, L=64, 30 kHz, "
                        f"{work / 1e6:.1f} Msamples·ch in all, voltage resident in HBM",
            "tiles": {"blocks_per_segment": tiles[0], "likelihood_windows_per_block": tiles[1]},
            "lanes": lanes,
            "autotune_lanes_msamples_ch_per_s": {str(k): round(v, 2) for k, v in lrates.items()},
            "autotune_tiles_msamples_ch_per_s": {f"{k[0]}x{k[1]}": round(v, 2) for k, v in trates.items()},
            "groups": stf["groups"],
            "streams_value": round(work / sts["seconds"] / 1e6, 3),
            "streams_workers": workers,
            "per_segment_value": round(work / st1["seconds"] / 1e6, 3),
            "fused_over_per_segment": round(st1["seconds"] / stf["seconds"], 3),
            "fused_over_streams": round(sts["seconds"] / stf["seconds"], 3),
            "one_lane_host_planned_value": round(work / sth["seconds"] / 1e6, 3),
            "one_lane_graph_value": None,  # hipGraph replays switched off in this build
            "identical_spike_set": bool(identical),
            "spikes": n_spk,
        },
        # job level: the first pass's dense FP32 likelihood work over the whole fused run (every
        # kernel, every pass) -- a lower bound on the likelihood kernel's own rate
        "roofline": {"bound": "fp32", "achieved": round(achieved, 3), "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(achieved / 157.3, 4), "traffic": None, "scope": "job"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_refine(args):
    """Differentiable variant: 32 ch, 30 kHz, 120 s, 60 units, 64-sample templates perturbed from the
    truth by 10 % Gaussian noise; spike times from one mixed-precision pursuit with the perturbed
    templates; 50 Adam steps on the residual energy; gradient throughput of forward + backward."""
    import torch
    from oracle import refine_fp64
    from spikesorting_fullpursuit_amd import refine, synthetic
    from spikesorting_fullpursuit_amd.overlap import binary_pursuit_parallel as bpp

    t0 = time.time()
    d = synthetic.generate_on_device("probe32", duration_s=args.duration)
    V = d["voltage"]
    C, N = V.shape
    w = d["workload"]
    truth = d["templates"].astype(np.float32)
    U, _, L = truth.shape
    rng = np.random.default_rng(w.seed + 1)
    Tp = (truth * (1.0 + 0.1 * rng.standard_normal(truth.shape))).astype(np.float32)
    SS, SS_by_chan, thr = synthetic.separability_inputs(Tp.astype(np.float64), w.fs)
    m = {"templates": Tp.reshape(U, C * L).astype(np.float64), "template_SS": SS,
         "template_SS_by_chan": SS_by_chan, "neuron_lower_thresholds": thr}
    ev, lab, _, _ = bpp.binary_pursuit(V, np.float32, d["sort_info"], m, n_max_shift_inds=d["n_max_shift_inds"])
    onsets = np.asarray(ev, np.int64) - d["clip_init_samples"]
    lab = np.asarray(lab, np.int64)
    log(f"generated probe32 and pursued with perturbed templates: {len(onsets)} spikes "
        f"({len(d['truth_onsets'])} planted) in {time.time() - t0:.1f}s")

    refine.refine_templates(V, Tp, onsets, lab, steps=2, truth=truth)  # warm-up
    torch.cuda.synchronize()
    t1 = time.time()
    Tref, amp, hist = refine.refine_templates(V, Tp, onsets, lab, steps=50, truth=truth)
    torch.cuda.synchronize()
    adam_s = time.time() - t1
    log(f"50 Adam steps in {adam_s:.2f}s: template error {hist[0]['template_error']:.4f} -> "
        f"{hist[-1]['template_error']:.4f}, energy {hist[0]['energy']:.6g} -> {hist[-1]['energy']:.6g}")

    Td = torch.tensor(Tp, device=V.device)
    table = refine.SpikeTable(onsets, lab, U, N, device=V.device)
    a = torch.ones(table.n, device=V.device, dtype=torch.float32)
    for _ in range(max(1, args.warmup)):
        refine.time_gradient(V, Td, a, table, reps=1)
    sec = refine.time_gradient(V, Td, a, table, reps=max(1, args.steps) * 10)
    # per-kernel times with events on the current stream (the stream the kernels run on)
    ev0, ev1, ev2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    reps = 20
    ev0.record()
    for _ in range(reps):
        R, E = refine.forward_residual(V, Td, a, table)
    ev1.record()
    for _ in range(reps):
        refine.backward_grads(R, Td, a, table)
    ev2.record()
    torch.cuda.synchronize()
    fwd_ms = ev0.elapsed_time(ev1) / reps
    bwd_ms = ev1.elapsed_time(ev2) / reps
    fwd_bytes = 2 * 4 * C * N                     # read V, write R
    bwd_bytes = 2 * 4 * table.n * C * L           # R at every spike window, once per gradient
    fwd_gbs = fwd_bytes / (fwd_ms * 1e-3) / 1e9

    # gradient check against the FP64 oracle on the whole recording (also the CPU baseline)
    ns = N
    sel = (onsets < ns)
    Vs = V[:, :ns].contiguous()
    tab_s = refine.SpikeTable(onsets[sel], lab[sel], U, ns, device=V.device)
    a_s = torch.ones(tab_s.n, device=V.device, dtype=torch.float32)
    Rs, Es = refine.forward_residual(Vs, Td, a_s, tab_s)
    gT, ga = refine.backward_grads(Rs, Td, a_s, tab_s)
    Vh = Vs.cpu().numpy()
    t2 = time.perf_counter()
    E64, gT64, ga64, _ = refine_fp64.gradients(Vh, Tp, onsets[sel], lab[sel], np.ones(int(sel.sum())))
    tc = time.perf_counter() - t2
    gT_err = float(np.max(np.abs(gT.cpu().numpy() - gT64)) / max(1e-30, np.max(np.abs(gT64))))
    ga_err = float(np.max(np.abs(tab_s.from_sorted(ga).cpu().numpy() - ga64)) / max(1e-30, np.max(np.abs(ga64))))
    E_err = abs(float(Es) - E64) / E64

    value = N * C / sec / 1e6
    line = {
        "metric": "template-refinement gradient throughput Msamples·ch/s at 32ch (residual energy forward + "
                  "backward w.r.t. templates and amplitudes)",
        "value": round(value, 3),
        "unit": "Msamples·ch/s",
        "n_gpus": 1,
        "steps": max(1, args.steps) * 10,
        "warmup": max(1, args.warmup),
        "ms_per_step": round(1000 * sec, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (energy sums in f64)",
        "data": "synthetic",
        "config": {
            "workload": f"probe32: {C} ch, {w.fs / 1000:g} kHz, {N / w.fs:g} s, {U} units, L={L}, templates = "
                        f"truth x (1 + 0.1 N(0,1)), spike times from one mixed-precision pursuit",
            "spikes": int(table.n),
            "spikes_per_s": round(table.n / sec, 1),
            "adam_steps": 50,
            "adam_s_for_50_steps": round(adam_s, 3),
            "template_error_before": round(hist[0]["template_error"], 6),
            "template_error_after": round(hist[-1]["template_error"], 6),
            "energy_before": hist[0]["energy"],
            "energy_after": hist[-1]["energy"],
            "vs_fp64": {"energy_rel": E_err, "grad_templates_rel_max": gT_err,
                                 "grad_amplitudes_rel_max": ga_err},
        },
        "roofline": {
            "kernel": "k_residual_energy (forward: residual and FP64 energy partials)",
            "bound": "hbm", "achieved": round(fwd_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(fwd_gbs / HBM_PEAK_GBS, 5), "traffic": None,
            "avg_launch_ms": round(fwd_ms, 4), "bytes_per_launch": fwd_bytes,
            "backward_ms": round(bwd_ms, 4),
            "backward_algorithmic_gbs": round(bwd_bytes / (bwd_ms * 1e-3) / 1e9, 2),
        },
        "cpu_baseline": {"value": round(ns * C / tc / 1e6, 6), "unit": "Msamples·ch/s", "cores": 1, "kind": "port",
                         "sample": f"FP64 NumPy residual energy and gradients, whole recording ({ns / w.fs:g} s, "
                                   f"{int(sel.sum())} spikes), {tc:.2f} s"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
