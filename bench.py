#!/usr/bin/env python3
"""Benchmark of the MI355X dedispersion hot path.

Workload (default): the 8-range AA plan on 4096 channels x 1,048,576 u8
samples (1760 -> 950 MHz, 64 us; ranges 0-150/0.1/1 ... 2000-3000/2.0/8;
6080 DM trials after AA's rounding), seeded N(128, 16^2) noise plus 16
dispersed pulses, input resident in HBM before timing starts.

One step = one pass of the hot path over the chunk: corner turn, x2 binning
cascade and dedispersion of all 8 ranges (aa_hip_ddtr_execute_chunk).
value = DM x samples per second = sum_r ndms_r * t_processed_r / step time.

Also reported:
  roofline     the dedisperse kernel alone, HIP events on its stream:
               algorithmic channel-adds (nchans * ndms_r * t_processed_r per
               launch) / kernel time, against MI355X's FP32 VALU add roofline
               (78.64 T adds/s = 256 CU x 128 packed adds/clk x 2.4 GHz);
  cpu_baseline the CPU oracle (OpenMP C) on the host cores, timed on the
               whole hot path over 1/24 of the chunk's output samples (every
               DM row of every range; --cpu-slice 1 runs the whole plane, the
               default on the single-range plan),
               best of 3 runs after a warm-up run; the slice is also
               compared bit for bit with the GPU output.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1|2]
                       [--kernel auto|direct|lds] [--no-cpu-baseline] [--cpu-slice N]
                       [--cpu-repeats R]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "astro-accelerate_amd"))

CONFIGS = {
    1: dict(name="plain_single_range_512ch_131072 (config 1)", nchans=512, nsamp=131072, fch1=1550.0,
            foff=-0.5859, tsamp=64e-6, ranges=[(0, 500, 0.5, 1, 1)], pulses="config1"),
    2: dict(name="aa_8_range_4096ch_1048576 (config 2)", nchans=4096, nsamp=1048576, fch1=1760.0,
            foff=-810.0 / 4096, tsamp=64e-6,
            ranges=[(0, 150, 0.1, 1, 1), (150, 300, 0.2, 1, 1), (300, 500, 0.25, 1, 1), (500, 900, 0.4, 2, 2),
                    (900, 1200, 0.6, 4, 4), (1200, 1500, 0.8, 4, 4), (1500, 2000, 1.0, 4, 4),
                    (2000, 3000, 2.0, 8, 8)], pulses="config2"),
}
ADD_ROOFLINE_T = 78.64   # T adds/s, packed FP32 VALU adds
# HBM bytes moved by the dedisperse launches of one step, from rocprofv3
# FETCH_SIZE (x2.000, calibrated on a 1 GiB device copy) + WRITE_SIZE (x1.000)
# summed over the launches of chunk shape 4 (pair rows on 2016-sample
# tiles, the automatic choice): the 8 launches of the 8-range plan
# (profiles/dedisperse_wide_8range_4096ch_1048576_hbm_traffic.txt) and the
# one of the single-range plan (519.3 MB read + 505.5 MB written,
# profiles/dedisperse_wide_1range_512ch_131072_kernel_trace.txt).
# Counters cannot be read live, so the figure is tied to the kernel it was
# taken on.
MEASURED_TRAFFIC = {2: {"bytes": 173.5e9, "kernel": "pair",
                        "profile": "profiles/dedisperse_wide_8range_4096ch_1048576_hbm_traffic.txt"},
                    1: {"bytes": 1.0248e9, "kernel": "pair",
                        "profile": "profiles/dedisperse_wide_1range_512ch_131072_kernel_trace.txt"}}
# The kernels add 8-bit samples with integer SWAR (packed 16-bit fields,
# v_perm + v_add3_u32), not FP32 adds, so the FP32 figure above is a nominal
# ratio.  Their own ceiling is the VALU issue rate: 256 CU x 4 SIMD x 16
# lanes x 2.4 GHz = 39.32 T lane-ops/s, divided by the lane-ops the kernel
# issues per channel-add (SQ_INSTS_VALU x 64 / channel-adds, rocprofv3).
VALU_LANE_OPS_T = 39.32
MEASURED_ISSUE = {2: {"lane_ops_per_add": 0.401, "kernel": "pair",
                      "profile": "profiles/dedisperse_wide_8range_4096ch_1048576_sq_counters.txt"}}
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 5; 200 on the single-range plan, whose step is ~1.3 ms)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 2; 200 on the single-range plan, so the clocks have "
                         "risen before the timed region)")
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--tune-cache", default=None, help="config 4: JSON file of autotuned chunk shapes")
    ap.add_argument("--beam-set", default="standard", choices=["standard", "small"],
                    help="config 4: the 48 beams (standard), or 256 small beams (64-256 channels, "
                         "2^14-2^15 samples, DM max 100/200 in 1-2 ranges) whose launches do not fill the GPU")
    ap.add_argument("--kernel", default="auto", choices=["auto", "direct", "lds", "c32", "c16", "pair", "quad"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-slice", type=int, default=None,
                    help="CPU baseline: time the oracle on 1/N of the chunk's output samples (all DM rows); "
                         "1 runs the whole plane (default: 24 on the 8-range plan, 1 on the single range)")
    ap.add_argument("--cpu-repeats", type=int, default=3,
                    help="CPU baseline: timed runs after one warm-up run; the best is reported")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--e2e", action="store_true",
                    help="end-to-end run instead: SIGPROC file -> reader -> FFI pipeline -> host DM-time "
                         "planes; prints the real-time factor (not the contract line)")
    ap.add_argument("--e2e-chunks", type=int, default=1,
                    help="--e2e: cap device memory so the observation runs in about this many time chunks")
    ap.add_argument("--zero-dm", action="store_true",
                    help="--e2e: with the pipeline's zero_dm option; alone: time the zero-DM kernel on the "
                         "config's [t][c] input (HBM roofline line, not the contract line)")
    ap.add_argument("--normalization", action="store_true",
                    help="--e2e: with the pipeline's input and output DDTR normalisation options; alone: time "
                         "the input normalisation on the config's [t][c] input and the output normalisation on "
                         "its largest range's plane (HBM roofline lines, not the contract line)")
    ap.add_argument("--spd", action="store_true",
                    help="time the first single-pulse pass (boxcar MSD + SNR, widths 1-32) on a plane of range 0's "
                         "size (not the contract line)")
    ap.add_argument("--zero-dm-outliers", action="store_true",
                    help="--e2e: with the pipeline's zero_dm_with_outliers option; alone: time zero-DM with "
                         "outlier rejection on the config's [t][c] input (not the contract line)")
    args = ap.parse_args()
    short = args.config == 1 and not (args.e2e or args.zero_dm or args.zero_dm_outliers or args.normalization)
    if args.steps is None:
        args.steps = 200 if short else 5
    if args.warmup is None:
        args.warmup = 200 if short else 2
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); "
              "dedispersion would shard by DM range or beam with no exchange", file=sys.stderr)
        sys.exit(2)

    import torch
    import py_astro_accelerate as aa

    if args.e2e:
        return bench_e2e(args, CONFIGS[args.config if args.config in CONFIGS else 2])
    if args.zero_dm:
        return bench_zero_dm(args, CONFIGS[args.config if args.config in CONFIGS else 2])
    if args.zero_dm_outliers:
        return bench_zero_dm(args, CONFIGS[args.config if args.config in CONFIGS else 2], outliers=True)
    if args.normalization:
        return bench_normalization(args, CONFIGS[args.config if args.config in CONFIGS else 2])
    if args.spd:
        return bench_spd(args, CONFIGS[2])
    if args.config == 3:
        return bench_frac(args)
    if args.config == 4:
        return bench_multibeam(args)
    if args.config == 5:
        return bench_refine(args)
    cfg = CONFIGS[args.config]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lib = aa.load_library()
    kernel = {"auto": aa.KERNEL_AUTO, "direct": aa.KERNEL_DIRECT, "lds": aa.KERNEL_LDS, "c32": aa.KERNEL_LDS_C32,
              "c16": aa.KERNEL_LDS_C16, "pair": aa.KERNEL_LDS_PAIR, "quad": aa.KERNEL_LDS_QUAD}[args.kernel]

    if args.copy_calibration:
        src = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 1 GiB read + 1 GiB written")
        del src, dst

    free_b, _ = torch.cuda.mem_get_info()
    plan = aa.make_plan(cfg["ranges"])
    md = aa.make_metadata(cfg["nchans"], cfg["nsamp"], cfg["fch1"], cfg["foff"], cfg["tsamp"])
    strat = aa.aa_py_ddtr_strategy(plan, md, free_b)
    assert strat.ready, "strategy not ready"
    assert strat.num_tchunks == 1, "benchmark workload is a single chunk"
    nr = strat.nranges
    ndms = [strat.ndms(r) for r in range(nr)]
    tps = [strat.t_processed(r, 0) for r in range(nr)]
    nchans, nsamp = cfg["nchans"], cfg["nsamp"]
    dm_x_samples = sum(n * t for n, t in zip(ndms, tps))
    adds = nchans * dm_x_samples
    log(f"[bench] {cfg['name']}: ranges={nr} ndms={ndms} t_processed={tps} maxshift={strat.maxshift}")

    # ---- synthetic input, resident in HBM
    stream = torch.cuda.Stream()
    dmshifts = strat.dmshifts()
    d_ds = torch.from_numpy(dmshifts).to(dev)
    d_in = torch.empty((nsamp, nchans), dtype=torch.uint8, device=dev)
    rng = np.random.default_rng(SEED)
    with torch.cuda.stream(stream):
        aa.synth_noise_u8(SEED, d_in, 128.0, 16.0, stream=stream)
        if cfg["pulses"] == "config1":
            pulses = [(250.0, 60000, 8)]
        else:
            top = strat.dm(nr - 1)[1]
            pulses = [(float(rng.uniform(0, top)), int(rng.integers(0, nsamp - strat.maxshift - 64)),
                       int(rng.integers(1, 65))) for _ in range(16)]
        for dm, t0, w in pulses:
            aa.inject_pulse_u8(d_in, nsamp, nchans, d_ds, dm, cfg["tsamp"], t0, w, 32, stream=stream)
    outs = [torch.empty((ndms[r], tps[r]), dtype=torch.float32, device=dev) for r in range(nr)]
    ddtr = aa.DdtrPlan.from_strategy(strat)
    ddtr.set_kernel(kernel)
    stream.synchronize()
    log(f"[bench] device bytes: input {d_in.numel() / 1e9:.2f} GB, outputs "
        f"{sum(o.numel() for o in outs) * 4 / 1e9:.2f} GB, plan {ddtr.device_bytes() / 1e9:.2f} GB")

    def step():
        ddtr.execute_chunk(d_in, tps, outs, stream=stream)

    with torch.cuda.stream(stream):
        for w in range(args.warmup):
            step()
            stream.synchronize()
            log(f"[bench] warmup {w + 1}/{args.warmup} done")
        # timed region: barrier (single rank) + synchronize on both sides
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_wall0 = time.perf_counter()
        e0.record(stream)
        for k in range(args.steps):
            step()
        e1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        t_wall = time.perf_counter() - t_wall0
    ms_step = e0.elapsed_time(e1) / args.steps
    log(f"[bench] {args.steps} steps: {ms_step:.3f} ms/step (wall {t_wall * 1e3 / args.steps:.3f} ms/step)")

    # ---- dominant kernel alone: every level is resident after the steps, so
    # run_range launches only the dedisperse kernel.  The launches run back to
    # back (one synchronisation after all of them), as inside the steps.
    reps = 3
    kt = [0.0] * nr
    marks = []
    with torch.cuda.stream(stream):
        for _ in range(reps):
            for r in range(nr):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                ddtr.run_range(r, tps[r], outs[r], stream=stream)
                b.record(stream)
                marks.append((r, a, b))
        stream.synchronize()
    for r, a, b in marks:
        kt[r] += a.elapsed_time(b) / reps
    k_total_ms = sum(kt)
    achieved = adds / (k_total_ms * 1e-3) / 1e12
    # algorithmic HBM bytes of the 8 launches: each reads its level's rows once
    # (t_processed + maxshift/inBin u8 samples per channel) and writes its plane once
    hbm_bytes = sum(nchans * (tps[r] + strat.maxshift // strat.dm(r)[3]) for r in range(nr)) + 4 * dm_x_samples
    log(f"[bench] dedisperse kernel per range (ms): {[round(x, 3) for x in kt]} total {k_total_ms:.3f} ms -> "
        f"{achieved:.2f} T adds/s ({achieved / ADD_ROOFLINE_T * 100:.1f}% of {ADD_ROOFLINE_T})")

    value = dm_x_samples / (ms_step * 1e-3)
    meas = MEASURED_TRAFFIC.get(args.config)
    traffic = meas["bytes"] if meas and args.kernel in ("auto", meas["kernel"]) else None
    result = {
        "metric": "DM×sample/s bit-exact, % FP32 add roofline; batched vs per-beam speedup",
        "value": value,
        "unit": "DM-trial samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic: seeded N(128,16^2) u8 noise + dispersed pulses, resident in HBM",
        "config": {"workload": cfg["name"], "nchans": nchans, "nsamp": nsamp, "ranges": len(cfg["ranges"]),
                   "dm_trials": sum(ndms), "dm_x_samples": dm_x_samples, "channel_adds": adds,
                   "kernel": args.kernel, "parallelism": "single GPU"},
        "channel_adds_per_s": adds / (ms_step * 1e-3),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": ADD_ROOFLINE_T, "unit": "Tadd/s",
                     "frac": achieved / ADD_ROOFLINE_T, "traffic": traffic,
                     "kernel": "dedisperse (all ranges)", "kernel_ms": k_total_ms,
                     "hbm_algorithmic_GBs": hbm_bytes / (k_total_ms * 1e-3) / 1e9,
                     "hbm_frac": hbm_bytes / (k_total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "hbm_algorithmic_bytes": hbm_bytes,
                     "traffic_source": meas["profile"] if traffic else None,
                     "peak_note": "nominal: FP32 packed-add rate; the kernels add integers (SWAR)"},
    }
    iss = MEASURED_ISSUE.get(args.config)
    if iss and args.kernel in ("auto", iss["kernel"]):
        bound = VALU_LANE_OPS_T / iss["lane_ops_per_add"]
        result["roofline"]["valu_issue_bound"] = {
            "lane_ops_per_add": iss["lane_ops_per_add"], "peak_lane_ops_T": VALU_LANE_OPS_T,
            "bound_Tadds": bound, "frac": achieved / bound, "source": iss["profile"]}

    if not args.no_cpu_baseline:
        result["cpu_baseline"], result["bit_exact_cpu_slice"] = cpu_baseline(
            args, cfg, strat, d_in, outs, ndms, tps, dm_x_samples)
    print(json.dumps(result), flush=True)


FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 (packed FMA); non-packed issue is half of it


def bench_e2e(args, cfg):
    """SIGPROC 8-bit file -> aa_sigproc_input.read_signal_u8 (pinned) -> FFI
    pipeline (H2D of each time chunk, corner turn, binning, dedispersion,
    D2H into pinned host planes).  The file is written first (untimed) from
    the same synthetic noise as the device benchmark; the timed part is the
    read and the pipeline, from bind to the last range.  Real-time factor =
    observation length / that time."""
    import tempfile
    import time
    import torch
    import py_astro_accelerate as aa

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nchans, nsamp = cfg["nchans"], cfg["nsamp"]
    d_in = torch.empty((nsamp, nchans), dtype=torch.uint8, device=dev)
    aa.synth_noise_u8(SEED, d_in, 128.0, 16.0)
    torch.cuda.synchronize()
    host = d_in.cpu().numpy()
    del d_in
    tmpdir = os.environ.get("TMPDIR", tempfile.gettempdir())
    path = os.path.join(tmpdir, f"aa_e2e_{os.getpid()}.fil")
    aa.write_sigproc(path, 8, nchans, host.tobytes(), fch1=cfg["fch1"], foff=cfg["foff"], tsamp=cfg["tsamp"])
    del host
    log(f"[e2e] wrote {path}: {nchans} ch x {nsamp} samples u8 ({os.path.getsize(path) / 1e9:.2f} GB)")
    plan = aa.make_plan(cfg["ranges"])
    cap = 0
    if args.e2e_chunks > 1:   # largest cap (1 GB steps down) that gives that many chunks
        md0 = aa.make_metadata(nchans, nsamp, cfg["fch1"], cfg["foff"], cfg["tsamp"])
        for gb in range(200, 1, -1):
            st0 = aa.aa_py_ddtr_strategy(plan, md0, gb << 30)
            if st0.ready and st0.num_tchunks >= args.e2e_chunks:
                cap = gb << 30
                break
    try:
        t0 = time.perf_counter()
        reader = aa.aa_sigproc_input(path)
        md = reader.metadata()
        buf = reader.read_signal_u8(pinned=True)
        t1 = time.perf_counter()
        comps = aa.aa_py_pipeline_components()
        comps.dedispersion = True
        opts = aa.aa_py_pipeline_component_options()
        opts.copy_ddtr_data_to_host = True
        opts.zero_dm = bool(args.zero_dm)
        opts.zero_dm_with_outliers = bool(args.zero_dm_outliers)
        opts.input_DDTR_normalization = bool(args.normalization)
        opts.output_DDTR_normalization = bool(args.normalization)
        pipe = aa.aa_py_pipeline(comps, opts, md, buf, 0, max_device_bytes=cap)
        assert pipe.bind_ddtr_plan(plan)
        steps = 0
        while pipe.run():
            steps += 1
        ok = pipe.status_code() == 0
        t2 = time.perf_counter()
        strat = aa.aa_py_ddtr_strategy(plan, md, cap if cap else torch.cuda.mem_get_info()[0])
        nchunks = strat.num_tchunks
        pipe.cleanUp()
        reader.close()
    finally:
        os.remove(path)
    obs_s = nsamp * cfg["tsamp"]
    total = t2 - t0
    log(f"[e2e] read {t1 - t0:.3f} s, pipeline {t2 - t1:.3f} s ({nchunks} chunks, {steps} run() calls), "
        f"observation {obs_s:.2f} s -> real-time factor {obs_s / total:.1f}")
    print(json.dumps({
        "metric": "end-to-end real-time factor (SIGPROC file -> host DM-time planes)",
        "value": obs_s / total, "unit": "x real time", "higher_is_better": True, "ok": ok,
        "read_s": t1 - t0, "pipeline_s": t2 - t1, "observation_s": obs_s, "time_chunks": nchunks,
        "config": {"workload": cfg["name"], "nchans": nchans, "nsamp": nsamp, "input": "SIGPROC 8-bit file",
                   "host_buffers": "pinned", "device_memory_cap_bytes": cap, "zero_dm": bool(args.zero_dm),
                   "zero_dm_with_outliers": bool(args.zero_dm_outliers),
                   "input_output_normalization": bool(args.normalization)},
    }), flush=True)


def bench_zero_dm(args, cfg, outliers=False):
    """The zero-DM kernel on the config's [t][c] u8 input (4096 x 1,048,576
    for config 2), in place, timed with HIP events on its stream.  HBM-bound:
    algorithmic bytes = the input read once and written once.  With
    `outliers`, zero-DM with outlier rejection (float64 statistics, up to six
    passes over each time sample: compute-bound, reported as samples/s)."""
    import torch
    import py_astro_accelerate as aa
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nchans, nsamp = cfg["nchans"], cfg["nsamp"]
    d = torch.empty((nsamp, nchans), dtype=torch.uint8, device=dev)
    aa.synth_noise_u8(SEED, d, 128.0, 16.0)
    bp = torch.full((nchans,), 127.5, dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream()
    fn = aa.zero_dm_outliers if outliers else aa.zero_dm
    for _ in range(args.warmup):
        fn(d, bp, nbits=8, stream=stream)
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        fn(d, bp, nbits=8, stream=stream)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    if outliers:
        sps = nchans * nsamp / (ms * 1e-3)
        log(f"[zero-dm-outliers] {nchans} ch x {nsamp} samples u8: {ms:.3f} ms per pass, {sps / 1e9:.1f} G "
            f"samples/s")
        print(json.dumps({"metric": "zero-DM with outliers throughput", "value": sps / 1e9,
                          "unit": "G samples/s", "ms_per_pass": ms,
                          "config": {"workload": cfg["name"], "nchans": nchans, "nsamp": nsamp}}), flush=True)
        return
    nbytes = 2.0 * nchans * nsamp
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"[zero-dm] {nchans} ch x {nsamp} samples u8: {ms:.3f} ms per pass, {gbs:.0f} GB/s "
        f"({gbs / HBM_PEAK_GBS * 100:.1f}% of {HBM_PEAK_GBS:.0f})")
    print(json.dumps({"metric": "zero-DM filter throughput", "value": gbs, "unit": "GB/s (input read + written)",
                      "ms_per_pass": ms, "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                                                      "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS},
                      "config": {"workload": cfg["name"], "nchans": nchans, "nsamp": nsamp}}), flush=True)


def bench_spd(args, cfg):
    """The first single-pulse pass (aa_hip_spd_first_pass, boxcar widths
    1..32) on a [1536][890624] float plane: the size of range 0's plane of
    the 8-range plan, seeded N(0, 1) noise.  Timed with HIP events on its
    stream; MSD only (one read of the plane) and MSD + SNR (the plane read
    twice, SNR written as float and the width as u16).  Work per start sample
    and width: one LDS read and about four VALU operations, so the pass is
    VALU-bound; reported beside the algorithmic HBM bytes."""
    import torch
    import py_astro_accelerate as aa
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ndms, nsamp, W = 1536, 890624, aa.SPD_MAX_WIDTH
    g = torch.Generator(device=dev).manual_seed(SEED)
    plane = torch.randn((ndms, nsamp), generator=g, device=dev, dtype=torch.float32)
    stream = torch.cuda.Stream()
    work = aa.spd_workspace(dev)
    msd = torch.empty((W, 2), dtype=torch.float32, device=dev)
    snr = torch.empty((ndms, nsamp), dtype=torch.float32, device=dev)
    width = torch.empty((ndms, nsamp), dtype=torch.int16, device=dev)
    torch.cuda.synchronize()

    def timed(with_snr):
        def once():
            aa.spd_first_pass_raw(plane.data_ptr(), nsamp, ndms, nsamp, W, msd, snr if with_snr else None,
                                  width if with_snr else None, work, stream.cuda_stream)
        for _ in range(max(1, args.warmup)):
            once()
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            once()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / args.steps

    ms_msd = timed(False)
    ms_all = timed(True)
    n = ndms * nsamp
    bytes_msd, bytes_all = 4.0 * n, 4.0 * n * 2 + 4.0 * n + 2.0 * n
    boxcars = float(n) * W                     # (start sample, width) pairs per pass
    log(f"[spd] {ndms} x {nsamp} plane, widths 1..{W}: MSD {ms_msd:.3f} ms "
        f"({bytes_msd / (ms_msd * 1e-3) / 1e9:.0f} GB/s), MSD + SNR {ms_all:.3f} ms "
        f"({bytes_all / (ms_all * 1e-3) / 1e9:.0f} GB/s); {2 * boxcars / (ms_all * 1e-3) / 1e12:.2f} T boxcar "
        f"terms/s over both passes")
    print(json.dumps({"metric": "single-pulse first pass throughput", "value": n / (ms_all * 1e-3) / 1e9,
                      "unit": "G start samples/s (32 widths, MSD + SNR)", "ms_msd": ms_msd, "ms_msd_snr": ms_all,
                      "hbm_bytes_msd": bytes_msd, "hbm_bytes_msd_snr": bytes_all,
                      "roofline": {"bound": "valu", "hbm_achieved_GBs": bytes_all / (ms_all * 1e-3) / 1e9,
                                   "hbm_frac": bytes_all / (ms_all * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "boxcar_terms_per_s": 2 * boxcars / (ms_all * 1e-3)},
                      "config": {"workload": "range 0 plane of the 8-range plan (1536 x 890624 f32), widths 1-32",
                                 "ndms": ndms, "nsamp": nsamp, "max_width": W}}), flush=True)


def bench_normalization(args, cfg):
    """DDTR input normalisation on the config's [t][c] u8 input (per channel
    over all samples) and output normalisation on a [1536][890624] float
    plane (range 0 of the 8-range plan; per time sample over DMs), in place,
    timed with HIP events on their stream.  HBM-bound: algorithmic bytes =
    the data read once and written once (the kernels read twice: sums, then
    the shift).  The input call includes its scratch allocation and its
    return after the stream has finished."""
    import torch
    import py_astro_accelerate as aa
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.Stream()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / args.steps

    nchans, nsamp = cfg["nchans"], cfg["nsamp"]
    d = torch.empty((nsamp, nchans), dtype=torch.uint8, device=dev)
    aa.synth_noise_u8(SEED, d, 128.0, 16.0)
    ms_in = timed(lambda: aa.input_ddtr_normalization(d, nbits=8, stream=stream))
    del d
    ndms, tp = 1536, 890624
    plane = torch.empty((ndms, tp), dtype=torch.float32, device=dev)
    with torch.cuda.stream(stream):
        plane.normal_(127.0, 10.0)
    ms_out = timed(lambda: aa.output_ddtr_normalization(plane, stream=stream))
    for name, ms, nbytes, shape in (("input", ms_in, 2.0 * nchans * nsamp, f"{nsamp} x {nchans} u8"),
                                    ("output", ms_out, 2.0 * 4 * ndms * tp, f"{ndms} x {tp} f32")):
        gbs = nbytes / (ms * 1e-3) / 1e9
        log(f"[normalization] {name} {shape}: {ms:.3f} ms per pass, {gbs:.0f} GB/s "
            f"({gbs / HBM_PEAK_GBS * 100:.1f}% of {HBM_PEAK_GBS:.0f})")
        print(json.dumps({"metric": f"DDTR {name} normalisation throughput", "value": gbs,
                          "unit": "GB/s (data read + written once)", "ms_per_pass": ms,
                          "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": gbs / HBM_PEAK_GBS},
                          "config": {"workload": cfg["name"], "shape": shape}}), flush=True)


def bench_frac(args):
    """Differentiable fractional-delay dedispersion: 256 candidate windows of
    4096 channels x 1024 float32 samples, 64 DM offsets in [-2, 2]; one step
    = forward (profiles), adjoint of the squared-profile loss w.r.t. the
    input and w.r.t. each candidate's DM (aa_hip_frac_forward/_adjoint)."""
    import torch
    import frac_ddtr

    B, C, T, K = 256, 4096, 1024, 64
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ds = frac_ddtr.delay_per_dm(C, 1760.0, -810.0 / C, 64e-6)
    trials = np.linspace(-2.0, 2.0, K).astype(np.float32)
    plan = frac_ddtr.FracPlan(ds, trials, T, max_batch=B)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        x, r, _ = frac_ddtr.synth_candidates(B, C, T, ds, seed=SEED, device=dev)
        dm = torch.zeros(B, device=dev)
        y = torch.empty((B, K, T), device=dev)
    stream.synchronize()
    log(f"[bench] frac: {B} windows x {C} ch x {T} samples, {K} DM offsets; grouping {plan.grouping()}")
    lerps = B * K * C * T
    flops = {"forward": 4.0 * lerps, "adjoint_x": 4.0 * lerps, "adjoint_dm": 2.125 * lerps}

    def step(times=None):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if times is not None else None
        if ev: ev[0].record(stream)
        plan.forward(x, dm, out=y)
        gy = y * 2.0
        if ev: ev[1].record(stream)
        gx, _ = plan.adjoint(x, gy, dm, want_x=True, want_dm=False)
        if ev: ev[2].record(stream)
        _, gdm = plan.adjoint(x, gy, dm, want_x=False, want_dm=True)
        if ev:
            ev[3].record(stream)
            ev[3].synchronize()
            times["forward"] += ev[0].elapsed_time(ev[1])
            times["adjoint_x"] += ev[1].elapsed_time(ev[2])
            times["adjoint_dm"] += ev[2].elapsed_time(ev[3])
        return gdm

    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms_step = e0.elapsed_time(e1) / args.steps
        times = {k: 0.0 for k in flops}
        for _ in range(3):
            step(times)
    times = {k: v / 3 for k, v in times.items()}
    per = {k: {"ms": times[k], "tflops": flops[k] / (times[k] * 1e-3) / 1e12} for k in flops}
    dom = max(per, key=lambda k: per[k]["ms"])
    log(f"[bench] frac step {ms_step:.3f} ms; per kernel: " +
        ", ".join(f"{k} {v['ms']:.3f} ms {v['tflops']:.1f} TFLOP/s" for k, v in per.items()))
    result = {
        "metric": "DM×sample/s bit-exact, % FP32 add roofline; batched vs per-beam speedup",
        "value": B * K * T / (ms_step * 1e-3),
        "unit": "DM-trial samples/s (forward + input and DM adjoints)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: N(0,1) windows + Gaussian pulse (sigma 4, amplitude 0.5) at residual DM",
        "config": {"workload": "frac_ddtr 256 windows x 4096 ch x 1024 samples x 64 DM offsets (config 3)",
                   "channel_lerps": lerps, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": per[dom]["tflops"], "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": per[dom]["tflops"] / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": dom, "kernels": per},
    }
    print(json.dumps(result), flush=True)


def bench_multibeam(args):
    """48 beams (1024/2048/4096 channels, 2^16..2^19 samples, DM max 500/1000/
    3000 in 1-4 AA-style ranges, seeded u8 noise + 2 pulses each): all beams
    batched and replayed from a hipGraph (DdtrBatch: one launch per kernel
    instantiation over all beams' ranges), against the same beams launched
    one after another (execute_chunk per beam) on the same tuned shapes
    (speedup_vs_per_beam_same_kernels, beside the 1.5x target), and on the
    byte-row shapes.  Chunk shapes are autotuned per shape class first.
    --beam-set small: 256 beams of 64-256 channels whose launches hold a few
    dozen tiles each, fewer than the GPU's 256 CUs."""
    import torch
    import py_astro_accelerate as aa
    import multibeam as mb

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    small = args.beam_set == "small"
    specs = mb.small_multibeam_specs(SEED) if small else mb.multibeam_specs(48, SEED)
    stream = torch.cuda.Stream()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        beams = mb.build_beams(specs, dev, stream=stream)
    stream.synchronize()
    dmxs = sum(b.dm_x_samples for b in beams)
    adds = sum(b.channel_adds for b in beams)
    nranges = sum(len(b.tps) for b in beams)
    log(f"[bench] multibeam: {len(beams)} beams, {nranges} ranges, {dmxs:.4e} DMxsamples, {adds:.4e} adds; "
        f"set-up {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    cache = mb.autotune(beams, cache_path=args.tune_cache, reps=1, stream=stream)
    stream.synchronize()
    shapes = {}
    for v in cache.values():
        shapes[v["variant"]] = shapes.get(v["variant"], 0) + 1
    log(f"[bench] autotune: {len(cache)} shape classes in {time.perf_counter() - t0:.1f} s, shapes chosen {shapes}")
    batch = aa.DdtrBatch([b.plan for b in beams])
    ins, tps, outs = [b.d_input for b in beams], [b.tps for b in beams], [b.outs for b in beams]

    def per_beam():
        mb.run_per_beam(beams, stream=stream)

    def batched():
        batch.execute(ins, tps, outs, use_graph=True, stream=stream)

    def timed(fn, steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    with torch.cuda.stream(stream):
        for _ in range(max(1, args.warmup)):
            per_beam()
        stream.synchronize()
        want = mb.output_digest(beams)
        for _ in range(max(1, args.warmup)):
            batched()                      # first call: eager pass + graph capture
        stream.synchronize()
        exact = mb.output_digest(beams) == want
        log(f"[bench] batched output bit-identical to per-beam launches: {exact}; stats {batch.stats()}")
        ms_per_beam = timed(per_beam, args.steps)
        log(f"[bench] per-beam launches: {ms_per_beam:.2f} ms/step")
        ms_batched = timed(batched, args.steps)
        log(f"[bench] batched (graph, persistent): {ms_batched:.2f} ms/step -> speedup {ms_per_beam / ms_batched:.3f}x")
        # The same per-beam launches on the byte-row shapes (each range's
        # faster of the 32- and 16-channel shapes where one fits; ranges no
        # byte-row shape fits keep the automatic choice): the path a port of
        # the reference's one-beam-at-a-time kernel would run.
        for b in beams:
            for r in range(len(b.tps)):
                ms = cache[repr(mb.shape_class(b, r))]["ms"]
                fit = [v for v in (0, 1) if ms[v] >= 0]
                b.plan.set_range_variant(r, min(fit, key=lambda v: ms[v]) if fit else -1)
        for _ in range(max(1, args.warmup)):
            per_beam()
        stream.synchronize()
        exact_bytes = mb.output_digest(beams) == want
        ms_per_beam_bytes = timed(per_beam, args.steps)
        log(f"[bench] per-beam launches, byte-row shapes: {ms_per_beam_bytes:.2f} ms/step (bit-identical: "
            f"{exact_bytes}) -> batched speedup {ms_per_beam_bytes / ms_batched:.3f}x")
        for b in beams:   # back to the tuned shapes
            for r in range(len(b.tps)):
                b.plan.set_range_variant(r, int(cache[repr(mb.shape_class(b, r))]["variant"]))
    achieved = adds / (ms_batched * 1e-3) / 1e12
    result = {
        "metric": "DM×sample/s bit-exact, % FP32 add roofline; batched vs per-beam speedup",
        "value": dmxs / (ms_batched * 1e-3),
        "unit": "DM-trial samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_batched,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic: seeded N(128,16^2) u8 noise + 2 dispersed pulses per beam, resident in HBM",
        "config": {"workload": ("multibeam small: 256 beams, 64-256 ch, 2^14-2^15 samples, DM max 100/200, 1-2 ranges"
                                if small else
                                "multibeam 48 beams, 1024-4096 ch, 2^16-2^19 samples, 1-4 AA ranges (config 4)"),
                   "beams": len(beams), "ranges": nranges, "dm_x_samples": dmxs, "channel_adds": adds,
                   "parallelism": "single GPU"},
        "per_beam_ms_per_step": ms_per_beam,
        "speedup_vs_per_beam_same_kernels": ms_per_beam / ms_batched,
        "speedup_target": 1.5,
        "bit_exact_vs_per_beam": exact,
        "batch_stats": batch.stats(),
        "per_beam_byte_rows_ms_per_step": ms_per_beam_bytes,
        "speedup_vs_per_beam_byte_rows": ms_per_beam_bytes / ms_batched,
        "bit_exact_vs_per_beam_byte_rows": exact_bytes,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": ADD_ROOFLINE_T, "unit": "Tadd/s",
                     "frac": achieved / ADD_ROOFLINE_T, "traffic": None,
                     "kernel": "batched step (corner turns + binning + dedispersion launches, one graph)"},
    }
    print(json.dumps(result), flush=True)


def bench_refine(args):
    """Batched DM refinement: 1024 candidate windows of uneven size (1024-4096
    channels x 256-4096 float32 samples, N(0,1) + Gaussian pulse at a residual
    DM uniform in +-1), 20 fused forward + adjoint + update steps per
    candidate in one launch; accuracy against the injected DM and against a
    dense DM-grid search (448 trials, step 0.01) through the plain u8 path."""
    import torch
    import frac_ddtr

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    cands = frac_ddtr.synth_refine_candidates(1024, SEED, dev)
    torch.cuda.synchronize()
    n = len(cands["nchans"])
    elems = int(sum(c * t for c, t in zip(cands["nchans"], cands["nsamp"])))
    log(f"[bench] refine: {n} windows, {elems * 4 / 1e9:.2f} GB, set-up {time.perf_counter() - t0:.1f} s")
    plan = frac_ddtr.RefinePlan(cands["nchans"], cands["nsamp"], cands["x_offsets"], cands["ds_offsets"])
    nsteps = 20
    SCAN = (9, -1.0, 0.25)   # coarse energy scan (forward only), then the gradient steps

    def state():
        return (torch.zeros(n, device=dev), torch.full((n,), 0.1, device=dev),
                torch.zeros(n, dtype=torch.float64, device=dev), torch.empty(n, dtype=torch.float64, device=dev),
                torch.empty(n, dtype=torch.float64, device=dev))

    for _ in range(max(1, args.warmup)):
        st = state()
        plan.run(cands["x"], cands["ds"], *st, nsteps=nsteps, scan=SCAN)
    torch.cuda.synchronize()
    ms = 0.0
    for _ in range(args.steps):
        st = state()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.run(cands["x"], cands["ds"], *st, nsteps=nsteps, scan=SCAN)
        e1.record()
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1) / args.steps
    u = st[0].cpu().numpy()
    err = np.abs(u - cands["r"])
    t0 = time.perf_counter()
    u_grid = frac_ddtr.dense_grid_search_plain(cands)
    torch.cuda.synchronize()
    t_grid = time.perf_counter() - t0
    err_grid = np.abs(u_grid - cands["r"])
    log(f"[bench] refine: {ms:.2f} ms for {nsteps} steps x {n} windows; |u - r| median {np.median(err):.4f} "
        f"p90 {np.percentile(err, 90):.4f}; dense grid (plain path) {t_grid:.2f} s, median {np.median(err_grid):.4f} "
        f"p90 {np.percentile(err_grid, 90):.4f}")
    # each of the nsteps + 1 evaluations reads every window twice (forward, recomputing
    # adjoint), each scan point once, and the channel powers of the loss once
    bytes_read = 4.0 * elems * (2 * (nsteps + 1) + SCAN[0] + 1)
    gbs = bytes_read / (ms * 1e-3) / 1e9
    result = {
        "metric": "DM×sample/s bit-exact, % FP32 add roofline; batched vs per-beam speedup",
        "value": float(sum(cands["nsamp"])) * (nsteps + 1 + SCAN[0]) / (ms * 1e-3),
        "unit": "DM-trial samples/s (scan evaluations + fused forward + adjoint evaluations)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: N(0,1) windows + Gaussian pulse (sigma 4, amplitude 0.5) at residual DM in +-1",
        "config": {"workload": "batched refinement 1024 windows 1024-4096 ch x 256-4096 samples, 20 steps (config 5)",
                   "windows": n, "window_elements": elems, "gradient_steps": nsteps, "coarse_scan": list(SCAN),
                   "loss": "profile energy minus the white-noise energy of the window's channel powers",
                   "parallelism": "single GPU"},
        "dm_error_refined": {"median": float(np.median(err)), "p90": float(np.percentile(err, 90)),
                             "max": float(err.max())},
        "dm_error_dense_grid_plain_path": {"median": float(np.median(err_grid)),
                                           "p90": float(np.percentile(err_grid, 90)), "max": float(err_grid.max()),
                                           "trials": 448, "step": 0.01, "seconds": t_grid},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "refine_kernel (fused forward + adjoint + update)"},
    }
    print(json.dumps(result), flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and the CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, usable


def cpu_baseline(args, cfg, strat, d_in, outs, ndms, tps, dm_x_samples):
    """CPU oracle (OpenMP C, checker only) timed on the whole hot path over a
    time slice of the same workload: the corner turn of the slice's input
    (its output samples plus maxshift), the x2 binning cascade and every DM
    row of every range over the slice's output samples.  The rate is measured
    on that slice, not extrapolated; every output value of the slice is also
    compared with the GPU's planes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ddtr_oracle as oracle

    if args.cpu_slice is None:
        args.cpu_slice = 24 if args.config == 2 else 1
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or cpu_model()[1]
    nchans = cfg["nchans"]
    nr = strat.nranges
    ibs = [strat.dm(r)[3] for r in range(nr)]
    ib_max = max(ibs)
    # output samples of range 0 in the slice: a whole number of the coarsest binning
    tp0 = max(ib_max, (tps[0] // max(1, args.cpu_slice)) // ib_max * ib_max)
    len0 = tp0 + strat.maxshift
    host = np.ascontiguousarray(d_in[:len0].cpu().numpy())

    def run_once():
        t_start = time.perf_counter()
        levels = {1: oracle.corner_turn(host, len0, nchans)}
        t_ct = time.perf_counter() - t_start
        b = 1
        while b < ib_max:
            levels[2 * b] = oracle.bin2(levels[b], levels[b].shape[1] // 2)
            b *= 2
        t_bin = time.perf_counter() - t_start - t_ct
        planes = []
        for r in range(nr):
            ms, st = strat.range_scalars(r)
            planes.append(oracle.dedisperse(levels[ibs[r]], strat.dmshifts(), ndms[r], tp0 // ibs[r], float(ms),
                                            float(st), ibs[r]))
        return time.perf_counter() - t_start, t_ct, t_bin, planes

    run_once()   # warm-up (page faults of the levels and planes, thread pool start)
    runs = [run_once() for _ in range(max(1, args.cpu_repeats))]
    total, t_ct, t_bin, planes = min(runs, key=lambda x: x[0])
    times = [round(x[0], 3) for x in runs]
    del host
    slice_dmxs = sum(ndms[r] * (tp0 // ibs[r]) for r in range(nr))
    exact = True
    for r in range(nr):
        got = outs[r][:, : tp0 // ibs[r]].cpu().numpy()
        if not np.array_equal(got, planes[r]):
            exact = False
            log(f"[bench] MISMATCH range {r}: {int((got != planes[r]).sum())} values differ")
    model, usable = cpu_model()
    log(f"[bench] cpu oracle on {tp0} of {tps[0]} output samples: corner turn {t_ct:.2f}s, binning {t_bin:.2f}s, "
        f"best of {len(runs)} {total:.2f}s (runs {times}) -> {slice_dmxs / total:.3e} DMxsamples/s on {threads} "
        f"threads ({model}, {usable} usable CPUs); slice bit-exact against the GPU planes: {exact}")
    return ({"value": slice_dmxs / total, "unit": "DM-trial samples/s", "cores": threads, "kind": "port",
             "sample": f"every DM row of all {nr} ranges over the first {tp0} of {tps[0]} output samples "
                       f"(1/{args.cpu_slice} of the chunk; {slice_dmxs} DMxsamples): corner turn of "
                       f"{len0} input samples, binning cascade and dedispersion, measured; best of "
                       f"{len(runs)} runs after a warm-up",
             "seconds": total, "run_seconds": times, "cpu_model": model, "usable_cpus": usable,
             "omp_threads": threads}, exact)


if __name__ == "__main__":
    main()
