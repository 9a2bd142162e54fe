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
  cpu_baseline the CPU oracle (OpenMP C) on the host cores over a bounded
               sample: full corner turn and binning, plus 16 DM rows per
               range, extrapolated to the whole plane; the sampled rows are
               also compared bit for bit with the GPU output.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1|2]
                       [--kernel auto|direct|lds] [--no-cpu-baseline]
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
# summed over the 8 launches; recorded in
# profiles/dedisperse_lds_8range_4096ch_1048576_hbm_traffic.txt. Counters
# cannot be read live, so the figure is tied to the kernel it was taken on.
MEASURED_TRAFFIC = {2: {"bytes": 78.70e9, "kernel": "lds",
                        "profile": "profiles/dedisperse_lds_8range_4096ch_1048576_hbm_traffic.txt"}}
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2])
    ap.add_argument("--kernel", default="auto", choices=["auto", "direct", "lds"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=16, help="DM rows per range in the CPU sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "dedispersion shards by DM range or beam with no exchange, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    import py_astro_accelerate as aa

    cfg = CONFIGS[args.config]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lib = aa.load_library()
    kernel = {"auto": aa.KERNEL_AUTO, "direct": aa.KERNEL_DIRECT, "lds": aa.KERNEL_LDS}[args.kernel]

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
    # run_range launches only the dedisperse kernel.
    reps = 3
    kt = [0.0] * nr
    with torch.cuda.stream(stream):
        for _ in range(reps):
            for r in range(nr):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                ddtr.run_range(r, tps[r], outs[r], stream=stream)
                b.record(stream)
                b.synchronize()
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
                     "traffic_source": meas["profile"] if traffic else None},
    }

    if not args.no_cpu_baseline:
        result["cpu_baseline"], result["bit_exact_sample"] = cpu_baseline(
            args, cfg, strat, d_in, outs, ndms, tps, dm_x_samples)
    print(json.dumps(result), flush=True)


def cpu_baseline(args, cfg, strat, d_in, outs, ndms, tps, dm_x_samples):
    """CPU oracle on a bounded sample of the same workload (checker only)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ddtr_oracle as oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nchans, nsamp = cfg["nchans"], cfg["nsamp"]
    host = d_in.cpu().numpy()
    len0 = tps[0] + strat.maxshift
    t0 = time.perf_counter()
    L = oracle.corner_turn(np.ascontiguousarray(host[:len0]), len0, nchans)
    t_ct = time.perf_counter() - t0
    levels = {1: L}
    t_bin = 0.0
    b = 1
    while b < max(strat.dm(r)[3] for r in range(strat.nranges)):
        t0 = time.perf_counter()
        levels[2 * b] = oracle.bin2(levels[b], levels[b].shape[1] // 2)
        t_bin += time.perf_counter() - t0
        b *= 2
    rng = np.random.default_rng(7)
    t_rows_extrap = 0.0
    exact = True
    sample_dmxs = 0
    for r in range(strat.nranges):
        ib = strat.dm(r)[3]
        k = min(args.cpu_rows, ndms[r])
        idx = np.unique(np.concatenate([[0, ndms[r] - 1], rng.integers(0, ndms[r], k)]))[:k]
        ms, st = strat.range_scalars(r)
        lv = levels[ib]
        t0 = time.perf_counter()
        rows = oracle.dedisperse_rows(lv, lv.shape[1], nchans, strat.dmshifts(), idx, tps[r], float(ms), float(st), ib)
        dt = time.perf_counter() - t0
        t_rows_extrap += dt * ndms[r] / len(idx)
        sample_dmxs += len(idx) * tps[r]
        got = outs[r][idx.tolist()].cpu().numpy()
        if not np.array_equal(got, rows):
            exact = False
            log(f"[bench] MISMATCH range {r}: {int((got != rows).sum())} values differ")
    total = t_ct + t_bin + t_rows_extrap
    log(f"[bench] cpu oracle: corner turn {t_ct:.2f}s, binning {t_bin:.2f}s, rows extrapolated "
        f"{t_rows_extrap:.1f}s -> {dm_x_samples / total:.3e} DMxsamples/s on {threads} threads; "
        f"sampled rows bit-exact: {exact}")
    return ({"value": dm_x_samples / total, "unit": "DM-trial samples/s", "cores": threads, "kind": "port",
             "sample": f"full corner turn + binning; {args.cpu_rows} DM rows per range "
                       f"({sample_dmxs} DMxsamples) extrapolated to all {sum(ndms)} rows",
             "seconds_extrapolated": total}, exact)


if __name__ == "__main__":
    main()
