"""Many beams at once: plan generation, device set-up, tile autotuning and the
batched launch (DdtrBatch) beside the per-beam launches it replaces.

No counterpart in the reference, which dedisperses one beam per process
(include/aa_pipeline_api.hpp:113-117).  Every beam keeps AA's own plan
semantics (aa_ddtr_plan ranges, aa_ddtr_strategy, the dmshift table of its
band); only the launch is shared.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

import py_astro_accelerate as aa

BAND_TOP_MHZ = 1760.0
BAND_WIDTH_MHZ = 810.0
TSAMP = 64e-6


def aa_style_ranges(dm_max: float, nranges: int):
    """AA-style DM plan to dm_max in 1..4 ranges: the DM step doubles with the
    binning factor from range to range (step / inBin constant), as in AA's
    own multi-range plans (examples/src/filterbank_dedispersion.cpp:26-33)."""
    base = {100: 0.5, 200: 0.5, 500: 0.25, 1000: 0.4, 3000: 0.5}[int(dm_max)]
    edges = {1: [0.0, 1.0], 2: [0.0, 1 / 3, 1.0], 3: [0.0, 1 / 6, 1 / 2, 1.0],
             4: [0.0, 1 / 10, 1 / 4, 1 / 2, 1.0]}[nranges]
    out = []
    for i in range(nranges):
        b = 1 << i
        out.append((round(edges[i] * dm_max, 3), round(edges[i + 1] * dm_max, 3), base * b, b, b))
    return out


def multibeam_specs(nbeams: int = 48, seed: int = 20261202,
                    nchans_set=(1024, 2048, 4096), nsamp_set=(65536, 131072, 262144, 524288),
                    dm_max_set=(500, 1000, 3000), max_ranges: int = 4):
    """Seeded beam specs: nchans, nsamp (powers of 2), DM max, 1-max_ranges
    AA-style ranges; combinations whose dispersion sweep does not fit the
    beam's length are drawn again (DM 3000 needs >= 262144 samples and >= 2
    ranges, DM 1000 >= 131072 samples)."""
    rng = np.random.default_rng(seed)
    specs = []
    while len(specs) < nbeams:
        nchans = int(rng.choice(nchans_set))
        nsamp = int(rng.choice(nsamp_set))
        dm_max = int(rng.choice(dm_max_set))
        nr = int(rng.integers(1, max_ranges + 1))
        if dm_max == 3000 and (nsamp < 262144 or nr < 2):
            continue
        if dm_max == 1000 and nsamp < 131072:
            continue
        specs.append(dict(nchans=nchans, nsamp=nsamp, dm_max=dm_max, ranges=aa_style_ranges(dm_max, nr),
                          fch1=BAND_TOP_MHZ, foff=-BAND_WIDTH_MHZ / nchans, tsamp=TSAMP,
                          seed=int(rng.integers(0, 2**31)),
                          pulses=[(float(rng.uniform(0.05, 0.9) * dm_max), float(rng.uniform(0.05, 0.6)),
                                   int(rng.integers(1, 33))) for _ in range(2)]))
    return specs


def small_multibeam_specs(seed: int = 20261202):
    """256 small beams: 64-256 channels, 2^14-2^15 samples, DM max 100/200 in
    1-2 ranges; a range holds a few dozen 16-DM tiles, fewer than the GPU's
    CUs, so launches per range leave most of the GPU idle."""
    return multibeam_specs(256, seed, nchans_set=(64, 128, 256), nsamp_set=(16384, 32768), dm_max_set=(100, 200),
                           max_ranges=2)


@dataclass
class Beam:
    spec: dict
    strategy: "aa.aa_py_ddtr_strategy"
    plan: "aa.DdtrPlan"
    d_input: object
    outs: list
    tps: list
    ndms: list

    @property
    def dm_x_samples(self) -> int:
        return sum(n * t for n, t in zip(self.ndms, self.tps))

    @property
    def channel_adds(self) -> int:
        return self.spec["nchans"] * self.dm_x_samples


def build_beams(specs: Sequence[dict], device, stream=None, free_memory: int = 64 << 30) -> List[Beam]:
    """Strategy, device plan, seeded u8 input ([t][c], N(128, 16^2) noise plus
    the spec's pulses of +32) and float outputs for every beam."""
    import torch
    beams = []
    for sp in specs:
        plan = aa.make_plan(sp["ranges"])
        md = aa.make_metadata(sp["nchans"], sp["nsamp"], sp["fch1"], sp["foff"], sp["tsamp"])
        st = aa.aa_py_ddtr_strategy(plan, md, free_memory)
        if not st.ready:
            raise aa.AAHipError(f"strategy not ready for beam spec {sp}")
        if st.num_tchunks != 1:
            raise aa.AAHipError(f"beam spec needs {st.num_tchunks} chunks; batched beams are single-chunk")
        nr = st.nranges
        ndms = [st.ndms(r) for r in range(nr)]
        tps = [st.t_processed(r, 0) for r in range(nr)]
        d_in = torch.empty((sp["nsamp"], sp["nchans"]), dtype=torch.uint8, device=device)
        aa.synth_noise_u8(sp["seed"], d_in, 128.0, 16.0, stream=stream)
        d_ds = torch.from_numpy(st.dmshifts()).to(device)
        for dm, t_frac, width in sp["pulses"]:
            t0 = int(t_frac * (sp["nsamp"] - st.maxshift))
            aa.inject_pulse_u8(d_in, sp["nsamp"], sp["nchans"], d_ds, dm, sp["tsamp"], t0, width, 32, stream=stream)
        outs = [torch.empty((ndms[r], tps[r]), dtype=torch.float32, device=device) for r in range(nr)]
        beams.append(Beam(sp, st, aa.DdtrPlan.from_strategy(st), d_in, outs, tps, ndms))
    return beams


def shape_class(beam: Beam, r: int):
    """Key under which a tuning result is shared: channels, binning, DM trials,
    DM step per binned sample (delay spread of a tile) and time tiles."""
    low, high, step, ib, _ = beam.strategy.dm(r)
    return (beam.spec["nchans"], ib, beam.ndms[r], round(step / ib, 4), math.ceil(beam.tps[r] / 2048))


def autotune(beams: Sequence[Beam], cache_path: Optional[str] = None, reps: int = 2, stream=None,
             log=None) -> Dict[str, dict]:
    """Picks the LDS chunk shape of every (beam, range) by timing each shape
    that fits (aa_hip_ddtr_autotune_range) once per shape class; results are
    kept in a JSON cache (`cache_path`) and applied to every beam."""
    cache: Dict[str, dict] = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    tuned = 0
    for b in beams:
        ingested = False
        for r in range(len(b.ndms)):
            key = repr(shape_class(b, r))
            if key in cache:
                b.plan.set_range_variant(r, int(cache[key]["variant"]))
                continue
            if not ingested:
                b.plan.ingest(b.d_input, b.tps[0], stream=stream)
                ingested = True
            best, ms = b.plan.autotune_range(r, b.tps[r], b.outs[r], reps=reps, stream=stream)
            cache[key] = {"variant": best, "ms": ms}
            tuned += 1
            if log:
                log(f"[autotune] {key}: shape {best} ({ms})")
    if cache_path:
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return cache


def run_per_beam(beams: Sequence[Beam], stream=None):
    """The baseline: each beam's own launches (execute_chunk), one after another."""
    for b in beams:
        b.plan.execute_chunk(b.d_input, b.tps, b.outs, stream=stream)


def output_digest(beams: Sequence[Beam]):
    """Per (beam, range) integer digest of the output bits (position-weighted
    sums of the float32 words), to compare two runs bit for bit on the device."""
    import torch
    out = []
    for b in beams:
        for o in b.outs:
            w = o.contiguous().view(torch.int32).to(torch.int64).flatten()
            pos = torch.arange(w.numel(), device=w.device, dtype=torch.int64) % 1021 + 1
            out.append((int(w.sum()), int((w * pos).sum())))
    return out
