"""GPU parity at the headline size: the 8-range AA plan on 4096 channels x
1,048,576 u8 samples (1760 -> 950 MHz, 64 us; ranges 0-150/0.1/1 ...
2000-3000/2.0/8, 6080 DM trials after AA's rounding), the plan of
examples/src/filterbank_dedispersion.cpp:26-33 in the reference.

One chunk is run through aa_hip_ddtr_execute_chunk with the automatic kernel
choice (the pair kernel: 256 channel chunks and 16 flushes of its 16-bit
accumulators per tile at this channel count), then checked three ways:

  * sampled rows: at least 64 DM rows per range (DM 0, the last DM, the
    first and last DM of tiles across the range, the trials of every
    injected pulse and their neighbours) bit for bit against the oracle on
    the oracle's own corner turn and binning cascade;
  * every row: sum_t out[d][t] * nchans * inBin equals
    sum_c (P_c[delay(c,d) + t_p] - P_c[delay(c,d)]), P_c the prefix sums of
    channel c's binned level, built with torch integer ops on the device
    (independent of the library's kernels) and the oracle's delay table;
  * range 0's DM-0 row equals the per-sample channel mean of the input
    (all delays are 0 there), and every injected pulse peaks at its trial
    and arrival sample.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCHANS, NSAMP, FCH1, FOFF, TSAMP = 4096, 1048576, 1760.0, -810.0 / 4096, 64e-6
RANGES = [(0, 150, 0.1, 1, 1), (150, 300, 0.2, 1, 1), (300, 500, 0.25, 1, 1), (500, 900, 0.4, 2, 2),
          (900, 1200, 0.6, 4, 4), (1200, 1500, 0.8, 4, 4), (1500, 2000, 1.0, 4, 4), (2000, 3000, 2.0, 8, 8)]
SEED = 20261203
NPULSES = 16
ROWS_PER_RANGE = 64


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_pulses(maxshift, top_dm, seed=SEED):
    """16 pulses: DM uniform in [0, top_dm), start sample uniform with the
    whole sweep inside the chunk, width 1-64 samples."""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(0, top_dm)), int(rng.integers(0, NSAMP - maxshift - 64)), int(rng.integers(1, 65)))
            for _ in range(NPULSES)]


def pulse_trial(dm_ranges, dm):
    """(range, nearest trial) of a DM; dm_ranges[r] = (low, step, ndms, inBin)."""
    for r, (low, step, ndms, _) in enumerate(dm_ranges):
        if low <= dm < low + ndms * step:
            return r, min(ndms - 1, max(0, int(round((dm - low) / step))))
    r = len(dm_ranges) - 1
    return r, dm_ranges[r][2] - 1


def off_trials(dm_ranges, r, width, dmshift_max, tsamp=TSAMP):
    """Trials away from a pulse's nearest trial at which its sweep is smeared
    by at least 4 widths + 32 samples in the lowest channel (at least 8)."""
    low, step, ndms, ib = dm_ranges[r]
    per_trial = dmshift_max * step / (np.float32(tsamp) * ib)   # samples of level ib per trial
    return max(8, int(np.ceil((4.0 * width / ib + 32.0) / per_trial)))


def sample_rows(ndms, pulse_rows, rng, k=ROWS_PER_RANGE):
    """DM 0, the last DM, the first and last DM of tiles of 16 across the
    range, the pulse rows (trial, +-1, +-off), then random rows up to k."""
    ntiles = (ndms + 15) // 16
    rows = {0, ndms - 1}
    for t in sorted({0, 1, ntiles // 4, ntiles // 2, (3 * ntiles) // 4, ntiles - 2, ntiles - 1}):
        if t >= 0:
            rows.update({16 * t, min(ndms - 1, 16 * t + 15)})
    for d, off in pulse_rows:
        rows.update(x for x in (d - off, d - 1, d, d + 1, d + off) if 0 <= x < ndms)
    while len(rows) < min(k, ndms):
        rows.add(int(rng.integers(0, ndms)))
    return np.array(sorted(rows), dtype=np.int32)


def pulse_peak_ok(row_at, row_plus, row_minus, t0, width, ib):
    """Checks one pulse on three DM rows of its range (the nearest trial and
    the trials off_trials() steps either side; None past the range's end).
    Returns (ok, message).

    In a window of +-256 samples around the arrival sample t0/inBin, the
    nearest trial's row peaks inside [t0/inBin - 2, (t0 + width - 1)/inBin + 8]
    (a trial off the true DM by up to half a step smears the pulse later by a
    few samples), the peak rises at least 10 robust sigma above the window's
    median, and the rows either side peak lower in the same window."""
    tp = row_at.size
    a, b = t0 // ib, (t0 + width - 1) // ib
    lo, hi = max(0, a - 256), min(tp, b + 257)
    win = row_at[lo:hi].astype(np.float64)
    k = int(np.argmax(win))
    tpk = lo + k
    med = float(np.median(win))
    mad = float(np.median(np.abs(win - med))) * 1.4826
    peak = float(win[k])
    others = [float(np.max(r[lo:hi])) for r in (row_plus, row_minus) if r is not None]
    msg = (f"t0={t0} w={width} ib={ib}: peak {peak:.3f} at {tpk} (window [{a - 2}, {b + 8}]), median {med:.3f}, "
           f"robust sigma {mad:.4f}, rows either side peak {others}")
    ok = (a - 2 <= tpk <= b + 8) and peak >= med + 10.0 * max(mad, 1e-3) and all(o < peak for o in others)
    return ok, msg


@pytest.fixture(scope="module")
def c2(aa, cuda):
    """Strategy, device input (noise + 16 pulses) and the automatic-kernel
    output planes of the 8-range plan, one chunk."""
    import torch
    say("config-2: strategy, synthetic input, execute_chunk (auto kernel)")
    free_b, _ = torch.cuda.mem_get_info()
    plan = aa.make_plan(RANGES)
    md = aa.make_metadata(NCHANS, NSAMP, FCH1, FOFF, TSAMP)
    strat = aa.aa_py_ddtr_strategy(plan, md, free_b)
    assert strat.ready and strat.num_tchunks == 1
    nr = strat.nranges
    ndms = [strat.ndms(r) for r in range(nr)]
    tps = [strat.t_processed(r, 0) for r in range(nr)]
    assert ndms == [1536, 768, 832, 1024, 512, 384, 512, 512] and strat.maxshift == 157920
    dm_ranges = []
    for r in range(nr):
        low, _, step, ib, _ = strat.dm(r)
        dm_ranges.append((low, step, ndms[r], ib))
    ds = strat.dmshifts()
    d_ds = torch.from_numpy(ds).to(cuda)
    d_in = torch.empty((NSAMP, NCHANS), dtype=torch.uint8, device=cuda)
    aa.synth_noise_u8(SEED, d_in, 128.0, 16.0)
    pulses = make_pulses(strat.maxshift, strat.dm(nr - 1)[1])
    for dm, t0, w in pulses:
        aa.inject_pulse_u8(d_in, NSAMP, NCHANS, d_ds, dm, TSAMP, t0, w, 32)
    outs = [torch.full((ndms[r], tps[r]), -1.0, dtype=torch.float32, device=cuda) for r in range(nr)]
    ddtr = aa.DdtrPlan.from_strategy(strat)
    ddtr.set_kernel(aa.KERNEL_AUTO)
    ddtr.execute_chunk(d_in, tps, outs)
    torch.cuda.synchronize()
    say(f"config-2: ran {sum(ndms)} DM trials x t_processed {tps}")
    yield dict(strat=strat, ndms=ndms, tps=tps, dm_ranges=dm_ranges, ds=ds, d_in=d_in, outs=outs,
               pulses=pulses, ddtr=ddtr)
    ddtr.close()
    del outs, d_in
    torch.cuda.empty_cache()


def _pulse_rows(c2):
    rows = {r: [] for r in range(len(c2["ndms"]))}
    for dm, t0, w in c2["pulses"]:
        r, d = pulse_trial(c2["dm_ranges"], dm)
        rows[r].append((d, off_trials(c2["dm_ranges"], r, w, float(c2["ds"][-1]))))
    return rows


def test_config2_sampled_rows_bit_exact(oracle, c2):
    """>= 64 rows per range, bit for bit against the oracle."""
    strat, ndms, tps = c2["strat"], c2["ndms"], c2["tps"]
    say("config-2: oracle corner turn + binning cascade on the host copy of the input")
    host = c2["d_in"].cpu().numpy()
    len0 = tps[0] + strat.maxshift
    levels = {1: oracle.corner_turn(np.ascontiguousarray(host[:len0]), len0, NCHANS)}
    del host
    b = 1
    while b < max(ib for *_, ib in c2["dm_ranges"]):
        levels[2 * b] = oracle.bin2(levels[b], levels[b].shape[1] // 2)
        b *= 2
    prow = _pulse_rows(c2)
    rng = np.random.default_rng(11)
    for r in range(len(ndms)):
        ib = c2["dm_ranges"][r][3]
        idx = sample_rows(ndms[r], prow[r], rng)
        assert idx.size >= min(ROWS_PER_RANGE, ndms[r])
        ms, st = strat.range_scalars(r)
        lv = levels[ib]
        assert lv.shape[1] >= tps[r] + strat.maxshift // ib
        want = oracle.dedisperse_rows(lv, lv.shape[1], NCHANS, c2["ds"], idx, tps[r], float(ms), float(st), ib)
        got = c2["outs"][r][idx.tolist()].cpu().numpy()
        bad = np.argwhere(got != want)
        say(f"config-2 range {r}: {idx.size} rows x {tps[r]} samples, {len(bad)} differ")
        assert len(bad) == 0, f"range {r}: first difference at row {idx[bad[0][0]]} sample {bad[0][1]}"


def test_config2_every_row_checksum(oracle, c2):
    """Every DM row of every range: its sum over time equals the prefix-sum
    identity on the binned input (exact integers throughout)."""
    import torch
    strat, ndms, tps = c2["strat"], c2["ndms"], c2["tps"]
    dev = c2["d_in"].device
    len0 = tps[0] + strat.maxshift
    level = c2["d_in"][:len0].t().contiguous()          # [c][t], the corner turn
    b = 1
    for r in range(len(ndms)):
        ib = c2["dm_ranges"][r][3]
        while b < ib:                                     # floor((a + b) / 2), row length halves
            n = level.shape[1] // 2
            level = ((level[:, 0:2 * n:2].to(torch.int16) + level[:, 1:2 * n:2].to(torch.int16)) >> 1).to(torch.uint8)
            b *= 2
        nrow = level.shape[1]
        pref = torch.zeros((NCHANS, nrow + 1), dtype=torch.int32, device=dev)
        torch.cumsum(level, dim=1, dtype=torch.int32, out=pref[:, 1:])
        ms, st = strat.range_scalars(r)
        delays = oracle.delay_table(c2["ds"], ndms[r], float(ms), float(st))    # [ndms][nchans]
        assert delays.min() >= 0 and delays.max() + tps[r] <= nrow
        dl = torch.from_numpy(delays).to(dev).to(torch.int64)
        base = torch.arange(NCHANS, device=dev, dtype=torch.int64) * (nrow + 1)
        flat = pref.view(-1)
        want = (flat[base + dl + tps[r]].to(torch.int64) - flat[base + dl].to(torch.int64)).sum(dim=1)
        got = torch.sum(c2["outs"][r], dim=1, dtype=torch.float64) * float(NCHANS * ib)
        diff = (got - want.to(torch.float64)).abs()
        nbad = int((diff != 0).sum().item())
        say(f"config-2 range {r}: {ndms[r]} row checksums, {nbad} differ")
        assert nbad == 0, f"range {r}: first bad row {int(torch.nonzero(diff)[0].item())}"
        del pref, dl, flat, want, got, diff


def test_config2_dm0_row_is_channel_mean(c2):
    """Range 0 starts at DM 0, where every delay is 0: its first row is the
    per-sample mean over the 4096 channels, over the whole chunk."""
    import torch
    tp0 = c2["tps"][0]
    assert c2["dm_ranges"][0][0] == 0.0
    mean = c2["d_in"][:tp0].to(torch.int32).sum(dim=1).to(torch.float32) / float(NCHANS)
    got = c2["outs"][0][0]
    nbad = int((got != mean).sum().item())
    say(f"config-2: DM-0 row vs channel mean over {tp0} samples: {nbad} differ")
    assert nbad == 0


def test_config2_pulses_peak_at_trial_and_sample(c2):
    """Each injected pulse peaks at its nearest DM trial, at its arrival
    sample (the top channel's), above the trials either side at which its
    sweep is smeared well past its width."""
    for dm, t0, w in c2["pulses"]:
        r, d = pulse_trial(c2["dm_ranges"], dm)
        ndm, ib = c2["ndms"][r], c2["dm_ranges"][r][3]
        off = off_trials(c2["dm_ranges"], r, w, float(c2["ds"][-1]))
        plane = c2["outs"][r]
        row = lambda x: plane[x].cpu().numpy() if 0 <= x < ndm else None  # noqa: E731
        ok, msg = pulse_peak_ok(row(d), row(d + off), row(d - off), t0, w, ib)
        say(f"config-2 pulse DM {dm:.2f} -> range {r} trial {d} (+-{off}): {msg}")
        assert ok, msg
