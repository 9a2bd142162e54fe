"""The strategy's time chunking (cases 3 and 4 of src/aa_ddtr_strategy.cpp:
209-274 and 324-388), pinned by what its arithmetic fixes rather than by the
oracle agreeing with the product:

  * invariants over randomised multi-chunk plans, for the product library's
    strategy and the oracle's: the chunks never pass the input
    (sum_k t_p[0][k] + maxshift <= nsamp) and leave less than one
    granule (224 * inBin_last samples) unprocessed; every t_p[0][k] is a
    multiple of that granule; t_p[r][k] = t_p[0][k] / inBin_r; one chunk's
    input and output fit the memory the strategy was given;
  * chunk-partition invariance: the oracle's DM-time planes of a
    multi-chunk strategy equal the single-chunk planes on their common
    columns (chunk k lands at column inc_k / inBin_r);
  * one case-3 plan worked out by hand from the reference's formulas.
"""
import numpy as np
import pytest

GRAN = 224          # SDIVINT * 2 * SNUMREG (include/aa_params.hpp)
RESERVE = 104857600  # bytes the strategy keeps back (src/aa_ddtr_strategy.cpp:77)


def _product(aa, ranges, nchans, nsamp, fch1, foff, tsamp, free):
    plan = aa.make_plan(ranges)
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    return aa.aa_py_ddtr_strategy(plan, md, free)


def _random_plans(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        nchans = int(rng.choice([16, 64, 96, 128, 256, 512, 1024, 4096]))
        nsamp = int(rng.integers(20_000, 3_000_000))
        fch1 = float(rng.uniform(300, 1800))
        foff = -float(rng.uniform(0.05, 0.5)) * fch1 / nchans
        tsamp = float(rng.choice([32e-6, 64e-6, 128e-6, 256e-6]))
        nr = int(rng.integers(1, 5))
        lo, ranges, ib = 0.0, [], 1
        for _ in range(nr):
            step = float(rng.choice([0.1, 0.25, 0.5, 1.0, 2.0]))
            hi = lo + step * int(rng.integers(32, 700))
            ranges.append((lo, hi, step, ib, ib))
            lo = hi
            ib *= int(rng.choice([1, 2]))
        free = int(rng.integers(RESERVE + 2_000_000, RESERVE + 400_000_000))
        out.append((ranges, nchans, nsamp, fch1, foff, tsamp, free))
    return out


def _check_invariants(name, ndms, tps, maxshift, nsamp, nchans, inbins, free):
    nch = len(tps[0])
    ib_last = inbins[-1]
    g = GRAN * ib_last
    total = sum(tps[0])
    assert total + maxshift <= nsamp, f"{name}: chunks pass the input"
    assert nsamp - maxshift - total < g, f"{name}: a whole granule left unprocessed"
    max_ndms = max(ndms)
    per_sample = 2 * nchans + 4 * max_ndms if nchans < max_ndms else 6 * nchans
    for k in range(nch):
        assert tps[0][k] > 0 and tps[0][k] % g == 0, f"{name}: chunk {k} not a whole number of granules"
        for r, ib in enumerate(inbins):
            assert tps[r][k] * ib == tps[0][k], f"{name}: range {r} chunk {k}"
        assert (tps[0][k] + maxshift) * per_sample <= free - RESERVE, f"{name}: chunk {k} exceeds the memory"
    # all chunks but the last have the same length; the last is not longer
    assert all(t == tps[0][0] for t in tps[0][:-1]) and tps[0][-1] <= tps[0][0]


def test_chunk_invariants_random_plans(aa, oracle):
    checked = multi = 0
    for i, (ranges, nchans, nsamp, fch1, foff, tsamp, free) in enumerate(_random_plans(400, 2024)):
        s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=free)
        p = _product(aa, ranges, nchans, nsamp, fch1, foff, tsamp, free)
        assert s.ok == p.ready, f"plan {i}"
        if not s.ok:
            continue
        inbins = [d[3] for d in s.dm]
        _check_invariants(f"oracle plan {i}", s.ndms, s.t_processed, s.maxshift, nsamp, nchans, inbins, free)
        ptp = [[p.t_processed(r, k) for k in range(p.num_tchunks)] for r in range(p.nranges)]
        _check_invariants(f"product plan {i}", [p.ndms(r) for r in range(p.nranges)], ptp, p.maxshift, nsamp,
                          nchans, inbins, free)
        checked += 1
        multi += s.num_tchunks > 1
    assert checked >= 150 and multi >= 100, (checked, multi)


def test_chunk_partition_invariance(oracle):
    """Multi-chunk planes = single-chunk planes on the common columns."""
    nchans, nsamp, tsamp, fch1, foff = 64, 300_000, 64e-6, 1400.0, -4.0
    ranges = [(0, 100, 1, 1, 1), (100, 200, 2, 2, 2), (200, 400, 4, 4, 4)]
    one = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp)
    assert one.ok and one.num_tchunks == 1
    many = None
    for extra in range(40_000_000, 1_000_000, -200_000):
        s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=RESERVE + extra)
        if s.ok and s.num_tchunks >= 4:
            many = s
            break
    assert many is not None
    assert many.maxshift == one.maxshift and many.ndms == one.ndms and many.dm == one.dm
    data = oracle.gen_noise(31, nsamp, nchans)
    oracle.inject_pulse(data, one.dmshifts, 150.0, tsamp, 120_000, 10, 40)
    a, b = oracle.run(one, data), oracle.run(many, data)
    for r in range(one.nranges):
        n = min(a[r].shape[1], b[r].shape[1])
        assert n >= b[r].shape[1] - 1
        np.testing.assert_array_equal(b[r][:, :n], a[r][:, :n], err_msg=f"range {r}")


def test_case3_worked_by_hand(aa, oracle):
    """64 channels, 1400 MHz down by 4 MHz per channel, 64 us, 200,000
    samples; ranges 0-100/1/bin 1 and 100-200/2/bin 2; free memory
    136,857,600 bytes.  From src/aa_ddtr_strategy.cpp:
      ndms = 64 * ((int(100/1) + 64) / 64) = 128 and 64 * ((int(100/2) + 64) / 64) = 64;
      redefined ranges 0-128/1 and 128-256/2;
      dmshift[63] = 4148.741601 * (1/1148^2 - 1/1400^2) = 1.0313e-3, so the
      last range's shift is ceil(256 * 1.0313e-3 / 64e-6) ~ 4125, then
      int((4125 / 2 + 224) / 224) = 10 granules of 224 * 2: 4480, plus 224:
      maxshift = 4704 (:112-140);
      nchans 64 < max_ndms 128: case 1 or 3; gpu_memory = 136,857,600 -
      104,857,600 = 32,000,000 bytes; max_tsamps = 32,000,000 / (2*64 +
      4*128) = 50,000 <= nsamp: case 3 (:166-209);
      samp_block_size = 50,000 - 4704 = 45,296; local_t_processed =
      floor(45,296 / 2 / 224) = 101 -> 101 * 224 * 2 = 45,248 (:216-224);
      num_blocks = int(195,296 / 45,248) = 4; remainder = 200,000 - 4 *
      45,248 - 4704 = 14,304 -> floor(14,304 / 2) / 224 = 31 -> 31 * 224 *
      2 = 13,888, one more block (:226-241);
      t_processed[0] = [45,248] * 4 + [13,888]; t_processed[1] = [22,624] * 4
      + [6,944] (:243-270)."""
    ranges = [(0, 100, 1, 1, 1), (100, 200, 2, 2, 2)]
    free = 136_857_600
    want_tp = [[45248] * 4 + [13888], [22624] * 4 + [6944]]
    s = oracle.strategy(ranges, 64, 200_000, 1400.0, -4.0, 64e-6, free_memory=free)
    assert s.ok and s.strategy_case == 3
    assert (s.ndms, s.maxshift, s.num_tchunks, s.t_processed) == ([128, 64], 4704, 5, want_tp)
    assert [d[:3] for d in s.dm] == [(0.0, 128.0, 1.0), (128.0, 256.0, 2.0)]
    assert s.dmshifts[63] == pytest.approx(1.0313e-3, rel=1e-3)
    p = _product(aa, ranges, 64, 200_000, 1400.0, -4.0, 64e-6, free)
    assert p.ready and p.maxshift == 4704 and p.num_tchunks == 5
    assert [[p.t_processed(r, k) for k in range(5)] for r in range(2)] == want_tp
