"""Zero-DM with outlier rejection (src/aa_device_zero_dm_outliers_kernel.cu:
190-233, src/aa_zero_dm_outliers.cpp:47-67; the pipeline's
zero_dm_with_outliers option, include/aa_permitted_pipelines_generic.hpp:
713-719).

CPU: the oracle (the reference's streaming statistics and merge order)
pinned by a hand-worked row, by the empty-selection case, and against a
textbook mean / population-sigma rejection loop on rows where no value sits
near a rejection limit (there the two must agree exactly, since the final
mean is the exact channel sum over the same set).  GPU: the kernels through
the C ABI and the pipeline option, bit for bit against the oracle.
"""
import sys

import numpy as np
import pytest


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def textbook_mean(row, margin=1e-3):
    """Mean of the channels left after iterated 3-sigma rejection (population
    sigma, strict limits, at most 5 rejection passes, stop once the float32
    mean drops by less than 0.001), or None when some value lies within
    `margin` of a limit (rounding of sigma could then decide)."""
    x = row.astype(np.float64)
    m, s = x.mean(), x.std()
    m32 = np.float32(m)
    for _ in range(5):
        old = m32
        lo, hi = m - 3 * s, m + 3 * s
        if np.min(np.abs(x - lo)) < margin or np.min(np.abs(x - hi)) < margin:
            return None
        sel = x[(x > lo) & (x < hi)]
        if sel.size == 0:
            return np.float32(np.nan)
        m, s = sel.sum() / sel.size, sel.std()
        m32 = np.float32(m)
        if np.float32(old - m32) < np.float32(0.001):
            break
    return m32


def apply_mean(row, mean, bp, nbits):
    top = (1 << nbits) - 1
    r = (row.astype(np.float32) - np.float32(mean)) + bp.astype(np.float32)
    r = np.where(np.isnan(r), 0.0, np.clip(r, 0, top))
    return np.trunc(r).astype(row.dtype)


# ------------------------------------------------------------ oracle (CPU)
def test_oracle_hand_worked_row(oracle):
    """128 channels 100,101,102,103 repeating, channel 7 = 250.  Pass 0: mean
    13139/128 = 102.6484375, sigma 13.1, limits 63.3 .. 142.0 drop the 250.
    Pass 1: mean 12889/127 = 101.48819, sigma 1.12 keeps all 127.  Pass 2:
    same set, the mean stops moving.  x' = trunc(x - 101.48819 + 127.5)."""
    row = (100 + np.arange(128) % 4).astype(np.uint8)
    row[7] = 250
    got = oracle.zero_dm_outliers(row[None, :])[0]
    want = (np.array([126, 127, 128, 129] * 32)).astype(np.uint8)
    want[7] = 255
    np.testing.assert_array_equal(got, want)
    # the plain zero-DM filter keeps the outlier in its mean: 100 -> 124
    assert oracle.zero_dm(row[None, :])[0][0] == 124


def test_oracle_empty_selection_gives_zero(oracle):
    """A constant row has sigma 0; the first rejection pass keeps nothing
    (strict limits), its mean is 0/0, and NaN converts to 0."""
    const = np.full((2, 64), 90, np.uint8)
    np.testing.assert_array_equal(oracle.zero_dm_outliers(const), np.zeros((2, 64), np.uint8))
    near = np.full((1, 128), 100, np.uint8)
    near[0, 3] = 101           # sigma 0.088 drops the 101, then sigma 0 drops the rest
    np.testing.assert_array_equal(oracle.zero_dm_outliers(near), np.zeros((1, 128), np.uint8))


def test_oracle_matches_textbook_rejection(oracle):
    rng = np.random.default_rng(11)
    checked = 0
    for nchans in (96, 128, 200, 1024, 4096):
        x = np.clip(np.rint(rng.normal(120.0, 12.0, size=(60, nchans))), 0, 255).astype(np.uint8)
        hot = rng.random(size=x.shape) < 0.02
        x[hot] = rng.integers(200, 256, size=int(hot.sum()))
        bp = rng.uniform(110.0, 140.0, size=nchans).astype(np.float32)
        got = oracle.zero_dm_outliers(x, bandpass=bp)
        for t in range(x.shape[0]):
            m = textbook_mean(x[t])
            if m is None:
                continue
            checked += 1
            np.testing.assert_array_equal(got[t], apply_mean(x[t], m, bp, 8), err_msg=f"{nchans} ch, row {t}")
    assert checked >= 250


def test_oracle_u16_and_4bit(oracle):
    rng = np.random.default_rng(12)
    x16 = rng.integers(20000, 40000, size=(20, 256)).astype(np.uint16)
    got16 = oracle.zero_dm_outliers(x16, nbits=16)
    x4 = rng.integers(0, 16, size=(20, 256)).astype(np.uint8)
    got4 = oracle.zero_dm_outliers(x4, nbits=4)
    assert got4.max() <= 15
    for t in range(20):
        m = textbook_mean(x16[t], margin=1e-2)
        if m is not None and np.abs(x16[t].astype(np.float64).sum()) < 2**24:
            np.testing.assert_array_equal(got16[t], apply_mean(x16[t], m, np.full(256, 32767.5, np.float32), 16))


# --------------------------------------------------------- device (GPU)
CASES = [
    # nsamp, nchans, dtype, nbits
    (2000, 4096, np.uint8, 8),
    (3000, 96, np.uint8, 8),
    (500, 1000, np.uint8, 8),
    (700, 40000, np.uint8, 8),      # counts beyond the factor table
    (1000, 256, np.uint8, 4),
    (400, 300, np.uint16, 16),
    (300, 128, np.uint16, 8),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}_{np.dtype(c[2]).name}_b{c[3]}" for c in CASES])
def test_zero_dm_outliers_kernel_matches_oracle(aa, oracle, cuda, case):
    import torch
    nsamp, nchans, dt, nbits = case
    say(f"zero-DM with outliers {case}")
    rng = np.random.default_rng(nsamp + 3 * nchans)
    top = (1 << nbits) - 1
    x = np.clip(np.rint(rng.normal(top * 0.45, top * 0.06, size=(nsamp, nchans))), 0, top).astype(dt)
    hot = rng.random(size=x.shape) < 0.03
    x[hot] = top
    x[0, :] = top // 2                     # a constant row (empty selection -> 0)
    bp = rng.uniform(0.4 * top, 0.6 * top, size=nchans).astype(np.float32)
    want = oracle.zero_dm_outliers(x, bandpass=bp, nbits=nbits)
    d = torch.from_numpy(x).to(cuda)
    aa.zero_dm_outliers(d, torch.from_numpy(bp).to(cuda), nbits=nbits)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


@pytest.mark.gpu
def test_pipeline_zero_dm_with_outliers_matches_oracle(aa, oracle, cuda):
    say("pipeline zero_dm_with_outliers, several chunks")
    nchans, nsamp, tsamp, fch1, foff = 96, 200000, 1.28e-4, 1200.0, -2.5
    ranges = [(0, 40, 0.5, 1, 1), (40, 80, 1.0, 2, 2), (80, 160, 2.0, 4, 4)]
    free = None
    for extra in range(200_000_000, 1_000_000, -250_000):
        s0 = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=104857600 + extra)
        if s0.ok and s0.num_tchunks >= 2:
            free = 104857600 + extra
            break
    assert free is not None
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=free)
    data = oracle.gen_noise(53, nsamp, nchans)
    oracle.inject_pulse(data, s.dmshifts, 60.0, tsamp, 70000, 8, 40)
    data[120000:120040, :] = np.minimum(data[120000:120040, :].astype(np.int32) + 60, 255).astype(np.uint8)
    data[:, 11] = 255                      # a saturated channel, rejected as an outlier
    want = oracle.run(s, oracle.zero_dm_outliers(data))
    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    opts.copy_ddtr_data_to_host = True
    opts.zero_dm_with_outliers = True
    plan = aa.make_plan(ranges)
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    pipe = aa.aa_py_pipeline(comps, opts, md, np.ascontiguousarray(data), 0, max_device_bytes=free)
    assert pipe.bind_ddtr_plan(plan)
    while pipe.run():
        pass
    assert pipe.status_code() == 0
    strat = aa.aa_py_ddtr_strategy(plan, md, free)
    assert strat.num_tchunks == s.num_tchunks >= 2
    planes = pipe.output_planes(strat)
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])
    pipe.cleanUp()
