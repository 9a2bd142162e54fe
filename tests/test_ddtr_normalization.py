"""DDTR input/output normalisation and the set_bandpass_average bandpass.

Reference: normalization_transposed_kernel (src/aa_device_zero_dm_kernel.cu:
100-164) behind zero_dm_normalization_dm / post_DDTR_normalization
(src/aa_zero_dm.cpp:49-74), run by the generic runner on each [t][c] chunk
(include/aa_permitted_pipelines_generic.hpp:731-735) and each [dm][t] plane
(:823-831); set_bandpass_average is src/aa_host_cpu_msd.cpp:7-39 (:533-551).

CPU: the oracle pinned by hand-worked answers, a float case whose value
depends on the reference's summation order, and exact brute force where
float arithmetic is exact.  GPU: the kernels through the C ABI and the
pipeline options, bit for bit against the oracle.
"""
import sys
from fractions import Fraction

import numpy as np
import pytest


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------ oracle (CPU)
def test_oracle_input_normalization_hand_worked(oracle):
    x = np.array([[10, 20], [30, 40], [50, 60]], np.uint8)
    # column 0: mean 30, s = 30 - 127.5 = -97.5 -> 107.5, 127.5, 147.5 -> truncated
    assert oracle.input_normalization(x).tolist() == [[107, 107], [127, 127], [147, 147]]
    # clipping: column mean 127.5 + 100 -> values shifted down by 100, 0 stays 0
    y = np.array([[0], [255], [255], [255]], np.uint8)      # mean 191.25, s = 63.75
    assert oracle.input_normalization(y).ravel().tolist() == [0, 191, 191, 191]
    # 4-bit: factor 7.5, clamp at 15
    z = np.array([[0], [15], [15], [2]], np.uint8)          # mean 8, s = 0.5
    assert oracle.input_normalization(z, nbits=4).ravel().tolist() == [0, 14, 14, 1]


def test_oracle_output_normalization_summation_order(oracle):
    """Column of 65 floats: row 0 = 2^24, rows 1..64 = 1.  The reference sums
    rows y, y+32, ... into partial y in row order (partial 0: 2^24 + 1 + 1
    stays 2^24), then p[y] += p[y+i] for i = 16..1, reaching 2^24 + 62 (the
    exact sum is 2^24 + 64; a plain running sum gives 2^24).  Worked by
    hand: s = fl(16777278 / 65) - 127.5 = 258111.96875 - 127.5."""
    col = np.ones((65, 1), np.float32)
    col[0, 0] = 16777216.0
    got = oracle.output_normalization(col)
    s = 257984.46875
    assert got[1, 0] == np.float32(1.0 - s) == np.float32(-257983.46875)
    assert got[0, 0] == np.float32(16519232.0)
    # every row moves by the same s
    np.testing.assert_array_equal(got[1:, 0], np.full(64, np.float32(-257983.46875)))


def test_oracle_input_normalization_exact_brute_force(oracle):
    """64 rows: sums are small integers and /64 is exact, so every float step
    is exact and x' = clamp(floor(x - S/64 + 127.5)) in rationals."""
    rng = np.random.default_rng(3)
    x = rng.integers(0, 256, size=(64, 7), dtype=np.uint8)
    got = oracle.input_normalization(x)
    for c in range(7):
        s = Fraction(int(x[:, c].astype(np.int64).sum()), 64) - Fraction(255, 2)
        for r in range(64):
            v = Fraction(int(x[r, c])) - s
            want = 0 if v < 0 else 255 if v > 255 else int(v)   # int() truncates toward zero (v >= 0 here)
            assert got[r, c] == want, (r, c)


def test_oracle_output_normalization_properties(oracle):
    rng = np.random.default_rng(4)
    plane = rng.normal(300.0, 40.0, size=(96, 50)).astype(np.float32)
    out = oracle.output_normalization(plane)
    # each time sample's mean over DMs lands on 127.5 (to float rounding)
    np.testing.assert_allclose(out.astype(np.float64).mean(axis=0), 127.5, atol=1e-3)
    # the shift is one value per column
    d = (plane.astype(np.float64) - out.astype(np.float64))
    np.testing.assert_allclose(d, np.broadcast_to(d[0], d.shape), atol=1e-3)
    # a constant column becomes exactly 127.5
    const = np.full((40, 3), 12.25, np.float32)
    np.testing.assert_array_equal(oracle.output_normalization(const), np.full((40, 3), 127.5, np.float32))


def test_oracle_bandpass_average(oracle):
    # 4 samples x 3 channels: value i averages the i-th run of 4 consecutive samples
    x = np.arange(12, dtype=np.uint16).reshape(4, 3)
    assert oracle.bandpass_average(x).tolist() == [1.5, 5.5, 9.5]
    rng = np.random.default_rng(5)
    y = rng.integers(0, 256, size=(1000, 24), dtype=np.uint16)
    want = y.reshape(-1)[:24 * 1000].reshape(24, 1000).astype(np.float64).mean(axis=1).astype(np.float32)
    np.testing.assert_array_equal(oracle.bandpass_average(y), want)


def test_bandpass_average_host_library_matches_oracle(aa, oracle):
    """The library's host bandpass (no GPU involved) against the oracle."""
    rng = np.random.default_rng(6)
    for dt in (np.uint8, np.uint16):
        x = rng.integers(0, 256, size=(3001, 40)).astype(dt)
        np.testing.assert_array_equal(aa.bandpass_average(x), oracle.bandpass_average(x))


# --------------------------------------------------------- device (GPU)
IN_CASES = [
    # nsamp, nchans, dtype, nbits
    (1000, 256, np.uint8, 8),        # dword path, integer partial sums
    (777, 99, np.uint8, 8),          # one column per lane
    (70000, 4096, np.uint8, 8),      # tall chunk, many row segments
    (4000, 64, np.uint8, 4),         # 4-bit values
    (300, 130, np.uint16, 8),        # u16 holding 8-bit values
    (20000, 40, np.uint16, 16),      # 16-bit sums can round: the literal-order kernel
    (1, 16, np.uint8, 8),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", IN_CASES, ids=[f"{c[0]}x{c[1]}_{np.dtype(c[2]).name}_b{c[3]}" for c in IN_CASES])
def test_input_normalization_kernel_matches_oracle(aa, oracle, cuda, case):
    import torch
    nsamp, nchans, dt, nbits = case
    say(f"input normalisation {case}")
    rng = np.random.default_rng(nsamp + nchans)
    top = (1 << nbits) - 1
    x = rng.integers(0, top + 1, size=(nsamp, nchans)).astype(dt)
    x[:, 0] = top // 3                        # a constant channel
    want = oracle.input_normalization(x, nbits=nbits)
    d = torch.from_numpy(x).to(cuda)
    aa.input_ddtr_normalization(d, nbits=nbits)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


OUT_CASES = [(1536, 5000), (64, 70), (97, 1), (1, 33), (512, 20480)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", OUT_CASES, ids=[f"{c[0]}dm_{c[1]}t" for c in OUT_CASES])
def test_output_normalization_kernel_matches_oracle(aa, oracle, cuda, case):
    import torch
    ndms, nsamp = case
    say(f"output normalisation {case}")
    rng = np.random.default_rng(ndms * 3 + nsamp)
    plane = rng.normal(127.0, 30.0, size=(ndms, nsamp)).astype(np.float32)
    plane[0, :] += 1.0e6                      # large values make the summation order visible
    want = oracle.output_normalization(plane)
    d = torch.from_numpy(plane).to(cuda)
    aa.output_ddtr_normalization(d)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


def _pipeline(aa, ranges, md, data, free, **opts_on):
    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    opts.copy_ddtr_data_to_host = True
    for k, v in opts_on.items():
        setattr(opts, k, v)
    plan = aa.make_plan(ranges)
    data = np.ascontiguousarray(data)
    pipe = aa.aa_py_pipeline(comps, opts, md, data, 0, max_device_bytes=free if data.dtype == np.uint8 else 0)
    assert pipe.bind_ddtr_plan(plan)
    while pipe.run():
        pass
    assert pipe.status_code() == 0
    strat = aa.aa_py_ddtr_strategy(plan, md, free)
    planes = pipe.output_planes(strat)
    pipe.cleanUp()
    return strat, planes


@pytest.mark.gpu
def test_pipeline_output_normalization_matches_oracle(aa, oracle, cuda):
    """Each (chunk, range) plane normalised over its DMs before the copy to
    the host; columns are time samples, so the whole plane's normalisation
    is the same as the per-chunk ones."""
    say("pipeline output_DDTR_normalization, several chunks")
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
    data = oracle.gen_noise(41, nsamp, nchans)
    oracle.inject_pulse(data, s.dmshifts, 60.0, tsamp, 70000, 8, 40)
    want = [oracle.output_normalization(p) for p in oracle.run(s, data)]
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    strat, planes = _pipeline(aa, ranges, md, data, free, output_DDTR_normalization=True)
    assert strat.num_tchunks == s.num_tchunks >= 2
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])


@pytest.mark.gpu
def test_pipeline_input_normalization_matches_oracle(aa, oracle, cuda):
    """Several time chunks: each chunk's t_processed + maxshift rows are
    normalised per channel on their own (the next chunk's copy arriving
    meanwhile), then corner-turned and dedispersed.  The oracle normalises
    the same rows of a copy of the input for each chunk and keeps that
    chunk's columns."""
    say("pipeline input_DDTR_normalization, several chunks")
    nchans, nsamp, tsamp, fch1, foff = 96, 200000, 1.28e-4, 1200.0, -2.5
    ranges = [(0, 40, 0.5, 1, 1), (40, 80, 1.0, 2, 2), (80, 160, 2.0, 4, 4)]
    free = None
    for extra in range(200_000_000, 1_000_000, -250_000):
        s0 = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=104857600 + extra)
        if s0.ok and s0.num_tchunks >= 3:
            free = 104857600 + extra
            break
    assert free is not None
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=free)
    data = oracle.gen_noise(43, nsamp, nchans)
    data[:, 5] = np.minimum(data[:, 5].astype(np.int32) + 50, 255).astype(np.uint8)   # a hot channel
    data[90000:, 17] = data[90000:, 17] // 2                                          # a channel that drops
    oracle.inject_pulse(data, s.dmshifts, 90.0, tsamp, 30000, 6, 40)
    tp = s.t_processed
    want = [np.zeros((s.ndms[r], sum(tp[r])), np.float32) for r in range(s.nranges)]
    for k in range(s.num_tchunks):
        inc, rows = sum(tp[0][:k]), tp[0][k] + s.maxshift
        dk = data.copy()
        dk[inc:inc + rows] = oracle.input_normalization(data[inc:inc + rows])
        out = oracle.run(s, dk)
        for r in range(s.nranges):
            c0 = sum(tp[r][:k])
            want[r][:, c0:c0 + tp[r][k]] = out[r][:, c0:c0 + tp[r][k]]
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    strat, planes = _pipeline(aa, ranges, md, data, free, input_DDTR_normalization=True)
    assert strat.num_tchunks == s.num_tchunks >= 3
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])


@pytest.mark.gpu
def test_pipeline_zero_dm_with_bandpass_average_matches_oracle(aa, oracle, cuda):
    """set_bandpass_average replaces the zero-DM bandpass with the means the
    reference computes from the host input."""
    say("pipeline zero_dm + set_bandpass_average")
    nchans, nsamp, tsamp, fch1, foff = 64, 50000, 64e-6, 1400.0, -4.0
    ranges = [(0, 60, 0.5, 1, 1), (60, 180, 1.0, 2, 2)]
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp)
    data = oracle.gen_noise(47, nsamp, nchans)
    oracle.inject_pulse(data, s.dmshifts, 70.0, tsamp, 20000, 6, 40)
    bp = oracle.bandpass_average(data)
    want = oracle.run(s, oracle.zero_dm(data, bandpass=bp))
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    _, planes = _pipeline(aa, ranges, md, data, 280 * 2**30, zero_dm=True, set_bandpass_average=True)
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])


@pytest.mark.gpu
@pytest.mark.parametrize("zdm", ["outliers", "plain"])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_pipeline_all_value_options_compose(aa, oracle, cuda, dtype, zdm):
    """Zero-DM (with outliers, or plain), input and output normalisation
    together, in the reference's order (zero-DM, input normalisation, corner
    turn, dedispersion, output normalisation), on u8 and on u16 host input."""
    say(f"pipeline {zdm} zero-DM + input + output normalisation, {np.dtype(dtype).name} input")
    nchans, nsamp, tsamp, fch1, foff = 128, 60000, 64e-6, 1400.0, -2.0
    ranges = [(0, 60, 0.5, 1, 1), (60, 180, 1.0, 2, 2)]
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp)
    assert s.ok and s.num_tchunks == 1
    data = oracle.gen_noise(59, nsamp, nchans)
    data[:, 9] = 250                      # a hot channel
    oracle.inject_pulse(data, s.dmshifts, 90.0, tsamp, 30000, 6, 40)
    rows = s.t_processed[0][0] + s.maxshift
    pre = data.copy()
    zdm_fn = oracle.zero_dm_outliers if zdm == "outliers" else oracle.zero_dm
    pre[:rows] = oracle.input_normalization(zdm_fn(data[:rows]))
    want = [oracle.output_normalization(p) for p in oracle.run(s, pre)]
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    zopt = {"zero_dm_with_outliers": True} if zdm == "outliers" else {"zero_dm": True}
    _, planes = _pipeline(aa, ranges, md, data.astype(dtype), 280 * 2**30, input_DDTR_normalization=True,
                          output_DDTR_normalization=True, **zopt)
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])
