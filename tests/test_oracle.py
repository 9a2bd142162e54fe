"""Pins the CPU oracle itself (CPU only), with checks that do not retype its
formula:

* delays: a numpy float64 restatement of fma + multiply + truncation;
* sums: brute force with Python integers on tiny inputs;
* invariants: the DM-0 row is the channel mean, constant input gives a
  constant plane, the plane is linear in the input;
* binning: floor-average, and a x4 cascade differs from floor(sum/4);
* the AA known-answer test: the reference's fake pulse (peak exactly 255,
  zero background) dedisperses to exactly 255 at the injected (DM, t)
  (tests/test_ddtr_fakesignal_single.cpp:111 of the reference);
* runner sequencing (chunks, binning levels, output columns) against an
  independent numpy pipeline, with small free memory forcing several chunks.
"""
import numpy as np
import pytest


def fma32(a, b, c):
    # exact product and sum in float64 for these magnitudes, one rounding to float32
    return np.float32(np.float64(np.float32(a)) * np.float64(np.float32(b)) + np.float64(np.float32(c)))


def delay_np(dmshift, d, mstartdm, mdmstep):
    shift = fma32(np.float32(d), mdmstep, mstartdm)
    prod = np.float32(np.float64(np.float32(dmshift)) * np.float64(shift))  # float32 multiply
    return int(np.trunc(prod))


def numpy_dedisperse(L, dmshifts, ndms, tp, mstartdm, mdmstep, inBin):
    nchans = L.shape[0]
    out = np.zeros((ndms, tp), dtype=np.float32)
    for d in range(ndms):
        acc = np.zeros(tp, dtype=np.int64)
        for c in range(nchans):
            dl = delay_np(dmshifts[c], d, mstartdm, mdmstep)
            acc += L[c, dl:dl + tp]
        out[d] = (acc.astype(np.float32) / np.float32(nchans)) / np.float32(inBin)
    return out


def test_delay_table_matches_closed_form(oracle):
    ds = oracle.dmshifts(512, 1550.0, -0.5859)
    s = oracle.strategy([(0, 500, 0.5, 1, 1)], 512, 131072, 1550.0, -0.5859, 64e-6)
    ms, st = s.range_scalars(0)
    tab = oracle.delay_table(ds, 1024, ms, st)
    rng = np.random.default_rng(1)
    for d, c in zip(rng.integers(0, 1024, 400), rng.integers(0, 512, 400)):
        assert tab[d, c] == delay_np(ds[c], d, ms, st)
    assert tab[0].max() == 0                      # DM 0: no delay
    assert tab[-1, -1] == max(tab.max(axis=1))    # lowest channel, highest DM is the largest
    # the pulse of config 1 sits at trial 500 exactly: fl(250/64us) = 500 * mdmstep
    assert np.float32(500) * st == np.float32(np.float32(250.0) / np.float32(64e-6))


def test_range_scalars_offset_start(oracle):
    s = oracle.strategy([(0, 150, 0.1, 1, 1), (150, 300, 0.2, 2, 2)], 256, 1 << 20, 1500.0, -1.0, 64e-6)
    ms, st = s.range_scalars(1)
    tsamp_r = np.float32(np.float32(64e-6) * np.float32(2))
    assert ms == np.float32(np.float32(s.dm[1][0]) / tsamp_r)
    assert st == np.float32(np.float32(0.2) / tsamp_r)


def test_bruteforce_tiny(oracle):
    rng = np.random.default_rng(7)
    nchans, n = 6, 600
    ds = oracle.dmshifts(nchans, 400.0, -20.0)
    L = rng.integers(0, 256, size=(nchans, n), dtype=np.uint8)
    ms, st = np.float32(3.0), np.float32(37.5)
    ndms, tp, ib = 5, 200, 2
    got = oracle.dedisperse(L, ds, ndms, tp, float(ms), float(st), ib)
    # pure Python integers, independent indexing
    for d in range(ndms):
        delays = [delay_np(ds[c], d, ms, st) for c in range(nchans)]
        for t in range(tp):
            s_ = sum(int(L[c][t + delays[c]]) for c in range(nchans))
            assert got[d][t] == np.float32(np.float32(np.float32(s_) / np.float32(nchans)) / np.float32(ib))


def test_invariants(oracle):
    rng = np.random.default_rng(3)
    nchans, n = 64, 5000
    ds = oracle.dmshifts(nchans, 1400.0, -4.0)
    L = rng.integers(0, 120, size=(nchans, n), dtype=np.uint8)
    ms, st = 0.0, 40.0
    out = oracle.dedisperse(L, ds, 16, 2000, ms, st, 1)
    # DM 0 row: per-sample channel mean
    mean = (L[:, :2000].astype(np.int64).sum(axis=0).astype(np.float32) / np.float32(nchans))
    np.testing.assert_array_equal(out[0], mean)
    # constant input -> constant plane
    const = np.full((nchans, n), 77, dtype=np.uint8)
    np.testing.assert_array_equal(oracle.dedisperse(const, ds, 16, 2000, ms, st, 1), np.float32(77))
    # linearity of the integer sums
    A = rng.integers(0, 100, size=(nchans, n), dtype=np.uint8)
    B = rng.integers(0, 100, size=(nchans, n), dtype=np.uint8)
    sa = oracle.dedisperse(A, ds, 16, 2000, ms, st, 1) * nchans
    sb = oracle.dedisperse(B, ds, 16, 2000, ms, st, 1) * nchans
    sab = oracle.dedisperse(A + B, ds, 16, 2000, ms, st, 1) * nchans
    np.testing.assert_array_equal(np.rint(sa + sb), np.rint(sab))


def test_binning(oracle):
    rng = np.random.default_rng(5)
    L = rng.integers(0, 256, size=(9, 1001), dtype=np.uint8)
    b1 = oracle.bin2(L, 500)
    np.testing.assert_array_equal(b1, ((L[:, 0:1000:2].astype(int) + L[:, 1:1000:2]) // 2).astype(np.uint8))
    b2 = oracle.bin2(b1, 250)
    assert b2.shape == (9, 250)
    # x4 is a cascade of two floor-averages, not floor(sum/4):
    # [1,0,2,1] -> [floor(1/2), floor(3/2)] = [0,1] -> floor(1/2) = 0, while floor(4/4) = 1
    y = np.array([[1, 0, 2, 1]], dtype=np.uint8)
    assert oracle.bin2(oracle.bin2(y, 2), 1)[0, 0] == 0
    assert (1 + 0 + 2 + 1) // 4 == 1


def test_corner_turn_is_transpose(oracle):
    rng = np.random.default_rng(11)
    x = rng.integers(0, 256, size=(333, 70), dtype=np.uint8)
    np.testing.assert_array_equal(oracle.corner_turn(x, 333, 70), x.T)
    y = rng.integers(0, 256, size=(50, 7), dtype=np.uint16)
    np.testing.assert_array_equal(oracle.corner_turn(y, 50, 7), y.T.astype(np.uint8))


def test_aa_known_answer_255(oracle):
    # Geometry of the reference's fake-signal test (128 channels, 300 MHz
    # band, 64 us, 2 s, DM 0-250 step 1), with seeded DM positions and starts.
    tsamp, nchans, fch1 = 6.4e-5, 128, 1550.0
    foff = -300.0 / nchans
    nsamples = int(2 / tsamp)
    s = oracle.strategy([(0, 250.0, 1, 1, 1)], nchans, nsamples, fch1, foff, tsamp)
    assert s.ok
    width = int(1 / (tsamp * 25))
    rng = np.random.default_rng(2018)
    for _ in range(6):
        dm_pos = float(rng.integers(0, 200))
        start = int(rng.integers(1, 10)) * 1000 + int(rng.integers(0, 10)) * 100 + int(rng.integers(0, 10)) * 10
        sig = oracle.aa_fake_signal(s.dmshifts, nsamples, tsamp, s.maxshift, dm_pos, start, width, 0.5)
        planes = oracle.run(s, np.ascontiguousarray(sig.astype(np.uint8)))
        d_idx = int(dm_pos) // int(s.dm[0][2])
        assert planes[0][d_idx, start] == 255.0, (dm_pos, start)
        # neighbouring DM trials do not reach the peak
        assert planes[0][d_idx + 3, start] < 255.0


def numpy_pipeline(s, data_tc):
    """Independent restatement of the runner: chunks, cascade, columns."""
    planes = [np.zeros((s.ndms[r], sum(s.t_processed[r])), np.float32) for r in range(s.nranges)]
    inc = 0
    for k in range(s.num_tchunks):
        tp0 = s.t_processed[0][k]
        L = data_tc[inc:inc + tp0 + s.maxshift].T.copy()
        levels = {1: L}
        for r in range(s.nranges):
            ib = s.dm[r][3]
            b = max(levels)
            while b < ib:
                prev = levels[b]
                n = prev.shape[1] // 2
                levels[2 * b] = ((prev[:, 0:2 * n:2].astype(int) + prev[:, 1:2 * n:2]) // 2).astype(np.uint8)
                b *= 2
            ms, st = s.range_scalars(r)
            tp = s.t_processed[r][k]
            planes[r][:, inc // ib: inc // ib + tp] = numpy_dedisperse(levels[ib], s.dmshifts, s.ndms[r], tp,
                                                                       ms, st, ib)
        inc += tp0
    return planes


def test_runner_sequencing_multi_chunk(oracle):
    nchans, nsamp, tsamp = 32, 60000, 2.56e-4
    ranges = [(0, 30, 1.0, 1, 1), (30, 60, 2.0, 2, 2), (60, 120, 4.0, 4, 4)]
    # small free memory -> several time chunks (strategy case 3)
    s = oracle.strategy(ranges, nchans, nsamp, 800.0, -2.0, tsamp, free_memory=104857600 + 1_600_000)
    assert s.ok and s.num_tchunks >= 2, (s.ok, s.num_tchunks)
    data = oracle.gen_noise(99, nsamp, nchans)
    oracle.inject_pulse(data, s.dmshifts, 45.0, tsamp, 20000, 4, 60)
    got = oracle.run(s, data)
    want = numpy_pipeline(s, data)
    for r in range(s.nranges):
        np.testing.assert_array_equal(got[r], want[r])


def test_noise_statistics(oracle):
    x = oracle.gen_noise(2026, 2000, 512).astype(np.float64)
    assert abs(x.mean() - 128.0) < 0.1
    assert abs(x.std() - 16.0) < 0.1
    assert oracle.gen_noise(2026, 10, 512).tobytes() == oracle.gen_noise(2026, 2000, 512)[:10].tobytes()
