"""The oracle's path for samples wider than 8 bits (u16 rows, float channel
sums in channel order), pinned without the reference's GPU code:

  * exact regime: while nchans x max sample < 2^24 every float partial sum
    is an exact integer, so the plane equals integer brute force over the
    oracle's delay table, whatever the order;
  * rounding regime, worked by hand: 257 channels of 65535 then 63 of 1 at
    DM 0.  256 x 65535 = 16,776,960 is exact; + 65535 = 16,842,495 is a tie
    between the floats 16,842,494 and 16,842,496 (spacing 2 above 2^24) and
    rounds to the even mantissa, 16,842,496; each + 1 is again a tie and
    rounds back to 16,842,496.  The exact sum is 16,842,558, and the
    reversed channel order (ones first) gives 16,842,558 too: only the
    reference's order c = 0 .. nchans-1 gives 16,842,496;
  * a pure-Python float32 loop on small inputs in the rounding regime;
  * binning and the transpose against numpy, and the runner against its
    parts.
"""
import numpy as np


def _delays(oracle, ds, ndms, ms, st):
    return oracle.delay_table(ds, ndms, ms, st)   # [ndms][nchans]


def test_exact_regime_equals_integer_brute_force(oracle):
    rng = np.random.default_rng(11)
    nchans, ndms, tp, ib = 96, 24, 700, 2
    ds = oracle.dmshifts(nchans, 1500.0, -2.0)
    ms, st = np.float32(10.0) / (np.float32(64e-6) * 2), np.float32(0.5) / (np.float32(64e-6) * 2)
    dl = _delays(oracle, ds, ndms, float(ms), float(st))
    n = tp + int(dl.max()) + 5
    L = rng.integers(0, 4096, size=(nchans, n)).astype(np.uint16)   # 96 x 4095 < 2^24: exact
    got = oracle.dedisperse_rows_u16(L, n, nchans, ds, np.arange(ndms), tp, float(ms), float(st), ib)
    want = np.zeros((ndms, tp), np.int64)
    for d in range(ndms):
        for c in range(nchans):
            want[d] += L[c, dl[d, c]:dl[d, c] + tp]
    np.testing.assert_array_equal(got, ((want.astype(np.float32) / np.float32(nchans)) / np.float32(ib)))


def test_rounding_regime_hand_worked_column(oracle):
    nchans = 320
    L = np.zeros((nchans, 4), np.uint16)
    L[:257] = 65535
    L[257:] = 1
    ds = np.zeros(nchans, np.float32)                     # DM 0: every delay 0
    got = oracle.dedisperse_rows_u16(L, 4, nchans, ds, [0], 4, 0.0, 1.0, 1)
    assert np.all(got == np.float32(16842496.0) / np.float32(320.0))
    rev = oracle.dedisperse_rows_u16(L[::-1].copy(), 4, nchans, ds, [0], 4, 0.0, 1.0, 1)
    assert np.all(rev == np.float32(16842558.0) / np.float32(320.0))


def test_rounding_regime_python_loop(oracle):
    rng = np.random.default_rng(5)
    nchans, ndms, tp = 300, 6, 40
    ds = oracle.dmshifts(nchans, 1400.0, -1.0)
    ms, st = 0.0, float(np.float32(2.0) / np.float32(1e-4))
    dl = _delays(oracle, ds, ndms, ms, st)
    n = tp + int(dl.max()) + 1
    L = rng.integers(40000, 65536, size=(nchans, n)).astype(np.uint16)   # sums pass 2^24 after ~300 channels
    got = oracle.dedisperse_rows_u16(L, n, nchans, ds, np.arange(ndms), tp, ms, st, 4)
    for d in range(ndms):
        for t in range(tp):
            acc = np.float32(0.0)
            for c in range(nchans):
                acc = np.float32(acc + np.float32(L[c, t + dl[d, c]]))
            assert got[d, t] == np.float32(np.float32(acc / np.float32(nchans)) / np.float32(4.0)), (d, t)
    # the regime does round here: the exact mean differs somewhere
    exact = np.array([[sum(int(L[c, t + dl[d, c]]) for c in range(nchans)) for t in range(tp)] for d in range(ndms)])
    assert np.any(got * nchans * 4 != exact)


def test_transpose_and_bin_u16(oracle):
    rng = np.random.default_rng(3)
    x = rng.integers(0, 65536, size=(777, 33)).astype(np.uint16)
    T = oracle.transpose_u16(x, 777, 33)
    np.testing.assert_array_equal(T, x.T)
    b = oracle.bin2_u16(T, 777 // 2)
    np.testing.assert_array_equal(b, ((T[:, 0:776:2].astype(np.int64) + T[:, 1:776:2]) // 2).astype(np.uint16))


def test_run_u16_equals_its_parts(oracle):
    rng = np.random.default_rng(9)
    nchans, nsamp, tsamp = 64, 20000, 64e-6
    ranges = [(0, 60, 1.0, 1, 1), (60, 180, 2.0, 2, 2), (180, 300, 4.0, 4, 4)]
    s = oracle.strategy(ranges, nchans, nsamp, 1500.0, -3.0, tsamp, nbits=16)
    assert s.ok and s.num_tchunks == 1
    x = rng.integers(0, 65536, size=(nsamp, nchans)).astype(np.uint16)
    planes = oracle.run(s, x)
    len0 = s.t_processed[0][0] + s.maxshift
    lv = oracle.transpose_u16(x[:len0], len0, nchans)
    b = 1
    for r in range(s.nranges):
        ib = s.dm[r][3]
        while b < ib:
            lv = oracle.bin2_u16(lv, lv.shape[1] // 2)
            b *= 2
        ms, st = s.range_scalars(r)
        want = oracle.dedisperse_rows_u16(lv, lv.shape[1], nchans, s.dmshifts, np.arange(s.ndms[r]),
                                          s.t_processed[r][0], ms, st, ib)
        np.testing.assert_array_equal(planes[r], want)
