"""The float64 first-pass oracle (oracle/spd_ref.py) pinned without the
library: brute-force loops on tiny planes, closed forms, invariants."""
import numpy as np
import pytest

import spd_ref


@pytest.mark.parametrize("ndms,nsamp,W", [(1, 1, 1), (3, 7, 3), (2, 40, 32), (5, 33, 32), (4, 20, 6)])
def test_oracle_matches_brute_force(ndms, nsamp, W):
    rng = np.random.default_rng(ndms * 100 + nsamp)
    x = rng.normal(3.0, 2.0, size=(ndms, nsamp)).astype(np.float32)
    msd, snr, width = spd_ref.first_pass(x, W)
    bm, bs, bw = spd_ref.brute_force(x, W)
    np.testing.assert_allclose(msd, bm, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(snr, bs, rtol=1e-10, atol=1e-10)
    np.testing.assert_array_equal(width, bw)


def test_alternating_plane_closed_form():
    """x[t] = (-1)^t: odd widths sum to (-1)^t, even widths to 0."""
    nsamp, W = 101, 32
    x = np.tile(np.where(np.arange(nsamp) % 2 == 0, 1.0, -1.0), (3, 1))
    msd = spd_ref.boxcar_msd(x, W)
    for w in range(1, W + 1):
        n = nsamp - w + 1
        if w % 2 == 0:
            assert msd[w - 1, 0] == 0.0 and msd[w - 1, 1] == 0.0
        else:
            mean = (n % 2) / n          # n alternating values starting at +1
            assert msd[w - 1, 0] == pytest.approx(mean, abs=1e-14)
            assert msd[w - 1, 1] == pytest.approx(np.sqrt(1.0 - mean * mean), abs=1e-14)
    snr, width = spd_ref.boxcar_snr(x, msd, W)
    # even widths have sd 0 (SNR 0): at an even start (+1) width 1 gives
    # (1 - m1)/sd1 > 0, which beats every other odd width (same sum, larger sd)
    assert np.all(width[:, 0:nsamp:2] == 1)


def test_ramp_closed_form():
    """x[t] = t: B_w(t) = w t + w(w-1)/2, uniform over n = nsamp - w + 1 starts."""
    nsamp, W = 500, 32
    x = np.tile(np.arange(nsamp, dtype=np.float64), (2, 1))
    msd = spd_ref.boxcar_msd(x, W)
    for w in range(1, W + 1):
        n = nsamp - w + 1
        assert msd[w - 1, 0] == pytest.approx(w * (n - 1) / 2.0 + w * (w - 1) / 2.0, rel=1e-12)
        assert msd[w - 1, 1] == pytest.approx(w * np.sqrt((n * n - 1) / 12.0), rel=1e-12)


def test_affine_invariance():
    """SNR and width do not change under x -> a x + b (a > 0); the MSD moves
    to (a mean_w + w b, a sd_w)."""
    rng = np.random.default_rng(3)
    x = rng.normal(0.0, 1.0, size=(6, 300))
    a, b = 3.5, -17.25
    m0, s0, w0 = spd_ref.first_pass(x, 32)
    m1, s1, w1 = spd_ref.first_pass(a * x + b, 32)
    w = np.arange(1, 33)
    np.testing.assert_allclose(m1[:, 0], a * m0[:, 0] + w * b, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(m1[:, 1], a * m0[:, 1], rtol=1e-10)
    np.testing.assert_allclose(s1, s0, rtol=1e-9, atol=1e-9)
    np.testing.assert_array_equal(w1, w0)


def test_spike_and_boxcar_pulse():
    """A single large sample is found at width 1; a flat-topped pulse of
    width 8 is found at its start with width 8."""
    rng = np.random.default_rng(9)
    x = rng.normal(0.0, 1.0, size=(4, 2000))
    x[1, 700] += 40.0
    x[3, 1200:1208] += 10.0
    msd, snr, width = spd_ref.first_pass(x, 32)
    assert width[1, 700] == 1 and np.unravel_index(np.argmax(snr[1]), snr[1].shape)[0] == 700
    k = int(np.argmax(snr[3]))
    assert k == 1200 and width[3, k] == 8


def test_planted_index_mistake_is_caught():
    """A boxcar sum taken one sample late (a plausible off-by-one) disagrees
    with the brute force."""
    rng = np.random.default_rng(5)
    x = rng.normal(0.0, 1.0, size=(2, 30))
    bm, _, _ = spd_ref.brute_force(x, 8)
    shifted = np.concatenate([x[:, 1:], x[:, :1]], axis=1)
    wrong = spd_ref.boxcar_msd(shifted, 8)
    assert not np.allclose(wrong, bm, rtol=1e-9, atol=1e-9)
