"""float64 restatement of the first single-pulse-search pass on a DM-time plane.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
aa_hip_spd_first_pass; the product never imports it.

What it restates (reference files cited):
  * the first boxcar plane of the single-pulse search,
    src/aa_device_SPS_long_kernel.cu:42-168 (SPDT_GPU_1st_plane): for every
    DM row and start sample t, boxcar sums B_w(t) = sum_{i<w} x[t+i] for
    widths w = 1..nBoxcars (nBoxcars = PD_MAXTAPS = 32,
    include/aa_params.hpp:37), SNR_w = (B_w - mean_w) / sd_w, and the output
    SNR(t) = max_w SNR_w with its width w (1-based, the "taps" plane);
  * the per-width (mean, sd) those SNRs divide by (d_MSD[w-1], float2),
    which the reference estimates in MSD_plane_profile
    (src/aa_device_MSD_plane_profile.cu) by decimation and interpolation.
    Here they are the exact mean and population standard deviation of every
    boxcar sum of width w that lies inside the row (t + w <= nsamp), over
    all rows: n_w = ndms * (nsamp - w + 1) values.

Conventions of this restatement (the library follows them):
  * near the row's end only widths that fit are compared (t + w <= nsamp);
  * a width whose sd is 0 contributes SNR 0;
  * ties between widths go to the smaller width.

Parity: pinned by brute force on tiny planes (tests/test_spd.py) and by
closed forms (a plane of i.i.d. +-1 values, a linear ramp).
"""
from __future__ import annotations

import numpy as np

MAX_WIDTH = 32


def boxcar_sums(plane: np.ndarray, width: int) -> np.ndarray:
    """[ndms][nsamp - width + 1] float64 sums of `width` consecutive samples."""
    x = np.asarray(plane, dtype=np.float64)
    c = np.zeros((x.shape[0], x.shape[1] + 1), dtype=np.float64)
    np.cumsum(x, axis=1, out=c[:, 1:])
    return c[:, width:] - c[:, :-width]


def boxcar_msd(plane: np.ndarray, max_width: int) -> np.ndarray:
    """[max_width][2] float64: mean and population sd of every in-row boxcar
    sum of width w = 1..max_width."""
    ndms, nsamp = plane.shape
    assert 1 <= max_width <= min(MAX_WIDTH, nsamp)
    out = np.zeros((max_width, 2), dtype=np.float64)
    x = np.asarray(plane, dtype=np.float64)
    for w in range(1, max_width + 1):
        b = boxcar_sums(x, w)
        m = b.mean()
        out[w - 1, 0] = m
        out[w - 1, 1] = np.sqrt(np.mean((b - m) ** 2))
    return out


def boxcar_snr(plane: np.ndarray, msd: np.ndarray, max_width: int):
    """(snr [ndms][nsamp] float64, width [ndms][nsamp] int): the best SNR over
    the widths that fit at each start sample, and that width (1-based)."""
    ndms, nsamp = plane.shape
    x = np.asarray(plane, dtype=np.float64)
    best = np.full((ndms, nsamp), -np.inf)
    arg = np.zeros((ndms, nsamp), dtype=np.int64)
    for w in range(1, max_width + 1):
        b = boxcar_sums(x, w)
        m, s = msd[w - 1]
        snr = (b - m) / s if s > 0 else np.zeros_like(b)
        cur = best[:, :nsamp - w + 1]
        better = snr > cur
        cur[better] = snr[better]
        arg[:, :nsamp - w + 1][better] = w
    return best, arg


def first_pass(plane: np.ndarray, max_width: int = MAX_WIDTH):
    """(msd, snr, width) of one DM-time plane."""
    msd = boxcar_msd(plane, max_width)
    snr, width = boxcar_snr(plane, msd, max_width)
    return msd, snr, width


def brute_force(plane: np.ndarray, max_width: int):
    """Plain loops (tiny planes only): the same three outputs."""
    ndms, nsamp = plane.shape
    x = [[float(v) for v in row] for row in np.asarray(plane, dtype=np.float64)]
    msd = []
    for w in range(1, max_width + 1):
        vals = [sum(x[d][t:t + w]) for d in range(ndms) for t in range(nsamp - w + 1)]
        m = sum(vals) / len(vals)
        msd.append((m, (sum((v - m) ** 2 for v in vals) / len(vals)) ** 0.5))
    snr = [[0.0] * nsamp for _ in range(ndms)]
    width = [[0] * nsamp for _ in range(ndms)]
    for d in range(ndms):
        for t in range(nsamp):
            best, bw = None, 0
            for w in range(1, max_width + 1):
                if t + w > nsamp:
                    break
                m, s = msd[w - 1]
                v = (sum(x[d][t:t + w]) - m) / s if s > 0 else 0.0
                if best is None or v > best:
                    best, bw = v, w
            snr[d][t], width[d][t] = best, bw
    return np.array(msd), np.array(snr), np.array(width)
