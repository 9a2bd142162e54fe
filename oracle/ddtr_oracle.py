"""ctypes loader for the CPU oracle (oracle/_build/libddtr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product library
never imports it.  Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libddtr_oracle.so")
MAX_RANGES = 64


class _Range(ctypes.Structure):
    _fields_ = [("low", ctypes.c_float), ("high", ctypes.c_float), ("step", ctypes.c_float),
                ("inBin", ctypes.c_int), ("outBin", ctypes.c_int)]


class _Strategy(ctypes.Structure):
    _fields_ = [("ok", ctypes.c_int), ("nranges", ctypes.c_int), ("nchans", ctypes.c_int),
                ("nsamp", ctypes.c_int), ("maxshift", ctypes.c_int), ("max_ndms", ctypes.c_int),
                ("total_ndms", ctypes.c_int), ("num_tchunks", ctypes.c_int), ("strategy_case", ctypes.c_int),
                ("ndms", ctypes.c_int * MAX_RANGES), ("dm", _Range * MAX_RANGES),
                ("t_processed", ctypes.POINTER(ctypes.c_int)), ("max_chunks", ctypes.c_int),
                ("dmshifts", ctypes.POINTER(ctypes.c_float)), ("tsamp", ctypes.c_double)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, f, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
        pf, pi = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
        sig = {
            "oracle_dmshifts": ([i, d, d, f, pf], None),
            "oracle_strategy_compute": ([ctypes.POINTER(_Range), i, i, i, d, d, d, i, f, sz, i,
                                         ctypes.POINTER(_Strategy)], i),
            "oracle_strategy_free": ([ctypes.POINTER(_Strategy)], None),
            "oracle_delay": ([f, i, f, f], i),
            "oracle_delay_table": ([pf, i, i, f, f, pi], None),
            "oracle_range_scalars": ([f, i, f, f, pf, pf], None),
            "oracle_corner_turn_u8": ([vp, i, i, vp, sz], None),
            "oracle_corner_turn_u16": ([vp, i, i, vp, sz], None),
            "oracle_bin_u8": ([vp, sz, i, i, vp, sz], None),
            "oracle_zero_dm_u8": ([vp, i, i, vp, i], None),
            "oracle_zero_dm_u16": ([vp, i, i, vp, i], None),
            "oracle_zero_dm_outliers_u8": ([vp, i, i, vp, i], None),
            "oracle_zero_dm_outliers_u16": ([vp, i, i, vp, i], None),
            "oracle_normalize_columns_u8": ([vp, sz, sz, f, i], None),
            "oracle_normalize_columns_u16": ([vp, sz, sz, f, i], None),
            "oracle_normalize_columns_f32": ([vp, sz, sz, f], None),
            "oracle_bandpass_average_u16": ([vp, sz, sz, vp], None),
            "oracle_dedisperse_range": ([vp, sz, i, i, pf, i, i, f, f, i, pf, sz], None),
            "oracle_dedisperse_rows": ([vp, sz, i, pf, pi, i, i, f, f, i, pf, sz], None),
            "oracle_ddtr_run_u8": ([ctypes.POINTER(_Strategy), vp, ctypes.POINTER(pf)], i),
            "oracle_transpose_u16": ([vp, i, i, vp, sz], None),
            "oracle_bin_u16": ([vp, sz, i, i, vp, sz], None),
            "oracle_dedisperse_rows_u16": ([vp, sz, i, pf, pi, i, i, f, f, i, pf, sz], None),
            "oracle_ddtr_run_u16": ([ctypes.POINTER(_Strategy), vp, ctypes.POINTER(pf)], i),
            "oracle_gen_noise_u8": ([ctypes.c_uint64, i, i, d, d, vp], None),
            "oracle_inject_pulse_u8": ([vp, i, i, pf, d, d, i, i, i], None),
            "oracle_aa_mask": ([f, i, pf], i),
            "oracle_aa_fake_signal_u16": ([pf, i, i, d, i, d, i, i, f, i, vp], i),
        }
        for name, (a, r) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def _pf(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _pi(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _vp(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


@dataclass
class Strategy:
    ok: bool
    nchans: int
    nsamp: int
    tsamp: float
    maxshift: int
    max_ndms: int
    num_tchunks: int
    strategy_case: int
    ndms: List[int]
    dm: List[tuple]                  # (low, high, step, inBin, outBin) float32 values
    t_processed: List[List[int]]     # [range][chunk]
    dmshifts: np.ndarray
    _raw: object = field(default=None, repr=False)

    @property
    def nranges(self) -> int:
        return len(self.ndms)

    def range_scalars(self, r: int):
        ms, st = ctypes.c_float(), ctypes.c_float()
        lo, _, step, ib, _ = self.dm[r]
        lib().oracle_range_scalars(float(np.float32(self.tsamp)), ib, lo, step, ctypes.byref(ms), ctypes.byref(st))
        return ms.value, st.value


def dmshifts(nchans: int, fch1: float, foff: float, power: float = 2.0) -> np.ndarray:
    out = np.zeros(nchans, dtype=np.float32)
    lib().oracle_dmshifts(nchans, fch1, foff, power, _pf(out))
    return out


def strategy(ranges: Sequence[Sequence[float]], nchans: int, nsamp: int, fch1: float, foff: float,
             tsamp: float, nbits: int = 8, power: float = 2.0, free_memory: int = 280 * 2**30,
             enable_analysis: bool = False) -> Strategy:
    n = len(ranges)
    arr = (_Range * max(n, 1))(*[_Range(r[0], r[1], r[2], int(r[3]), int(r[4]) if len(r) > 4 else int(r[3]))
                                 for r in ranges])
    s = _Strategy()
    lib().oracle_strategy_compute(arr, n, nchans, nsamp, fch1, foff, tsamp, nbits, power, int(free_memory),
                                  int(enable_analysis), ctypes.byref(s))
    nch = max(s.num_tchunks, 0)
    tp = [[s.t_processed[r * s.max_chunks + k] for k in range(nch)] for r in range(n)] if s.t_processed else []
    ds = np.ctypeslib.as_array(s.dmshifts, shape=(nchans,)).copy() if s.dmshifts else np.zeros(0, np.float32)
    out = Strategy(bool(s.ok), nchans, nsamp, tsamp, s.maxshift, s.max_ndms, s.num_tchunks, s.strategy_case,
                   [s.ndms[r] for r in range(n)],
                   [(s.dm[r].low, s.dm[r].high, s.dm[r].step, s.dm[r].inBin, s.dm[r].outBin) for r in range(n)],
                   tp, ds, s)
    return out


def run(s: Strategy, input_tc: np.ndarray) -> List[np.ndarray]:
    """Whole DDTR of a u8 [t][c] input (or a u16 one: samples wider than 8
    bits, float channel sums in channel order): one float32 [ndms][total_t]
    plane per range."""
    assert input_tc.dtype in (np.uint8, np.uint16) and input_tc.flags.c_contiguous
    planes = [np.zeros((s.ndms[r], sum(s.t_processed[r])), dtype=np.float32) for r in range(s.nranges)]
    ptrs = (ctypes.POINTER(ctypes.c_float) * s.nranges)(*[_pf(p) for p in planes])
    fn = lib().oracle_ddtr_run_u8 if input_tc.dtype == np.uint8 else lib().oracle_ddtr_run_u16
    st = fn(ctypes.byref(s._raw), _vp(input_tc), ptrs)
    if st != 0:
        raise RuntimeError(f"oracle run failed ({st})")
    return planes


def corner_turn(input_tc: np.ndarray, nsamp: int, nchans: int) -> np.ndarray:
    out = np.zeros((nchans, nsamp), dtype=np.uint8)
    if input_tc.dtype == np.uint8:
        lib().oracle_corner_turn_u8(_vp(input_tc), nsamp, nchans, _vp(out), nsamp)
    else:
        lib().oracle_corner_turn_u16(_vp(np.ascontiguousarray(input_tc, np.uint16)), nsamp, nchans, _vp(out), nsamp)
    return out


def zero_dm(data_tc: np.ndarray, bandpass=None, nbits: int = 8) -> np.ndarray:
    """Zero-DM filtered copy of [nsamp, nchans] u8 or u16 samples (default
    bandpass (2^nbits - 1) / 2 per channel, as the strategy sets it)."""
    out = np.ascontiguousarray(data_tc).copy()
    nsamp, nchans = out.shape
    bp = np.full(nchans, ((1 << nbits) - 1.0) / 2.0, dtype=np.float32) if bandpass is None else \
        np.ascontiguousarray(bandpass, dtype=np.float32)
    fn = lib().oracle_zero_dm_u8 if out.dtype == np.uint8 else lib().oracle_zero_dm_u16
    fn(_vp(out), nsamp, nchans, _vp(bp), nbits)
    return out


def zero_dm_outliers(data_tc: np.ndarray, bandpass=None, nbits: int = 8) -> np.ndarray:
    """Zero-DM with outlier rejection, filtered copy of [nsamp, nchans] u8 or
    u16 samples (default bandpass (2^nbits - 1) / 2 per channel)."""
    out = np.ascontiguousarray(data_tc).copy()
    nsamp, nchans = out.shape
    bp = np.full(nchans, ((1 << nbits) - 1.0) / 2.0, dtype=np.float32) if bandpass is None else \
        np.ascontiguousarray(bandpass, dtype=np.float32)
    fn = lib().oracle_zero_dm_outliers_u8 if out.dtype == np.uint8 else lib().oracle_zero_dm_outliers_u16
    fn(_vp(out), nsamp, nchans, _vp(bp), nbits)
    return out


def normalize_columns(data: np.ndarray, factor: float, nbits: int = 32) -> np.ndarray:
    """Column-normalised copy of a [rows, cols] u8, u16 or float32 array
    (the reference's normalization_transposed_kernel)."""
    out = np.ascontiguousarray(data).copy()
    rows, cols = out.shape
    if out.dtype == np.uint8:
        lib().oracle_normalize_columns_u8(_vp(out), rows, cols, factor, nbits)
    elif out.dtype == np.uint16:
        lib().oracle_normalize_columns_u16(_vp(out), rows, cols, factor, nbits)
    elif out.dtype == np.float32:
        lib().oracle_normalize_columns_f32(_vp(out), rows, cols, factor)
    else:
        raise TypeError(out.dtype)
    return out


def input_normalization(data_tc: np.ndarray, nbits: int = 8) -> np.ndarray:
    """Input DDTR normalisation of a [t][c] chunk: each channel's mean over
    the chunk moved to (2^nbits - 1)/2."""
    return normalize_columns(data_tc, float(np.float32(((1 << nbits) - 1) / 2.0)), nbits)


def output_normalization(plane: np.ndarray) -> np.ndarray:
    """Output DDTR normalisation of a [dm][t] float plane: each time sample's
    mean over DMs moved to 127.5."""
    return normalize_columns(np.asarray(plane, dtype=np.float32), 127.5, 32)


def bandpass_average(data_tc: np.ndarray) -> np.ndarray:
    """The set_bandpass_average bandpass of a [nsamp, nchans] input."""
    x = np.ascontiguousarray(data_tc, dtype=np.uint16)
    nsamp, nchans = x.shape
    bp = np.zeros(nchans, dtype=np.float32)
    lib().oracle_bandpass_average_u16(_vp(x), nchans, nsamp, _vp(bp))
    return bp


def bin2(level: np.ndarray, n_out: int) -> np.ndarray:
    level = np.ascontiguousarray(level)
    out = np.zeros((level.shape[0], n_out), dtype=np.uint8)
    lib().oracle_bin_u8(_vp(level), level.shape[1], level.shape[0], n_out, _vp(out), n_out)
    return out


def dedisperse(level: np.ndarray, dmshifts_: np.ndarray, ndms: int, t_processed: int, mstartdm: float,
               mdmstep: float, inBin: int) -> np.ndarray:
    level = np.ascontiguousarray(level, dtype=np.uint8)
    ds = np.ascontiguousarray(dmshifts_, dtype=np.float32)
    out = np.zeros((ndms, t_processed), dtype=np.float32)
    lib().oracle_dedisperse_range(_vp(level), level.shape[1], level.shape[1], level.shape[0], _pf(ds), ndms,
                                  t_processed, mstartdm, mdmstep, inBin, _pf(out), t_processed)
    return out


def dedisperse_rows(level: np.ndarray, pitch: int, nchans: int, dmshifts_: np.ndarray, dm_idx: Sequence[int],
                    t_processed: int, mstartdm: float, mdmstep: float, inBin: int) -> np.ndarray:
    ds = np.ascontiguousarray(dmshifts_, dtype=np.float32)
    idx = np.ascontiguousarray(dm_idx, dtype=np.int32)
    out = np.zeros((idx.size, t_processed), dtype=np.float32)
    lib().oracle_dedisperse_rows(_vp(level), pitch, nchans, _pf(ds), _pi(idx), idx.size, t_processed,
                                 mstartdm, mdmstep, inBin, _pf(out), t_processed)
    return out


def transpose_u16(input_tc: np.ndarray, nsamp: int, nchans: int) -> np.ndarray:
    """[t][c] u16 -> [c][t] u16 (samples wider than 8 bits keep their width)."""
    x = np.ascontiguousarray(input_tc, dtype=np.uint16)
    out = np.zeros((nchans, nsamp), dtype=np.uint16)
    lib().oracle_transpose_u16(_vp(x), nsamp, nchans, _vp(out), nsamp)
    return out


def bin2_u16(level: np.ndarray, n_out: int) -> np.ndarray:
    level = np.ascontiguousarray(level, dtype=np.uint16)
    out = np.zeros((level.shape[0], n_out), dtype=np.uint16)
    lib().oracle_bin_u16(_vp(level), level.shape[1], level.shape[0], n_out, _vp(out), n_out)
    return out


def dedisperse_rows_u16(level: np.ndarray, pitch: int, nchans: int, dmshifts_: np.ndarray, dm_idx: Sequence[int],
                        t_processed: int, mstartdm: float, mdmstep: float, inBin: int) -> np.ndarray:
    """DM rows of u16 [c][t] rows (pitch in samples): float sums in channel order."""
    lv = np.ascontiguousarray(level, dtype=np.uint16)
    ds = np.ascontiguousarray(dmshifts_, dtype=np.float32)
    idx = np.ascontiguousarray(dm_idx, dtype=np.int32)
    out = np.zeros((idx.size, t_processed), dtype=np.float32)
    lib().oracle_dedisperse_rows_u16(_vp(lv), pitch, nchans, _pf(ds), _pi(idx), idx.size, t_processed,
                                     mstartdm, mdmstep, inBin, _pf(out), t_processed)
    return out


def delay_table(dmshifts_: np.ndarray, ndms: int, mstartdm: float, mdmstep: float) -> np.ndarray:
    ds = np.ascontiguousarray(dmshifts_, dtype=np.float32)
    out = np.zeros((ndms, ds.size), dtype=np.int32)
    lib().oracle_delay_table(_pf(ds), ds.size, ndms, mstartdm, mdmstep, _pi(out))
    return out


def gen_noise(seed: int, nsamp: int, nchans: int, mean: float = 128.0, sigma: float = 16.0) -> np.ndarray:
    out = np.zeros((nsamp, nchans), dtype=np.uint8)
    lib().oracle_gen_noise_u8(seed & 0xFFFFFFFFFFFFFFFF, nsamp, nchans, mean, sigma, _vp(out))
    return out


def inject_pulse(data_tc: np.ndarray, dmshifts_: np.ndarray, dm: float, tsamp: float, t0: int, width: int,
                 amplitude: int) -> None:
    ds = np.ascontiguousarray(dmshifts_, dtype=np.float32)
    lib().oracle_inject_pulse_u8(_vp(data_tc), data_tc.shape[0], data_tc.shape[1], _pf(ds), dm, tsamp, t0, width,
                                 amplitude)


def aa_mask(sigma: float, length: int):
    m = np.zeros(length, dtype=np.float32)
    pos = lib().oracle_aa_mask(sigma, length, _pf(m))
    return m, pos


def aa_fake_signal(dmshifts_: np.ndarray, nsamples: int, tsamp: float, maxshift: int, dm_pos: float,
                   signal_start: int, width: int, sigma: float, period_ms: int = 0) -> np.ndarray:
    ds = np.ascontiguousarray(dmshifts_, dtype=np.float32)
    out = np.zeros((nsamples, ds.size), dtype=np.uint16)
    r = lib().oracle_aa_fake_signal_u16(_pf(ds), ds.size, nsamples, tsamp, maxshift, dm_pos, signal_start, width,
                                        sigma, period_ms, _vp(out))
    if r < 0:
        raise ValueError("fake signal too short for the plan")
    return out
