"""ctypes front end of libaa_ddtr_hip.so (MI355X dedispersion).

Two parts:

* The reference's Python interface for the dedispersion path
  (cmake/py_astro_accelerate.py.in): ``aa_py_dm``, ``aa_py_ddtr_plan``,
  ``aa_py_filterbank_metadata``, ``filterbank_metadata_struct``,
  ``aa_py_pipeline_components``, ``aa_py_pipeline_component_options`` and
  ``aa_py_pipeline`` keep their names, constructor arguments and methods, so a
  script written for the reference runs against this library.  Plus
  ``aa_py_ddtr_strategy`` (host-only strategy for explicit free memory).

* Device-level helpers over torch tensors (``DdtrPlan``, ``corner_turn_u8``,
  ``bin2_u8``, ``dedisperse_u8``, ``synth_noise_u8``, ``inject_pulse_u8``)
  calling the aa_hip_* entry points with raw device pointers and the current
  torch stream.  torch only provides device memory and streams.

The library must be built (``make -C astro-accelerate_amd``); loading fails
loudly when it is missing.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AA_DDTR_HIP_LIB", os.path.join(_HERE, "lib", "libaa_ddtr_hip.so"))

_lib = None


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    """Load (once) and return the HIP dedispersion library; raise if absent.

    A process must hold a single HIP runtime.  torch ships its own
    libamdhip64.so.7; when torch is importable it is imported first, so that
    the library's libamdhip64.so.7 dependency binds to the runtime torch
    already mapped (a second runtime in the same process does not see the
    device).
    """
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libaa_ddtr_hip.so not found at {p}; build it with `make -C astro-accelerate_amd`")
    try:
        import torch  # noqa: F401  (binds the HIP runtime before the library loads)
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    _declare(lib)
    if path is None:
        _lib = lib
    return lib


PFLOAT = ctypes.POINTER(ctypes.c_float)
PPFLOAT = ctypes.POINTER(PFLOAT)
PPPFLOAT = ctypes.POINTER(PPFLOAT)


class filterbank_metadata_struct(ctypes.Structure):
    """Same fields as the reference's filterbank_metadata_struct."""
    _fields_ = [
        ("m_tstart", ctypes.c_double),
        ("m_tsamp", ctypes.c_double),
        ("m_fch1", ctypes.c_double),
        ("m_foff", ctypes.c_double),
        ("m_nbits", ctypes.c_int),
        ("m_nsamples", ctypes.c_int),
        ("m_nchans", ctypes.c_int),
        ("m_src_raj", ctypes.c_double),
    ]


class aa_py_pipeline_components(ctypes.Structure):
    _fields_ = [("dedispersion", ctypes.c_bool), ("analysis", ctypes.c_bool),
                ("periodicity", ctypes.c_bool), ("fdas", ctypes.c_bool)]


class aa_py_pipeline_component_options(ctypes.Structure):
    _fields_ = [(n, ctypes.c_bool) for n in (
        "zero_dm", "zero_dm_with_outliers", "input_DDTR_normalization", "output_DDTR_normalization",
        "set_bandpass_average", "old_rfi", "copy_ddtr_data_to_host", "msd_baseline_noise", "output_dmt",
        "output_ffdot_plan", "output_fdas_list", "candidate_algorithm", "fdas_custom_fft", "fdas_inbin",
        "fdas_norm")]


class filterbank_header(ctypes.Structure):
    """aa_hip_filterbank_header: the SIGPROC header fields the reader keeps."""
    _fields_ = [(n, ctypes.c_int) for n in ("telescope_id", "machine_id", "data_type", "barycentric",
                                            "pulsarcentric", "nbits", "nsamples", "nchans", "nifs")] + \
               [(n, ctypes.c_double) for n in ("az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp",
                                               "fch1", "foff", "fchannel", "refdm", "period")]


def _declare(lib: ctypes.CDLL) -> None:
    vp, i, f, d, sz, b = (ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double,
                          ctypes.c_size_t, ctypes.c_bool)
    u64, i64 = ctypes.c_uint64, ctypes.c_int64
    pi = ctypes.POINTER(ctypes.c_int)
    sig = {
        "aa_hip_last_error": ([], ctypes.c_char_p),
        "aa_hip_version": ([], ctypes.c_char_p),
        "aa_py_ddtr_plan": ([], vp),
        "aa_py_ddtr_plan_delete": ([vp], None),
        "aa_py_ddtr_plan_add_dm": ([vp, f, f, f, i, i], b),
        "aa_py_ddtr_plan_range": ([vp], sz),
        "aa_py_ddtr_plan_set_power": ([vp, f], b),
        "aa_py_ddtr_plan_power": ([vp], f),
        "aa_py_ddtr_plan_set_enable_msd_baseline": ([vp, b], b),
        "aa_py_ddtr_plan_enable_msd_baseline_noise": ([vp], b),
        "aa_py_ddtr_plan_bind_bandpass_normalization": ([vp, PFLOAT, i], b),
        "aa_py_filterbank_metadata": ([d, d, i, i, d, d, i, d], vp),
        "aa_py_filterbank_metadata_delete": ([vp], None),
        "aa_py_filterbank_metadata_tstart": ([vp], d),
        "aa_py_pipeline_api": ([aa_py_pipeline_components, aa_py_pipeline_component_options,
                                filterbank_metadata_struct, ctypes.POINTER(ctypes.c_ushort), i], vp),
        "aa_py_pipeline_api_u8": ([aa_py_pipeline_components, aa_py_pipeline_component_options,
                                   filterbank_metadata_struct, ctypes.POINTER(ctypes.c_ubyte), i], vp),
        "aa_hip_py_pipeline_api_u8_capped": ([aa_py_pipeline_components, aa_py_pipeline_component_options,
                                              filterbank_metadata_struct, ctypes.POINTER(ctypes.c_ubyte), i, sz],
                                             vp),
        "aa_py_pipeline_api_delete": ([vp], None),
        "aa_py_pipeline_api_bind_ddtr_plan": ([vp, vp], b),
        "aa_py_pipeline_api_ddtr_strategy": ([vp], vp),
        "aa_py_ddtr_strategy_delete": ([vp], None),
        "aa_py_get_ddtr_nRanges": ([vp], sz),
        "aa_py_get_ndms_array": ([vp], pi),
        "aa_py_get_t_proc": ([vp, i, i], i),
        "aa_py_dm_low": ([vp, i], i),
        "aa_py_total_computed_samples": ([vp], i),
        "aa_py_pipeline_api_run": ([vp, pi], b),
        "aa_py_buffer": ([vp], PPPFLOAT),
        "aa_py_current_range": ([vp], i),
        "aa_py_current_time_chunk": ([vp], i),
        "aa_py_current_inc": ([vp], ctypes.c_long),
        "aa_py_cleanup": ([vp], b),
        "aa_hip_pipeline_set_ddtr_consumer": ([vp, vp, vp], b),
        "aa_py_sigproc_input": ([ctypes.c_char_p], vp),
        "aa_py_sigproc_input_del": ([vp], None),
        "aa_py_sigproc_input_read_metadata": ([vp], filterbank_metadata_struct),
        "aa_py_sigproc_input_read_signal": ([vp], b),
        "aa_py_sigproc_input_input_buffer": ([vp], ctypes.POINTER(ctypes.c_ushort)),
        "aa_hip_spd_workspace_bytes": ([], sz),
        "aa_hip_spd_first_pass": ([vp, sz, i, i, i, vp, vp, sz, vp, sz, vp, sz, vp], i),
        "aa_hip_spd_noalloc_scratch_bytes": ([], sz),
        "aa_hip_spd_first_pass_noalloc": ([vp, sz, i, i, i, vp, vp, sz, vp, sz, vp], i),
        "aa_py_ddtr_strategy_create": ([vp, filterbank_metadata_struct, sz, b], vp),
        "aa_py_ddtr_strategy_ready": ([vp], b),
        "aa_py_ddtr_strategy_maxshift": ([vp], i),
        "aa_py_ddtr_strategy_num_tchunks": ([vp], i),
        "aa_py_ddtr_strategy_max_ndms": ([vp], i),
        "aa_py_ddtr_strategy_nranges": ([vp], sz),
        "aa_py_ddtr_strategy_ndms": ([vp, i], i),
        "aa_py_ddtr_strategy_t_processed": ([vp, i, i], i),
        "aa_py_ddtr_strategy_dm": ([vp, i, PFLOAT, PFLOAT, PFLOAT, pi, pi], b),
        "aa_py_ddtr_strategy_dmshifts": ([vp, PFLOAT, i], b),
        "aa_hip_ddtr_create": ([i, i, PFLOAT, i, PFLOAT, PFLOAT, pi, pi, d, i, i,
                                ctypes.POINTER(vp)], i),
        "aa_hip_ddtr_create_u16": ([i, i, PFLOAT, i, PFLOAT, PFLOAT, pi, pi, d, i, i,
                                    ctypes.POINTER(vp)], i),
        "aa_hip_ddtr_destroy": ([vp], i),
        "aa_hip_ddtr_set_kernel": ([vp, i], i),
        "aa_hip_ddtr_device_bytes": ([vp], sz),
        "aa_hip_ddtr_execute_chunk": ([vp, vp, i, pi, ctypes.POINTER(vp), ctypes.POINTER(sz), vp], i),
        "aa_hip_ddtr_ingest": ([vp, vp, i, i, vp], i),
        "aa_hip_ddtr_run_range": ([vp, i, i, vp, sz, vp], i),
        "aa_hip_corner_turn_u8": ([vp, i, i, vp, sz, vp], i),
        "aa_hip_corner_turn_u16": ([vp, i, i, vp, sz, vp, vp], i),
        "aa_hip_bin2_u8": ([vp, sz, i, i, vp, sz, vp], i),
        "aa_hip_corner_turn_levels": ([vp, i, i, i, ctypes.POINTER(vp), ctypes.POINTER(sz), i, vp, pi, vp], i),
        "aa_hip_dedisperse_u8": ([vp, sz, i, i, PFLOAT, vp, i, i, f, f, i, vp, sz, i, vp], i),
        "aa_hip_synth_noise_u8": ([u64, i64, d, d, vp, vp], i),
        "aa_hip_inject_pulse_u8": ([vp, i, i, vp, d, d, i, i, i, vp], i),
        "aa_hip_ddtr_set_range_variant": ([vp, i, i], i),
        "aa_hip_ddtr_autotune_range": ([vp, i, i, vp, sz, i, vp, pi, PFLOAT], i),
        "aa_hip_ddtr_batch_create": ([i, ctypes.POINTER(vp), ctypes.POINTER(vp)], i),
        "aa_hip_ddtr_batch_destroy": ([vp], i),
        "aa_hip_ddtr_batch_execute": ([vp, ctypes.POINTER(vp), pi, ctypes.POINTER(vp), ctypes.POINTER(sz), i, vp],
                                      i),
        "aa_hip_ddtr_batch_stats": ([vp, pi, ctypes.POINTER(ctypes.c_long), pi], i),
        "aa_hip_ddtr_batch_ct_stats": ([vp, pi, pi], i),
        "aa_hip_refine_create": ([i, i, ctypes.POINTER(i64), ctypes.POINTER(i64), pi, pi, ctypes.POINTER(vp)], i),
        "aa_hip_refine_destroy": ([vp], i),
        "aa_hip_refine_set_loss": ([vp, i], i),
        "aa_hip_refine_run": ([vp, vp, vp, vp, vp, vp, vp, vp, i, f, f, i, f, f, vp], i),
        "aa_hip_frac_create": ([i, i, PFLOAT, i, PFLOAT, i, i, f, ctypes.POINTER(vp)], i),
        "aa_hip_frac_destroy": ([vp], i),
        "aa_hip_frac_grouping": ([vp, pi, pi], i),
        "aa_hip_frac_forward": ([vp, vp, i, vp, vp, vp], i),
        "aa_hip_frac_adjoint": ([vp, vp, vp, i, vp, vp, vp, vp], i),
        "aa_hip_zero_dm_u8": ([vp, i, i, vp, i, vp], i),
        "aa_hip_zero_dm_u16": ([vp, i, i, vp, i, vp], i),
        "aa_hip_zero_dm_outliers_u8": ([vp, ctypes.c_long, i, vp, i, vp], i),
        "aa_hip_zero_dm_outliers_u16": ([vp, ctypes.c_long, i, vp, i, vp], i),
        "aa_hip_input_ddtr_normalization_u8": ([vp, ctypes.c_long, i, i, vp], i),
        "aa_hip_input_ddtr_normalization_u16": ([vp, ctypes.c_long, i, i, vp], i),
        "aa_hip_output_ddtr_normalization": ([vp, i, ctypes.c_long, vp], i),
        "aa_hip_bandpass_average": ([vp, i, ctypes.c_size_t, ctypes.c_size_t, vp], i),
        "aa_hip_sigproc_open": ([ctypes.c_char_p, ctypes.POINTER(vp), ctypes.POINTER(filterbank_header)], i),
        "aa_hip_sigproc_read_u8": ([vp, i, ctypes.POINTER(ctypes.POINTER(ctypes.c_ubyte)), ctypes.POINTER(sz)], i),
        "aa_hip_sigproc_read_u16": ([vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_ushort)), ctypes.POINTER(sz)], i),
        "aa_hip_sigproc_close": ([vp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def _lib_or_raise() -> ctypes.CDLL:
    return load_library()


# aa_hip_status (include/aa_ddtr_hip.h)
AA_HIP_SUCCESS = 0
AA_HIP_ERROR_INVALID_ARGUMENT = -1
AA_HIP_ERROR_DEVICE = -2
AA_HIP_ERROR_OUT_OF_MEMORY = -3
AA_HIP_ERROR_UNSUPPORTED = -4
AA_HIP_ERROR_INPUT_RANGE = -5
AA_HIP_ERROR_NO_DEVICE = -6


class AAHipError(RuntimeError):
    """A negative aa_hip_status (``.status``), with the library's message."""

    def __init__(self, message: str, status: Optional[int] = None):
        super().__init__(message)
        self.status = status


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = _lib_or_raise().aa_hip_last_error().decode(errors="replace")
        raise AAHipError(f"{what} failed with status {status}: {msg}", status)


# ---------------------------------------------------------------- FFI mirror
class aa_py_dm:
    def __init__(self, low: float, high: float, step: float, inBin: int, outBin: int):
        self.m_low, self.m_high, self.m_step = low, high, step
        self.m_inBin, self.m_outBin = inBin, outBin

    def low(self):
        return self.m_low

    def high(self):
        return self.m_high

    def step(self):
        return self.m_step

    def inBin(self):
        return self.m_inBin

    def outBin(self):
        return self.m_outBin


class aa_py_ddtr_plan:
    """Same constructor as the reference: an array of aa_py_dm."""

    def __init__(self, dm):
        lib = _lib_or_raise()
        self.m_obj = lib.aa_py_ddtr_plan()
        self.m_dm = list(dm)
        for r in self.m_dm:
            if not isinstance(r, aa_py_dm):
                raise TypeError(f"Supplied dm is the wrong type, {type(r).__name__}")
            lib.aa_py_ddtr_plan_add_dm(self.m_obj, r.low(), r.high(), r.step(), int(r.inBin()), int(r.outBin()))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self.m_obj:
            _lib_or_raise().aa_py_ddtr_plan_delete(self.m_obj)
            self.m_obj = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pointer(self):
        return self.m_obj

    def set_power(self, power: float):
        return _lib_or_raise().aa_py_ddtr_plan_set_power(self.m_obj, power)

    def power(self):
        return _lib_or_raise().aa_py_ddtr_plan_power(self.m_obj)

    def set_enable_msd_baseline_noise(self, flag: bool):
        return _lib_or_raise().aa_py_ddtr_plan_set_enable_msd_baseline(self.m_obj, bool(flag))

    def enable_msd_baseline_noise(self):
        return _lib_or_raise().aa_py_ddtr_plan_enable_msd_baseline_noise(self.m_obj)

    def bind_bandpass_normalization(self, data, data_size: int):
        arr = np.ascontiguousarray(data, dtype=np.float32)
        return _lib_or_raise().aa_py_ddtr_plan_bind_bandpass_normalization(
            self.m_obj, arr.ctypes.data_as(PFLOAT), int(data_size))

    def range(self):
        return _lib_or_raise().aa_py_ddtr_plan_range(self.m_obj)

    def print_info(self):
        print("AA_PY_DDTR_PLAN INFORMATION:")
        for r in self.m_dm:
            print(f"     aa_py_ddtr_plan range: low {r.low()}, high {r.high()}, step {r.step()}, "
                  f"inBin {r.inBin()}, outBin {r.outBin()}")


class aa_py_filterbank_metadata:
    def __init__(self, tstart: float, tsamp: float, nbits: int, nsamples: int, fch1: float, foff: float,
                 nchans: int, src_raj: float = 0.0):
        self.m_obj = _lib_or_raise().aa_py_filterbank_metadata(tstart, tsamp, nbits, nsamples, fch1, foff,
                                                               nchans, src_raj)
        self.m_struct = filterbank_metadata_struct(tstart, tsamp, fch1, foff, nbits, nsamples, nchans, src_raj)

    def pointer(self):
        return self.m_obj

    def tstart(self):
        return _lib_or_raise().aa_py_filterbank_metadata_tstart(self.m_obj)

    def struct(self) -> filterbank_metadata_struct:
        return self.m_struct

    def __del__(self):
        try:
            if self.m_obj:
                _lib_or_raise().aa_py_filterbank_metadata_delete(self.m_obj)
                self.m_obj = None
        except Exception:
            pass


class aa_hip_ddtr_plane(ctypes.Structure):
    """One range's device DM-time plane handed to a pipeline's consumer
    (aa_hip_ddtr_plane in include/aa_ddtr_hip.h)."""
    _fields_ = [("range", ctypes.c_int), ("time_chunk", ctypes.c_int), ("d_plane", ctypes.c_void_p),
                ("ndms", ctypes.c_int), ("nsamp", ctypes.c_int), ("pitch", ctypes.c_size_t),
                ("dm_low", ctypes.c_float), ("dm_step", ctypes.c_float), ("inBin", ctypes.c_int),
                ("tsamp", ctypes.c_double), ("inc", ctypes.c_long), ("device", ctypes.c_int),
                ("stream", ctypes.c_void_p)]


DDTR_CONSUMER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(aa_hip_ddtr_plane), ctypes.c_void_p)


class aa_py_pipeline:
    """The reference's aa_py_pipeline for the dedispersion component.

    ``input_data`` is a ctypes pointer (or numpy array) of u16 samples in
    [t][c] order, as the reference's SIGPROC reader produces; a numpy uint8
    array selects the u8 entry point (aa_sigproc_input.read_signal_u8's
    buffer, pinned or not, is one).  ``max_device_bytes`` (u8 only, no
    reference counterpart) caps the device memory the strategy plans for,
    which sets the time-chunk size.
    """

    def __init__(self, pipeline: aa_py_pipeline_components, pipeline_options: aa_py_pipeline_component_options,
                 metadata: filterbank_metadata_struct, input_data, card_number: int, max_device_bytes: int = 0):
        lib = _lib_or_raise()
        self._keepalive = input_data
        if isinstance(input_data, np.ndarray):
            if input_data.dtype == np.uint8:
                ptr = input_data.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
                if max_device_bytes:
                    self.m_obj = lib.aa_hip_py_pipeline_api_u8_capped(pipeline, pipeline_options, metadata, ptr,
                                                                      int(card_number), int(max_device_bytes))
                else:
                    self.m_obj = lib.aa_py_pipeline_api_u8(pipeline, pipeline_options, metadata, ptr,
                                                           int(card_number))
                return
            if max_device_bytes:
                raise TypeError("max_device_bytes needs uint8 input")
            if input_data.dtype != np.uint16:
                raise TypeError("input_data must be uint16 or uint8")
            input_data = input_data.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort))
        self.m_obj = lib.aa_py_pipeline_api(pipeline, pipeline_options, metadata, input_data, int(card_number))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self.m_obj:
            _lib_or_raise().aa_py_pipeline_api_delete(self.m_obj)
            self.m_obj = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bind_ddtr_plan(self, plan: aa_py_ddtr_plan):
        return _lib_or_raise().aa_py_pipeline_api_bind_ddtr_plan(self.m_obj, plan.pointer())

    def ddtr_strategy(self):
        return _lib_or_raise().aa_py_pipeline_api_ddtr_strategy(self.m_obj)

    def ddtr_range(self):
        return _lib_or_raise().aa_py_get_ddtr_nRanges(self.m_obj)

    def ddtr_ndms(self):
        lib = _lib_or_raise()
        n = self.ddtr_range()
        p = lib.aa_py_get_ndms_array(self.m_obj)
        return [p[k] for k in range(n)]

    def run(self):
        self.m_status_code_c_int = ctypes.c_int()
        return bool(_lib_or_raise().aa_py_pipeline_api_run(self.m_obj, ctypes.byref(self.m_status_code_c_int)))

    def status_code(self):
        return self.m_status_code_c_int.value

    def get_buffer(self):
        lib = _lib_or_raise()
        return lib.aa_py_total_computed_samples(self.m_obj), lib.aa_py_buffer(self.m_obj)

    def dm_low(self, pos: int):
        return _lib_or_raise().aa_py_dm_low(self.m_obj, int(pos))

    def t_processed(self, r: int, chunk: int):
        return _lib_or_raise().aa_py_get_t_proc(self.m_obj, int(r), int(chunk))

    def current_range(self):
        return _lib_or_raise().aa_py_current_range(self.m_obj)

    def current_time_chunk(self):
        return _lib_or_raise().aa_py_current_time_chunk(self.m_obj)

    def current_inc(self):
        return _lib_or_raise().aa_py_current_inc(self.m_obj)

    def cleanUp(self):
        return bool(_lib_or_raise().aa_py_cleanup(self.m_obj))

    def set_ddtr_consumer(self, fn):
        """fn(plane: aa_hip_ddtr_plane) -> int, called once per (time chunk,
        range) with the range's device plane, where the reference's runner
        calls analysis_GPU; non-zero stops the pipeline.  None clears it."""
        lib = _lib_or_raise()
        if fn is None:
            self._consumer = None
            return bool(lib.aa_hip_pipeline_set_ddtr_consumer(self.m_obj, None, None))

        def trampoline(pl, _user):
            try:
                return int(fn(pl.contents))
            except Exception:   # an exception cannot cross the C frames: report it as a failure
                import traceback
                traceback.print_exc()
                return -1

        self._consumer = DDTR_CONSUMER(trampoline)   # kept alive with the pipeline
        return bool(lib.aa_hip_pipeline_set_ddtr_consumer(self.m_obj, ctypes.cast(self._consumer, ctypes.c_void_p),
                                                          None))

    def output_planes(self, strategy: "aa_py_ddtr_strategy") -> List[np.ndarray]:
        """Copies of the host output as numpy [ndms][total_t] arrays per range."""
        _, buf = self.get_buffer()
        if not buf:
            raise AAHipError("no host output (enable copy_ddtr_data_to_host)")
        planes = []
        for r in range(strategy.nranges):
            total = sum(strategy.t_processed(r, k) for k in range(strategy.num_tchunks))
            nd = strategy.ndms(r)
            # rows of one range are contiguous in this library
            base = np.ctypeslib.as_array(buf[r][0], shape=(nd * total,))
            planes.append(base.reshape(nd, total).copy())
        return planes


class aa_py_ddtr_strategy:
    """Strategy of a plan for explicit metadata and free device memory (host only)."""

    def __init__(self, plan: aa_py_ddtr_plan, metadata: filterbank_metadata_struct, free_memory: int,
                 enable_analysis: bool = False):
        lib = _lib_or_raise()
        self.m_obj = lib.aa_py_ddtr_strategy_create(plan.pointer(), metadata, int(free_memory), bool(enable_analysis))
        self.nchans = metadata.m_nchans
        self.tsamp = metadata.m_tsamp
        if not self.m_obj:
            raise AAHipError("strategy allocation failed")

    def __del__(self):
        try:
            if self.m_obj:
                _lib_or_raise().aa_py_ddtr_strategy_delete(self.m_obj)
                self.m_obj = None
        except Exception:
            pass

    @property
    def ready(self) -> bool:
        return bool(_lib_or_raise().aa_py_ddtr_strategy_ready(self.m_obj))

    @property
    def maxshift(self) -> int:
        return _lib_or_raise().aa_py_ddtr_strategy_maxshift(self.m_obj)

    @property
    def num_tchunks(self) -> int:
        return _lib_or_raise().aa_py_ddtr_strategy_num_tchunks(self.m_obj)

    @property
    def max_ndms(self) -> int:
        return _lib_or_raise().aa_py_ddtr_strategy_max_ndms(self.m_obj)

    @property
    def nranges(self) -> int:
        return _lib_or_raise().aa_py_ddtr_strategy_nranges(self.m_obj)

    def ndms(self, r: int) -> int:
        return _lib_or_raise().aa_py_ddtr_strategy_ndms(self.m_obj, int(r))

    def t_processed(self, r: int, chunk: int) -> int:
        return _lib_or_raise().aa_py_ddtr_strategy_t_processed(self.m_obj, int(r), int(chunk))

    def dm(self, r: int):
        low, high, step = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        ib, ob = ctypes.c_int(), ctypes.c_int()
        ok = _lib_or_raise().aa_py_ddtr_strategy_dm(self.m_obj, int(r), ctypes.byref(low), ctypes.byref(high),
                                                    ctypes.byref(step), ctypes.byref(ib), ctypes.byref(ob))
        if not ok:
            raise IndexError(r)
        return low.value, high.value, step.value, ib.value, ob.value

    def dmshifts(self) -> np.ndarray:
        out = np.zeros(self.nchans, dtype=np.float32)
        if not _lib_or_raise().aa_py_ddtr_strategy_dmshifts(self.m_obj, out.ctypes.data_as(PFLOAT), self.nchans):
            raise AAHipError("dmshifts unavailable")
        return out

    def range_scalars(self, r: int):
        """(mstartdm, mdmstep) of range r as float32, derived as the reference does."""
        low, _, step, ib, _ = self.dm(r)
        tsamp_r = np.float32(np.float32(self.tsamp) * np.float32(ib))
        return np.float32(np.float32(low) / tsamp_r), np.float32(np.float32(step) / tsamp_r)


class aa_py_sigproc_input:
    """The reference's Python SIGPROC reader (cmake/py_astro_accelerate.py.in,
    class aa_py_sigproc_input, over src/aa_py_sigproc_input.cpp):
    read_metadata() -> filterbank_metadata_struct, read_signal() -> bool,
    input_buffer() -> ctypes POINTER(c_ushort) to nsamples x nchans u16
    samples in [t][c] order, ready for aa_py_pipeline."""

    def __init__(self, path: str):
        self.m_obj = _lib_or_raise().aa_py_sigproc_input(ctypes.c_char_p(path.encode("utf-8")))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if getattr(self, "m_obj", None):
            _lib_or_raise().aa_py_sigproc_input_del(self.m_obj)
            self.m_obj = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def read_metadata(self) -> filterbank_metadata_struct:
        return _lib_or_raise().aa_py_sigproc_input_read_metadata(self.m_obj)

    def read_signal(self) -> bool:
        return bool(_lib_or_raise().aa_py_sigproc_input_read_signal(self.m_obj))

    def input_buffer(self):
        return _lib_or_raise().aa_py_sigproc_input_input_buffer(self.m_obj)


class aa_sigproc_input:
    """SIGPROC filterbank reader, the reference's aa_sigproc_input
    (include/aa_sigproc_input.hpp:17-79) over aa_hip_sigproc_*.  read_signal()
    widens to u16 as the reference does; read_signal_u8() keeps 1/2/4/8-bit
    data as u8 ([t][c], the dedispersion input), in pinned host memory when
    asked.  Buffers are numpy views valid while this object lives."""

    def __init__(self, path: str):
        self.path = path
        self._h = ctypes.c_void_p()
        self._hdr = filterbank_header()
        self._data = None

    def open(self) -> bool:
        if self._h:
            return True
        st = _lib_or_raise().aa_hip_sigproc_open(self.path.encode(), ctypes.byref(self._h), ctypes.byref(self._hdr))
        if st != 0:
            self._h = ctypes.c_void_p()
            self._error = AAHipError(f"aa_hip_sigproc_open: {_lib_or_raise().aa_hip_last_error().decode()}", st)
            return False
        return True

    def close(self) -> bool:
        if self._h:
            _lib_or_raise().aa_hip_sigproc_close(self._h)
            self._h = ctypes.c_void_p()
            self._data = None
        return True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def read_metadata(self) -> dict:
        """The header fields (a dict); raises AAHipError when the file cannot be read."""
        if not self.open():
            raise self._error
        return {n: getattr(self._hdr, n) for n, _ in filterbank_header._fields_}

    def metadata(self):
        """filterbank_metadata_struct of the header (geometry fields), as the pipeline takes it."""
        h = self.read_metadata()
        return make_metadata(h["nchans"], h["nsamples"], h["fch1"], h["foff"], h["tsamp"], nbits=h["nbits"],
                             tstart=h["tstart"])

    def read_signal(self) -> np.ndarray:
        """[nsamples, nchans] u16, as the reference's input_buffer()."""
        ptr, n = ctypes.POINTER(ctypes.c_ushort)(), ctypes.c_size_t()
        check(_lib_or_raise().aa_hip_sigproc_read_u16(self._h, ctypes.byref(ptr), ctypes.byref(n)), "read_u16")
        self._data = np.ctypeslib.as_array(ptr, shape=(n.value,)).reshape(self._hdr.nsamples, self._hdr.nchans)
        return self._data

    def read_signal_u8(self, pinned: bool = False) -> np.ndarray:
        """[nsamples, nchans] u8 (nbits <= 8); pinned host memory needs a GPU."""
        ptr, n = ctypes.POINTER(ctypes.c_ubyte)(), ctypes.c_size_t()
        check(_lib_or_raise().aa_hip_sigproc_read_u8(self._h, int(bool(pinned)), ctypes.byref(ptr), ctypes.byref(n)),
              "read_u8")
        self._data = np.ctypeslib.as_array(ptr, shape=(n.value,)).reshape(self._hdr.nsamples, self._hdr.nchans)
        return self._data


def write_sigproc(path: str, nbits: int, nchans: int, payload: bytes, nifs: int = 1, strings=None,
                  **fields) -> None:
    """Writes a SIGPROC filterbank file: HEADER_START, optional (keyword,
    string) records, the int and double keywords aa_sigproc_input reads (from
    ``fields``, with defaults), HEADER_END, then ``payload`` (records of
    nchans samples at nbits each, [t][c])."""
    import struct

    def rec(text: str) -> bytes:
        b = text.encode()
        return struct.pack("<i", len(b)) + b

    ints = dict(telescope_id=4, machine_id=10, data_type=1, nbits=nbits, nchans=nchans, nifs=nifs,
                nsamples=(len(payload) * 8) // max(1, nchans * nbits))
    dbls = dict(tstart=60000.5, tsamp=64e-6, fch1=1550.0, foff=-0.5859, src_raj=0.0, src_dej=0.0,
                az_start=0.0, za_start=0.0, refdm=0.0, period=0.0)
    for k, v in fields.items():
        if k in ints:
            ints[k] = v
        elif k in dbls:
            dbls[k] = v
        else:
            raise KeyError(k)
    h = rec("HEADER_START")
    for k, v in (strings or {}).items():
        h += rec(k) + rec(v)
    for k, v in ints.items():
        h += rec(k) + struct.pack("<i", int(v))
    for k, v in dbls.items():
        h += rec(k) + struct.pack("<d", float(v))
    h += rec("HEADER_END")
    with open(path, "wb") as fh:
        fh.write(h)
        fh.write(payload)


def make_plan(ranges: Sequence[Sequence[float]]) -> aa_py_ddtr_plan:
    """Plan from (low, high, step, inBin[, outBin]) tuples."""
    return aa_py_ddtr_plan([aa_py_dm(r[0], r[1], r[2], int(r[3]), int(r[4]) if len(r) > 4 else int(r[3]))
                            for r in ranges])


def make_metadata(nchans: int, nsamples: int, fch1: float, foff: float, tsamp: float, nbits: int = 8,
                  tstart: float = 0.0) -> filterbank_metadata_struct:
    return filterbank_metadata_struct(tstart, tsamp, fch1, foff, nbits, nsamples, nchans, 0.0)


# ------------------------------------------------------ device-level helpers
def _stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class DdtrPlan:
    """Device plan (aa_hip_ddtr handle) for a strategy's geometry.

    ``execute_chunk`` runs corner turn + binning cascade + every range on
    torch device tensors: the input is [t][c] uint8 (or int16/uint16-holding
    8-bit samples), outputs are float32 [ndms_r][>= t_processed_r] tensors.
    With ``sample_bits=16`` the plan takes [t][c] 16-bit samples of any value
    (an int16 tensor holding their bit patterns) and sums them in float in
    channel order, as the reference does for data wider than 8 bits.
    """

    def __init__(self, dmshifts: np.ndarray, ranges: Sequence[tuple], tsamp: float, maxshift: int,
                 max_t_processed0: int, device: int = 0, sample_bits: int = 8):
        lib = _lib_or_raise()
        self.dmshifts = np.ascontiguousarray(dmshifts, dtype=np.float32)
        self.nchans = int(self.dmshifts.size)
        self.ranges = [tuple(r) for r in ranges]  # (dm_low, dm_step, ndms, inBin)
        n = len(self.ranges)
        low = (ctypes.c_float * n)(*[r[0] for r in self.ranges])
        step = (ctypes.c_float * n)(*[r[1] for r in self.ranges])
        ndms = (ctypes.c_int * n)(*[int(r[2]) for r in self.ranges])
        inbin = (ctypes.c_int * n)(*[int(r[3]) for r in self.ranges])
        h = ctypes.c_void_p()
        if sample_bits not in (8, 16):
            raise ValueError("sample_bits must be 8 or 16")
        create = lib.aa_hip_ddtr_create if sample_bits == 8 else lib.aa_hip_ddtr_create_u16
        check(create(int(device), self.nchans, self.dmshifts.ctypes.data_as(PFLOAT), n, low, step,
                     ndms, inbin, float(tsamp), int(maxshift), int(max_t_processed0),
                     ctypes.byref(h)), "aa_hip_ddtr_create")
        self.sample_bits = sample_bits
        self.handle = h
        self.tsamp = tsamp
        self.maxshift = int(maxshift)

    @classmethod
    def from_strategy(cls, s: aa_py_ddtr_strategy, device: int = 0, sample_bits: int = 8) -> "DdtrPlan":
        ranges = []
        for r in range(s.nranges):
            low, _, step, ib, _ = s.dm(r)
            ranges.append((low, step, s.ndms(r), ib))
        max_tp0 = max(s.t_processed(0, k) for k in range(s.num_tchunks))
        return cls(s.dmshifts(), ranges, s.tsamp, s.maxshift, max_tp0, device, sample_bits)

    def close(self):
        if getattr(self, "handle", None):
            _lib_or_raise().aa_hip_ddtr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kernel(self, kernel: int):
        check(_lib_or_raise().aa_hip_ddtr_set_kernel(self.handle, int(kernel)), "set_kernel")

    def device_bytes(self) -> int:
        return _lib_or_raise().aa_hip_ddtr_device_bytes(self.handle)

    def execute_chunk(self, d_input, t_processed: Sequence[int], outs, stream=None):
        lib = _lib_or_raise()
        n = len(self.ranges)
        bits = 8 if d_input.element_size() == 1 else 16
        tp = (ctypes.c_int * n)(*[int(v) for v in t_processed])
        ptrs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        pitches = (ctypes.c_size_t * n)(*[o.stride(0) for o in outs])
        check(lib.aa_hip_ddtr_execute_chunk(self.handle, _ptr(d_input), bits, tp, ptrs, pitches,
                                            _stream_ptr(stream)), "execute_chunk")

    def ingest(self, d_input, t_processed0: int, stream=None):
        bits = 8 if d_input.element_size() == 1 else 16
        check(_lib_or_raise().aa_hip_ddtr_ingest(self.handle, _ptr(d_input), bits, int(t_processed0),
                                                 _stream_ptr(stream)), "ingest")

    def run_range(self, r: int, t_processed: int, out, stream=None):
        check(_lib_or_raise().aa_hip_ddtr_run_range(self.handle, int(r), int(t_processed), _ptr(out),
                                                    out.stride(0), _stream_ptr(stream)), "run_range")

    def set_range_variant(self, r: int, variant: int):
        """Preferred LDS chunk shape of range r (0: 32 channels, 1: 16 channels, 2: 16 channels in
        pre-summed pairs on 1008-sample tiles, 3: 16 channels in pre-summed groups of four, 4: shape 2
        on 2016-sample tiles, 5: shape 2 on 1512-sample tiles, 6: shape 2 with 32 DM trials per tile,
        -1: first that fits)."""
        check(_lib_or_raise().aa_hip_ddtr_set_range_variant(self.handle, int(r), int(variant)), "set_range_variant")

    def autotune_range(self, r: int, t_processed: int, out, reps: int = 2, stream=None):
        """Times each LDS chunk shape that fits range r; keeps the fastest.
        Returns (best variant or -1, [ms per variant, -1 = does not fit])."""
        best = ctypes.c_int()
        ms = (ctypes.c_float * LDS_SHAPES)()
        check(_lib_or_raise().aa_hip_ddtr_autotune_range(self.handle, int(r), int(t_processed), _ptr(out),
                                                         out.stride(0), int(reps), _stream_ptr(stream),
                                                         ctypes.byref(best), ms), "autotune_range")
        return best.value, [float(v) for v in ms]


class DdtrBatch:
    """Many beams (DdtrPlan handles) dedispersed together: byte-row ranges in one persistent
    launch, the others in launches of their own, optionally replayed from a hipGraph
    (aa_hip_ddtr_batch_*)."""

    def __init__(self, plans: Sequence["DdtrPlan"]):
        lib = _lib_or_raise()
        self.plans = list(plans)   # keep the beams alive as long as the batch
        arr = (ctypes.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        h = ctypes.c_void_p()
        check(lib.aa_hip_ddtr_batch_create(len(self.plans), arr, ctypes.byref(h)), "aa_hip_ddtr_batch_create")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            _lib_or_raise().aa_hip_ddtr_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, inputs, t_processed, outs, use_graph: bool = True, stream=None):
        """inputs[b]: [t][c] uint8 tensor; t_processed[b][r]; outs[b][r]: float32 [ndms][>= t_processed]."""
        lib = _lib_or_raise()
        flat_tp = [int(v) for tp in t_processed for v in tp]
        flat_out = [o for ob in outs for o in ob]
        n = len(flat_out)
        ins = (ctypes.c_void_p * len(inputs))(*[x.data_ptr() for x in inputs])
        tp = (ctypes.c_int * n)(*flat_tp)
        ptrs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in flat_out])
        pitches = (ctypes.c_size_t * n)(*[o.stride(0) for o in flat_out])
        check(lib.aa_hip_ddtr_batch_execute(self.handle, ins, tp, ptrs, pitches, int(bool(use_graph)),
                                            _stream_ptr(stream)), "aa_hip_ddtr_batch_execute")

    def stats(self):
        nj, nd = ctypes.c_int(), ctypes.c_int()
        nt = ctypes.c_long()
        check(_lib_or_raise().aa_hip_ddtr_batch_stats(self.handle, ctypes.byref(nj), ctypes.byref(nt),
                                                      ctypes.byref(nd)), "aa_hip_ddtr_batch_stats")
        cb, co = ctypes.c_int(), ctypes.c_int()
        check(_lib_or_raise().aa_hip_ddtr_batch_ct_stats(self.handle, ctypes.byref(cb), ctypes.byref(co)),
              "aa_hip_ddtr_batch_ct_stats")
        return {"jobs": nj.value, "tiles": nt.value, "direct_jobs": nd.value,
                "corner_turn_launches": cb.value, "corner_turns_alone": co.value}


def corner_turn_u8(d_in_tc, nsamp: int, nchans: int, d_out, out_pitch: int, stream=None):
    check(_lib_or_raise().aa_hip_corner_turn_u8(_ptr(d_in_tc), int(nsamp), int(nchans), _ptr(d_out),
                                                int(out_pitch), _stream_ptr(stream)), "corner_turn_u8")


def corner_turn_u16(d_in_tc, nsamp: int, nchans: int, d_out, out_pitch: int, d_err, stream=None):
    check(_lib_or_raise().aa_hip_corner_turn_u16(_ptr(d_in_tc), int(nsamp), int(nchans), _ptr(d_out),
                                                 int(out_pitch), _ptr(d_err), _stream_ptr(stream)),
          "corner_turn_u16")


def corner_turn_levels(d_in_tc, sample_bytes: int, nsamp: int, nchans: int, d_levels, pitches, d_err=None,
                       stream=None) -> int:
    """Corner turn plus len(d_levels) - 1 x2 binned levels in one pass; returns
    the number of binned levels written (0 when the shape takes the byte path)."""
    n = len(d_levels)
    ptrs = (ctypes.c_void_p * n)(*[x.data_ptr() for x in d_levels])
    pit = (ctypes.c_size_t * n)(*[int(p) for p in pitches])
    fused = ctypes.c_int(-1)
    check(_lib_or_raise().aa_hip_corner_turn_levels(_ptr(d_in_tc), int(sample_bytes), int(nsamp), int(nchans), ptrs,
                                                    pit, n - 1, _ptr(d_err) if d_err is not None else None,
                                                    ctypes.byref(fused), _stream_ptr(stream)), "corner_turn_levels")
    return fused.value


def bin2_u8(d_in, in_pitch: int, nchans: int, n_out: int, d_out, out_pitch: int, stream=None):
    check(_lib_or_raise().aa_hip_bin2_u8(_ptr(d_in), int(in_pitch), int(nchans), int(n_out), _ptr(d_out),
                                         int(out_pitch), _stream_ptr(stream)), "bin2_u8")


def dedisperse_u8(d_in, pitch: int, row_len: int, dmshifts: np.ndarray, d_dmshifts, ndms: int,
                  t_processed: int, mstartdm: float, mdmstep: float, inBin: int, d_out, out_pitch: int,
                  kernel: int = 0, stream=None):
    h = np.ascontiguousarray(dmshifts, dtype=np.float32)
    check(_lib_or_raise().aa_hip_dedisperse_u8(_ptr(d_in), int(pitch), int(row_len), int(h.size),
                                               h.ctypes.data_as(PFLOAT), _ptr(d_dmshifts), int(ndms),
                                               int(t_processed), float(mstartdm), float(mdmstep), int(inBin),
                                               _ptr(d_out), int(out_pitch), int(kernel), _stream_ptr(stream)),
          "dedisperse_u8")


def zero_dm(d_data_tc, d_bandpass, nbits: int = 8, stream=None):
    """Zero-DM filter in place on a [nsamp, nchans] uint8 or uint16 device tensor
    (d_bandpass: nchans float32 device values)."""
    nsamp, nchans = d_data_tc.shape
    fn = _lib_or_raise().aa_hip_zero_dm_u8 if d_data_tc.element_size() == 1 else _lib_or_raise().aa_hip_zero_dm_u16
    check(fn(_ptr(d_data_tc), int(nsamp), int(nchans), _ptr(d_bandpass), int(nbits), _stream_ptr(stream)), "zero_dm")


def zero_dm_outliers(d_data_tc, d_bandpass, nbits: int = 8, stream=None):
    """Zero-DM with outlier rejection in place on a [nsamp, nchans] uint8 or
    uint16 device tensor (d_bandpass: nchans float32 device values)."""
    nsamp, nchans = d_data_tc.shape
    L = _lib_or_raise()
    fn = L.aa_hip_zero_dm_outliers_u8 if d_data_tc.element_size() == 1 else L.aa_hip_zero_dm_outliers_u16
    check(fn(_ptr(d_data_tc), int(nsamp), int(nchans), _ptr(d_bandpass), int(nbits), _stream_ptr(stream)),
          "zero_dm_outliers")


def input_ddtr_normalization(d_data_tc, nbits: int = 8, stream=None):
    """Input DDTR normalisation in place on a [nsamp, nchans] uint8 or uint16
    device tensor: each channel's mean over the nsamp samples moved to
    (2^nbits - 1)/2 (the reference's zero_dm_normalization_dm)."""
    nsamp, nchans = d_data_tc.shape
    L = _lib_or_raise()
    fn = L.aa_hip_input_ddtr_normalization_u8 if d_data_tc.element_size() == 1 else \
        L.aa_hip_input_ddtr_normalization_u16
    check(fn(_ptr(d_data_tc), int(nsamp), int(nchans), int(nbits), _stream_ptr(stream)), "input_ddtr_normalization")


def output_ddtr_normalization(d_plane, stream=None):
    """Output DDTR normalisation in place on a [ndms, nsamp] float32 device
    tensor: each time sample's mean over DMs moved to 127.5 (the reference's
    post_DDTR_normalization)."""
    ndms, nsamp = d_plane.shape
    check(_lib_or_raise().aa_hip_output_ddtr_normalization(_ptr(d_plane), int(ndms), int(nsamp),
                                                           _stream_ptr(stream)), "output_ddtr_normalization")


SPD_MAX_WIDTH = 32   # AA_HIP_SPD_MAX_WIDTH


def spd_workspace(device=None):
    """A device workspace for spd_first_pass (uint8 tensor, 256-byte aligned)."""
    import torch
    n = int(_lib_or_raise().aa_hip_spd_workspace_bytes())
    return torch.empty(n, dtype=torch.uint8, device=device if device is not None else "cuda")


def spd_first_pass_raw(d_plane_ptr: int, pitch: int, ndms: int, nsamp: int, max_width: int, d_msd, d_snr=None,
                       d_width=None, work=None, stream_ptr=None):
    """aa_hip_spd_first_pass on a raw device plane pointer (e.g. a consumer's
    aa_hip_ddtr_plane); outputs are torch tensors: d_msd float32 [max_width][2],
    d_snr float32 [ndms][>= nsamp], d_width int16 [ndms][>= nsamp]."""
    if work is None:
        work = spd_workspace(d_msd.device)
    st = ctypes.c_void_p(stream_ptr) if stream_ptr is not None else _stream_ptr(None)
    check(_lib_or_raise().aa_hip_spd_first_pass(
        ctypes.c_void_p(d_plane_ptr), int(pitch), int(ndms), int(nsamp), int(max_width), _ptr(d_msd),
        _ptr(d_snr) if d_snr is not None else None, d_snr.stride(0) if d_snr is not None else 0,
        _ptr(d_width) if d_width is not None else None, d_width.stride(0) if d_width is not None else 0,
        _ptr(work), work.numel(), st), "spd_first_pass")


def spd_first_pass(d_plane, max_width: int = SPD_MAX_WIDTH, with_snr: bool = True, work=None, stream=None):
    """First single-pulse-search pass on a float32 [ndms][nsamp] device plane:
    returns (msd [max_width][2] = per-width mean and sd of the boxcar sums,
    snr [ndms][nsamp] float32, width [ndms][nsamp] int16) -- snr and width
    None when with_snr is False."""
    import torch
    ndms, nsamp = d_plane.shape
    dev = d_plane.device
    msd = torch.empty((max_width, 2), dtype=torch.float32, device=dev)
    snr = torch.empty((ndms, nsamp), dtype=torch.float32, device=dev) if with_snr else None
    width = torch.empty((ndms, nsamp), dtype=torch.int16, device=dev) if with_snr else None
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    spd_first_pass_raw(d_plane.data_ptr(), d_plane.stride(0), ndms, nsamp, max_width, msd, snr, width, work,
                       s.cuda_stream)
    return msd, snr, width


def spd_first_pass_noalloc(d_plane, msd, snr=None, width=None, max_width: int = SPD_MAX_WIDTH, stream=None):
    """aa_hip_spd_first_pass_noalloc: the first pass into caller-given output
    tensors (msd float32 [max_width][2], snr float32 [ndms][>= nsamp] or None,
    width int16 [ndms][>= nsamp] or None), with the library's static device
    workspace; nothing is allocated."""
    import torch
    ndms, nsamp = d_plane.shape
    s = stream if stream is not None else torch.cuda.current_stream(d_plane.device)
    check(_lib_or_raise().aa_hip_spd_first_pass_noalloc(
        ctypes.c_void_p(d_plane.data_ptr()), d_plane.stride(0), int(ndms), int(nsamp), int(max_width), _ptr(msd),
        _ptr(snr) if snr is not None else None, snr.stride(0) if snr is not None else 0,
        _ptr(width) if width is not None else None, width.stride(0) if width is not None else 0,
        ctypes.c_void_p(s.cuda_stream)), "spd_first_pass_noalloc")


def bandpass_average(data_tc: np.ndarray) -> np.ndarray:
    """The set_bandpass_average bandpass of a [nsamp, nchans] uint8/uint16
    host array: value i is the mean of the i-th run of nsamp consecutive
    samples of the [t][c] buffer, as the reference computes it."""
    x = np.ascontiguousarray(data_tc)
    if x.dtype not in (np.uint8, np.uint16):
        raise TypeError("bandpass_average takes uint8 or uint16 samples")
    nsamp, nchans = x.shape
    out = np.zeros(nchans, dtype=np.float32)
    check(_lib_or_raise().aa_hip_bandpass_average(x.ctypes.data, x.itemsize, nchans, nsamp, out.ctypes.data),
          "bandpass_average")
    return out


def synth_noise_u8(seed: int, d_out, mean: float = 128.0, sigma: float = 16.0, stream=None):
    check(_lib_or_raise().aa_hip_synth_noise_u8(int(seed) & 0xFFFFFFFFFFFFFFFF, int(d_out.numel()), float(mean),
                                                float(sigma), _ptr(d_out), _stream_ptr(stream)), "synth_noise_u8")


def inject_pulse_u8(d_data_tc, nsamp: int, nchans: int, d_dmshifts, dm: float, tsamp: float, t0: int,
                    width: int, amplitude: int, stream=None):
    check(_lib_or_raise().aa_hip_inject_pulse_u8(_ptr(d_data_tc), int(nsamp), int(nchans), _ptr(d_dmshifts),
                                                 float(dm), float(tsamp), int(t0), int(width), int(amplitude),
                                                 _stream_ptr(stream)), "inject_pulse_u8")


KERNEL_AUTO, KERNEL_DIRECT, KERNEL_LDS, KERNEL_LDS_C32, KERNEL_LDS_C16, KERNEL_LDS_PAIR, KERNEL_LDS_QUAD = range(7)
LDS_SHAPES = 7   # LDS chunk shapes 0-6 (set_range_variant, autotune_range)
