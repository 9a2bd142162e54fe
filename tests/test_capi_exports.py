"""The built library exports every entry point its headers declare.

CPU-only: loads astro-accelerate_amd/lib/libaa_ddtr_hip.so and checks symbols
(ctypes for the extern "C" surface of include/aa_ddtr_hip.h, nm for the C++
drop-in functions of include/aa_*.hpp), then exercises the host-only entry
points; nothing here launches a kernel.
"""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "astro-accelerate_amd", "lib", "libaa_ddtr_hip.so")


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _c_functions(header):
    body = _strip_comments(open(os.path.join(INCLUDE, header)).read())
    body = re.sub(r"typedef\s+[^;{}]*\(\s*\*[^;]*;", "", body)   # function-pointer types are not entry points
    # a declaration: return type, name, '(' -- skip typedef'd struct names
    names = re.findall(r"^[ \t]*(?:const\s+)?[A-Za-z_][\w \t\*]*?\b([A-Za-z_]\w*)\s*\(", body, flags=re.M)
    return sorted({n for n in names if n not in ("if", "for", "while", "return", "sizeof", "defined")})


def _cpp_free_functions(header):
    body = _strip_comments(open(os.path.join(INCLUDE, header)).read())
    return re.findall(r"^\s*(?:void|int|bool)\s+([a-z_]\w*)\s*\(", body, flags=re.M)


def test_c_abi_exports_every_declared_function(aa):
    lib = aa.load_library()
    declared = _c_functions("aa_ddtr_hip.h")
    assert len(declared) >= 50, declared
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, f"declared in include/aa_ddtr_hip.h but not exported: {missing}"


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm not available")
def test_cpp_drop_in_functions_exported(aa):
    aa.load_library()
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    headers = ["aa_dedisperse.hpp", "aa_device_load_data.hpp", "aa_corner_turn.hpp", "aa_bin_gpu.hpp",
               "aa_pipeline_wrapper_functions.hpp"]
    names = [n for h in headers for n in _cpp_free_functions(h)]
    assert set(names) >= {"dedisperse", "load_chunk_data", "set_dedispersion_constants",
                          "corner_turn_SM_inplace", "bin_gpu"}, names
    missing = [n for n in names if f"astroaccelerate::{n}(" not in out]
    assert not missing, f"C++ drop-in functions not exported: {missing}"
    for sym in ("astroaccelerate::aa_ddtr_strategy::aa_ddtr_strategy(astroaccelerate::aa_ddtr_plan const&, "
                "astroaccelerate::aa_filterbank_metadata const&, unsigned long const&, bool const&, "
                "astroaccelerate::aa_device_info*)",):
        assert sym in out, sym


def test_version_and_last_error(aa):
    lib = aa.load_library()
    version = lib.aa_hip_version()
    assert isinstance(version, bytes) and version.strip()
    assert isinstance(lib.aa_hip_last_error(), bytes)


def test_host_only_plan_and_strategy_through_c_abi(aa):
    """Plan -> strategy over the C ABI without touching a device."""
    plan = aa.make_plan([(0.0, 150.0, 0.1, 1, 1), (150.0, 300.0, 0.2, 1, 1)])
    meta = aa.make_metadata(nchans=4096, nsamples=1048576, fch1=1760.0, foff=-810.0 / 4096, tsamp=64e-6)
    strat = aa.aa_py_ddtr_strategy(plan, meta, free_memory=280 << 30)
    assert strat.ready
    assert [strat.ndms(r) for r in range(strat.nranges)] == [1536, 768]
    assert strat.nranges == 2


def test_engine_create_without_device_reports_no_device(aa):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: the no-device path cannot be exercised")
    plan = aa.make_plan([(0.0, 100.0, 1.0, 1, 1)])
    meta = aa.make_metadata(nchans=64, nsamples=65536, fch1=1500.0, foff=-1.0, tsamp=64e-6)
    strat = aa.aa_py_ddtr_strategy(plan, meta, free_memory=8 << 30)
    assert strat.ready
    with pytest.raises(aa.AAHipError) as ei:
        aa.DdtrPlan.from_strategy(strat)
    assert ei.value.status == aa.AA_HIP_ERROR_NO_DEVICE
