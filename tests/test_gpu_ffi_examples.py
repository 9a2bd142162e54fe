"""The reference's own callers, run against this library on a SIGPROC file
the test writes (8-bit, 128 channels, 1500 MHz down by 2 MHz per channel,
64 us, 60,000 samples, seeded noise plus two dispersed pulses):

  * the call sequence of the reference's python/dedisp_with_aa.py:31-80
    (aa_py_sigproc_input -> read_metadata/read_signal/input_buffer ->
    aa_py_dm ranges as a numpy array -> aa_py_ddtr_plan -> pipeline with
    output_dmt and copy_ddtr_data_to_host -> run() loop -> get_buffer);
  * the compiled C++ callers in astro-accelerate_amd/examples: one in the
    shape of examples/src/dedispersion.cpp (aa_pipeline_api<unsigned short>,
    planes written to a file) and one calling dedisperse_telescope_data as
    examples/src/filterbank_dedispersion.cpp does.

The DM-time planes must equal the oracle's bit for bit.
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "astro-accelerate_amd", "examples", "bin")
NCHANS, NSAMP, FCH1, FOFF, TSAMP = 128, 60000, 1500.0, -2.0, 64e-6
# the reference script's plan: start DMs, steps, downsampling, 24 DMs per call x subcalls
DDMS, STARTS, DOWNSAMPS, SUBCALLS = [0.2, 0.3, 0.5, 1.0], [0.0, 38.4, 60.0, 108.0], [1, 2, 4, 8], [8, 3, 4, 4]
RANGES = [(STARTS[i], STARTS[i] + 24 * SUBCALLS[i] * DDMS[i], DDMS[i], DOWNSAMPS[i], 1) for i in range(4)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fil(aa, oracle, tmp_path_factory):
    path = str(tmp_path_factory.mktemp("fil") / "aa_dm90.fil")
    s = oracle.strategy([(r[0], r[1], r[2], r[3], r[3]) for r in RANGES], NCHANS, NSAMP, FCH1, FOFF, TSAMP)
    assert s.ok and s.num_tchunks == 1
    data = oracle.gen_noise(90, NSAMP, NCHANS)
    oracle.inject_pulse(data, s.dmshifts, 25.0, TSAMP, 20000, 4, 40)
    oracle.inject_pulse(data, s.dmshifts, 150.0, TSAMP, 35000, 16, 40)
    aa.write_sigproc(path, 8, NCHANS, data.tobytes(), fch1=FCH1, foff=FOFF, tsamp=TSAMP)
    want = oracle.run(s, data)
    return path, s, want


def test_reference_python_script_sequence(aa, cuda, fil):
    path, s, want = fil
    say("dedisp_with_aa.py call sequence")
    sigproc_input = aa.aa_py_sigproc_input(path)
    metadata = sigproc_input.read_metadata()
    assert (metadata.m_nchans, metadata.m_nsamples, metadata.m_nbits) == (NCHANS, NSAMP, 8)
    assert sigproc_input.read_signal()
    input_buffer = sigproc_input.input_buffer()
    temp_list = [aa.aa_py_dm(*r) for r in RANGES]
    dm_list = np.array(temp_list, dtype=aa.aa_py_dm)
    ddtr_plan = aa.aa_py_ddtr_plan(dm_list)
    ddtr_plan.set_enable_msd_baseline_noise(False)
    ddtr_plan.print_info()
    pipeline_components = aa.aa_py_pipeline_components()
    pipeline_components.dedispersion = True
    pipeline_options = aa.aa_py_pipeline_component_options()
    pipeline_options.output_dmt = True
    pipeline_options.copy_ddtr_data_to_host = True
    pipeline = aa.aa_py_pipeline(pipeline_components, pipeline_options, metadata, input_buffer, 0)
    assert pipeline.bind_ddtr_plan(ddtr_plan)
    while pipeline.run():
        assert pipeline.status_code() != -1
    assert pipeline.status_code() == 0
    ts_inc, ddtr_output = pipeline.get_buffer()
    assert ts_inc == s.t_processed[0][0]
    assert pipeline.ddtr_range() == s.nranges
    assert pipeline.ddtr_ndms() == s.ndms
    for r in range(pipeline.ddtr_range()):
        assert pipeline.dm_low(r) == int(s.dm[r][0])
        n = ts_inc // DOWNSAMPS[r]
        got = np.array([np.ctypeslib.as_array(ddtr_output[r][d], shape=(n,)) for d in range(s.ndms[r])])
        np.testing.assert_array_equal(got, want[r][:, :n], err_msg=f"range {r}")
    pipeline.cleanUp()
    sigproc_input.close()


def _run(args):
    say("running " + " ".join(os.path.basename(a) for a in args))
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    say(p.stdout[-2000:])
    if p.returncode != 0:
        say(p.stderr[-2000:])
    return p


def test_cpp_example_dedispersion(cuda, fil, tmp_path):
    path, s, want = fil
    exe = os.path.join(BIN, "examples_dedispersion")
    assert os.path.exists(exe), "C++ examples not built (run __graft_entry__.build())"
    out = str(tmp_path / "planes.bin")
    args = [exe, path, out]
    for r in RANGES:
        args += [repr(r[0]), repr(r[1]), repr(r[2]), str(r[3])]
    p = _run(args)
    assert p.returncode == 0 and "NOTICE: Finished." in p.stdout
    raw = open(out, "rb").read()
    pos = 0
    for r in range(s.nranges):
        ndms, nsamp, ib = struct.unpack_from("<3i", raw, pos)
        low, step = struct.unpack_from("<2f", raw, pos + 12)
        pos += 20
        assert (ndms, ib) == (s.ndms[r], s.dm[r][3])
        assert nsamp == s.t_processed[0][0] // ib
        assert low == np.float32(s.dm[r][0]) and step == np.float32(s.dm[r][2])
        plane = np.frombuffer(raw, dtype=np.float32, count=ndms * nsamp, offset=pos).reshape(ndms, nsamp)
        pos += 4 * ndms * nsamp
        np.testing.assert_array_equal(plane, want[r][:, :nsamp], err_msg=f"range {r}")
    assert pos == len(raw)


def test_cpp_example_dedisperse_telescope_data(cuda, fil):
    path, _, _ = fil
    exe = os.path.join(BIN, "examples_filterbank_dedispersion")
    assert os.path.exists(exe), "C++ examples not built (run __graft_entry__.build())"
    args = [exe, path]
    for r in RANGES:
        args += [repr(r[0]), repr(r[1]), repr(r[2]), str(r[3])]
    p = _run(args)
    assert p.returncode == 0 and "NOTICE: Finished." in p.stdout
    assert "The pipeline finished successfully." in p.stdout + p.stderr
