"""SIGPROC file -> u8 reader (pinned) -> FFI pipeline over several time
chunks (the next chunk's input copy overlapping the current chunk's ranges)
-> pinned host DM-time output, bit for bit against the CPU oracle.
"""
import sys

import numpy as np
import pytest

from test_sigproc import write_fil

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("pinned", [True, False])
def test_sigproc_to_pipeline_multi_chunk(aa, oracle, cuda, tmp_path, pinned):
    say(f"sigproc -> pipeline, several chunks, pinned={pinned}")
    nchans, nsamp, tsamp, fch1, foff = 96, 200000, 1.28e-4, 1200.0, -2.5
    ranges = [(0, 40, 0.5, 1, 1), (40, 80, 1.0, 2, 2), (80, 160, 2.0, 4, 4), (160, 400, 4.0, 8, 8)]
    free = None
    for extra in range(200_000_000, 1_000_000, -250_000):   # largest budget that gives >= 3 chunks
        s0 = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=104857600 + extra)
        if s0.ok and s0.num_tchunks >= 3:
            free = 104857600 + extra
            break
    assert free is not None
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=free)
    data = oracle.gen_noise(21, nsamp, nchans)
    for dm, t0, w in [(30.0, 50000, 4), (120.0, 90000, 16), (300.0, 20000, 40)]:
        oracle.inject_pulse(data, s.dmshifts, dm, tsamp, t0, w, 32)
    want = oracle.run(s, data)

    path = str(tmp_path / "obs.fil")
    write_fil(path, 8, nchans, np.ascontiguousarray(data).tobytes(), fch1=fch1, foff=foff, tsamp=tsamp)
    reader = aa.aa_sigproc_input(path)
    meta = reader.metadata()
    buf = reader.read_signal_u8(pinned=pinned)
    assert buf.shape == (nsamp, nchans) and np.array_equal(buf, data)

    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    opts.copy_ddtr_data_to_host = True
    plan = aa.make_plan(ranges)
    pipe = aa.aa_py_pipeline(comps, opts, meta, buf, 0, max_device_bytes=free)
    assert pipe.bind_ddtr_plan(plan)
    strat = aa.aa_py_ddtr_strategy(plan, meta, free)
    assert strat.num_tchunks == s.num_tchunks >= 3
    steps = 0
    while pipe.run():
        steps += 1
    assert pipe.status_code() == 0
    assert steps == s.num_tchunks * s.nranges + 1   # one call per (chunk, range), then finished_component
    planes = pipe.output_planes(strat)
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])
    pipe.cleanUp()
    reader.close()


@pytest.mark.parametrize("nbits", [1, 2, 4])
def test_sub_byte_sigproc_to_pipeline(aa, oracle, cuda, tmp_path, nbits):
    """1/2/4-bit files: the reader unpacks them to one byte per sample and the
    same kernels dedisperse them (exact sums)."""
    say(f"{nbits}-bit SIGPROC -> pipeline")
    nchans, nsamp, tsamp, fch1, foff = 64, 60000, 64e-6, 1400.0, -4.0
    ranges = [(0, 60, 0.5, 1, 1), (60, 180, 1.0, 2, 2)]
    rng = np.random.default_rng(nbits)
    packed = rng.integers(0, 256, size=nsamp * nchans * nbits // 8, dtype=np.uint8).tobytes()
    path = str(tmp_path / f"b{nbits}.fil")
    write_fil(path, nbits, nchans, packed, fch1=fch1, foff=foff, tsamp=tsamp)
    reader = aa.aa_sigproc_input(path)
    meta = reader.metadata()
    data = reader.read_signal_u8(pinned=True)
    assert data.max() <= (1 << nbits) - 1
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp)
    want = oracle.run(s, np.ascontiguousarray(data))
    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    opts.copy_ddtr_data_to_host = True
    plan = aa.make_plan(ranges)
    pipe = aa.aa_py_pipeline(comps, opts, meta, data, 0)
    assert pipe.bind_ddtr_plan(plan)
    while pipe.run():
        pass
    assert pipe.status_code() == 0
    strat = aa.aa_py_ddtr_strategy(plan, meta, 280 * 2**30)
    planes = pipe.output_planes(strat)
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])
    pipe.cleanUp()
    reader.close()
