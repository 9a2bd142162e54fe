"""The reference's device-level C++ entry points, called in the order its
runner calls them (include/aa_permitted_pipelines_generic.hpp:687-812):
load_chunk_data -> corner_turn_SM_inplace -> per range
set_dedispersion_constants, bin_gpu when the range bins further, dedisperse.
The symbols are the C++ ones a runner compiled against the reference's
headers links to (astroaccelerate::..., called here through ctypes); the
DM-time planes must equal the CPU oracle's bit for bit.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SYM = {
    "load_chunk_data": "_ZN15astroaccelerate15load_chunk_dataEPtPKtiiiPfS3_i",
    "corner_turn_SM_inplace": "_ZN15astroaccelerate22corner_turn_SM_inplaceEPtS0_mm",
    "set_dedispersion_constants": "_ZN15astroaccelerate26set_dedispersion_constantsEii",
    "bin_gpu": "_ZN15astroaccelerate7bin_gpuEPtPfii",
    "dedisperse": "_ZN15astroaccelerate10dedisperseEiiPiPfPtS1_S1_iS1_S1_S1_PKiii",
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fns(aa):
    lib = aa.load_library()
    vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    f = {k: getattr(lib, v) for k, v in SYM.items()}
    f["load_chunk_data"].argtypes = [vp, vp, i, i, i, vp, vp, i]
    f["corner_turn_SM_inplace"].argtypes = [vp, vp, sz, sz]
    f["set_dedispersion_constants"].argtypes = [i, i]
    f["bin_gpu"].argtypes = [vp, vp, i, i]
    f["dedisperse"].argtypes = [i, i, vp, vp, vp, vp, vp, i, vp, vp, vp, vp, i, i]
    f["dedisperse"].restype = i
    for k in ("load_chunk_data", "corner_turn_SM_inplace", "set_dedispersion_constants", "bin_gpu"):
        f[k].restype = None
    return f


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _run_sequence(aa, oracle, cuda, f, ranges, foff=-1.0, seed=9):
    """One chunk through the runner's call sequence, each range checked."""
    import torch
    nchans, nsamp, tsamp, fch1 = 256, 131072, 64e-6, 1500.0
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp)
    assert s.ok and s.num_tchunks == 1
    data = oracle.gen_noise(seed, nsamp, nchans)
    oracle.inject_pulse(data, s.dmshifts, 100.0, tsamp, 40000, 6, 40)
    want = oracle.run(s, data)

    tp = [s.t_processed[r][0] for r in range(s.nranges)]
    maxshift0 = s.maxshift
    length = tp[0] + maxshift0
    host16 = np.ascontiguousarray(data[:length].astype(np.uint16))
    d_in = torch.zeros(nchans * length, dtype=torch.int16, device=cuda)
    big = max(max(s.ndms[r] * tp[r] for r in range(s.nranges)), nchans * length)
    d_out = torch.zeros(big, dtype=torch.float32, device=cuda)
    dmshifts = np.ascontiguousarray(s.dmshifts, dtype=np.float32)
    d_ds = torch.from_numpy(dmshifts).to(cuda)
    inbin = np.array([s.dm[r][3] for r in range(s.nranges)], dtype=np.int32)
    dm_low = np.array([s.dm[r][0] for r in range(s.nranges)], dtype=np.float32)
    dm_step = np.array([s.dm[r][2] for r in range(s.nranges)], dtype=np.float32)
    ndms = np.array(s.ndms, dtype=np.int32)

    f["load_chunk_data"](d_in.data_ptr(), _p(host16), tp[0], maxshift0, nchans, _p(dmshifts), d_ds.data_ptr(), 8)
    f["corner_turn_SM_inplace"](d_in.data_ptr(), d_out.data_ptr(), nchans, length)
    old_bin = 1
    for r in range(s.nranges):
        ts = np.array([np.float32(tsamp) * np.float32(inbin[r])], dtype=np.float32)
        maxshift = maxshift0 // int(inbin[r])
        f["set_dedispersion_constants"](tp[r], maxshift)
        if inbin[r] > old_bin:
            f["bin_gpu"](d_in.data_ptr(), d_out.data_ptr(), nchans, tp[r - 1] + maxshift * int(inbin[r]))
            old_bin = int(inbin[r])
        st = f["dedisperse"](r, tp[r], _p(inbin), _p(dmshifts), d_in.data_ptr(), d_out.data_ptr(), d_ds.data_ptr(),
                             nchans, _p(ts), _p(dm_low), _p(dm_step), _p(ndms), 8, 0)
        assert st == 0
        torch.cuda.synchronize()
        got = d_out[:s.ndms[r] * tp[r]].cpu().numpy().reshape(s.ndms[r], tp[r])
        np.testing.assert_array_equal(got, want[r][:, :tp[r]], err_msg=f"foff {foff} range {r}")


@pytest.mark.parametrize("ranges", [
    [(0, 120, 0.5, 1, 1), (120, 360, 1.0, 2, 2)],                       # pair shapes, then one halving
    [(0, 60, 0.25, 1, 1)],                                               # one range, no binning
])
def test_reference_runner_sequence_matches_oracle(aa, oracle, cuda, ranges):
    say(f"C++ drop-in sequence {ranges}")
    _run_sequence(aa, oracle, cuda, _fns(aa), ranges)


def test_runner_sequence_two_observations_same_geometry(aa, oracle, cuda):
    """A second observation with the same channel count, DM plan and tsamp
    but another band (foff) must not reuse the first one's delays."""
    say("C++ drop-in sequence: two observations, foff -1.0 then -0.75")
    f = _fns(aa)
    ranges = [(0, 120, 0.5, 1, 1), (120, 360, 1.0, 2, 2)]
    _run_sequence(aa, oracle, cuda, f, ranges, foff=-1.0, seed=9)
    _run_sequence(aa, oracle, cuda, f, ranges, foff=-0.75, seed=10)
