"""Samples wider than 8 bits on the GPU (16-bit filterbanks; 32-bit files
arrive scaled to u16): u16 levels and float channel sums in channel order,
bit for bit against the oracle's restatement of the reference's cache kernel
(src/aa_device_dedispersion_kernel.cu:422-450), in the regime where the float
sums round (sums past 2^24) and through every entry point: the handle
(16-bit LDS kernel and direct kernel), the C++ dedisperse() drop-in with
nbits 16, and the pipeline from a 16-bit SIGPROC file.
"""
import sys

import numpy as np
import pytest

from test_gpu_cpp_shims import _fns, _p
from test_sigproc import write_fil

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _u16_noise(seed, nsamp, nchans, lo=20000):
    return np.random.default_rng(seed).integers(lo, 65536, size=(nsamp, nchans)).astype(np.uint16)


GEOMS = [
    # nchans, nsamp, fch1, foff, tsamp, ranges
    (512, 40000, 1500.0, -0.5, 64e-6, [(0, 100, 0.5, 1, 1), (100, 300, 1.0, 2, 2), (300, 700, 2.0, 4, 4)]),
    (300, 9000, 1400.0, -1.0, 1e-4, [(0, 50, 0.25, 1, 1)]),                 # 300 channels: 8-channel chunk ragged
    (1024, 20000, 1600.0, -0.25, 64e-6, [(0, 40, 0.2, 1, 1), (40, 120, 0.4, 2, 2)]),
]


@pytest.mark.parametrize("g", range(len(GEOMS)))
@pytest.mark.parametrize("kernel", [0, 1])   # AUTO (16-bit LDS kernel), DIRECT
def test_wide_plan_matches_oracle(aa, oracle, cuda, g, kernel):
    import torch
    nchans, nsamp, fch1, foff, tsamp, ranges = GEOMS[g]
    say(f"16-bit plan geometry {g} ({nchans} ch x {nsamp}) kernel {kernel}")
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, nbits=16)
    assert s.ok and s.num_tchunks == 1
    x = _u16_noise(g + 10 * kernel, nsamp, nchans)
    want = oracle.run(s, x)
    # the regime rounds: some float sums differ from the exact integer sums
    assert nchans * 65535 > 2 ** 24
    plan = aa.DdtrPlan(s.dmshifts, [(s.dm[r][0], s.dm[r][2], s.ndms[r], s.dm[r][3]) for r in range(s.nranges)],
                       tsamp, s.maxshift, s.t_processed[0][0], sample_bits=16)
    plan.set_kernel(kernel)
    d_in = torch.from_numpy(x.view(np.int16)).to(cuda)
    tps = [s.t_processed[r][0] for r in range(s.nranges)]
    outs = [torch.full((s.ndms[r], tps[r] + 3), -1.0, dtype=torch.float32, device=cuda) for r in range(s.nranges)]
    plan.execute_chunk(d_in, tps, outs)
    torch.cuda.synchronize()
    for r in range(s.nranges):
        got = outs[r].cpu().numpy()
        np.testing.assert_array_equal(got[:, :tps[r]], want[r], err_msg=f"range {r}")
        assert np.all(got[:, tps[r]:] == -1.0), "written past t_processed"


def test_wide_plan_wide_spread_falls_back_to_direct(aa, oracle, cuda):
    """A DM step whose 16-DM tiles spread beyond the 16-bit kernel's rows
    (510 samples) runs on the direct kernel under AUTO, still bit-exact."""
    import torch
    nchans, nsamp, fch1, foff, tsamp = 64, 30000, 1400.0, -8.0, 64e-6
    ranges = [(0, 64, 4.0, 1, 1)]   # about 190 samples of delay per trial in the lowest channel
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, nbits=16)
    assert s.ok
    say(f"16-bit plan, wide spread: ndms {s.ndms}, maxshift {s.maxshift}")
    x = _u16_noise(3, nsamp, nchans, lo=0)
    want = oracle.run(s, x)
    plan = aa.DdtrPlan(s.dmshifts, [(s.dm[0][0], s.dm[0][2], s.ndms[0], 1)], tsamp, s.maxshift,
                       s.t_processed[0][0], sample_bits=16)
    out = torch.zeros((s.ndms[0], s.t_processed[0][0]), dtype=torch.float32, device=cuda)
    plan.execute_chunk(torch.from_numpy(x.view(np.int16)).to(cuda), [s.t_processed[0][0]], [out])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want[0])


def test_cpp_dedisperse_nbits16(aa, oracle, cuda):
    """The reference's runner sequence with nbits 16 and samples above 255:
    load_chunk_data -> corner_turn_SM_inplace -> set_dedispersion_constants ->
    bin_gpu -> dedisperse, every range bit-exact."""
    import torch
    f = _fns(aa)
    nchans, nsamp, tsamp, fch1, foff = 512, 60000, 64e-6, 1500.0, -0.5
    ranges = [(0, 100, 0.5, 1, 1), (100, 300, 1.0, 2, 2)]
    say("C++ drop-in sequence, nbits 16")
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, nbits=16)
    assert s.ok and s.num_tchunks == 1
    tp = [s.t_processed[r][0] for r in range(s.nranges)]
    length = tp[0] + s.maxshift
    x = _u16_noise(77, nsamp, nchans)
    want = oracle.run(s, x)
    host16 = np.ascontiguousarray(x[:length])
    d_in = torch.zeros(nchans * length, dtype=torch.int16, device=cuda)
    big = max(max(s.ndms[r] * tp[r] for r in range(s.nranges)), nchans * length)
    d_out = torch.zeros(big, dtype=torch.float32, device=cuda)
    dmshifts = np.ascontiguousarray(s.dmshifts, dtype=np.float32)
    d_ds = torch.from_numpy(dmshifts).to(cuda)
    inbin = np.array([s.dm[r][3] for r in range(s.nranges)], dtype=np.int32)
    dm_low = np.array([s.dm[r][0] for r in range(s.nranges)], dtype=np.float32)
    dm_step = np.array([s.dm[r][2] for r in range(s.nranges)], dtype=np.float32)
    ndms = np.array(s.ndms, dtype=np.int32)
    f["load_chunk_data"](d_in.data_ptr(), _p(host16), tp[0], s.maxshift, nchans, _p(dmshifts), d_ds.data_ptr(), 16)
    f["corner_turn_SM_inplace"](d_in.data_ptr(), d_out.data_ptr(), nchans, length)
    old_bin = 1
    for r in range(s.nranges):
        ts = np.array([np.float32(tsamp) * np.float32(inbin[r])], dtype=np.float32)
        maxshift = s.maxshift // int(inbin[r])
        f["set_dedispersion_constants"](tp[r], maxshift)
        if inbin[r] > old_bin:
            f["bin_gpu"](d_in.data_ptr(), d_out.data_ptr(), nchans, tp[r - 1] + maxshift * int(inbin[r]))
            old_bin = int(inbin[r])
        st = f["dedisperse"](r, tp[r], _p(inbin), _p(dmshifts), d_in.data_ptr(), d_out.data_ptr(), d_ds.data_ptr(),
                             nchans, _p(ts), _p(dm_low), _p(dm_step), _p(ndms), 16, 0)
        assert st == 0
        torch.cuda.synchronize()
        got = d_out[:s.ndms[r] * tp[r]].cpu().numpy().reshape(s.ndms[r], tp[r])
        np.testing.assert_array_equal(got, want[r][:, :tp[r]], err_msg=f"range {r}")


def test_sixteen_bit_sigproc_to_pipeline(aa, oracle, cuda, tmp_path):
    """A 16-bit SIGPROC file through the reader (u16, as the reference widens
    it) and the pipeline: the planes on the host equal the oracle's."""
    say("16-bit SIGPROC -> pipeline")
    nchans, nsamp, tsamp, fch1, foff = 512, 50000, 64e-6, 1500.0, -0.5
    ranges = [(0, 100, 0.5, 1, 1), (100, 300, 1.0, 2, 2), (300, 600, 2.0, 4, 4)]
    x = _u16_noise(5, nsamp, nchans)
    path = str(tmp_path / "b16.fil")
    write_fil(path, 16, nchans, x.astype("<u2").tobytes(), fch1=fch1, foff=foff, tsamp=tsamp)
    reader = aa.aa_sigproc_input(path)
    meta = reader.metadata()
    data = reader.read_signal()
    assert np.array_equal(data, x)
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, nbits=16)
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
        np.testing.assert_array_equal(planes[r], want[r], err_msg=f"range {r}")
    pipe.cleanUp()
    reader.close()
