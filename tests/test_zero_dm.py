"""Zero-DM filter: the oracle pinned by hand-worked answers and invariants
(CPU), then the device kernels (aa_hip_zero_dm_u8/u16) and the pipeline's
zero_dm option bit for bit against it (GPU).

Rule (src/aa_device_zero_dm_kernel.cu:37-88): per time sample,
x' = trunc(clamp((x - mean over channels) + bandpass[c], 0, 2^nbits - 1)).
"""
import sys

import numpy as np
import pytest


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------ oracle (CPU)
def test_oracle_hand_worked(oracle):
    # mean 25, default bandpass 127.5: 10-25+127.5 = 112.5 -> 112, ...
    assert oracle.zero_dm(np.array([[10, 20, 30, 40]], np.uint8)).tolist() == [[112, 122, 132, 142]]
    # clipping: mean 127.5, bandpass 0: -127.5 -> 0, 127.5 -> 127
    assert oracle.zero_dm(np.array([[0, 255]], np.uint8), bandpass=[0.0, 0.0]).tolist() == [[0, 127]]
    # clipping at the top: mean 0.5, bandpass 300: 299.5 / 300.5 -> 255
    assert oracle.zero_dm(np.array([[0, 1]], np.uint8), bandpass=[300.0, 300.0]).tolist() == [[255, 255]]
    # 4-bit data clamps at 15: mean 1, bandpass 7.5 -> 6.5 / 8.5 / 20.5 -> 6, 8, 15
    got = oracle.zero_dm(np.array([[0, 2, 1, 1]], np.uint8), bandpass=[7.5, 7.5, 20.5, 7.5], nbits=4)
    assert got.tolist() == [[6, 8, 15, 7]]
    # u16 with 16-bit clamp: mean 30000, bandpass 40000 -> 10000 + 40000 ... 50000 + 40000 -> 65535
    got = oracle.zero_dm(np.array([[10000, 50000]], np.uint16), bandpass=[40000.0, 40000.0], nbits=16)
    assert got.tolist() == [[20000, 60000]]
    got = oracle.zero_dm(np.array([[0, 60000]], np.uint16), bandpass=[40000.0, 40000.0], nbits=16)
    assert got.tolist() == [[10000, 65535]]


def test_oracle_invariants(oracle):
    rng = np.random.default_rng(5)
    x = rng.integers(60, 190, size=(50, 64), dtype=np.uint8)
    bp = rng.uniform(100.0, 150.0, size=64).astype(np.float32)
    base = oracle.zero_dm(x, bandpass=bp)
    # a constant offset per time sample cancels in (x - mean) while nothing clips
    # (64 channels: the mean is exact in float, so x - mean is the same number)
    shift = rng.integers(-50, 50, size=(50, 1))
    shifted = (x.astype(np.int64) + shift).astype(np.uint8)
    np.testing.assert_array_equal(oracle.zero_dm(shifted, bandpass=bp), base)
    # a constant row gives trunc(bandpass)
    const = np.full((3, 64), 77, np.uint8)
    np.testing.assert_array_equal(oracle.zero_dm(const, bandpass=bp), np.broadcast_to(np.trunc(bp), (3, 64)))
    # every time sample is filtered on its own: permuting rows permutes the output
    perm = rng.permutation(50)
    np.testing.assert_array_equal(oracle.zero_dm(x[perm], bandpass=bp), base[perm])


# --------------------------------------------------------- device (GPU)
CASES = [
    # nsamp, nchans, dtype, nbits, bandpass kind
    (1000, 256, np.uint8, 8, "default"),      # dword path
    (777, 100, np.uint8, 8, "custom"),         # dword path, odd rows
    (513, 99, np.uint8, 8, "custom"),          # byte path (nchans % 4 != 0)
    (300, 4096, np.uint8, 8, "clipping"),      # dword path, heavy clipping both ways
    (400, 64, np.uint8, 4, "default"),         # 4-bit values, clamp at 15
    (350, 130, np.uint16, 8, "custom"),        # u16 holding 8-bit values
    (200, 64, np.uint16, 16, "custom"),        # 16-bit values
    (0, 64, np.uint8, 8, "default"),           # empty
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}_{np.dtype(c[2]).name}_b{c[3]}_{c[4]}" for c in CASES])
def test_zero_dm_kernel_matches_oracle(aa, oracle, cuda, case):
    import torch
    nsamp, nchans, dt, nbits, kind = case
    say(f"zero-DM kernel {case}")
    rng = np.random.default_rng(nsamp * 7 + nchans)
    top = (1 << nbits) - 1
    x = rng.integers(0, top + 1, size=(nsamp, nchans)).astype(dt)
    if kind == "default":
        bp = np.full(nchans, top / 2.0, np.float32)
    elif kind == "custom":
        bp = rng.uniform(0.25 * top, 0.75 * top, size=nchans).astype(np.float32)
    else:
        bp = rng.uniform(-200.0, 450.0, size=nchans).astype(np.float32)
    want = oracle.zero_dm(x, bandpass=bp, nbits=nbits) if nsamp else x
    d = torch.from_numpy(x).to(cuda) if nsamp else torch.empty((0, nchans), dtype=torch.from_numpy(x).dtype,
                                                                  device=cuda)
    d_bp = torch.from_numpy(bp).to(cuda)
    aa.zero_dm(d, d_bp, nbits=nbits)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("custom_bandpass", [False, True])
def test_pipeline_zero_dm_matches_oracle(aa, oracle, cuda, custom_bandpass):
    """The pipeline's zero_dm option: each chunk filtered before its corner
    turn, over several time chunks (capped device memory)."""
    say(f"pipeline zero_dm custom_bandpass={custom_bandpass}")
    nchans, nsamp, tsamp, fch1, foff = 96, 200000, 1.28e-4, 1200.0, -2.5
    ranges = [(0, 40, 0.5, 1, 1), (40, 80, 1.0, 2, 2), (80, 160, 2.0, 4, 4)]
    free = None
    for extra in range(200_000_000, 1_000_000, -250_000):
        s0 = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=104857600 + extra)
        if s0.ok and s0.num_tchunks >= 2:
            free = 104857600 + extra
            break
    assert free is not None
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=free)
    data = oracle.gen_noise(33, nsamp, nchans)
    oracle.inject_pulse(data, s.dmshifts, 60.0, tsamp, 70000, 8, 40)
    # a broadband (zero-DM) burst that the filter removes
    data[120000:120040, :] = np.minimum(data[120000:120040, :].astype(np.int32) + 60, 255).astype(np.uint8)
    bp = (np.random.default_rng(2).uniform(110.0, 140.0, nchans).astype(np.float32) if custom_bandpass
          else np.full(nchans, 127.5, np.float32))
    # the pipeline filters each chunk's rows (t_processed + maxshift) independently; rows are
    # filtered one time sample at a time, so filtering the whole input once is the same
    want = oracle.run(s, oracle.zero_dm(data, bandpass=bp))

    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    opts.copy_ddtr_data_to_host = True
    opts.zero_dm = True
    plan = aa.make_plan(ranges)
    if custom_bandpass:
        assert plan.bind_bandpass_normalization(bp, nchans)
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    pipe = aa.aa_py_pipeline(comps, opts, md, np.ascontiguousarray(data), 0, max_device_bytes=free)
    assert pipe.bind_ddtr_plan(plan)
    while pipe.run():
        pass
    assert pipe.status_code() == 0
    strat = aa.aa_py_ddtr_strategy(plan, md, free)
    assert strat.num_tchunks == s.num_tchunks >= 2
    planes = pipe.output_planes(strat)
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])
    pipe.cleanUp()
