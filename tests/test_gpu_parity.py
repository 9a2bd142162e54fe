"""GPU parity: every HIP path bit-exact against the CPU oracle.

All tests call the library through its C ABI (ctypes), on torch-allocated
device memory.  Sizes are chosen so the oracle finishes in seconds; the
full-size config-1 plane is compared whole.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# AA_HIP_KERNEL_DIRECT, _LDS (first chunk shape that fits), _LDS_C32, _LDS_C16, _LDS_PAIR
KERNELS = [1, 2, 3, 4, 5, 6]
KERNELS_FULL = [1, 2, 4, 5, 6]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def to_dev(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("nsamp,nchans", [(1000, 77), (4109, 130), (64, 64), (5, 3)])
def test_corner_turn_u8(aa, oracle, cuda, nsamp, nchans):
    import torch
    say(f"corner_turn_u8 nsamp={nsamp} nchans={nchans}")
    rng = np.random.default_rng(nsamp * 7 + nchans)
    x = rng.integers(0, 256, size=(nsamp, nchans), dtype=np.uint8)
    pitch = nsamp + 37
    out = torch.full((nchans, pitch), 0xAB, dtype=torch.uint8, device=cuda)
    aa.corner_turn_u8(to_dev(x, cuda), nsamp, nchans, out, pitch)
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got[:, :nsamp], oracle.corner_turn(x, nsamp, nchans))
    assert np.all(got[:, nsamp:] == 0xAB)  # nothing written past the row


def test_corner_turn_u16_and_flag(aa, oracle, cuda):
    import torch
    say("corner_turn_u16")
    rng = np.random.default_rng(4)
    x = rng.integers(0, 256, size=(3000, 100), dtype=np.uint16)
    out = torch.zeros((100, 3000), dtype=torch.uint8, device=cuda)
    err = torch.zeros(1, dtype=torch.int32, device=cuda)
    xd = torch.from_numpy(x.view(np.int16)).to(cuda)
    aa.corner_turn_u16(xd, 3000, 100, out, 3000, err)
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.corner_turn(x, 3000, 100))
    assert int(err.item()) == 0
    x[1234, 56] = 300
    xd = torch.from_numpy(x.view(np.int16)).to(cuda)
    aa.corner_turn_u16(xd, 3000, 100, out, 3000, err)
    assert int(err.item()) == 1


@pytest.mark.parametrize("nsamp,nchans,nlev,sample_bytes", [
    (4109, 144, 3, 1),     # ragged time tile (4109 = 16 x 256 + 13), partial channel tile (144 = 128 + 16)
    (256, 128, 1, 1),      # exactly one tile
    (1000, 16, 2, 1),      # one 16-channel group
    (70001, 4096, 3, 1),   # 4096 channels (32 channel tiles), odd length
    (3001, 112, 3, 2),     # u16 input holding 8-bit samples
    (777, 100, 2, 1),      # 100 channels: not a multiple of 16 -> byte path, levels left to bin2
])
def test_corner_turn_levels(aa, oracle, cuda, nsamp, nchans, nlev, sample_bytes):
    """The corner turn fused with the x2 cascade: level 0 is the transpose and
    level k the floor-average of level k-1, bit for bit with the oracle, and
    nothing is written past each level's nsamp >> k samples."""
    import torch
    say(f"corner_turn_levels nsamp={nsamp} nchans={nchans} nlev={nlev} bytes={sample_bytes}")
    rng = np.random.default_rng(nsamp + 31 * nchans)
    dt = np.uint8 if sample_bytes == 1 else np.uint16
    x = rng.integers(0, 256, size=(nsamp, nchans), dtype=dt)
    xd = to_dev(x if sample_bytes == 1 else x.view(np.int16), cuda)
    pitches = [((nsamp >> k) + 4096 + 255) // 256 * 256 for k in range(nlev + 1)]
    levels = [torch.full((nchans, p), 0xAB, dtype=torch.uint8, device=cuda) for p in pitches]
    err = torch.zeros(1, dtype=torch.int32, device=cuda)
    fused = aa.corner_turn_levels(xd, sample_bytes, nsamp, nchans, levels, pitches, err)
    assert fused == (nlev if nchans % 16 == 0 else 0)
    want = oracle.corner_turn(x, nsamp, nchans)
    for k in range(fused + 1):
        got = levels[k].cpu().numpy()
        n = nsamp >> k
        np.testing.assert_array_equal(got[:, :n], want, err_msg=f"level {k}")
        assert np.all(got[:, n:] == 0xAB), f"level {k} written past its {n} samples"
        if k < nlev:
            want = oracle.bin2(want, want.shape[1] // 2)
    for k in range(fused + 1, nlev + 1):
        assert np.all(levels[k].cpu().numpy() == 0xAB), "levels beyond the fused ones untouched"
    assert int(err.item()) == 0
    if sample_bytes == 2:
        x[nsamp // 2, nchans - 1] = 256
        aa.corner_turn_levels(to_dev(x.view(np.int16), cuda), 2, nsamp, nchans, levels, pitches, err)
        assert int(err.item()) == 1


@pytest.mark.parametrize("n_out,in_pitch,out_pitch", [(1000, 2048, 1024), (1001, 2003, 1001), (3, 8, 4)])
def test_bin2(aa, oracle, cuda, n_out, in_pitch, out_pitch):
    import torch
    say(f"bin2 n_out={n_out}")
    rng = np.random.default_rng(n_out)
    L = rng.integers(0, 256, size=(33, in_pitch), dtype=np.uint8)
    out = torch.zeros((33, out_pitch), dtype=torch.uint8, device=cuda)
    aa.bin2_u8(to_dev(L, cuda), in_pitch, 33, n_out, out, out_pitch)
    np.testing.assert_array_equal(out.cpu().numpy()[:, :n_out], oracle.bin2(L[:, :2 * n_out], n_out))


GEOMS = [
    # nchans, fch1, foff, tsamp, dm_low, dm_step, ndms, tp, inBin
    (64, 1400.0, -4.0, 64e-6, 0.0, 1.0, 64, 3000, 1),
    (200, 1550.0, -1.5, 64e-6, 153.6, 0.2, 128, 2240, 1),
    (512, 1550.0, -0.5859, 64e-6, 0.0, 0.5, 192, 4000, 1),
    (96, 900.0, -2.0, 1e-4, 300.0, 2.0, 70, 1001, 4),
    (77, 1200.0, -3.0, 64e-6, 0.0, 1.0, 48, 2500, 1),      # odd channel count, partial DM tile
    (33, 1400.0, -8.0, 64e-6, 20.0, 0.5, 16, 7000, 2),     # 33 channels: one full chunk + 1
    (64, 1400.0, -4.0, 64e-6, 0.0, 1.0, 64, 4033, 1),      # 2 x 2016 + 1 samples: one-sample last tile
    (128, 1400.0, -2.0, 64e-6, 0.0, 2.0, 100, 3500, 1),    # window offsets ~490: pair shape on 1008-sample tiles
    (96, 1200.0, -3.0, 64e-6, 0.0, 1.5, 70, 2017, 1),      # offsets ~730, near that shape's 772 limit
    (16384, 1500.0, -0.0244140625, 64e-6, 0.0, 0.2, 20, 2500, 1),  # above the reference's 8192-channel cap
]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("geom", GEOMS, ids=[f"{g[0]}ch_{g[6]}dm_{g[7]}t" for g in GEOMS])
def test_dedisperse_range(aa, oracle, cuda, geom, kernel):
    import torch
    nchans, fch1, foff, tsamp, low, step, ndms, tp, ib = geom
    say(f"dedisperse kernel={kernel} geom={geom}")
    ds = oracle.dmshifts(nchans, fch1, foff)
    tsr = np.float32(np.float32(tsamp) * np.float32(ib))
    ms, st = np.float32(np.float32(low) / tsr), np.float32(np.float32(step) / tsr)
    maxd = int(np.trunc(np.float32(ds[-1]) * np.float32(np.float64(ndms - 1) * st + ms))) + 1
    row_len = tp + maxd
    pitch = ((row_len + 4096 + 255) // 256) * 256
    rng = np.random.default_rng(nchans + ndms)
    L = np.zeros((nchans, pitch), dtype=np.uint8)
    L[:, :row_len] = rng.integers(0, 256, size=(nchans, row_len), dtype=np.uint8)
    want = oracle.dedisperse(L[:, :row_len], ds, ndms, tp, float(ms), float(st), ib)
    d_ds = to_dev(ds, cuda)
    out_pitch = tp + 13
    out = torch.full((ndms, out_pitch), -1.0, dtype=torch.float32, device=cuda)
    try:
        aa.dedisperse_u8(to_dev(L, cuda), pitch, row_len, ds, d_ds, ndms, tp, float(ms), float(st), ib, out,
                         out_pitch, kernel=kernel)
    except aa.AAHipError as e:
        # a forced chunk shape may not hold this geometry's windows: refused on the host
        assert kernel in (3, 5, 6) and e.status == aa.AA_HIP_ERROR_UNSUPPORTED, e
        return
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got[:, :tp], want)
    assert np.all(got[:, tp:] == -1.0)


def test_dedisperse_rejects_short_rows(aa, oracle, cuda):
    import torch
    say("dedisperse short rows rejected on the host")
    ds = oracle.dmshifts(64, 1400.0, -4.0)
    L = torch.zeros((64, 4096), dtype=torch.uint8, device=cuda)
    out = torch.zeros((64, 2000), dtype=torch.float32, device=cuda)
    with pytest.raises(aa.AAHipError):
        aa.dedisperse_u8(L, 4096, 2001, ds, to_dev(ds, cuda), 64, 2000, 0.0, 500.0, 1, out, 2000)


def _chunked_reference(oracle, ranges, nchans, nsamp, fch1, foff, tsamp, free, seed, pulses):
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=free)
    assert s.ok
    data = oracle.gen_noise(seed, nsamp, nchans)
    for dm, t0, w in pulses:
        oracle.inject_pulse(data, s.dmshifts, dm, tsamp, t0, w, 32)
    return s, data, oracle.run(s, data)


@pytest.mark.parametrize("kernel", KERNELS_FULL)
def test_execute_chunk_config1_full(aa, oracle, cuda, kernel):
    """512 ch x 131072 samples, DM 0-500/0.5: the whole plane, bit-exact."""
    import torch
    say(f"config-1 full plane kernel={kernel}")
    tsamp = 64e-6
    s, data, want = _chunked_reference(oracle, [(0, 500, 0.5, 1, 1)], 512, 131072, 1550.0, -0.5859, tsamp,
                                       280 * 2**30, 1, [(250.0, 60000, 8)])
    plan = aa.DdtrPlan(s.dmshifts, [(s.dm[0][0], s.dm[0][2], s.ndms[0], 1)], tsamp, s.maxshift, s.t_processed[0][0])
    plan.set_kernel(kernel)
    out = torch.empty((s.ndms[0], s.t_processed[0][0]), dtype=torch.float32, device=cuda)
    plan.execute_chunk(to_dev(data, cuda), [s.t_processed[0][0]], [out])
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got, want[0])
    # the injected pulse (DM 250 = trial 500) peaks at its arrival sample
    d, t = np.unravel_index(np.argmax(got), got.shape)
    assert d == 500 and 60000 <= t < 60008


@pytest.mark.parametrize("kernel", KERNELS_FULL)
def test_execute_chunks_multi_range_multi_chunk(aa, oracle, cuda, kernel):
    """Several chunks (small free memory), 4 ranges with x2/x4/x8 binning."""
    import torch
    say(f"multi-chunk multi-range kernel={kernel}")
    nchans, nsamp, tsamp = 96, 200000, 1.28e-4
    ranges = [(0, 40, 0.5, 1, 1), (40, 80, 1.0, 2, 2), (80, 160, 2.0, 4, 4), (160, 400, 4.0, 8, 8)]
    free = None
    for extra in range(200_000_000, 1_000_000, -250_000):  # largest budget that gives >= 3 chunks
        s0 = oracle.strategy(ranges, nchans, nsamp, 1200.0, -2.5, tsamp, free_memory=104857600 + extra)
        if s0.ok and s0.num_tchunks >= 3:
            free = 104857600 + extra
            break
    assert free is not None
    s, data, want = _chunked_reference(oracle, ranges, nchans, nsamp, 1200.0, -2.5, tsamp, free, 3,
                                       [(30.0, 50000, 4), (120.0, 90000, 16), (300.0, 20000, 40)])
    assert s.num_tchunks >= 3
    plan = aa.DdtrPlan(s.dmshifts, [(s.dm[r][0], s.dm[r][2], s.ndms[r], s.dm[r][3]) for r in range(s.nranges)],
                       tsamp, s.maxshift, max(s.t_processed[0]))
    plan.set_kernel(kernel)
    d_all = to_dev(data, cuda)
    inc = 0
    for k in range(s.num_tchunks):
        tps = [s.t_processed[r][k] for r in range(s.nranges)]
        outs = [torch.empty((s.ndms[r], tps[r]), dtype=torch.float32, device=cuda) for r in range(s.nranges)]
        plan.execute_chunk(d_all[inc:], tps, outs)
        for r in range(s.nranges):
            col = inc // s.dm[r][3]
            np.testing.assert_array_equal(outs[r].cpu().numpy(), want[r][:, col:col + tps[r]])
        inc += tps[0]


def _three_row_pair_chunks(oracle, s, r, ndm):
    """(DM tile, 16-channel chunk) records of range r in which a pair's relative delay takes three
    values over the tile's ndm trials while the chunk's rows leave room for a third pair row (two
    rows per pair of two or more values, 16 rows per chunk): the plan builds those pairs as three
    pair rows instead of two single-channel rows.  Delays from the oracle's table."""
    ms, st = s.range_scalars(r)
    delay = oracle.delay_table(s.dmshifts, s.ndms[r], ms, st).astype(np.int64)
    nch = delay.shape[1] // 2 * 2
    count = 0
    for d0 in range(0, s.ndms[r], ndm):
        rows = delay[np.minimum(np.arange(d0, d0 + ndm), s.ndms[r] - 1)]
        v = rows[:, 1:nch:2] - rows[:, 0:nch:2]
        nv = v.max(axis=0) - v.min(axis=0) + 1
        for k in range(0, nv.size, 8):
            chunk = nv[k:k + 8]
            count += int(np.any(chunk == 3) and np.minimum(chunk, 2).sum() < 16)
    return count


@pytest.mark.parametrize("variant", [2, 4, 5, 6])
def test_pair_shape_forced_with_saturated_stretch(aa, oracle, cuda, variant):
    """Each 16-byte-read pair shape (2: 1008-sample tiles, 4: 2016, 5: 1512, 6: 1008 with 32 DM
    trials per tile) on every range of a
    2-range plan over 512 channels (32 chunks: two windows of the 16-bit accumulators), on noise
    with a stretch of all-255 samples, where every 16-bit field reaches its 65,280 bound at each
    flush, and a pulse."""
    import torch
    say(f"pair shape {variant} forced on every range, saturated stretch")
    nchans, nsamp, tsamp = 512, 40000, 64e-6
    ranges = [(0, 100, 0.5, 1, 1), (100, 200, 1.0, 2, 2)]
    s = oracle.strategy(ranges, nchans, nsamp, 1550.0, -0.5859, tsamp)
    assert s.ok and s.num_tchunks == 1
    # both ranges hold pairs built as three pair rows
    assert all(_three_row_pair_chunks(oracle, s, r, 32 if variant == 6 else 16) > 0 for r in range(s.nranges))
    data = oracle.gen_noise(5 + variant, nsamp, nchans)
    data[9000:21000, :] = 255
    oracle.inject_pulse(data, s.dmshifts, 60.0, tsamp, 4000, 6, 40)
    want = oracle.run(s, data)
    plan = aa.DdtrPlan(s.dmshifts, [(s.dm[r][0], s.dm[r][2], s.ndms[r], s.dm[r][3]) for r in range(s.nranges)],
                       tsamp, s.maxshift, s.t_processed[0][0])
    tps = [s.t_processed[r][0] for r in range(s.nranges)]
    d_in = to_dev(data, cuda)
    outs = [torch.empty((s.ndms[r], tps[r]), dtype=torch.float32, device=cuda) for r in range(s.nranges)]
    plan.execute_chunk(d_in, tps, outs)   # ingest: the levels the ranges read
    for r in range(s.nranges):
        _, ms = plan.autotune_range(r, tps[r], outs[r], reps=1)
        assert ms[variant] > 0, f"shape {variant} does not fit range {r}: {ms}"
        plan.set_range_variant(r, variant)
        outs[r].fill_(-1.0)
    plan.execute_chunk(d_in, tps, outs)
    for r in range(s.nranges):
        got = outs[r].cpu().numpy()
        np.testing.assert_array_equal(got, want[r], err_msg=f"shape {variant} range {r}")
        ib = s.dm[r][3]   # inside the stretch every sum is 512 x 255: (sum / nchans) / inBin
        assert np.all(got[:, 9000 // ib:(21000 - s.maxshift) // ib] == np.float32(255.0) / np.float32(ib))


@pytest.mark.parametrize("ndms,nchans", [(70, 96), (48, 77), (33, 96)])
def test_shape6_ragged_dm_tiles(aa, oracle, cuda, ndms, nchans):
    """Shape 6 (32 DM trials per tile, two per wave) with a DM count that leaves the last tile
    partly empty (trials past ndms are neither summed into another row nor written), and a
    channel count whose last chunk holds 6 pairs and a single channel."""
    import torch
    say(f"shape 6 with {ndms} DM trials, {nchans} channels")
    fch1, foff, tsamp, low, step, tp = 1500.0, -1.0, 64e-6, 0.0, 0.5, 5000
    ds = oracle.dmshifts(nchans, fch1, foff)
    tsr = np.float32(tsamp)
    ms, st = np.float32(np.float32(low) / tsr), np.float32(np.float32(step) / tsr)
    maxd = int(np.trunc(np.float32(ds[-1]) * np.float32(np.float64(ndms - 1) * st + ms))) + 1
    nsamp = tp + maxd
    data = oracle.gen_noise(ndms, nsamp, nchans)
    want = oracle.dedisperse(oracle.corner_turn(data, nsamp, nchans), ds, ndms, tp, float(ms), float(st), 1)
    plan = aa.DdtrPlan(ds, [(low, step, ndms, 1)], tsamp, maxd, tp)
    d_in = to_dev(data, cuda)
    out = torch.full((ndms, tp), -1.0, dtype=torch.float32, device=cuda)
    plan.execute_chunk(d_in, [tp], [out])
    _, msv = plan.autotune_range(0, tp, out, reps=1)
    assert msv[6] > 0, f"shape 6 does not fit: {msv}"
    plan.set_range_variant(0, 6)
    out.fill_(-1.0)
    plan.execute_chunk(d_in, [tp], [out])
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def _metadata(aa, nchans, nsamp, fch1, foff, tsamp):
    return aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)


def test_ffi_pipeline_u16_matches_oracle(aa, oracle, cuda):
    """The reference FFI path: plan -> pipeline -> run() loop -> float*** buffer."""
    say("FFI pipeline u16")
    nchans, nsamp, tsamp, fch1, foff = 128, 120000, 64e-6, 1500.0, -2.0
    ranges = [(0, 100, 0.5, 1, 1), (100, 300, 1.0, 2, 2)]
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp)
    data = oracle.gen_noise(77, nsamp, nchans)
    oracle.inject_pulse(data, s.dmshifts, 170.0, tsamp, 30000, 12, 40)
    want = oracle.run(s, data)
    data16 = data.astype(np.uint16)

    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    opts.copy_ddtr_data_to_host = True
    plan = aa.make_plan(ranges)
    pipe = aa.aa_py_pipeline(comps, opts, _metadata(aa, nchans, nsamp, fch1, foff, tsamp), data16, 0)
    assert pipe.bind_ddtr_plan(plan)
    statuses = []
    while pipe.run():
        statuses.append(pipe.status_code())
    statuses.append(pipe.status_code())
    assert statuses[-2:] == [2, 0] and all(x == 1 for x in statuses[:-2])
    strat = aa.aa_py_ddtr_strategy(plan, _metadata(aa, nchans, nsamp, fch1, foff, tsamp), 280 * 2**30)
    planes = pipe.output_planes(strat)
    for r in range(s.nranges):
        np.testing.assert_array_equal(planes[r], want[r])
    assert pipe.ddtr_ndms() == s.ndms
    pipe.cleanUp()


def test_ffi_pipeline_rejects_out_of_range_u16(aa, cuda):
    say("FFI pipeline rejects samples > 255")
    nchans, nsamp, tsamp = 64, 40000, 64e-6
    data = np.full((nsamp, nchans), 100, dtype=np.uint16)
    data[5, 5] = 4000
    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    pipe = aa.aa_py_pipeline(comps, opts, _metadata(aa, nchans, nsamp, 1400.0, -4.0, tsamp), data, 0)
    assert pipe.bind_ddtr_plan(aa.make_plan([(0, 50, 1, 1, 1)]))
    assert not pipe.run()
    assert pipe.status_code() == -1


def test_ffi_pipeline_rejects_unsupported_option(aa, cuda):
    say("FFI pipeline rejects old RFI")
    data = np.zeros((40000, 64), dtype=np.uint16)
    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    opts.old_rfi = True
    pipe = aa.aa_py_pipeline(comps, opts, _metadata(aa, 64, 40000, 1400.0, -4.0, 64e-6), data, 0)
    assert pipe.bind_ddtr_plan(aa.make_plan([(0, 50, 1, 1, 1)]))
    assert not pipe.run()


def test_gpu_known_answer_255(aa, oracle, cuda):
    """The reference's fake-signal KAT through the FFI pipeline on the GPU."""
    say("GPU KAT 255")
    tsamp, nchans, fch1 = 6.4e-5, 128, 1550.0
    foff = -300.0 / nchans
    nsamples = int(2 / tsamp)
    s = oracle.strategy([(0, 250.0, 1, 1, 1)], nchans, nsamples, fch1, foff, tsamp)
    width = int(1 / (tsamp * 25))
    rng = np.random.default_rng(4242)
    for _ in range(3):
        dm_pos = float(rng.integers(0, 200))
        start = int(rng.integers(1, 10)) * 1000 + int(rng.integers(0, 10)) * 100 + int(rng.integers(0, 10)) * 10
        sig = oracle.aa_fake_signal(s.dmshifts, nsamples, tsamp, s.maxshift, dm_pos, start, width, 0.5)
        comps = aa.aa_py_pipeline_components()
        comps.dedispersion = True
        opts = aa.aa_py_pipeline_component_options()
        opts.copy_ddtr_data_to_host = True
        plan = aa.make_plan([(0, 250.0, 1, 1, 1)])
        pipe = aa.aa_py_pipeline(comps, opts, _metadata(aa, nchans, nsamples, fch1, foff, tsamp), sig, 0)
        assert pipe.bind_ddtr_plan(plan)
        while pipe.run():
            pass
        _, buf = pipe.get_buffer()
        assert buf[0][int(dm_pos)][start] == 255.0, (dm_pos, start)
        pipe.cleanUp()
