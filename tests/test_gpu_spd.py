"""GPU: the first single-pulse-search pass (aa_hip_spd_first_pass) against
the float64 oracle (oracle/spd_ref.py), and the pipeline's hand-off of each
range's device plane to a consumer (aa_hip_pipeline_set_ddtr_consumer).

Tolerances (float32 sums on the device, float64 in the oracle): MSD mean
and sd within 2e-6 relative (+1e-6 absolute); SNR within 2e-3 absolute;
the width must match wherever the oracle's best SNR beats the runner-up by
more than 2e-3.
"""
import sys

import numpy as np
import pytest

import spd_ref

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _all_snr(x, msd, W):
    ndms, nsamp = x.shape
    out = np.full((W, ndms, nsamp), -np.inf)
    for w in range(1, W + 1):
        b = spd_ref.boxcar_sums(x, w)
        m, s = msd[w - 1]
        out[w - 1, :, :nsamp - w + 1] = (b - m) / s if s > 0 else 0.0
    return out


def check_first_pass(x, W, msd, snr, width):
    """Compares device outputs (numpy) with the oracle on the float32 plane x."""
    x64 = x.astype(np.float64)
    want_msd = spd_ref.boxcar_msd(x64, W)
    np.testing.assert_allclose(msd, want_msd, rtol=2e-6, atol=1e-6)
    if snr is None:
        return
    all_snr = _all_snr(x64, want_msd, W)
    want_snr = all_snr.max(axis=0)
    want_w = all_snr.argmax(axis=0) + 1
    np.testing.assert_allclose(snr, want_snr, rtol=0, atol=2e-3)
    if W > 1:
        srt = np.sort(all_snr, axis=0)
        clear = (srt[-1] - srt[-2]) > 2e-3
    else:
        clear = np.ones(x.shape, dtype=bool)
    bad = np.argwhere(clear & (width != want_w))
    assert len(bad) == 0, f"{len(bad)} widths differ, first at {bad[0]}"
    assert np.all(width >= 1) and np.all(width <= W)
    # near the row end only the widths that fit: start sample t allows nsamp - t
    t = np.arange(x.shape[1])
    assert np.all(width <= np.minimum(W, x.shape[1] - t)[None, :])


PLANES = [
    # ndms, nsamp, W, kind
    (37, 1000, 32, "normal"),
    (5, 8192 + 31, 32, "normal"),        # exactly one full tile plus its halo
    (3, 3 * 8192 + 5, 7, "normal"),      # several tiles, short last tile
    (64, 20000, 32, "ddtr"),             # dedispersion-like values: 128 + small noise
    (16, 9000, 1, "normal"),
    (8, 50, 32, "pulse"),
    (1, 32, 32, "normal"),               # one start sample holds every width
]


@pytest.mark.parametrize("plane", PLANES, ids=[f"{p[0]}x{p[1]}_w{p[2]}_{p[3]}" for p in PLANES])
def test_first_pass_matches_oracle(aa, cuda, plane):
    import torch
    ndms, nsamp, W, kind = plane
    say(f"spd first pass {plane}")
    rng = np.random.default_rng(ndms * 7919 + nsamp)
    if kind == "ddtr":
        x = (128.0 + 0.25 * rng.normal(size=(ndms, nsamp))).astype(np.float32)
        x[10, 5000:5012] += 3.0
    else:
        x = rng.normal(1.5, 2.0, size=(ndms, nsamp)).astype(np.float32)
        if kind == "pulse":
            x[3, 10:18] += 6.0
    pitch = nsamp + 29
    d = torch.zeros((ndms, pitch), dtype=torch.float32, device=cuda)
    d[:, :nsamp] = torch.from_numpy(x).to(cuda)
    msd, snr, width = aa.spd_first_pass(d[:, :nsamp], W)
    torch.cuda.synchronize()
    check_first_pass(x, W, msd.cpu().numpy(), snr.cpu().numpy(), width.cpu().numpy().astype(np.int64))
    if kind == "pulse":
        assert int(width[3, 10].item()) == 8


def test_first_pass_msd_only_and_repeatable(aa, cuda):
    """MSD without the SNR planes; two calls give identical bits (the
    reduction order is fixed)."""
    import torch
    say("spd first pass: MSD only, twice")
    rng = np.random.default_rng(1)
    x = rng.normal(0.0, 1.0, size=(300, 40000)).astype(np.float32)
    d = torch.from_numpy(x).to(cuda)
    m1, s1, w1 = aa.spd_first_pass(d, 32, with_snr=False)
    m2, _, _ = aa.spd_first_pass(d, 32, with_snr=False)
    torch.cuda.synchronize()
    assert s1 is None and w1 is None
    assert torch.equal(m1, m2)
    check_first_pass(x, 32, m1.cpu().numpy(), None, None)


def test_first_pass_rejects_bad_arguments(aa, cuda):
    import torch
    say("spd first pass: argument checks")
    d = torch.zeros((4, 100), dtype=torch.float32, device=cuda)
    for W in (0, 33):
        with pytest.raises(aa.AAHipError):
            aa.spd_first_pass(d, W)
    with pytest.raises(aa.AAHipError):
        aa.spd_first_pass(d[:, :10], 11)   # widest boxcar longer than the row


def test_pipeline_hands_every_plane_to_the_consumer(aa, oracle, cuda):
    """Multi-chunk, multi-range FFI pipeline with a consumer that runs the
    first pass on each plane it is handed: the metadata match the strategy,
    and the first pass of each handed plane equals the oracle's first pass of
    the same (chunk, range) columns of the oracle's DM-time planes."""
    import torch
    say("pipeline -> consumer hand-off")
    nchans, nsamp, tsamp, fch1, foff = 128, 160000, 64e-6, 1500.0, -2.0
    ranges = [(0, 60, 0.5, 1, 1), (60, 180, 1.0, 2, 2)]
    data = oracle.gen_noise(2024, nsamp, nchans)
    s_full = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp)
    oracle.inject_pulse(data, s_full.dmshifts, 40.0, tsamp, 50000, 6, 40)
    md = aa.make_metadata(nchans, nsamp, fch1, foff, tsamp)
    plan = aa.make_plan(ranges)
    cap = None
    for mb in range(400, 20, -2):   # a device budget that gives at least two chunks
        st = aa.aa_py_ddtr_strategy(plan, md, mb << 20)
        if st.ready and st.num_tchunks >= 2:
            cap = mb << 20
            break
    assert cap is not None
    s = oracle.strategy(ranges, nchans, nsamp, fch1, foff, tsamp, free_memory=cap)
    assert s.ok and s.num_tchunks >= 2
    want = oracle.run(s, data)

    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    pipe = aa.aa_py_pipeline(comps, opts, md, data, 0, max_device_bytes=cap)
    work = aa.spd_workspace(cuda)
    seen = []

    def consumer(pl):
        W = 16
        msd = torch.empty((W, 2), dtype=torch.float32, device=cuda)
        snr = torch.empty((pl.ndms, pl.nsamp), dtype=torch.float32, device=cuda)
        wid = torch.empty((pl.ndms, pl.nsamp), dtype=torch.int16, device=cuda)
        aa.spd_first_pass_raw(pl.d_plane, pl.pitch, pl.ndms, pl.nsamp, W, msd, snr, wid, work, pl.stream)
        seen.append(dict(range=pl.range, chunk=pl.time_chunk, ndms=pl.ndms, nsamp=pl.nsamp, pitch=pl.pitch,
                         inBin=pl.inBin, inc=pl.inc, dm_low=pl.dm_low, dm_step=pl.dm_step, tsamp=pl.tsamp,
                         msd=msd, snr=snr, width=wid))
        return 0

    assert pipe.set_ddtr_consumer(consumer)
    assert pipe.bind_ddtr_plan(plan)
    while pipe.run():
        pass
    assert pipe.status_code() == 0
    torch.cuda.synchronize()
    assert [(x["chunk"], x["range"]) for x in seen] == [(k, r) for k in range(s.num_tchunks)
                                                       for r in range(s.nranges)]
    inc = 0
    for x in seen:
        k, r = x["chunk"], x["range"]
        if r == 0 and k > 0:
            inc += s.t_processed[0][k - 1]
        ib = s.dm[r][3]
        assert (x["ndms"], x["nsamp"], x["pitch"], x["inBin"], x["inc"]) == \
            (s.ndms[r], s.t_processed[r][k], s.t_processed[r][k], ib, inc)
        assert x["dm_low"] == pytest.approx(s.dm[r][0]) and x["dm_step"] == pytest.approx(s.dm[r][2])
        assert x["tsamp"] == pytest.approx(float(np.float32(tsamp) * np.float32(ib)))
        cols = want[r][:, inc // ib: inc // ib + x["nsamp"]]
        check_first_pass(cols, 16, x["msd"].cpu().numpy(), x["snr"].cpu().numpy(),
                         x["width"].cpu().numpy().astype(np.int64))
    pipe.cleanUp()


def test_pipeline_consumer_error_stops_the_run(aa, oracle, cuda):
    say("pipeline consumer error")
    nchans, nsamp, tsamp = 64, 40000, 64e-6
    data = oracle.gen_noise(5, nsamp, nchans)
    comps = aa.aa_py_pipeline_components()
    comps.dedispersion = True
    opts = aa.aa_py_pipeline_component_options()
    pipe = aa.aa_py_pipeline(comps, opts, aa.make_metadata(nchans, nsamp, 1400.0, -4.0, tsamp), data, 0)
    assert pipe.set_ddtr_consumer(lambda pl: 7)
    assert pipe.bind_ddtr_plan(aa.make_plan([(0, 50, 1, 1, 1)]))
    assert not pipe.run()
    assert pipe.status_code() == -1
    pipe.cleanUp()
