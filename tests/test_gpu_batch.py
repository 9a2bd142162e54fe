"""Batched multi-beam dedispersion (aa_hip_ddtr_batch_*) on the GPU.

Every beam of a batch is checked bit for bit against the CPU oracle run on
the same seeded input, and the batched launches (eager and replayed from
their hipGraph) against the per-beam launches.  The beams mix channel
counts, lengths, DM plans, forced byte-row chunk shapes (the persistent
byte-row kernel), ranges left to the automatic choice (pair shapes, one
batched launch per tile and row width) and a range no shape fits (the
direct kernel, launched on its own).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_specs():
    """Seven small beams: ds up to ~2.6 samples/DM (100 MHz below 1760 MHz) or
    the full 810 MHz band for two of them (large delay spreads); the last
    has 40 channels, which the batched corner turn does not take."""
    narrow = dict(fch1=1760.0, tsamp=64e-6)
    specs = [
        dict(nchans=64, nsamp=16384, ranges=[(0, 200, 0.5, 1, 1)], bw=100.0, seed=11,
             pulses=[(120.0, 0.3, 4)]),
        dict(nchans=128, nsamp=32768, ranges=[(0, 150, 0.5, 1, 1), (150, 450, 1.0, 2, 2)], bw=100.0, seed=12,
             pulses=[(90.0, 0.2, 2), (400.0, 0.5, 8)]),
        dict(nchans=96, nsamp=16384, ranges=[(0, 50, 0.25, 1, 1), (50, 150, 0.5, 2, 2), (150, 300, 1.0, 4, 4)],
             bw=100.0, seed=13, pulses=[(200.0, 0.1, 16)]),
        dict(nchans=256, nsamp=32768, ranges=[(0, 300, 1.0, 1, 1)], bw=100.0, seed=14, pulses=[(33.0, 0.4, 1)]),
        # full band: 16-channel shape needed for range 1 (step/inBin 0.6: 15 trials span ~440 samples)
        dict(nchans=64, nsamp=65536, ranges=[(0, 100, 0.25, 1, 1), (100, 300, 1.2, 2, 2)], bw=810.0, seed=15,
             pulses=[(250.0, 0.2, 3)]),
        # full band, DM step 8: no LDS shape holds a tile's window -> direct kernel
        dict(nchans=48, nsamp=32768, ranges=[(0, 300, 8.0, 1, 1)], bw=810.0, seed=16, pulses=[(160.0, 0.3, 6)]),
        # 40 channels (not a multiple of 16): corner-turned by a launch of its own
        dict(nchans=40, nsamp=16384, ranges=[(0, 120, 0.5, 1, 1), (120, 240, 1.0, 2, 2)], bw=100.0, seed=17,
             pulses=[(70.0, 0.4, 5)]),
    ]
    for s in specs:
        s.update(narrow)
        s["foff"] = -s.pop("bw") / s["nchans"]
        s["dm_max"] = s["ranges"][-1][1]
    return specs


def oracle_planes(oracle, spec):
    s = oracle.strategy(spec["ranges"], spec["nchans"], spec["nsamp"], spec["fch1"], spec["foff"], spec["tsamp"],
                        free_memory=64 << 30)
    assert s.ok and s.num_tchunks == 1
    data = oracle.gen_noise(spec["seed"], spec["nsamp"], spec["nchans"])
    for dm, t_frac, w in spec["pulses"]:
        oracle.inject_pulse(data, s.dmshifts, dm, spec["tsamp"], int(t_frac * (spec["nsamp"] - s.maxshift)), w, 32)
    return oracle.run(s, data)


@pytest.fixture(scope="module")
def beams(aa, cuda):
    import multibeam as mb
    bs = mb.build_beams(small_specs(), cuda)
    bs[4].plan.set_range_variant(0, 1)   # 16-channel chunks where 32 also fit
    bs[1].plan.set_range_variant(1, 1)
    bs[3].plan.set_range_variant(0, 6)   # 32 DM trials per tile (its own batched launch)
    bs[2].plan.set_range_variant(2, 6)
    return bs


def check_against_oracle(oracle, beams, tag):
    for i, b in enumerate(beams):
        want = oracle_planes(oracle, b.spec)
        for r, o in enumerate(b.outs):
            got = o.cpu().numpy()
            assert np.array_equal(got, want[r][:, :b.tps[r]]), f"{tag}: beam {i} range {r}"


def test_per_beam_launches_match_oracle(aa, oracle, beams):
    import multibeam as mb
    say("per-beam launches vs oracle")
    mb.run_per_beam(beams)
    check_against_oracle(oracle, beams, "per-beam")


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_launch_matches_oracle(aa, oracle, cuda, beams, use_graph):
    import torch
    say(f"batched launch graph={use_graph}")
    for b in beams:
        for o in b.outs:
            o.fill_(-7.0)
    batch = aa.DdtrBatch([b.plan for b in beams])
    batch.execute([b.d_input for b in beams], [b.tps for b in beams], [b.outs for b in beams],
                  use_graph=use_graph)
    torch.cuda.synchronize()
    st = batch.stats()
    # batched launches: the byte-row ranges (at least the two forced ones)
    # and the pair-shape ranges; the direct-kernel range on its own
    assert st["jobs"] >= 3 and st["direct_jobs"] >= 1
    assert st["jobs"] + st["direct_jobs"] == sum(len(b.tps) for b in beams)
    # corner turns: one batched launch per number of binned levels written
    # (0, 1 and 2 among the six beams of 48-256 channels), the 40-channel beam alone
    assert st["corner_turn_launches"] == 3 and st["corner_turns_alone"] == 1
    check_against_oracle(oracle, beams, f"batched graph={use_graph}")


def test_graph_replays_follow_new_input(aa, oracle, cuda, beams):
    """Capture once, then replay several times with fresh input written into
    the same buffers (and eager device work between replays): every replay
    equals the per-beam launches on that input."""
    import torch
    import multibeam as mb
    say("graph replays with changing input")
    batch = aa.DdtrBatch([b.plan for b in beams])
    args = ([b.d_input for b in beams], [b.tps for b in beams], [b.outs for b in beams])
    batch.execute(*args, use_graph=True)          # eager pass + capture
    for rep in range(3):
        for i, b in enumerate(beams):
            aa.synth_noise_u8(1000 * rep + i, b.d_input, 128.0, 16.0)
        torch.cuda.synchronize()
        mb.run_per_beam(beams)
        want = mb.output_digest(beams)
        for b in beams:
            for o in b.outs:
                o.zero_()
        batch.execute(*args, use_graph=True)      # replay
        torch.cuda.synchronize()
        assert mb.output_digest(beams) == want, f"replay {rep}"


def test_autotune_keeps_results_bit_exact(aa, oracle, cuda, beams, tmp_path):
    import torch
    import multibeam as mb
    say("autotune then batched launch")
    cache = mb.autotune(beams, cache_path=str(tmp_path / "tune.json"), reps=1)
    assert cache and all(v["variant"] in (-1, 0, 1, 2, 3, 4, 5, 6) for v in cache.values())
    for v in cache.values():   # the fastest shape that fits; -1 only when none fits
        fits = [ms for ms in v["ms"] if ms >= 0]
        assert (v["variant"] == -1) == (not fits)
        if fits:
            assert v["ms"][v["variant"]] == min(fits)
    # the full-band beam's second range does not fit the 32-channel shape
    key = repr(mb.shape_class(beams[4], 1))
    assert cache[key]["ms"][0] == -1.0 and cache[key]["variant"] in (1, 2, 3, 4, 5, 6)
    mb.run_per_beam(beams)
    want = mb.output_digest(beams)
    batch = aa.DdtrBatch([b.plan for b in beams])
    batch.execute([b.d_input for b in beams], [b.tps for b in beams], [b.outs for b in beams], use_graph=False)
    torch.cuda.synchronize()
    assert mb.output_digest(beams) == want
