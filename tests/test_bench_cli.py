"""bench.py's command line on the CPU: it measures one GPU and says so when
asked for more; its modes are listed in --help."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_help_lists_the_modes():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--config", "--e2e", "--zero-dm", "--zero-dm-outliers", "--normalization"):
        assert flag in r.stdout


def test_multibeam_sets_plan_as_drawn(oracle):
    """Both beam sets of the multi-beam benchmark draw plans the strategy
    accepts in one time chunk; the small set's ranges hold fewer 16-DM
    tiles (2016-sample tiles) than the GPU has CUs, the case batching is for."""
    sys.path.insert(0, os.path.join(ROOT, "astro-accelerate_amd"))
    import multibeam as mb
    import bench
    sets = {"standard": mb.multibeam_specs(48, bench.SEED), "small": mb.small_multibeam_specs(bench.SEED)}
    tiles = []
    for name, specs in sets.items():
        for s in specs:
            st = oracle.strategy(s["ranges"], s["nchans"], s["nsamp"], s["fch1"], s["foff"], s["tsamp"])
            assert st.ok and st.num_tchunks == 1, (name, s)
            if name == "small":
                tiles += [-(-st.ndms[r] // 16) * -(-st.t_processed[r][0] // 2016) for r in range(st.nranges)]
    assert len(sets["small"]) == 256 and sorted(tiles)[len(tiles) // 2] < 256
