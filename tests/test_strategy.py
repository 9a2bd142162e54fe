"""Strategy (plan -> DM grid, maxshift, t_processed, chunks) on the CPU.

Pins, for the oracle and for the product library independently:
  * the reference's own strategy values (tests/golden/strategy_fixtures.json);
  * the dmshift table against a numpy restatement of its closed form;
  * agreement of oracle and product over randomised plans, including the
    chunked cases 3/4 (small free memory) and the 4-bit granularity;
  * the reference's rejection rules.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "strategy_fixtures.json")))["cases"]


def product_strategy(aa, case, free_memory=280 * 2**30, nbits=8, analysis=False, power=None):
    plan = aa.make_plan(case["ranges"])
    if power is not None:
        plan.set_power(power)
    md = aa.make_metadata(case["nchans"], case["nsamp"], case["fch1"], case["foff"], case["tsamp"], nbits=nbits)
    return aa.aa_py_ddtr_strategy(plan, md, free_memory, analysis)


def dmshift_closed_form(nchans, fch1, foff):
    """(float)(4148.741601f * (1/f^2 - 1/fch1^2)), f = f32(fch1) + f32(foff)*c in float32."""
    f1 = np.float32(fch1)
    f = f1 + np.float32(foff) * np.arange(nchans, dtype=np.float32)  # float32 mul, then float32 add
    k = np.float64(np.float32(4148.741601))
    return (k * (1.0 / f.astype(np.float64) ** 2 - 1.0 / np.float64(f1) ** 2)).astype(np.float32)


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_oracle_matches_reference_fixture(oracle, case):
    s = oracle.strategy(case["ranges"], case["nchans"], case["nsamp"], case["fch1"], case["foff"], case["tsamp"])
    assert s.ok
    assert s.maxshift == case["maxshift"]
    assert s.num_tchunks == case["num_tchunks"]
    assert s.ndms == case["ndms"]
    assert s.t_processed == case["t_processed"]
    np.testing.assert_allclose([d[0] for d in s.dm], case["dm_low"], rtol=0, atol=1e-4)
    if "dmshift_last" in case:
        assert s.dmshifts[-1] == pytest.approx(case["dmshift_last"], rel=1e-8)
    if "max_ndms" in case:
        assert s.max_ndms == case["max_ndms"]


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_product_matches_reference_fixture(aa, case):
    s = product_strategy(aa, case)
    assert s.ready
    assert s.maxshift == case["maxshift"]
    assert s.num_tchunks == case["num_tchunks"]
    assert [s.ndms(r) for r in range(s.nranges)] == case["ndms"]
    assert [[s.t_processed(r, k) for k in range(s.num_tchunks)] for r in range(s.nranges)] == case["t_processed"]
    np.testing.assert_allclose([s.dm(r)[0] for r in range(s.nranges)], case["dm_low"], rtol=0, atol=1e-4)


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_dmshift_closed_form(aa, oracle, case):
    want = dmshift_closed_form(case["nchans"], case["fch1"], case["foff"])
    got_o = oracle.dmshifts(case["nchans"], case["fch1"], case["foff"])
    got_p = product_strategy(aa, case).dmshifts()
    np.testing.assert_array_equal(got_o, want)
    np.testing.assert_array_equal(got_p, want)
    assert want[0] == 0.0 and np.all(np.diff(want.astype(np.float64)) > 0)


def test_config2_is_case4_single_chunk(aa, oracle):
    # The 16 GiB output cap makes max_tsamps == nsamp, so the "does not fit"
    # branch runs and still yields exactly one chunk.
    case = FIXTURES[1]
    s = oracle.strategy(case["ranges"], case["nchans"], case["nsamp"], case["fch1"], case["foff"], case["tsamp"])
    assert s.strategy_case == 4 and s.num_tchunks == 1


def _random_case(rng):
    nchans = int(rng.choice([16, 64, 100, 128, 512, 1000, 2048, 4096]))
    fch1 = float(rng.uniform(300, 2000))
    foff = -float(rng.uniform(0.01, 0.8 * fch1 / nchans))
    tsamp = float(rng.choice([32e-6, 64e-6, 1e-4, 2.56e-4]))
    nr = int(rng.integers(1, 6))
    ranges, low, inb = [], 0.0, 1
    for _ in range(nr):
        step = float(rng.choice([0.05, 0.1, 0.25, 0.5, 1.0, 2.0]))
        high = low + float(rng.uniform(5, 400))
        ranges.append([low, high, step, inb, inb])
        low = high
        if rng.random() < 0.5:
            inb *= 2
    nsamp = int(rng.choice([1 << 14, 1 << 16, 1 << 18, 1 << 20, 3_000_000]))
    free = int(rng.choice([280 * 2**30, 2**30, 300 * 2**20, 150 * 2**20, 64 * 2**20]))
    return dict(nchans=nchans, fch1=fch1, foff=foff, tsamp=tsamp, ranges=ranges, nsamp=nsamp), free


def test_oracle_and_product_agree_on_random_plans(aa, oracle):
    rng = np.random.default_rng(20261202)
    n_ready = n_chunked = 0
    for _ in range(300):
        case, free = _random_case(rng)
        nbits = int(rng.choice([8, 8, 4]))
        analysis = bool(rng.random() < 0.2)
        so = oracle.strategy(case["ranges"], case["nchans"], case["nsamp"], case["fch1"], case["foff"],
                             case["tsamp"], nbits=nbits, free_memory=free, enable_analysis=analysis)
        sp = product_strategy(aa, case, free_memory=free, nbits=nbits, analysis=analysis)
        assert so.ok == sp.ready, case
        if not so.ok:
            continue
        n_ready += 1
        n_chunked += so.num_tchunks > 1
        assert so.maxshift == sp.maxshift
        assert so.num_tchunks == sp.num_tchunks
        assert so.ndms == [sp.ndms(r) for r in range(sp.nranges)]
        assert so.t_processed == [[sp.t_processed(r, k) for k in range(sp.num_tchunks)] for r in range(sp.nranges)]
        for r in range(sp.nranges):
            assert np.float32(so.dm[r][0]) == np.float32(sp.dm(r)[0])
            assert np.float32(so.dm[r][1]) == np.float32(sp.dm(r)[1])
        np.testing.assert_array_equal(so.dmshifts, sp.dmshifts())
    assert n_ready > 50 and n_chunked > 5, (n_ready, n_chunked)


def test_rejections(aa, oracle):
    base = dict(nchans=512, nsamp=131072, fch1=1550.0, foff=-0.5859, tsamp=64e-6, ranges=[[0, 500, 0.5, 1, 1]])
    # maxshift >= nsamp
    short = dict(base, nsamp=7000)
    assert not oracle.strategy(short["ranges"], 512, 7000, 1550.0, -0.5859, 64e-6).ok
    assert not product_strategy(aa, short).ready
    # not enough memory for maxshift samples
    assert not product_strategy(aa, base, free_memory=104857600 + 1000).ready
    # empty plan
    md = aa.make_metadata(512, 131072, 1550.0, -0.5859, 64e-6)
    empty = aa.aa_py_ddtr_plan([])
    assert not aa.aa_py_ddtr_strategy(empty, md, 2**34).ready


def test_non_power2_dispersion_index(aa, oracle):
    case = dict(nchans=256, nsamp=1 << 18, fch1=1400.0, foff=-1.0, tsamp=64e-6, ranges=[[0, 300, 1, 1, 1]])
    so = oracle.strategy(case["ranges"], 256, 1 << 18, 1400.0, -1.0, 64e-6, power=2.2)
    sp = product_strategy(aa, case, power=2.2)
    assert so.ok and sp.ready
    np.testing.assert_array_equal(so.dmshifts, sp.dmshifts())
    assert so.maxshift == sp.maxshift
