"""GPU: aa_hip_spd_first_pass_noalloc (the first single-pulse pass with its
workspace in the library's static device array) against
aa_hip_spd_first_pass (workspace passed in) on the same planes.

Two no-allocation calls run back to back, the second on a smaller plane and
a smaller width limit, so the second reads a scratch array that still holds
the first call's moments; each must equal the workspace entry point's output
bit for bit (same kernels, same reduction order).  Width limits outside
1..32 are refused before anything is launched.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _plane(torch, dev, seed, ndms, nsamp, pitch):
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn((ndms, pitch), generator=g, dtype=torch.float32) * 3.0 + 10.0
    return x.to(dev)[:, :nsamp]


def _outputs(torch, dev, ndms, nsamp, W, fill):
    msd = torch.full((W, 2), fill, dtype=torch.float32, device=dev)
    snr = torch.full((ndms, nsamp + 5), fill, dtype=torch.float32, device=dev)[:, :nsamp]
    width = torch.full((ndms, nsamp + 3), -1, dtype=torch.int16, device=dev)[:, :nsamp]
    return msd, snr, width


def test_noalloc_matches_workspace_entry_point(aa, cuda):
    import torch
    cases = [(7, 96, 20000, 20011, 32), (8, 40, 5000, 5000, 12)]   # seed, ndms, nsamp, pitch, W: second smaller
    planes = [_plane(torch, cuda, s, n, t, p) for s, n, t, p, _ in cases]
    scratch = int(aa.load_library().aa_hip_spd_noalloc_scratch_bytes())
    assert scratch >= int(aa.load_library().aa_hip_spd_workspace_bytes())
    say(f"no-alloc first pass: static scratch {scratch} bytes; calls on {[c[1:] for c in cases]}")
    got = []
    for (seed, ndms, nsamp, _, W), x in zip(cases, planes):   # two no-allocation calls in a row
        msd, snr, width = _outputs(torch, cuda, ndms, nsamp, W, float("nan"))
        aa.spd_first_pass_noalloc(x, msd, snr, width, max_width=W)
        got.append((msd, snr, width))
    torch.cuda.synchronize()
    for (seed, ndms, nsamp, _, W), x, (msd, snr, width) in zip(cases, planes, got):
        want_msd, want_snr, want_width = aa.spd_first_pass(x, max_width=W)
        torch.cuda.synchronize()
        say(f"  ndms {ndms} nsamp {nsamp} W {W}: msd {msd[:2].tolist()}")
        assert torch.equal(msd, want_msd)
        assert torch.equal(snr, want_snr)
        assert torch.equal(width, want_width)
        assert bool(torch.isfinite(snr).all())


def test_noalloc_msd_only_and_refusals(aa, cuda):
    import torch
    x = _plane(torch, cuda, 9, 16, 3000, 3000)
    msd = torch.empty((32, 2), dtype=torch.float32, device=cuda)
    aa.spd_first_pass_noalloc(x, msd, None, None, max_width=32)
    want, _, _ = aa.spd_first_pass(x, max_width=32, with_snr=False)
    torch.cuda.synchronize()
    assert torch.equal(msd, want)
    for bad_w in (0, 33):
        with pytest.raises(Exception, match="max_width"):
            aa.spd_first_pass_noalloc(x, torch.empty((max(bad_w, 1), 2), dtype=torch.float32, device=cuda),
                                      None, None, max_width=bad_w)
