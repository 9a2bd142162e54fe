"""Reference for the fractional-delay dedispersion and its gradients.

TEST INFRASTRUCTURE ONLY: the checker that the HIP kernels (aa_hip_frac_*)
are compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
baseline leg may import it; the product never does.

The transform has no counterpart in the reference (AstroAccelerate has no
fractional delays or autograd).  Its delays are built on the reference's
dispersion table (dmshift[c], src/aa_ddtr_strategy.cpp:81-92) divided by
tsamp.  For candidate b, trial k and output sample t (x zero outside [0, T)):

    u = trials[k] + dm[b];  tau = delay_per_dm[c] * u
    y[b, k, t] = sum_c (1 - f) x[b, c, t + i0] + f x[b, c, t + i0 + 1],
    i0 = floor(tau), f = tau - i0.

Two implementations, pinned against each other and against mathematics in
tests/test_frac_ddtr.py:
  * ``frac_ddtr_loops``: scalar float64 loops (tiny inputs only);
  * ``frac_ddtr_torch``: vectorised PyTorch (gather), differentiable in x and
    dm through autograd, in any float dtype (float64 in the tests; the HIP
    kernels evaluate the delays in double too and sum in float32).
"""
from __future__ import annotations

import math

import numpy as np


def frac_ddtr_loops(x, dm, trials, delay_per_dm):
    """float64 scalar loops over numpy arrays; x [B, C, T] -> y [B, K, T]."""
    x = np.asarray(x, dtype=np.float64)
    B, C, T = x.shape
    K = len(trials)
    y = np.zeros((B, K, T))
    for b in range(B):
        for k in range(K):
            u = float(trials[k]) + float(dm[b])
            for c in range(C):
                tau = float(delay_per_dm[c]) * u
                i0 = math.floor(tau)
                f = tau - i0
                for t in range(T):
                    j = t + i0
                    a = x[b, c, j] if 0 <= j < T else 0.0
                    n = x[b, c, j + 1] if 0 <= j + 1 < T else 0.0
                    y[b, k, t] += (1.0 - f) * a + f * n
    return y


def frac_ddtr_torch(x, dm, trials, delay_per_dm):
    """Vectorised reference: x [B, C, T], dm [B] (same device) -> y [B, K, T].

    Differentiable in x and dm.  The floor is taken without gradient, as the
    derivative of the piecewise-linear interpolation in the DM is the slope
    of the cell it lies in.
    """
    import torch
    import torch.nn.functional as F

    B, C, T = x.shape
    dev = x.device
    tr = torch.as_tensor(np.asarray(trials), device=dev)
    ds = torch.as_tensor(np.asarray(delay_per_dm), device=dev)
    tpos = torch.arange(T, device=dev)
    out = []
    for k in range(tr.numel()):
        u = tr[k].to(x.dtype) + dm.to(x.dtype)                 # [B]
        tau = ds.to(x.dtype)[None, :] * u[:, None]            # [B, C]
        i0 = torch.floor(tau).detach()
        f = tau - i0
        i0 = i0.to(torch.int64)
        lo = int(max(0, -int(i0.min()))) + 1
        hi = int(max(0, int(i0.max()))) + 2
        xp = F.pad(x, (lo, hi))
        idx = (tpos[None, None, :] + i0[:, :, None] + lo).expand(B, C, T)
        x0 = torch.gather(xp, 2, idx)
        x1 = torch.gather(xp, 2, idx + 1)
        out.append(((1.0 - f)[:, :, None] * x0 + f[:, :, None] * x1).sum(dim=1))
    return torch.stack(out, dim=1)


def noise_energy_torch(x, u, delay_per_dm):
    """Expected profile energy of white noise with the channel powers of x,
    at residual DMs u [B] (one trial at offset 0):

        N(u) = sum_c v_c ((1 - f_c)^2 (T - |i_c|)+ + f_c^2 (T - |i_c + 1|)+),
        v_c = mean_t x[c][t]^2,  tau_c = delay_per_dm[c] * u,  i_c = floor(tau_c).

    For independent zero-mean channels of variances v_c, E[sum_t y(t)^2]
    equals N: the cross-channel terms vanish and channel c's two taps (weights
    1 - f_c and f_c) cover (T - |i_c|)+ and (T - |i_c + 1|)+ output samples
    of [0, T).  Differentiable in u (the floor carries no gradient); x [B, C, T]."""
    import torch
    B, C, T = x.shape
    ds = torch.as_tensor(np.asarray(delay_per_dm), device=x.device).to(u.dtype)
    v = (x.detach().to(u.dtype) ** 2).mean(dim=2)                     # [B, C]
    tau = ds[None, :] * u[:, None]                                   # [B, C]
    i0 = torch.floor(tau).detach()
    f = tau - i0
    n0 = torch.clamp(T - torch.abs(i0), min=0.0)
    n1 = torch.clamp(T - torch.abs(i0 + 1.0), min=0.0)
    return (v * ((1.0 - f) ** 2 * n0 + f ** 2 * n1)).sum(dim=1)
