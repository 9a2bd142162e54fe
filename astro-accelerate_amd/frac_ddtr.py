"""Differentiable fractional-delay dedispersion on MI355X (PyTorch binding).

The forward pass and its adjoint are hand-written HIP kernels in
libaa_ddtr_hip.so (``aa_hip_frac_*``, include/aa_ddtr_hip.h); this module
wraps them in a ``torch.autograd.Function`` so that the dedispersed profiles
are differentiable with respect to the input windows and to a continuous DM
per candidate.  No counterpart exists in the reference, which has neither
autograd nor fractional delays; the delays use the reference's dispersion
table (dmshift[c], src/aa_ddtr_strategy.cpp:81-92) divided by tsamp.

For candidate b, trial k and output sample t (x is zero outside [0, T)):

    u   = trials[k] + dm[b]             tau = delay_per_dm[c] * u   (float32)
    y[b, k, t] = sum_c (1 - f) x[b, c, t + i0] + f x[b, c, t + i0 + 1],
    i0 = floor(tau), f = tau - i0.

The backward pass recomputes the interpolation from x (nothing from the
forward pass is kept besides its inputs).  torch provides device memory and
streams only.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

import py_astro_accelerate as aa


def aa_dmshifts(nchans: int, fch1: float, foff: float, tsamp: float, power: float = 2.0) -> np.ndarray:
    """The reference's dmshift table for this band (via the library's strategy)."""
    plan = aa.make_plan([(0.0, 1.0, 1.0, 1, 1)])
    if power != 2.0:
        plan.set_power(power)
    meta = aa.make_metadata(nchans=nchans, nsamples=1 << 20, fch1=fch1, foff=foff, tsamp=tsamp)
    strat = aa.aa_py_ddtr_strategy(plan, meta, free_memory=1 << 40)
    if not strat.ready:
        raise aa.AAHipError("strategy for the dmshift table was not ready")
    return strat.dmshifts()


def delay_per_dm(nchans: int, fch1: float, foff: float, tsamp: float) -> np.ndarray:
    """Delay in samples per unit DM, float32(dmshift[c] / float32(tsamp))."""
    ds = aa_dmshifts(nchans, fch1, foff, tsamp)
    return (ds / np.float32(tsamp)).astype(np.float32)


def _stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class FracPlan:
    """Device plan (aa_hip_frac handle): delay table, trial offsets, window length."""

    def __init__(self, delay_per_dm: np.ndarray, trials: Sequence[float], nsamp: int, max_batch: int,
                 dm_absmax: float = 16.0, device: int = 0):
        lib = aa.load_library()
        self.delay_per_dm = np.ascontiguousarray(delay_per_dm, dtype=np.float32)
        self.trials = np.ascontiguousarray(np.asarray(trials, dtype=np.float32))
        self.nchans = int(self.delay_per_dm.size)
        self.ntrials = int(self.trials.size)
        self.nsamp = int(nsamp)
        self.max_batch = int(max_batch)
        h = ctypes.c_void_p()
        aa.check(lib.aa_hip_frac_create(int(device), self.nchans,
                                        self.delay_per_dm.ctypes.data_as(aa.PFLOAT), self.ntrials,
                                        self.trials.ctypes.data_as(aa.PFLOAT), self.nsamp, self.max_batch,
                                        float(dm_absmax), ctypes.byref(h)), "aa_hip_frac_create")
        self.handle = h

    def grouping(self):
        tq, cq = ctypes.c_int(), ctypes.c_int()
        aa.check(aa.load_library().aa_hip_frac_grouping(self.handle, ctypes.byref(tq), ctypes.byref(cq)),
                 "aa_hip_frac_grouping")
        return tq.value, cq.value

    def close(self):
        if getattr(self, "handle", None):
            aa.load_library().aa_hip_frac_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_x(self, x, name="x"):
        import torch
        if not (x.is_cuda and x.dtype == torch.float32 and x.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")
        return x

    def _check_batch(self, x, dm, rows: int):
        import torch
        if x.dim() != 3 or x.shape[1] != rows or x.shape[2] != self.nsamp:
            raise ValueError(f"expected [B, {rows}, {self.nsamp}], got {tuple(x.shape)}")
        b = x.shape[0]
        if b > self.max_batch:
            raise ValueError(f"batch {b} exceeds the plan's {self.max_batch}")
        if not (dm.is_cuda and dm.dtype == torch.float32 and dm.is_contiguous() and dm.shape == (b,)):
            raise ValueError("dm must be a contiguous float32 CUDA tensor of shape [B]")
        return b

    def forward(self, x, dm, out=None):
        """y[B, K, T] for windows x[B, C, T] at continuous DM offsets dm[B]."""
        import torch
        self._check_x(x)
        b = self._check_batch(x, dm, self.nchans)
        y = out if out is not None else torch.empty((b, self.ntrials, self.nsamp), dtype=torch.float32,
                                                    device=x.device)
        self._check_x(y, "out")
        aa.check(aa.load_library().aa_hip_frac_forward(self.handle, ctypes.c_void_p(x.data_ptr()), b,
                                                        ctypes.c_void_p(dm.data_ptr()),
                                                        ctypes.c_void_p(y.data_ptr()), _stream_handle()),
                 "aa_hip_frac_forward")
        return y

    def adjoint(self, x, gy, dm, want_x: bool = True, want_dm: bool = True):
        """(gx [B, C, T] float32 or None, gdm [B] float64 or None) for the adjoint gy [B, K, T]."""
        import torch
        self._check_x(gy, "gy")
        b = self._check_batch(gy, dm, self.ntrials)
        gx = torch.empty((b, self.nchans, self.nsamp), dtype=torch.float32, device=gy.device) if want_x else None
        gdm = torch.empty((b,), dtype=torch.float64, device=gy.device) if want_dm else None
        if want_dm:
            self._check_x(x)
            if tuple(x.shape) != (b, self.nchans, self.nsamp):
                raise ValueError("x does not match gy")
        aa.check(aa.load_library().aa_hip_frac_adjoint(
            self.handle, ctypes.c_void_p(x.data_ptr() if x is not None else 0), ctypes.c_void_p(gy.data_ptr()), b,
            ctypes.c_void_p(dm.data_ptr()), ctypes.c_void_p(gx.data_ptr() if gx is not None else 0),
            ctypes.c_void_p(gdm.data_ptr() if gdm is not None else 0), _stream_handle()), "aa_hip_frac_adjoint")
        return gx, gdm


def _autograd_function():
    import torch

    class FracDedisperse(torch.autograd.Function):
        """y = fractional-delay dedispersion of x at DM offsets dm (see module docstring)."""

        @staticmethod
        def forward(ctx, x, dm, plan: FracPlan):
            x = x.contiguous()
            dm = dm.contiguous()
            y = plan.forward(x, dm)
            ctx.save_for_backward(x, dm)
            ctx.plan = plan
            return y

        @staticmethod
        def backward(ctx, gy):
            x, dm = ctx.saved_tensors
            want_x, want_dm = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
            if not (want_x or want_dm):
                return None, None, None
            gx, gdm = ctx.plan.adjoint(x, gy.contiguous(), dm, want_x=want_x, want_dm=want_dm)
            return gx, (gdm.to(dm.dtype) if gdm is not None else None), None

    return FracDedisperse


_FN = None


def frac_dedisperse(x, dm, plan: FracPlan):
    """Differentiable y[B, K, T] = FracDedisperse(x[B, C, T], dm[B]) on the plan's device."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(x, dm, plan)


def synth_candidates(batch: int, nchans: int, nsamp: int, delay_per_dm, seed: int, device,
                     residual_max: float = 1.5, sigma: float = 4.0, amplitude: float = 0.5,
                     dm_range=(50.0, 2000.0)):
    """Candidate windows: N(0, 1) float32 noise plus a Gaussian pulse (sigma samples,
    amplitude) dispersed at a residual DM r_b = DM_true - DM_candidate, uniform in
    [-residual_max, residual_max], in the frame of the candidate's DM (the window
    follows the candidate's delay track): pulse centre in channel c at
    nsamp / 2 + delay_per_dm[c] * r_b.  Returns (x [B, C, T], r [B], dm_true [B])."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    x = torch.randn((batch, nchans, nsamp), generator=g, device=device, dtype=torch.float32)
    r = (torch.rand((batch,), generator=g, device=device, dtype=torch.float64) * 2.0 - 1.0) * residual_max
    dm_true = dm_range[0] + torch.rand((batch,), generator=g, device=device, dtype=torch.float64) * (
        dm_range[1] - dm_range[0])
    ds = torch.as_tensor(np.asarray(delay_per_dm, dtype=np.float64), device=device)
    t = torch.arange(nsamp, device=device, dtype=torch.float64)
    for b in range(batch):
        centre = nsamp / 2.0 + ds * r[b]                               # [C]
        pulse = amplitude * torch.exp(-0.5 * ((t[None, :] - centre[:, None]) / sigma) ** 2)
        x[b] += pulse.to(torch.float32)
    return x, r, dm_true
