"""Differentiable fractional-delay dedispersion on MI355X (PyTorch binding).

The forward pass and its adjoint are hand-written HIP kernels in
libaa_ddtr_hip.so (``aa_hip_frac_*``, include/aa_ddtr_hip.h); this module
wraps them in a ``torch.autograd.Function`` so that the dedispersed profiles
are differentiable with respect to the input windows and to a continuous DM
per candidate.  No counterpart exists in the reference, which has neither
autograd nor fractional delays; the delays use the reference's dispersion
table (dmshift[c], src/aa_ddtr_strategy.cpp:81-92) divided by tsamp.

For candidate b, trial k and output sample t (x is zero outside [0, T)):

    u   = trials[k] + dm[b]             tau = delay_per_dm[c] * u   (in double)
    y[b, k, t] = sum_c (1 - f) x[b, c, t + i0] + f x[b, c, t + i0 + 1],
    i0 = floor(tau), f = (float)(tau - i0).

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


# ------------------------------------------------------------ refinement
REFINE_LOSSES = {"energy": 0, "excess": 1}   # AA_HIP_REFINE_LOSS_ENERGY, _EXCESS_ENERGY


class RefinePlan:
    """Batched DM refinement of candidate windows of uneven size
    (aa_hip_refine_*): windows packed in one float32 buffer, [C_b][T_b] at
    x_offsets[b], delay tables at ds_offsets[b] of a second buffer.

    loss "excess" (default): profile energy minus the energy white noise of
    the window's channel powers gives at the same delays; "energy": the
    profile energy sum_t y(t)^2 alone (peaks at u = 0 on weak candidates,
    where no channel is interpolated)."""

    def __init__(self, nchans: Sequence[int], nsamp: Sequence[int], x_offsets: Sequence[int],
                 ds_offsets: Sequence[int], device: int = 0, loss: str = "excess"):
        if loss not in REFINE_LOSSES:
            raise ValueError(f"loss must be one of {sorted(REFINE_LOSSES)}")
        lib = aa.load_library()
        n = len(nchans)
        self.ncand = n
        xo = (ctypes.c_int64 * n)(*[int(v) for v in x_offsets])
        do = (ctypes.c_int64 * n)(*[int(v) for v in ds_offsets])
        nc = (ctypes.c_int * n)(*[int(v) for v in nchans])
        ns = (ctypes.c_int * n)(*[int(v) for v in nsamp])
        h = ctypes.c_void_p()
        aa.check(lib.aa_hip_refine_create(int(device), n, xo, do, nc, ns, ctypes.byref(h)), "aa_hip_refine_create")
        self.handle = h
        aa.check(lib.aa_hip_refine_set_loss(h, REFINE_LOSSES[loss]), "aa_hip_refine_set_loss")
        self.loss = loss
        # extents the buffers must cover (checked before every launch)
        self.x_extent = max(int(o) + int(c) * int(t) for o, c, t in zip(x_offsets, nchans, nsamp))
        self.ds_extent = max(int(o) + int(c) for o, c in zip(ds_offsets, nchans))

    def close(self):
        if getattr(self, "handle", None):
            aa.load_library().aa_hip_refine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, x, ds, u, step, prev_grad, loss, grad, nsteps: int, step_max: float = 0.5,
            step_min: float = 1e-4, scan=None):
        """Optional coarse energy scan (scan = (n, lo, step): offsets relative
        to u), then nsteps fused forward + adjoint + update steps for every
        candidate (state tensors updated in place; loss/grad at the final u)."""
        import torch
        for t, dt in ((x, torch.float32), (ds, torch.float32), (u, torch.float32), (step, torch.float32),
                      (prev_grad, torch.float64), (loss, torch.float64), (grad, torch.float64)):
            if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
                raise ValueError("refine buffers: contiguous CUDA tensors of the documented dtypes")
        for t in (u, step, prev_grad, loss, grad):
            if t.numel() != self.ncand:
                raise ValueError("state tensors need one entry per candidate")
        if x.numel() < self.x_extent or ds.numel() < self.ds_extent:
            raise ValueError("window or delay-table buffer smaller than the candidate table needs")
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        sn, slo, sst = scan if scan is not None else (0, 0.0, 0.0)
        aa.check(aa.load_library().aa_hip_refine_run(self.handle, p(x), p(ds), p(u), p(step), p(prev_grad), p(loss),
                                                      p(grad), int(sn), float(slo), float(sst), int(nsteps),
                                                      float(step_max), float(step_min), _stream_handle()),
                 "aa_hip_refine_run")


def synth_refine_candidates(ncand: int, seed: int, device, nchans_set=(1024, 1536, 2048, 2560, 3072, 3584, 4096),
                            nsamp_set=(256, 512, 1024, 2048, 4096), residual_max: float = 1.0,
                            sigma: float = 4.0, amplitude: float = 0.5):
    """Uneven candidate windows, seeded: N(0, 1) noise plus a Gaussian pulse at
    residual DM r_b (uniform in +-residual_max) in the frame of the
    candidate's DM, centred in the window; band 1760 MHz down 810 MHz, 64 us.
    Returns dict(x, ds, x_offsets, ds_offsets, nchans, nsamp, r, tables)."""
    import torch
    rng = np.random.default_rng(seed)
    nchans = [int(rng.choice(nchans_set)) for _ in range(ncand)]
    nsamp = [int(rng.choice(nsamp_set)) for _ in range(ncand)]
    r = rng.uniform(-residual_max, residual_max, ncand)
    tables, ds_off, parts = {}, {}, []
    off = 0
    for C in sorted(set(nchans)):
        tables[C] = delay_per_dm(C, 1760.0, -810.0 / C, 64e-6)
        ds_off[C] = off
        parts.append(tables[C])
        off += C
    ds = torch.from_numpy(np.concatenate(parts)).to(device)
    sizes = [C * T for C, T in zip(nchans, nsamp)]
    x_offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    x = torch.empty(int(sum(sizes)), dtype=torch.float32, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    for b in range(ncand):
        C, T = nchans[b], nsamp[b]
        w = x[int(x_offsets[b]):int(x_offsets[b]) + C * T].view(C, T)
        w.normal_(generator=g)
        t = torch.arange(T, device=device, dtype=torch.float64)
        centre = T / 2.0 + torch.as_tensor(tables[C], device=device, dtype=torch.float64) * float(r[b])
        w += (amplitude * torch.exp(-0.5 * ((t[None, :] - centre[:, None]) / sigma) ** 2)).to(torch.float32)
    return dict(x=x, ds=ds, x_offsets=x_offsets, ds_offsets=np.array([ds_off[C] for C in nchans], dtype=np.int64),
                nchans=nchans, nsamp=nsamp, r=r, tables=tables)


def dense_grid_search_plain(cands, u_lo: float = -2.24, du: float = 0.01, ndms: int = 448, tsamp: float = 64e-6):
    """Best residual DM of every candidate on a dense grid, with the plain u8
    dedispersion path (aa_hip_dedisperse_u8): each window is quantised to u8
    (round(128 + 16 x), clipped), each channel row is shifted by the plain
    path's own delay at u = 0 so that grid trial k is residual u_lo + k du,
    and the trial with the largest profile energy sum_t (out - 128)^2 wins.
    Returns the residual DM per candidate."""
    import torch
    dev = cands["x"].device
    best = np.zeros(len(cands["nchans"]))
    dmshift_cache = {}
    for b, (C, T) in enumerate(zip(cands["nchans"], cands["nsamp"])):
        if C not in dmshift_cache:
            dms = aa_dmshifts(C, 1760.0, -810.0 / C, tsamp)
            mst = np.float32(0.0)                                   # plain-path trial 0 = residual u_lo
            mstep = np.float32(np.float32(du) / np.float32(tsamp))
            k0 = int(round(-u_lo / du))                             # the trial at residual 0
            o = np.trunc(dms * np.float32(np.float32(k0) * mstep)).astype(np.int64)
            maxd = int(np.trunc(np.float32(dms[-1] * np.float32(np.float32(ndms - 1) * mstep)))) + 1
            dmshift_cache[C] = (dms, torch.from_numpy(dms).to(dev), torch.from_numpy(o).to(dev), float(mst),
                                float(mstep), maxd)
        dms, d_dms, o, mst, mstep, maxd = dmshift_cache[C]
        w = cands["x"][int(cands["x_offsets"][b]):int(cands["x_offsets"][b]) + C * T].view(C, T)
        q = torch.clamp(torch.round(128.0 + 16.0 * w), 0, 255).to(torch.uint8)
        row_len = T + maxd
        pitch = ((row_len + 4096 + 255) // 256) * 256
        L = torch.full((C, pitch), 128, dtype=torch.uint8, device=dev)
        tt = torch.arange(row_len, device=dev)
        src = tt[None, :] - o[:, None]                                         # L[c][t'] = q[c][t' - o_c]
        ok = (src >= 0) & (src < T)
        L[:, :row_len] = torch.where(ok, torch.gather(q, 1, src.clamp(0, T - 1)), torch.full_like(src, 128,
                                                                                               dtype=torch.uint8))
        out = torch.empty((ndms, T), dtype=torch.float32, device=dev)
        aa.dedisperse_u8(L, pitch, row_len, dms, d_dms, ndms, T, mst, mstep, 1, out, T)
        energy = ((out.double() - 128.0) ** 2).sum(dim=1)
        best[b] = u_lo + du * int(torch.argmax(energy))
    return best
