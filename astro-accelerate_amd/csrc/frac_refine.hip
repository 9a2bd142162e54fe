// Batched DM refinement of candidate windows of uneven size, for gfx950.
// No counterpart in the reference.  Each candidate b is a float32 window
// x_b[c][t] (C_b channels x T_b samples) cut along the candidate's
// dispersion track; its continuous residual DM u_b is refined by gradient
// ascent on the energy of the dedispersed profile,
//     E(u) = sum_t y(t)^2,  y(t) = sum_c (1 - f_c) x[c][t + i_c] + f_c x[c][t + i_c + 1],
//     tau_c = ds[c] * u (double), i_c = floor(tau_c), f_c = tau_c - i_c,
// with x = 0 outside [0, T) and ds[c] the delay per unit DM in samples.
//     dE/du = sum_c ds[c] sum_t 2 y(t) (x[c][t + i_c + 1] - x[c][t + i_c]).
//
// Two losses (refine_args::loss_kind):
//  0  L = E.
//  1  L = E - N, the energy in excess of what white noise of the window's
//     channel powers v_c = mean_t x[c][t]^2 would give at the same delays:
//         N(u) = sum_c v_c ((1 - f_c)^2 (T - |i_c|)+ + f_c^2 (T - |i_c + 1|)+),
//     the expected sum_t y(t)^2 of independent channels of variances v_c
//     (each channel's two taps cover (T - |i|)+ and (T - |i + 1|)+ output
//     samples).  Interpolation averages neighbouring noise samples, so E of
//     pure noise is largest where every f_c is 0 (u = 0) and about a third
//     lower in between: on weak candidates that spike outweighs the pulse.
//     dN/du = sum_c v_c ds[c] (2 f_c (T - |i_c + 1|)+ - 2 (1 - f_c) (T - |i_c|)+).
//     v_c is computed once per run (refine_power_kernel).
//
// An optional coarse scan first evaluates L (forward pass only) at scan_n
// offsets u + scan_lo + k * scan_step and starts the steps from the best.
//
// One persistent launch: a work-group (1024 threads) takes candidates from a
// queue (largest first) and runs every step of that candidate itself, so the
// forward pass, the adjoint pass and the update are fused.  Thread layout: a
// group of S lanes covers the window's samples, 4 consecutive samples per
// lane (S = ceil(T / 4) rounded up to whole waves), and the work-group holds
// G = 1024 / S such groups, each summing every G-th channel; so short windows
// keep every lane busy.  A lane reads its 4 samples x[c][4s + i_c ..
// 4s + i_c + 3] as one 16-byte load, 4 channels in flight; the taps that
// reach into the next lane's samples are exchanged once per pass through
// LDS (pass 1) or folded into per-lane weights (pass 2).
// Pass 1 sums the groups' partial profiles in double, adds them in LDS in
// group order, pass 2 re-reads the windows (recomputation: no shifted tile is
// stored) and accumulates the DM derivative, and thread 0 applies a
// sign-based adaptive step (Rprop: the step grows by 1.2 while the sign of
// dL/du holds and halves when it flips); the result is the best point
// visited.  Every sum runs in a fixed order: results do not depend on
// scheduling.  Windows are read straight from global memory: with one trial
// per candidate there is no reuse to stage, and the passes are bound by HBM
// bandwidth.
#include "frac_refine.h"

#include <cmath>

namespace aa_hip {

namespace {

constexpr int kThreads = 1024;
constexpr int kSpt = 4;                       // consecutive samples per lane
constexpr int kU = 4;                         // channels in flight per lane
constexpr int kYSlots = kThreads * kSpt;      // the groups' partial profiles (G * T <= 4096)
static_assert(kRefineMaxSamples <= kYSlots, "a window's samples fit one lane span each");

__shared__ int2 r_tab[kRefineMaxChannels];   // (i_c, f_c bits) of the current u
__shared__ double r_y[kYSlots];
__shared__ double r_red[32];
__shared__ float r_u;
__shared__ long r_item;

__device__ __forceinline__ double block_sum(double v, int tid)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();                        // r_red free (previous reduction read)
    if ((tid & 63) == 0) r_red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) s += r_red[i];
    return s;
}

struct lane_layout {
    int G;        // groups in the work-group
    int g, s;     // this lane's group and position in it
    bool active;  // g < G
};

__device__ __forceinline__ lane_layout layout_of(int T, int tid)
{
    const int S = ((T + kSpt - 1) / kSpt + 63) / 64 * 64;   // <= 1024 for T <= 4096
    lane_layout L;
    L.G = kThreads / S;
    L.g = tid / S;
    L.s = tid - L.g * S;
    L.active = L.g < L.G;
    return L;
}

// w[m] = row[j + m], m = 0..3, zero outside [0, T).
__device__ __forceinline__ void load4(const float* __restrict__ row, int j, int T, float (&w)[4])
{
    if (j >= 0 && j + 3 < T) {
        float4 q;
        __builtin_memcpy(&q, row + j, sizeof(q));   // dword-aligned 16-byte load
        w[0] = q.x;
        w[1] = q.y;
        w[2] = q.z;
        w[3] = q.w;
    } else {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int i = j + m;
            w[m] = (i >= 0 && i < T) ? row[i] : 0.0f;
        }
    }
}

__device__ __forceinline__ float load1(const float* __restrict__ row, int i, int T)
{
    return (i >= 0 && i < T) ? row[i] : 0.0f;
}

// (L, dL/du) of candidate `cd` at residual DM u; every thread gets both
// (dL/du only when want_grad).  power: the candidate's channel powers v_c
// (loss 1) or null (loss 0).
__device__ void evaluate(const refine_args& a, const refine_cand& cd, const double* __restrict__ power, float u,
                         int tid, bool want_grad, double* loss, double* grad)
{
    const float* __restrict__ x = a.x + cd.x_off;
    const float* __restrict__ ds = a.ds + cd.ds_off;
    const int C = cd.C, T = cd.T;
    double npart = 0.0, dnpart = 0.0;   // this thread's channels of N and dN/du
    __syncthreads();  // r_tab and r_y free
    for (int c = tid; c < C; c += kThreads) {
        const double d = (double)ds[c];
        const double tau = d * (double)u;
        const double fl = fmin(fmax(floor(tau), -1073741824.0), 1073741824.0);
        const float f32 = (float)(tau - fl);
        const int i = (int)fl;
        r_tab[c] = make_int2(i, __float_as_int(f32));
        if (power) {
            const double f = (double)f32;       // the weight the passes use
            const double n0 = (double)max(0, T - abs(i)), n1 = (double)max(0, T - abs(i + 1));
            const double v = power[c];
            npart = __builtin_fma(v, (1.0 - f) * (1.0 - f) * n0 + f * f * n1, npart);
            dnpart = __builtin_fma(v * d, 2.0 * f * n1 - 2.0 * (1.0 - f) * n0, dnpart);
        }
    }
    __syncthreads();

    // pass 1: this group's channels of the profile at the lane's 4 samples.
    // A lane reads only its own 4 samples x[c][t0 + i_c + m]; the second tap
    // of sample t0 - 1 (f x[c][t0 + i_c]) is its left neighbour's, kept in yn
    // and added to the neighbour's sample in LDS after the pass.  The last
    // lane of a group also reads x[c][4S + i_c] when its last sample, 4S - 1,
    // is inside the window (T = 4S).
    const lane_layout L = layout_of(T, tid);
    const int t0 = kSpt * L.s;
    const int S = kThreads / L.G;
    const bool edge = L.active && L.s == S - 1 && kSpt * S == T;
    double y[kSpt], yn = 0.0;
#pragma unroll
    for (int m = 0; m < kSpt; ++m) y[m] = 0.0;
    if (L.active) {
        int c = L.g;
        for (; c + (kU - 1) * L.G < C; c += kU * L.G) {
            float w[kU][4];
            double f[kU];
            int i0[kU];
#pragma unroll
            for (int k = 0; k < kU; ++k) {
                const int cc = c + k * L.G;
                const int2 e = r_tab[cc];
                f[k] = (double)__int_as_float(e.y);
                i0[k] = e.x;
                load4(x + (size_t)cc * T, t0 + e.x, T, w[k]);
            }
#pragma unroll
            for (int k = 0; k < kU; ++k) {
#pragma unroll
                for (int m = 0; m < kSpt; ++m) y[m] = __builtin_fma(1.0 - f[k], (double)w[k][m], y[m]);
#pragma unroll
                for (int m = 1; m < kSpt; ++m) y[m - 1] = __builtin_fma(f[k], (double)w[k][m], y[m - 1]);
                yn = __builtin_fma(f[k], (double)w[k][0], yn);
                if (edge) y[kSpt - 1] = __builtin_fma(f[k], (double)load1(x + (size_t)(c + k * L.G) * T, T + i0[k], T),
                                                      y[kSpt - 1]);
            }
        }
        for (; c < C; c += L.G) {
            float w[4];
            const int2 e = r_tab[c];
            const double f = (double)__int_as_float(e.y);
            load4(x + (size_t)c * T, t0 + e.x, T, w);
#pragma unroll
            for (int m = 0; m < kSpt; ++m) y[m] = __builtin_fma(1.0 - f, (double)w[m], y[m]);
#pragma unroll
            for (int m = 1; m < kSpt; ++m) y[m - 1] = __builtin_fma(f, (double)w[m], y[m - 1]);
            yn = __builtin_fma(f, (double)w[0], yn);
            if (edge) y[kSpt - 1] = __builtin_fma(f, (double)load1(x + (size_t)c * T, T + e.x, T), y[kSpt - 1]);
        }
#pragma unroll
        for (int m = 0; m < kSpt; ++m)
            if (t0 + m < T) r_y[L.g * T + t0 + m] = y[m];
    }
    __syncthreads();
    if (L.active && t0 >= 1 && t0 - 1 < T) r_y[L.g * T + t0 - 1] += yn;   // one lane per slot
    __syncthreads();
    // the whole profile at the lane's samples and at t0 - 1: groups added in order
    double l2 = 0.0, yp = 0.0;
    if (L.active) {
#pragma unroll
        for (int m = 0; m < kSpt; ++m) {
            double s = 0.0;
            if (t0 + m < T)
                for (int g = 0; g < L.G; ++g) s += r_y[g * T + t0 + m];
            y[m] = s;                                   // 0 outside 0 <= t < T
            if (L.g == 0) l2 = __builtin_fma(s, s, l2);
        }
        if (t0 >= 1 && t0 - 1 < T)
            for (int g = 0; g < L.G; ++g) yp += r_y[g * T + t0 - 1];
    }

    if (!want_grad) {
        *loss = block_sum(l2 - npart, tid);
        *grad = 0.0;
        return;
    }
    // pass 2: dE/du over this group's channels, recomputing the window reads:
    // sum_t y(t) (x(t + i + 1) - x(t + i)) = sum_t (y(t - 1) - y(t)) x(t + i)
    // + y(4S - 1) x(4S + i), so each lane weighs its own 4 samples by
    // D[m] = y(t0 + m - 1) - y(t0 + m) (and the last lane, when T = 4S, its
    // sample 4S + i by y(T - 1)).
    double D[kSpt];
    D[0] = yp - y[0];
#pragma unroll
    for (int m = 1; m < kSpt; ++m) D[m] = y[m - 1] - y[m];
    double gs = 0.0;
    if (L.active) {
        int c = L.g;
        for (; c + (kU - 1) * L.G < C; c += kU * L.G) {
            float w[kU][4];
            int i0[kU];
#pragma unroll
            for (int k = 0; k < kU; ++k) {
                const int cc = c + k * L.G;
                i0[k] = r_tab[cc].x;
                load4(x + (size_t)cc * T, t0 + i0[k], T, w[k]);
            }
#pragma unroll
            for (int k = 0; k < kU; ++k) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < kSpt; ++m) s = __builtin_fma(D[m], (double)w[k][m], s);
                if (edge) s = __builtin_fma(y[kSpt - 1], (double)load1(x + (size_t)(c + k * L.G) * T, T + i0[k], T), s);
                gs = __builtin_fma((double)ds[c + k * L.G], s, gs);
            }
        }
        for (; c < C; c += L.G) {
            float w[4];
            const int i0 = r_tab[c].x;
            load4(x + (size_t)c * T, t0 + i0, T, w);
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < kSpt; ++m) s = __builtin_fma(D[m], (double)w[m], s);
            if (edge) s = __builtin_fma(y[kSpt - 1], (double)load1(x + (size_t)c * T, T + i0, T), s);
            gs = __builtin_fma((double)ds[c], s, gs);
        }
    }
    *loss = block_sum(l2 - npart, tid);
    *grad = block_sum(2.0 * gs - dnpart, tid);
}

__global__ __launch_bounds__(kThreads) void refine_kernel(refine_args a)
{
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) r_item = (long)atomicAdd(a.counter, 1);
        __syncthreads();
        const long item = r_item;
        __syncthreads();
        if (item >= a.ncand) break;
        const int b = a.order[item];
        const refine_cand cd = a.cands[b];
        const double* power = a.loss_kind == 1 ? a.power + cd.power_off : nullptr;
        if (tid == 0) r_u = a.u[b];
        __syncthreads();
        float step = a.step[b];
        double prev = a.prev_grad[b];
        double loss = 0.0, grad = 0.0;
        if (a.scan_n > 0) {  // coarse scan around u; the best offset starts the steps
            const float u0 = r_u;
            float best_u = u0;
            double best_l = -INFINITY;
            for (int k = 0; k < a.scan_n; ++k) {
                const float uk = u0 + a.scan_lo + a.scan_step * (float)k;
                evaluate(a, cd, power, uk, tid, false, &loss, &grad);
                if (loss > best_l) {
                    best_l = loss;
                    best_u = uk;
                }
            }
            if (tid == 0) r_u = best_u;
            __syncthreads();
        }
        float best_u = r_u;
        double best_l = -INFINITY, best_g = 0.0;
        for (int s = 0; s <= a.nsteps; ++s) {
            const float u = r_u;
            evaluate(a, cd, power, u, tid, true, &loss, &grad);
            if (loss > best_l) {  // the result is the best point visited
                best_l = loss;
                best_u = u;
                best_g = grad;
            }
            if (s == a.nsteps) break;
            if (tid == 0) {
                if (grad * prev > 0.0) step = fminf(step * 1.2f, a.step_max);
                else if (grad * prev < 0.0) step = fmaxf(step * 0.5f, a.step_min);
                const float du = grad > 0.0 ? step : (grad < 0.0 ? -step : 0.0f);
                r_u = u + du;
                prev = grad;
            }
            __syncthreads();
        }
        if (tid == 0) {
            a.u[b] = best_u;
            a.step[b] = step;
            a.prev_grad[b] = prev;
            a.loss[b] = best_l;
            a.grad[b] = best_g;
        }
        __syncthreads();
    }
}

// v_c = mean_t x[c][t]^2 of every candidate's channels: work-group (b, y)
// of candidate b takes channels 4 y + wave + 64 k, one wave per channel at a
// time with four 256-byte loads in flight per lane; fixed reduction order.
constexpr int kPowerBlocksY = 16;
__global__ __launch_bounds__(256) void refine_power_kernel(refine_args a)
{
    const refine_cand cd = a.cands[blockIdx.x];
    const float* __restrict__ x = a.x + cd.x_off;
    double* __restrict__ v = a.power + cd.power_off;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int T = cd.T;
    for (int c = 4 * (int)blockIdx.y + wave; c < cd.C; c += 4 * kPowerBlocksY) {
        const float* row = x + (size_t)c * T;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        int t = lane;
        for (; t + 192 < T; t += 256) {
            float q[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) q[m] = row[t + 64 * m];
#pragma unroll
            for (int m = 0; m < 4; ++m) s[m] = __builtin_fma((double)q[m], (double)q[m], s[m]);
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {   // fewer than 256 samples left
            if (t + 64 * m < T) {
                const double q = (double)row[t + 64 * m];
                s[m] = __builtin_fma(q, q, s[m]);
            }
        }
        double r = (s[0] + s[1]) + (s[2] + s[3]);
        for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off);
        if (lane == 0) v[c] = r / (double)T;
    }
}

__global__ void refine_reset_kernel(int* counter)
{
    if (threadIdx.x == 0) *counter = 0;
}

}  // namespace

int refine_blocks_per_cu()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, refine_kernel, kThreads, 0) != hipSuccess || n < 1) n = 1;
    return n;
}

hipError_t launch_refine(const refine_args& a, int grid, hipStream_t s)
{
    if (a.ncand <= 0) return hipSuccess;
    if (grid <= 0 || (a.loss_kind != 0 && a.loss_kind != 1) || (a.loss_kind == 1 && !a.power))
        return hipErrorInvalidValue;
    if (a.loss_kind == 1)
        hipLaunchKernelGGL(refine_power_kernel, dim3((unsigned)a.ncand, kPowerBlocksY), dim3(256), 0, s, a);
    hipLaunchKernelGGL(refine_reset_kernel, dim3(1), dim3(64), 0, s, a.counter);
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)grid), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace aa_hip
