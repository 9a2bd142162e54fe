// Batched DM refinement of candidate windows of uneven size, for gfx950.
// No counterpart in the reference.  Each candidate b is a float32 window
// x_b[c][t] (C_b channels x T_b samples) cut along the candidate's
// dispersion track; its continuous residual DM u_b is refined by gradient
// ascent on the energy of the dedispersed profile,
//     L(u) = sum_t y(t)^2,  y(t) = sum_c (1 - f_c) x[c][t + i_c] + f_c x[c][t + i_c + 1],
//     tau_c = ds[c] * u (double), i_c = floor(tau_c), f_c = tau_c - i_c,
// with x = 0 outside [0, T) and ds[c] the delay per unit DM in samples.
//     dL/du = sum_c ds[c] sum_t 2 y(t) (x[c][t + i_c + 1] - x[c][t + i_c]).
//
// An optional coarse scan first evaluates L (forward pass only) at scan_n
// offsets u + scan_lo + k * scan_step and starts the steps from the best.
//
// One persistent launch: a work-group (1024 threads) takes candidates from a
// queue (largest first) and runs every step of that candidate itself, so the
// forward pass, the adjoint pass and the update are fused: pass 1 sums the
// profile in double (each thread owns up to 4 samples, t = tid + 1024 j, and
// keeps y in registers), pass 2 re-reads the window (recomputation: no shifted tile is
// stored) and accumulates the DM derivative, and thread 0 applies a
// sign-based adaptive step (Rprop: the step grows by 1.2 while the sign of
// dL/du holds and halves when it flips); the result is the best point
// visited.  Windows are read straight from
// global memory, coalesced along t: with one trial per candidate there is no
// reuse to stage, and the passes are bound by HBM bandwidth.
#include "frac_refine.h"

#include <cmath>

namespace aa_hip {

namespace {

constexpr int kThreads = 1024;
constexpr int kPerThread = kRefineMaxSamples / kThreads;   // 4
constexpr int kUnroll = 4;                                  // channels in flight per thread

__shared__ int2 r_tab[kRefineMaxChannels];   // (i_c, f_c bits) of the current u
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

__device__ __forceinline__ float load_or_zero(const float* __restrict__ row, int idx, int T)
{
    return (idx >= 0 && idx < T) ? row[idx] : 0.0f;
}

// (L, dL/du) of candidate `cd` at residual DM u; every thread gets both
// (dL/du only when want_grad).
__device__ void evaluate(const refine_args& a, const refine_cand& cd, float u, int tid, bool want_grad,
                         double* loss, double* grad)
{
    const float* __restrict__ x = a.x + cd.x_off;
    const float* __restrict__ ds = a.ds + cd.ds_off;
    const int C = cd.C, T = cd.T;
    __syncthreads();  // r_tab free
    for (int c = tid; c < C; c += kThreads) {
        const double tau = (double)ds[c] * (double)u;
        const double fl = floor(tau);
        r_tab[c] = make_int2((int)fl, __float_as_int((float)(tau - fl)));
    }
    __syncthreads();

    // The profile is accumulated in double: at an aligned pulse y reaches
    // ~C/2 and float32 rounding of y would cost ~1e-5 of dL/du.
    double y[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) y[j] = 0.0;
    int c = 0;
    for (; c + kUnroll <= C; c += kUnroll) {
        float x0[kUnroll][kPerThread], x1[kUnroll][kPerThread];
        double w[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
            const int2 e = r_tab[c + k];
            w[k] = (double)__int_as_float(e.y);
            const float* row = x + (size_t)(c + k) * T;
#pragma unroll
            for (int j = 0; j < kPerThread; ++j) {
                const int i = tid + kThreads * j + e.x;
                x0[k][j] = load_or_zero(row, i, T);
                x1[k][j] = load_or_zero(row, i + 1, T);
            }
        }
#pragma unroll
        for (int k = 0; k < kUnroll; ++k)
#pragma unroll
            for (int j = 0; j < kPerThread; ++j)
                y[j] = __builtin_fma(w[k], (double)x1[k][j], __builtin_fma(1.0 - w[k], (double)x0[k][j], y[j]));
    }
    for (; c < C; ++c) {
        const int2 e = r_tab[c];
        const double f = (double)__int_as_float(e.y);
        const float* row = x + (size_t)c * T;
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
            const int i = tid + kThreads * j + e.x;
            y[j] = __builtin_fma(f, (double)load_or_zero(row, i + 1, T),
                                 __builtin_fma(1.0 - f, (double)load_or_zero(row, i, T), y[j]));
        }
    }
    double l2 = 0.0;
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        if (tid + kThreads * j >= T) y[j] = 0.0;    // outputs only for 0 <= t < T
        l2 += y[j] * y[j];
    }

    if (!want_grad) {
        *loss = block_sum(l2, tid);
        *grad = 0.0;
        return;
    }
    // pass 2: dL/du, recomputing the window reads
    double g = 0.0;
    c = 0;
    for (; c + kUnroll <= C; c += kUnroll) {
        double d[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
            const int2 e = r_tab[c + k];
            const float* row = x + (size_t)(c + k) * T;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < kPerThread; ++j) {
                const int i = tid + kThreads * j + e.x;
                s = __builtin_fma(y[j], (double)(load_or_zero(row, i + 1, T) - load_or_zero(row, i, T)), s);
            }
            d[k] = s;
        }
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) g = __builtin_fma((double)ds[c + k], d[k], g);
    }
    for (; c < C; ++c) {
        const int2 e = r_tab[c];
        const float* row = x + (size_t)c * T;
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
            const int i = tid + kThreads * j + e.x;
            s = __builtin_fma(y[j], (double)(load_or_zero(row, i + 1, T) - load_or_zero(row, i, T)), s);
        }
        g = __builtin_fma((double)ds[c], s, g);
    }
    *loss = block_sum(l2, tid);
    *grad = 2.0 * block_sum(g, tid);
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
        if (tid == 0) r_u = a.u[b];
        __syncthreads();
        float step = a.step[b];
        double prev = a.prev_grad[b];
        double loss = 0.0, grad = 0.0;
        if (a.scan_n > 0) {  // coarse energy scan around u; the best offset starts the steps
            const float u0 = r_u;
            float best_u = u0;
            double best_l = -1.0;
            for (int k = 0; k < a.scan_n; ++k) {
                const float uk = u0 + a.scan_lo + a.scan_step * (float)k;
                evaluate(a, cd, uk, tid, false, &loss, &grad);
                if (loss > best_l) {
                    best_l = loss;
                    best_u = uk;
                }
            }
            if (tid == 0) r_u = best_u;
            __syncthreads();
        }
        float best_u = r_u;
        double best_l = -1.0, best_g = 0.0;
        for (int s = 0; s <= a.nsteps; ++s) {
            const float u = r_u;
            evaluate(a, cd, u, tid, true, &loss, &grad);
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
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(refine_reset_kernel, dim3(1), dim3(64), 0, s, a.counter);
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)grid), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace aa_hip
