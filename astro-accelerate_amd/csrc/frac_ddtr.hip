// Fractional-delay dedispersion of FP32 candidate windows and its adjoint,
// for gfx950 (MI355X).  No counterpart in the reference: it dedisperses with
// the reference's delay table (dmshift[c] / tsamp samples per unit DM) but at
// continuous DMs, with linear interpolation between neighbouring samples, so
// that the output is differentiable in the input and in the DM.
//
// Work-group: 1024 threads = 16 waves; a tile of 1024 output samples, 16
// consecutive samples per lane.  Rows (input channels for the forward pass
// and the DM adjoint, trial profiles for the input adjoint) stream through
// LDS kFracChunkRows at a time, double-buffered: the global loads of chunk
// n+1 are in flight while chunk n is used.  A staged row is the tile's
// window plus the spread of delays over the rows' partners (trials or
// channels), starting 16-byte aligned; its float4 quads are stored in four
// phase planes (quad q at plane q % 4, slot q / 4), so that a lane run's five
// quads 4*lane + e/4 + m are stride-1 across the wave for every m: each of
// the five ds_read_b128 is free of bank conflicts.  The remaining sub-quad
// shift e % 4 is the same for the whole wave and selects one of four
// unrolled FMA bodies.  Per (row, partner, lane): 5 LDS reads, 32 FMAs.
//
// Delays are evaluated in double precision (u = trials[k] + dm[b],
// tau = ds[c] * u), once per (row, partner) and chunk, into a small LDS table
// of (window offset, weight); the weights are then rounded to float.
//
// A window beyond the staged row (the host sizes groups so that this does
// not happen for the DM spans it is given) is read from global memory
// instead, so results never depend on the grouping.
#include "frac_ddtr.h"

#include <algorithm>
#include <cmath>

namespace aa_hip {

namespace {

constexpr int kThreads = 1024;
constexpr int kSpl = 16;                   // samples per lane
constexpr int kRQ = kFracRowQuads;
constexpr int kPlane = kRQ / 4;            // slots per phase plane
constexpr int kCC = kFracChunkRows;
constexpr int kSlots = kCC * kRQ;
constexpr int kSlotsPerThread = (kSlots + kThreads - 1) / kThreads;
static_assert(kRQ % 4 == 0, "row length must be whole phase planes");
static_assert(kFracTile == 64 * kSpl, "tile = 64 lanes x kSpl samples");

__shared__ float4 f_rows0[kSlots];
__shared__ float4 f_rows1[kSlots];
__shared__ double f_red[16];

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// v[0..19] = the five quads of a lane run starting e samples into the row.
// With e/4 = 4 s + P, quad m of the run sits at plane (P + m) & 3, slot
// s + ((P + m) >> 2) + lane: one address per entry (base = row + s + lane)
// and, per phase P, five reads at offsets fixed at compile time.  Each quad
// stays one 16-byte read: the empty asm uses all four words, so the
// compiler can neither narrow a read to the words one shift needs nor move
// it into the shift's branch (words read at a 16-byte lane stride would
// conflict on the LDS banks).
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <int P>
__device__ __forceinline__ void read_window_p(const float4* __restrict__ base, float (&v)[20])
{
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const f32x4 t = *reinterpret_cast<const f32x4*>(base + ((P + m) & 3) * kPlane + ((P + m) >> 2));
        v[4 * m + 0] = t.x;
        v[4 * m + 1] = t.y;
        v[4 * m + 2] = t.z;
        v[4 * m + 3] = t.w;
    }
}

__device__ __forceinline__ void read_window(const float4* __restrict__ row, int e, int lane, float (&v)[20])
{
    const int e4 = e >> 2;
    const float4* base = static_cast<const float4*>(__builtin_assume_aligned(row + (e4 >> 2) + lane, 16));
    switch (e4 & 3) {
    case 0: read_window_p<0>(base, v); break;
    case 1: read_window_p<1>(base, v); break;
    case 2: read_window_p<2>(base, v); break;
    default: read_window_p<3>(base, v); break;
    }
#pragma unroll
    for (int m = 0; m < 5; ++m)
        asm volatile("" ::"v"(v[4 * m + 0]), "v"(v[4 * m + 1]), "v"(v[4 * m + 2]), "v"(v[4 * m + 3]));
}

// v[0..16] = src[start + m] (zero outside 0 <= index < n): the slow path.
__device__ __forceinline__ void read_window_global(const float* __restrict__ src, int n, int start, float (&v)[20])
{
#pragma unroll
    for (int m = 0; m < 17; ++m) {
        const int i = start + m;
        v[m] = (i >= 0 && i < n) ? src[i] : 0.0f;
    }
    v[17] = v[18] = v[19] = 0.0f;
}

// acc[j] += wa * v[S + j] + wb * v[S + j + 1]
template <int S>
__device__ __forceinline__ void lerp_acc(float (&acc)[kSpl], const float (&v)[20], float wa, float wb)
{
#pragma unroll
    for (int j = 0; j < kSpl; ++j) {
        acc[j] = __builtin_fmaf(wa, v[S + j], acc[j]);
        acc[j] = __builtin_fmaf(wb, v[S + j + 1], acc[j]);
    }
}

__device__ __forceinline__ void lerp_acc_s(int s, float (&acc)[kSpl], const float (&v)[20], float wa, float wb)
{
    switch (s) {
    case 0: lerp_acc<0>(acc, v, wa, wb); break;
    case 1: lerp_acc<1>(acc, v, wa, wb); break;
    case 2: lerp_acc<2>(acc, v, wa, wb); break;
    default: lerp_acc<3>(acc, v, wa, wb); break;
    }
}

// sum_m D[m] * v[S + m], m = 0..16
template <int S>
__device__ __forceinline__ float dot17(const float (&D)[17], const float (&v)[20])
{
    float p = 0.0f;
#pragma unroll
    for (int m = 0; m < 17; ++m) p = __builtin_fmaf(D[m], v[S + m], p);
    return p;
}

__device__ __forceinline__ float dot17_s(int s, const float (&D)[17], const float (&v)[20])
{
    switch (s) {
    case 0: return dot17<0>(D, v);
    case 1: return dot17<1>(D, v);
    case 2: return dot17<2>(D, v);
    default: return dot17<3>(D, v);
    }
}

__device__ __forceinline__ int lds_index(int r, int q) { return r * kRQ + (q & 3) * kPlane + (q >> 2); }

// Delay of a (row, partner) pair for the current chunk: the lane-run offset
// e into the staged row (samples) and the interpolation weight f.  Computed
// in double precision, once per chunk, by the staging threads: with float32
// delays the weights would carry errors of ~1e-5, which the DM gradient of a
// squared-profile loss turns into relative errors of ~1e-4.
constexpr int kPartners = 16 * kFracMaxPerWave;   // trials or channels of a group
constexpr int kTab = kCC * kPartners;
static_assert(kTab <= kThreads, "one table entry per thread");
__shared__ int2 f_tab0[kTab];
__shared__ int2 f_tab1[kTab];

__device__ __forceinline__ int ifloor(double v) { return (int)floor(v); }

// Forward (GRAD = false): y[b][k][t] for the 16*KQ trials of group g.
// DM adjoint (GRAD = true): partial sum over this tile and trial group of
//   ds[c] * sum_t gy[b][k][t] * (x[b][c][t + i0 + 1] - x[b][c][t + i0]),
// which is d(loss)/d(dm[b]) restricted to the group (d tau / d dm = ds[c]).
// Rows are channels; a row's staged window starts at the smallest delay of
// the group's trials.
template <int KQ, bool GRAD>
__global__ __launch_bounds__(kThreads) void frac_rows_kernel(frac_args a)
{
    const int tid = threadIdx.x, lane = tid & 63, w = uni(tid >> 6);   // wave-uniform: DM or channel slots branch on it
    int blk = blockIdx.x;
    const int g = blk % a.n_groups;
    blk /= a.n_groups;
    const int tt = blk % a.n_ttiles;
    const int b = blk / a.n_ttiles;
    const int C = a.C, T = a.T, K = a.K;
    const int t0 = tt * kFracTile;
    const double dmb = (double)a.dm[b];
    const int kbase = g * 16 * KQ;
    const int kend = min(K, kbase + 16 * KQ);

    double ulo = INFINITY, uhi = -INFINITY;
    for (int k = kbase; k < kend; ++k) {
        const double u = (double)a.trials[k] + dmb;
        ulo = fmin(ulo, u);
        uhi = fmax(uhi, u);
    }

    bool act[KQ];
#pragma unroll
    for (int q = 0; q < KQ; ++q) act[q] = kbase + w + 16 * q < K;

    float acc[KQ][kSpl];
    float D[KQ][17];
    double accd = 0.0;
    const int tl = t0 + kSpl * lane;  // first sample of this lane
    if constexpr (!GRAD) {
#pragma unroll
        for (int q = 0; q < KQ; ++q)
#pragma unroll
            for (int j = 0; j < kSpl; ++j) acc[q][j] = 0.0f;
    } else {
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            float G[kSpl];
            const int k = kbase + w + 16 * q;
            const float* gk = a.gy + ((size_t)b * K + (act[q] ? k : 0)) * T;
#pragma unroll
            for (int j4 = 0; j4 < kSpl / 4; ++j4) {
                const int t = tl + 4 * j4;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (act[q] && t < T) v = *reinterpret_cast<const float4*>(gk + t);
                G[4 * j4 + 0] = v.x;
                G[4 * j4 + 1] = v.y;
                G[4 * j4 + 2] = v.z;
                G[4 * j4 + 3] = v.w;
            }
            // sum_j G[j] (v[j+1] - v[j]) = sum_m (G[m-1] - G[m]) v[m]
            D[q][0] = -G[0];
#pragma unroll
            for (int m = 1; m < kSpl; ++m) D[q][m] = G[m - 1] - G[m];
            D[q][kSpl] = G[kSpl - 1];
        }
    }

    const float* xb = a.x + (size_t)b * C * T;
    // 4-aligned first sample of channel c's staged window
    auto row_start = [&](int c) -> int {
        const double d = (double)a.ds[c];
        return (t0 + ifloor(d * (d >= 0.0 ? ulo : uhi))) & ~3;
    };

    float4 pre[kSlotsPerThread];
    int2 ent = make_int2(0, 0);
    auto load_chunk = [&](int c0) {
#pragma unroll
        for (int i = 0; i < kSlotsPerThread; ++i) {
            const int s = tid + i * kThreads;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (s < kSlots) {
                const int r = s / kRQ, q = s - r * kRQ, c = c0 + r;
                if (c < C) {
                    const int st = row_start(c) + 4 * q;
                    if (st >= 0 && st < T) v = *reinterpret_cast<const float4*>(xb + (size_t)c * T + st);
                }
            }
            pre[i] = v;
        }
        if (tid < kTab) {
            const int r = tid / kPartners, p = tid - r * kPartners;
            const int c = c0 + r, k = kbase + p;
            ent = make_int2(0, 0);
            if (c < C && p < 16 * KQ && k < K) {
                const double tau = (double)a.ds[c] * ((double)a.trials[k] + dmb);
                const double fl = floor(tau);
                ent = make_int2(t0 + (int)fl - row_start(c), __float_as_int((float)(tau - fl)));
            }
        }
    };
    auto store_chunk = [&](float4* rows, int2* tab) {
#pragma unroll
        for (int i = 0; i < kSlotsPerThread; ++i) {
            const int s = tid + i * kThreads;
            if (s < kSlots) {
                const int r = s / kRQ, q = s - r * kRQ;
                rows[lds_index(r, q)] = pre[i];
            }
        }
        if (tid < kTab) tab[tid] = ent;
    };
    auto compute_chunk = [&](const float4* rows, const int2* tab, int c0) {
        const int nr = min(kCC, C - c0);
        // this wave's (row, partner) entries, one per lane (lane r * KQ + q),
        // read once per chunk and taken from their lanes in the loop
        static_assert(kCC * KQ <= 64, "a wave's entries fit its lanes");
        const int2 mine = lane < kCC * KQ ? tab[(lane / KQ) * kPartners + w + 16 * (lane % KQ)] : make_int2(0, 0);
        for (int r = 0; r < nr; ++r) {
            const int c = c0 + r;
            const float4* row = rows + r * kRQ;
#pragma unroll
            for (int q = 0; q < KQ; ++q) {
                if (!act[q]) continue;
                const int e = __builtin_amdgcn_readlane(mine.x, r * KQ + q);
                const float f = __int_as_float(__builtin_amdgcn_readlane(mine.y, r * KQ + q));
                float v[20];
                if ((unsigned)e <= (unsigned)kFracMaxOffset) {  // also rejects e < 0 (non-finite dm)
                    read_window(row, e, lane, v);
                    if constexpr (GRAD) {
                        const float p = dot17_s(e & 3, D[q], v);
                        accd = __builtin_fma((double)a.ds[c], (double)p, accd);
                    } else {
                        lerp_acc_s(e & 3, acc[q], v, 1.0f - f, f);
                    }
                } else {
                    read_window_global(xb + (size_t)c * T, T, kSpl * lane + e + row_start(c), v);
                    if constexpr (GRAD) {
                        const float p = dot17<0>(D[q], v);
                        accd = __builtin_fma((double)a.ds[c], (double)p, accd);
                    } else {
                        lerp_acc<0>(acc[q], v, 1.0f - f, f);
                    }
                }
            }
        }
    };

    const int nch = (C + kCC - 1) / kCC;
    load_chunk(0);
    store_chunk(f_rows0, f_tab0);
    __syncthreads();
    for (int ch = 0; ch < nch; ch += 2) {
        if (ch + 1 < nch) load_chunk((ch + 1) * kCC);
        compute_chunk(f_rows0, f_tab0, ch * kCC);
        if (ch + 1 < nch) store_chunk(f_rows1, f_tab1);
        __syncthreads();
        if (ch + 1 >= nch) break;
        if (ch + 2 < nch) load_chunk((ch + 2) * kCC);
        compute_chunk(f_rows1, f_tab1, (ch + 1) * kCC);
        if (ch + 2 < nch) store_chunk(f_rows0, f_tab0);
        __syncthreads();
    }

    if constexpr (!GRAD) {
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            if (!act[q]) continue;
            const int k = kbase + w + 16 * q;
            float* yk = a.y + ((size_t)b * K + k) * T;
#pragma unroll
            for (int j4 = 0; j4 < kSpl / 4; ++j4) {
                const int t = tl + 4 * j4;
                if (t < T)
                    *reinterpret_cast<float4*>(yk + t) =
                        make_float4(acc[q][4 * j4], acc[q][4 * j4 + 1], acc[q][4 * j4 + 2], acc[q][4 * j4 + 3]);
            }
        }
    } else {
        for (int off = 32; off > 0; off >>= 1) accd += __shfl_xor(accd, off);
        if (lane == 0) f_red[w] = accd;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int i = 0; i < 16; ++i) sum += f_red[i];
            a.gdm_part[blockIdx.x] = sum;
        }
    }
}

// gdm[b] = sum of the work-group partials of candidate b, in a fixed order.
__global__ void frac_gdm_finalize(const double* __restrict__ part, int per_b, int B, double* __restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int i = 0; i < per_b; ++i) s += part[(size_t)b * per_b + i];
    out[b] = s;
}

// Input adjoint: gx[b][c][s] = sum_k (1 - f) gy[b][k][s - i0] + f gy[b][k][s - i0 - 1]
// for the 16*CQ channels of group g.  Rows are trial profiles gy[b][k]; a
// row's staged window starts at the largest delay of the group's channels.
template <int CQ>
__global__ __launch_bounds__(kThreads) void frac_grad_x_kernel(frac_args a)
{
    const int tid = threadIdx.x, lane = tid & 63, w = uni(tid >> 6);   // wave-uniform: DM or channel slots branch on it
    int blk = blockIdx.x;
    const int g = blk % a.n_groups;
    blk /= a.n_groups;
    const int tt = blk % a.n_ttiles;
    const int b = blk / a.n_ttiles;
    const int C = a.C, T = a.T, K = a.K;
    const int s0 = tt * kFracTile;
    const double dmb = (double)a.dm[b];
    const int cbase = g * 16 * CQ;
    const int cend = min(C, cbase + 16 * CQ);

    double dlo = INFINITY, dhi = -INFINITY;
    for (int c = cbase; c < cend; ++c) {
        const double d = (double)a.ds[c];
        dlo = fmin(dlo, d);
        dhi = fmax(dhi, d);
    }

    bool act[CQ];
    float acc[CQ][kSpl];
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
        act[q] = cbase + w + 16 * q < C;
#pragma unroll
        for (int j = 0; j < kSpl; ++j) acc[q][j] = 0.0f;
    }
    const int sl = s0 + kSpl * lane;
    const float* gb = a.gy + (size_t)b * K * T;

    // 4-aligned first sample of trial k's staged window
    auto row_start = [&](int k) -> int {
        const double u = (double)a.trials[k] + dmb;
        return (s0 - 1 - ifloor(fmax(dlo * u, dhi * u))) & ~3;
    };

    float4 pre[kSlotsPerThread];
    int2 ent = make_int2(0, 0);
    auto load_chunk = [&](int k0) {
#pragma unroll
        for (int i = 0; i < kSlotsPerThread; ++i) {
            const int s = tid + i * kThreads;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (s < kSlots) {
                const int r = s / kRQ, q = s - r * kRQ, k = k0 + r;
                if (k < K) {
                    const int st = row_start(k) + 4 * q;
                    if (st >= 0 && st < T) v = *reinterpret_cast<const float4*>(gb + (size_t)k * T + st);
                }
            }
            pre[i] = v;
        }
        if (tid < kTab) {
            const int r = tid / kPartners, p = tid - r * kPartners;
            const int k = k0 + r, c = cbase + p;
            ent = make_int2(0, 0);
            if (k < K && p < 16 * CQ && c < C) {
                const double tau = (double)a.ds[c] * ((double)a.trials[k] + dmb);
                const double fl = floor(tau);
                ent = make_int2(s0 - 1 - (int)fl - row_start(k), __float_as_int((float)(tau - fl)));
            }
        }
    };
    auto store_chunk = [&](float4* rows, int2* tab) {
#pragma unroll
        for (int i = 0; i < kSlotsPerThread; ++i) {
            const int s = tid + i * kThreads;
            if (s < kSlots) {
                const int r = s / kRQ, q = s - r * kRQ;
                rows[lds_index(r, q)] = pre[i];
            }
        }
        if (tid < kTab) tab[tid] = ent;
    };
    auto compute_chunk = [&](const float4* rows, const int2* tab, int k0) {
        const int nr = min(kCC, K - k0);
        // this wave's (row, partner) entries, one per lane (lane r * CQ + q),
        // read once per chunk and taken from their lanes in the loop
        static_assert(kCC * CQ <= 64, "a wave's entries fit its lanes");
        const int2 mine = lane < kCC * CQ ? tab[(lane / CQ) * kPartners + w + 16 * (lane % CQ)] : make_int2(0, 0);
        for (int r = 0; r < nr; ++r) {
            const int k = k0 + r;
            const float4* row = rows + r * kRQ;
#pragma unroll
            for (int q = 0; q < CQ; ++q) {
                if (!act[q]) continue;
                const int e = __builtin_amdgcn_readlane(mine.x, r * CQ + q);
                const float f = __int_as_float(__builtin_amdgcn_readlane(mine.y, r * CQ + q));
                float v[20];
                // v[s + j] = gy[k][sl + j - i0 - 1] (weight f), v[s + j + 1] = gy[k][sl + j - i0] (weight 1 - f)
                if ((unsigned)e <= (unsigned)kFracMaxOffset) {  // also rejects e < 0 (non-finite dm)
                    read_window(row, e, lane, v);
                    lerp_acc_s(e & 3, acc[q], v, f, 1.0f - f);
                } else {
                    read_window_global(gb + (size_t)k * T, T, kSpl * lane + e + row_start(k), v);
                    lerp_acc<0>(acc[q], v, f, 1.0f - f);
                }
            }
        }
    };

    const int nch = (K + kCC - 1) / kCC;
    load_chunk(0);
    store_chunk(f_rows0, f_tab0);
    __syncthreads();
    for (int ch = 0; ch < nch; ch += 2) {
        if (ch + 1 < nch) load_chunk((ch + 1) * kCC);
        compute_chunk(f_rows0, f_tab0, ch * kCC);
        if (ch + 1 < nch) store_chunk(f_rows1, f_tab1);
        __syncthreads();
        if (ch + 1 >= nch) break;
        if (ch + 2 < nch) load_chunk((ch + 2) * kCC);
        compute_chunk(f_rows1, f_tab1, (ch + 1) * kCC);
        if (ch + 2 < nch) store_chunk(f_rows0, f_tab0);
        __syncthreads();
    }

#pragma unroll
    for (int q = 0; q < CQ; ++q) {
        if (!act[q]) continue;
        const int c = cbase + w + 16 * q;
        float* gxc = a.gx + ((size_t)b * C + c) * T;
#pragma unroll
        for (int j4 = 0; j4 < kSpl / 4; ++j4) {
            const int s = sl + 4 * j4;
            if (s < T)
                *reinterpret_cast<float4*>(gxc + s) =
                    make_float4(acc[q][4 * j4], acc[q][4 * j4 + 1], acc[q][4 * j4 + 2], acc[q][4 * j4 + 3]);
        }
    }
}

template <bool GRAD>
hipError_t launch_rows(const frac_args& a, hipStream_t s)
{
    const dim3 grid((unsigned)((size_t)a.B * a.n_ttiles * a.n_groups));
    switch (a.group_q) {
    case 2: hipLaunchKernelGGL((frac_rows_kernel<2, GRAD>), grid, dim3(kThreads), 0, s, a); break;
    case 1: hipLaunchKernelGGL((frac_rows_kernel<1, GRAD>), grid, dim3(kThreads), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

int frac_trials_per_wave(const float* h_trials, int K, float ds_absmax)
{
    for (int q = kFracMaxPerWave; q >= 1; q >>= 1) {
        bool ok = true;
        for (int k0 = 0; k0 < K && ok; k0 += 16 * q) {
            float lo = h_trials[k0], hi = h_trials[k0];
            for (int k = k0; k < std::min(K, k0 + 16 * q); ++k) {
                lo = std::min(lo, h_trials[k]);
                hi = std::max(hi, h_trials[k]);
            }
            // delay spread of the group plus floor/alignment slack
            const double spread = (double)ds_absmax * ((double)hi - (double)lo) * (1.0 + 1e-5) + 8.0;
            ok = spread <= (double)kFracMaxOffset;
        }
        if (ok) return q;
    }
    return 0;
}

hipError_t launch_frac_forward(const frac_args& a, hipStream_t s) { return launch_rows<false>(a, s); }

size_t frac_grad_dm_parts(const frac_args& a) { return (size_t)a.B * a.n_ttiles * a.n_groups; }

hipError_t launch_frac_grad_dm(const frac_args& a, hipStream_t s)
{
    const hipError_t e = launch_rows<true>(a, s);
    if (e != hipSuccess) return e;
    const int per_b = a.n_ttiles * a.n_groups;
    hipLaunchKernelGGL(frac_gdm_finalize, dim3((a.B + 255) / 256), dim3(256), 0, s, a.gdm_part, per_b, a.B, a.gdm);
    return hipGetLastError();
}

hipError_t launch_frac_grad_x(const frac_args& a, hipStream_t s)
{
    const dim3 grid((unsigned)((size_t)a.B * a.n_ttiles * a.n_groups));
    switch (a.group_q) {
    case 2: hipLaunchKernelGGL((frac_grad_x_kernel<2>), grid, dim3(kThreads), 0, s, a); break;
    case 1: hipLaunchKernelGGL((frac_grad_x_kernel<1>), grid, dim3(kThreads), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace aa_hip
