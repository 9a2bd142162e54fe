// First pass of the single-pulse search on one DM-time plane, the stage the
// reference's runner hands each range's device plane to
// (include/aa_permitted_pipelines_generic.hpp:848-889 -> analysis_GPU,
// src/aa_device_analysis.cu:126-133 reading [nDMs][nTimesamples]):
//
//  * boxcar MSD (the role of MSD_plane_profile, src/aa_device_MSD_plane_profile.cu):
//    mean and population sd of every in-row boxcar sum B_w(t) = sum_{i<w}
//    x[d][t+i], t + w <= nsamp, for w = 1..W (W <= 32), over the whole
//    plane.  Exact per width (the reference interpolates between decimated
//    widths).  One read of the plane.
//  * boxcar SNR (the first plane of SPDT, src/aa_device_SPS_long_kernel.cu:42-168):
//    snr[d][t] = max over widths that fit of (B_w(t) - mean_w) / sd_w and
//    width[d][t] = that w (ties: the smaller w; sd_w == 0 gives SNR 0).
//    One more read of the plane, one write of snr (float) and width (u16).
//
// Layout and numerics.  A tile is kTile consecutive start samples of one DM
// row plus the W-1 samples after them, staged in LDS.  Sums run in float on
// centred values (x - c), so a boxcar sum holds the noise and not the
// plane's offset: the MSD kernel centres each tile on its own mean, turns
// the tile's per-width float sums into (count, mean, centred second moment)
// in double and merges them into its workgroup's running moments with
// Chan's pairwise update (a persistent grid of kBlocks workgroups, tiles
// assigned by stride); one finalize workgroup merges the workgroups' moments
// in a fixed order, so the result does not depend on scheduling.  The SNR
// kernel centres on the plane mean c = mean_1 and divides
// (B'_w - (mean_w - w c)) by sd_w.
//
// Cost per start sample and width: 2 (MSD at W = 32, else 3) or 5 (SNR)
// VALU operations.  At W = 32 both passes read a run of consecutive start
// samples' samples into registers once (8 in the MSD pass, 4 in the SNR
// pass), else one LDS read per start sample and width.  VALU-bound; the
// plane is read twice from HBM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "ddtr_kernels.h"

namespace aa_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kK = 32;                    // start samples per thread per tile
constexpr int kTile = kThreads * kK;      // 8192 start samples of one row
constexpr int kMaxW = SPD_MAX_WIDTH;      // 32
constexpr int kBlocks = 1024;             // persistent MSD grid

struct spd_aux {              // finalize -> SNR kernel
    float c;                  // centring value: mean of width-1 sums (the plane mean)
    float pad[3];
    float off[kMaxW];         // mean_w - w * c
    float rsd[kMaxW];         // 1 / sd_w (0 when sd_w == 0)
};

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// v plus the value `ctrl` selects from another lane (DPP); lanes the row
// mask leaves out, or whose source is out of range, add 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

// Sum over the wave's 64 lanes, in lane 63 only (other lanes hold partial
// sums): quad sums (quad_perm 1,0,3,2 then 2,3,0,1), row sums (row_shr 4
// then 8, lanes 12-15 of each row), then row_bcast15 and row_bcast31.  A
// fixed order of additions, in registers (no LDS traffic).
__device__ __forceinline__ float wave_sum_to_lane63(float v)
{
    v = dpp_add<0xB1, 0xf>(v);
    v = dpp_add<0x4E, 0xf>(v);
    v = dpp_add<0x114, 0xf>(v);
    v = dpp_add<0x118, 0xf>(v);
    v = dpp_add<0x142, 0xa>(v);
    v = dpp_add<0x143, 0xc>(v);
    return v;
}

// Chan et al.'s pairwise update: (n_a, m_a, M2_a) += (n_b, m_b, M2_b).
__device__ __forceinline__ void chan_merge(double& na, double& ma, double& m2a, double nb, double mb, double m2b)
{
    if (nb <= 0.0) return;
    const double n = na + nb;
    const double delta = mb - ma;
    ma += delta * (nb / n);
    m2a += m2b + delta * delta * (na * nb / n);
    na = n;
}

// Per-workgroup moments over its tiles: partial[blk][w] = (n_w, mean of B_w,
// sum of squared deviations of B_w).
// WC: the width limit when it is known at compile time (32, the default),
// else 0 and W_arg.  A runtime limit costs a scalar condition per width
// (32 of them live across the loops, spilled to vector lanes).
// Four waves per SIMD (128 registers a lane): the persistent grid of
// kBlocks work-groups is resident at once only at four work-groups per CU.
template <int WC>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void spd_msd_kernel(
    const float* __restrict__ plane, long pitch, int ndms, int nsamp, int W_arg, double* __restrict__ partial)
{
    const int W = WC > 0 ? WC : W_arg;
    __shared__ __align__(16) float s[kTile + kMaxW];
    __shared__ float red[kThreads / 64][2 * kMaxW];
    __shared__ float wsum[kThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const long per_row = ((long)nsamp + kTile - 1) / kTile;
    const long ntiles = (long)ndms * per_row;
    double accn = 0.0, accm = 0.0, acc2 = 0.0;   // thread tid < W: width tid + 1
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long d = tile / per_row;
        const long t0 = (tile - d * per_row) * kTile;
        const float* row = plane + d * pitch;
        const long rest = (long)nsamp - t0;
        const int nout = (int)(rest < kTile ? rest : kTile);                 // start samples in the tile
        const int nload = (int)(rest < kTile + W - 1 ? rest : kTile + W - 1);  // samples staged
        __syncthreads();   // the previous tile's readers of s and red are done
        float part = 0.0f;
        for (int i = tid; i < kTile + kMaxW; i += kThreads) {
            const float v = i < nload ? row[t0 + i] : 0.0f;
            s[i] = v;
            if (i < nout) part += v;
        }
        part = wave_sum(part);
        if (lane == 0) wsum[wv] = part;
        __syncthreads();
        const float c = (wsum[0] + wsum[1] + wsum[2] + wsum[3]) / (float)nout;
        for (int i = tid; i < nload; i += kThreads) s[i] -= c;
        __syncthreads();
        float a1[kMaxW], a2[kMaxW];
#pragma unroll
        for (int w = 0; w < kMaxW; ++w) a1[w] = a2[w] = 0.0f;
        // All 32 widths fit at every start sample (a wave-uniform case): the
        // sums of B_w over the tile follow from the tile's sum C_0 and its
        // edges (below), so the loop keeps only the running sum and the
        // square sums; runs of kRun consecutive start samples per thread,
        // their kRun + 31 samples read once into registers with 16-byte
        // reads (instead of one LDS read per start sample and width).
        const bool fast = WC == kMaxW && nload == kTile + W - 1;
        if (fast) {
            constexpr int kRun = 8, kWin = kRun + kMaxW;   // 40 registers: s[j0 .. j0 + 39]
            static_assert(kTile % (kRun * kThreads) == 0, "whole runs per thread");
            float cs = 0.0f;   // this thread's part of C_0 = sum of s[0 .. kTile)
#pragma unroll 1
            for (int k = 0; k < kTile / (kRun * kThreads); ++k) {
                const int j0 = kRun * (tid + kThreads * k);
                float v[kWin];
#pragma unroll
                for (int m = 0; m < kWin / 4; ++m) {
                    const float4 q = *reinterpret_cast<const float4*>(s + j0 + 4 * m);
                    v[4 * m] = q.x;
                    v[4 * m + 1] = q.y;
                    v[4 * m + 2] = q.z;
                    v[4 * m + 3] = q.w;
                }
#pragma unroll
                for (int i = 0; i < kRun; ++i) {
                    cs += v[i];
                    float B = 0.0f;
#pragma unroll
                    for (int w = 0; w < kMaxW; ++w) {
                        B += v[i + w];
                        a2[w] = __builtin_fmaf(B, B, a2[w]);
                    }
                }
            }
            a1[0] = cs;
        } else if (nload == kTile + W - 1) {   // every width fits at every start sample
#pragma unroll 2
            for (int k = 0; k < kK; ++k) {
                const int j = tid + kThreads * k;
                float B = 0.0f;
#pragma unroll
                for (int w = 0; w < kMaxW; ++w) {
                    if (w < W) {
                        B += s[j + w];
                        a1[w] += B;
                        a2[w] = __builtin_fmaf(B, B, a2[w]);
                    }
                }
            }
        } else {   // the row's last tile: widths that fit
            for (int k = 0; k < kK; ++k) {
                const int j = tid + kThreads * k;
                if (j >= nout) break;
                float B = 0.0f;
#pragma unroll
                for (int w = 0; w < kMaxW; ++w) {
                    if (w < W && j + w < nload) {
                        B += s[j + w];
                        a1[w] += B;
                        a2[w] = __builtin_fmaf(B, B, a2[w]);
                    }
                }
            }
        }
#pragma unroll
        for (int w = 0; w < kMaxW; ++w) {
            if (w < W) {
                const float r1 = fast && w > 0 ? 0.0f : wave_sum_to_lane63(a1[w]);
                const float r2 = wave_sum_to_lane63(a2[w]);
                if (lane == 63) {
                    red[wv][w] = r1;
                    red[wv][kMaxW + w] = r2;
                }
            }
        }
        __syncthreads();
        if (tid < W) {
            const int w = tid + 1;
            double S1;
            if (fast) {
                // sum over the tile's starts j of B_w(j) = sum_{m<w} C_m, with
                // C_m = sum_{j<kTile} s[j + m] = C_{m-1} + s[kTile - 1 + m] - s[m - 1]:
                // S1 = w C_0 + sum_{i=1}^{w-1} (w - i) (s[kTile - 1 + i] - s[i - 1])
                S1 = (double)w * ((double)red[0][0] + (double)red[1][0] + (double)red[2][0] + (double)red[3][0]);
                for (int i = 1; i < w; ++i) S1 += (double)(w - i) * ((double)s[kTile - 1 + i] - (double)s[i - 1]);
            } else {
                S1 = (double)red[0][tid] + (double)red[1][tid] + (double)red[2][tid] + (double)red[3][tid];
            }
            const double S2 = (double)red[0][kMaxW + tid] + (double)red[1][kMaxW + tid] +
                              (double)red[2][kMaxW + tid] + (double)red[3][kMaxW + tid];
            const long fit = (long)nload - w + 1;               // start samples whose width-w boxcar is staged
            const double n = (double)(fit < nout ? (fit > 0 ? fit : 0) : nout);
            if (n > 0.0) {   // the tile's moments of B_w = B'_w + w c
                const double mt = S1 / n;
                const double m2t = S2 - S1 * mt;
                chan_merge(accn, accm, acc2, n, mt + (double)w * (double)c, m2t > 0.0 ? m2t : 0.0);
            }
        }
    }
    if (tid < W) {
        double* p = partial + ((long)blockIdx.x * kMaxW + tid) * 3;
        p[0] = accn;
        p[1] = accm;
        p[2] = acc2;
    }
}

// Merges the workgroups' moments in a fixed order; writes (mean_w, sd_w)
// and the SNR kernel's centring constants.  kFinLanes threads per width:
// lane i merges the partials of blocks i, i + kFinLanes, ... in turn, then
// the lanes' moments are merged pairwise in a fixed tree.
constexpr int kFinLanes = 32;
constexpr int kFinThreads = kMaxW * kFinLanes;
__global__ __launch_bounds__(kFinThreads) void spd_msd_finalize_kernel(const double* __restrict__ partial,
                                                                       int nblocks, int W, float* __restrict__ msd,
                                                                       spd_aux* __restrict__ aux)
{
    __shared__ double tn[kMaxW][kFinLanes], tm[kMaxW][kFinLanes], t2[kMaxW][kFinLanes];
    __shared__ double mean[kMaxW], sd[kMaxW];
    const int w = threadIdx.x / kFinLanes, i = threadIdx.x % kFinLanes;
    double n = 0.0, m = 0.0, m2 = 0.0;
    if (w < W)
        for (int b = i; b < nblocks; b += kFinLanes) {
            const double* p = partial + ((long)b * kMaxW + w) * 3;
            chan_merge(n, m, m2, p[0], p[1], p[2]);
        }
    tn[w][i] = n;
    tm[w][i] = m;
    t2[w][i] = m2;
    for (int h = kFinLanes / 2; h > 0; h >>= 1) {
        __syncthreads();
        if (i < h) {
            chan_merge(n, m, m2, tn[w][i + h], tm[w][i + h], t2[w][i + h]);
            tn[w][i] = n;
            tm[w][i] = m;
            t2[w][i] = m2;
        }
    }
    if (w < W && i == 0) {
        const double v = n > 0.0 ? m2 / n : 0.0;
        mean[w] = m;
        sd[w] = v > 0.0 ? sqrt(v) : 0.0;
        msd[2 * w] = (float)mean[w];
        msd[2 * w + 1] = (float)sd[w];
    }
    __syncthreads();
    if (w < W && i == 0) {
        const double c = (double)(float)mean[0];
        if (w == 0) aux->c = (float)c;
        aux->off[w] = (float)(mean[w] - (double)(w + 1) * c);
        aux->rsd[w] = sd[w] > 0.0 ? (float)(1.0 / sd[w]) : 0.0f;
    }
}

template <int WC>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void spd_snr_kernel(
    const float* __restrict__ plane, long pitch, int nsamp, int W_arg, const spd_aux* __restrict__ aux,
    float* __restrict__ snr, long snr_pitch, uint16_t* __restrict__ width, long width_pitch)
{
    const int W = WC > 0 ? WC : W_arg;
    __shared__ __align__(16) float s[kTile + kMaxW];
    __shared__ float off[kMaxW], rsd[kMaxW];
    const int tid = threadIdx.x;
    const long per_row = ((long)nsamp + kTile - 1) / kTile;
    const long d = blockIdx.x / per_row;
    const long t0 = (blockIdx.x - d * per_row) * kTile;
    const float* row = plane + d * pitch;
    const long rest = (long)nsamp - t0;
    const int nout = (int)(rest < kTile ? rest : kTile);
    const int nload = (int)(rest < kTile + W - 1 ? rest : kTile + W - 1);
    const float c = aux->c;
    if (tid < kMaxW) {
        off[tid] = aux->off[tid];
        rsd[tid] = aux->rsd[tid];
    }
    for (int i = tid; i < kTile + kMaxW; i += kThreads) s[i] = i < nload ? row[t0 + i] - c : 0.0f;
    __syncthreads();
    float* so = snr ? snr + d * snr_pitch + t0 : nullptr;
    uint16_t* wo = width ? width + d * width_pitch + t0 : nullptr;
    if (nload == kTile + W - 1) {   // every width fits at every start sample
        // SNR_w = B'_w * rsd_w - off_w * rsd_w: one FMA per width
        float p[kMaxW], r[kMaxW];
#pragma unroll
        for (int w = 0; w < kMaxW; ++w) {
            r[w] = rsd[w];
            p[w] = -off[w] * r[w];
        }
        if constexpr (WC == kMaxW) {
            // runs of kRun consecutive start samples per thread, their
            // kRun + 31 samples read into registers once
            constexpr int kRun = 4, kWin = kRun + kMaxW;   // 36 registers: s[j0 .. j0 + 35]
            static_assert(kTile % (kRun * kThreads) == 0, "whole runs per thread");
#pragma unroll 1
            for (int k = 0; k < kTile / (kRun * kThreads); ++k) {
                const int j0 = kRun * (tid + kThreads * k);
                float v[kWin];
#pragma unroll
                for (int m = 0; m < kWin / 4; ++m) {
                    const float4 q = *reinterpret_cast<const float4*>(s + j0 + 4 * m);
                    v[4 * m] = q.x;
                    v[4 * m + 1] = q.y;
                    v[4 * m + 2] = q.z;
                    v[4 * m + 3] = q.w;
                }
#pragma unroll
                for (int i = 0; i < kRun; ++i) {
                    float B = v[i];
                    float best = __builtin_fmaf(B, r[0], p[0]);
                    int bw = 1;
#pragma unroll
                    for (int w = 1; w < kMaxW; ++w) {
                        B += v[i + w];
                        const float x = __builtin_fmaf(B, r[w], p[w]);
                        if (x > best) {
                            best = x;
                            bw = w + 1;
                        }
                    }
                    if (so) so[j0 + i] = best;
                    if (wo) wo[j0 + i] = (uint16_t)bw;
                }
            }
            return;
        }
#pragma unroll 2
        for (int k = 0; k < kK; ++k) {
            const int j = tid + kThreads * k;
            float B = s[j];
            float best = __builtin_fmaf(B, r[0], p[0]);
            int bw = 1;
#pragma unroll
            for (int w = 1; w < kMaxW; ++w) {
                if (w < W) {
                    B += s[j + w];
                    const float v = __builtin_fmaf(B, r[w], p[w]);
                    if (v > best) {
                        best = v;
                        bw = w + 1;
                    }
                }
            }
            if (so) so[j] = best;
            if (wo) wo[j] = (uint16_t)bw;
        }
        return;
    }
    for (int k = 0; k < kK; ++k) {
        const int j = tid + kThreads * k;
        if (j >= nout) break;
        float B = 0.0f, best = 0.0f;
        int bw = 0;
#pragma unroll
        for (int w = 0; w < kMaxW; ++w) {
            if (w < W && j + w < nload) {
                B += s[j + w];
                const float v = (B - off[w]) * rsd[w];
                if (bw == 0 || v > best) {
                    best = v;
                    bw = w + 1;
                }
            }
        }
        if (so) so[j] = best;
        if (wo) wo[j] = (uint16_t)bw;
    }
}

}  // namespace

size_t spd_workspace_bytes() { return (size_t)kBlocks * kMaxW * 3 * sizeof(double) + 256 + sizeof(spd_aux); }

hipError_t launch_spd_first_pass(const float* plane, size_t pitch, int ndms, int nsamp, int W, float* msd,
                                 float* snr, size_t snr_pitch, uint16_t* width, size_t width_pitch, void* work,
                                 hipStream_t s)
{
    if (!plane || !msd || !work || ndms <= 0 || nsamp <= 0 || W < 1 || W > kMaxW || W > nsamp ||
        pitch < (size_t)nsamp)
        return hipErrorInvalidValue;
    if ((snr && snr_pitch < (size_t)nsamp) || (width && width_pitch < (size_t)nsamp)) return hipErrorInvalidValue;
    double* partial = static_cast<double*>(work);
    spd_aux* aux = reinterpret_cast<spd_aux*>(static_cast<char*>(work) + (size_t)kBlocks * kMaxW * 3 * sizeof(double));
    const long per_row = ((long)nsamp + kTile - 1) / kTile;
    const long ntiles = (long)ndms * per_row;
    if (ntiles > 0x7fffffffL) return hipErrorInvalidValue;
    if (W == kMaxW)
        hipLaunchKernelGGL(spd_msd_kernel<kMaxW>, dim3(kBlocks), dim3(kThreads), 0, s, plane, (long)pitch, ndms, nsamp,
                           W, partial);
    else
        hipLaunchKernelGGL(spd_msd_kernel<0>, dim3(kBlocks), dim3(kThreads), 0, s, plane, (long)pitch, ndms, nsamp, W,
                           partial);
    hipLaunchKernelGGL(spd_msd_finalize_kernel, dim3(1), dim3(kFinThreads), 0, s, (const double*)partial, kBlocks, W, msd, aux);
    if (snr || width) {
        if (W == kMaxW)
            hipLaunchKernelGGL(spd_snr_kernel<kMaxW>, dim3((unsigned)ntiles), dim3(kThreads), 0, s, plane, (long)pitch,
                               nsamp, W, (const spd_aux*)aux, snr, (long)snr_pitch, width, (long)width_pitch);
        else
            hipLaunchKernelGGL(spd_snr_kernel<0>, dim3((unsigned)ntiles), dim3(kThreads), 0, s, plane, (long)pitch,
                               nsamp, W, (const spd_aux*)aux, snr, (long)snr_pitch, width, (long)width_pitch);
    }
    return hipGetLastError();
}

// ------------------------------------------------ allocation-free variant
// Workspace of launch_spd_first_pass_static: a file-scope device array, one
// copy per device, held until the process ends.  The first pass needs the
// same fixed workspace for every plane (the persistent MSD grid's per-width
// moments and the finalize constants), so this size covers every call whose
// width limit is at most SPD_MAX_WIDTH.  Every byte a call reads is written
// earlier in the same call, so nothing needs clearing between calls.
constexpr size_t kSpdScratchBytes = (size_t)kBlocks * kMaxW * 3 * sizeof(double) + 256 + sizeof(spd_aux);
static __device__ __align__(256) unsigned char scratch[kSpdScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kScratchDevices] = {};   // device address of `scratch`, per device

// The current device's address of `scratch`, looked up once per device.
hipError_t scratch_address(void** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

size_t spd_static_scratch_bytes() { return kSpdScratchBytes; }

hipError_t launch_spd_first_pass_static(const float* plane, size_t pitch, int ndms, int nsamp, int W, float* msd,
                                        float* snr, size_t snr_pitch, uint16_t* width, size_t width_pitch,
                                        hipStream_t s)
{
    if (spd_workspace_bytes() > kSpdScratchBytes) return hipErrorInvalidValue;
    void* work = nullptr;
    const hipError_t e = scratch_address(&work);
    if (e != hipSuccess) return e;
    return launch_spd_first_pass(plane, pitch, ndms, nsamp, W, msd, snr, snr_pitch, width, width_pitch, work, s);
}

}  // namespace aa_hip
