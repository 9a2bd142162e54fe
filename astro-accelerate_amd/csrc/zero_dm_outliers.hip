// Zero-DM with outlier rejection on [t][c] samples, in place: the
// reference's zero_dm_outliers_kernel_channels
// (src/aa_device_zero_dm_outliers_kernel.cu:190-233, launched by
// src/aa_zero_dm_outliers.cpp:47-67 with 128 threads per time sample and
// outlier_sigma 3).  Per time sample:
//   1. mean and standard deviation over all channels;
//   2. up to 5 times: the same over the channels strictly inside
//      mean +- 3 sigma of the previous estimate, stopping once
//      old_mean - mean < 0.001 (signed, as written);
//   3. x' = (x - mean) + bandpass[c] in float, clamped to [0, 2^nbits - 1],
//      truncated; a NaN (no channel inside the limits) becomes 0.
// The statistics are the reference's streaming ones, and their rounding is
// kept: 128 accumulators (accumulator v takes channels v, v+128, ... in
// order; Initiate/Add_one at :15-29), merged pairwise (Merge at :31-40) as
// its block reduction does (v += v+64, v += v+32 for v < 64 and v < 32, then
// v += v+q for q = 16, 8, 4, 2, 1), then mean = M/j and stdev = sqrt(S/j)
// in double.  The operations mix float and double exactly as declared
// there; products feeding an add are fused (one rounding), which is how
// CUDA's default contraction compiles those lines.
//
// MI355X mapping: one wave per time sample, lane l holding accumulators l and
// l + 64 (so the first merge level is inside the lane and the rest are
// shuffles down to lane 0).  The 1/(j(j-1)) factors of Add_one are a table in
// LDS (correctly rounded doubles, the same values the division gives).
// Compute-bound on the float64 VALU (about 15 instructions per channel and
// pass, up to 6 passes); the row is re-read from L1/L2 for each pass, eight
// channels per lane at a time.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "ddtr_kernels.h"

namespace aa_hip {
namespace {

constexpr int kCoefTab = 256;   // Add_one factors for counts up to 255

struct welford {
    float M;
    float S;
    int j;
};

// SMALL: j * x and M stay below 2^24 (8-bit samples), so j * x - M, an
// integer, is exact in float and needs no float64 work.
template <bool SMALL>
__device__ __forceinline__ void wf_add(welford& a, float x, const double* coef)
{
    a.j += 1;
    const double rj = (double)a.j;
    a.M = a.M + x;
    const float ft = SMALL ? (float)a.j * x - a.M : (float)fma(rj, (double)x, -(double)a.M);
    const double c = a.j < kCoefTab ? coef[a.j] : 1.0 / (rj * (rj - 1.0));   // coef[1] = 0
    a.S = (float)fma(c * (double)ft, (double)ft, (double)a.S);
}

// Add_one on an empty accumulator with factor 0 (coef[1]) gives M = x,
// S = 0, j = 1 exactly, i.e. Initiate: one code path for every channel.
template <bool SMALL>
__device__ __forceinline__ void wf_push(welford& a, float x, const double* coef)
{
    wf_add<SMALL>(a, x, coef);
}

// a <- Merge(a, b), with the reference's handling of empty accumulators.
__device__ __forceinline__ void wf_merge(welford& a, float bM, float bS, int bj)
{
    if (bj <= 0) return;
    if (a.j == 0) {
        a.M = bM;
        a.S = bS;
        a.j = bj;
        return;
    }
    const double rb = (double)bj, ra = (double)a.j;
    const float ft = (float)fma(rb / ra, (double)a.M, -(double)bM);
    const double coef = ra / (rb * (ra + rb));
    a.S = (float)fma(coef * (double)ft, (double)ft, (double)(a.S + bS));
    a.M = a.M + bM;
    a.j += bj;
}

// Lane 0's reduction of the wave's 128 accumulators, broadcast: mean, stdev.
__device__ __forceinline__ void wf_reduce(welford v0, const welford& v1, int lane, float& mean, float& stdev)
{
    wf_merge(v0, v1.M, v1.S, v1.j);   // accumulator l += l + 64
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
        const float bM = __shfl_down(v0.M, q);
        const float bS = __shfl_down(v0.S, q);
        const int bj = __shfl_down(v0.j, q);
        if (lane < q) wf_merge(v0, bM, bS, bj);
    }
    float m = 0.0f, s = 0.0f;
    if (lane == 0) {
        m = (float)((double)v0.M / (double)v0.j);
        s = (float)sqrt((double)v0.S / (double)v0.j);
    }
    mean = __shfl(m, 0);
    stdev = __shfl(s, 0);
}

template <class T, bool SMALL>
__global__ __launch_bounds__(256) void zero_dm_outliers_kernel(T* data, long nsamp, int nchans,
                                                               const float* __restrict__ bp, float maxval)
{
    __shared__ double coef[kCoefTab];
    for (int i = threadIdx.x; i < kCoefTab; i += 256)
        coef[i] = i >= 2 ? 1.0 / ((double)i * ((double)i - 1.0)) : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long waves = (long)gridDim.x * 4;
    for (long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6); t < nsamp; t += waves) {
        T* row = data + t * (long)nchans;
        welford v0{0.0f, 0.0f, 0}, v1{0.0f, 0.0f, 0};
        float mean = 0.0f, stdev = 0.0f;
        // Pass 0 takes every channel (limits -inf, +inf); then up to 5
        // rejection passes.  Channels come in batches of 8 per lane
        // (lane + 64 k), all loads issued before the first add.
        float lo = -INFINITY, hi = INFINITY;
        for (int f = 0; f < 6; ++f) {
            const float old_mean = mean;
            v0 = welford{0.0f, 0.0f, 0};
            v1 = welford{0.0f, 0.0f, 0};
            for (int c = lane; c < nchans; c += 512) {
                float x[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) x[k] = c + 64 * k < nchans ? (float)row[c + 64 * k] : NAN;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (x[k] > lo && x[k] < hi) wf_push<SMALL>((k & 1) ? v1 : v0, x[k], coef);
            }
            wf_reduce(v0, v1, lane, mean, stdev);
            if (f > 0 && old_mean - mean < 0.001f) break;   // wave-uniform
            lo = fmaf(-3.0f, stdev, mean);
            hi = fmaf(3.0f, stdev, mean);
        }
        for (int c = lane; c < nchans; c += 64) {
            const float r = ((float)row[c] - mean) + bp[c];
            const float v = r < 0.0f ? 0.0f : (r > maxval ? maxval : (r == r ? r : 0.0f));
            row[c] = (T)(uint32_t)v;
        }
    }
}

template <class T>
hipError_t launch_outliers(T* data, long nsamp, int nchans, const float* bp, int nbits, hipStream_t s)
{
    if (!data || !bp || nsamp < 0 || nchans <= 0 || nbits < 1 || nbits > 16) return hipErrorInvalidValue;
    if (nsamp == 0) return hipSuccess;
    const long blocks = (nsamp + 3) / 4;
    const unsigned grid = (unsigned)(blocks < 16384 ? blocks : 16384);
    const float maxval = (float)((1u << nbits) - 1u);
    // Counts per accumulator reach ceil(nchans / 128); with 8-bit samples
    // j * x and M stay below 2^24 while that count stays below 65793.
    const bool small = sizeof(T) == 1 && (nchans + 127) / 128 < 65793;
    if (small)
        hipLaunchKernelGGL((zero_dm_outliers_kernel<T, true>), dim3(grid), dim3(256), 0, s, data, nsamp, nchans, bp,
                           maxval);
    else
        hipLaunchKernelGGL((zero_dm_outliers_kernel<T, false>), dim3(grid), dim3(256), 0, s, data, nsamp, nchans,
                           bp, maxval);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_zero_dm_outliers_u8(uint8_t* data_tc, long nsamp, int nchans, const float* bp, int nbits,
                                      hipStream_t s)
{
    if (nbits > 8) return hipErrorInvalidValue;
    return launch_outliers<uint8_t>(data_tc, nsamp, nchans, bp, nbits, s);
}

hipError_t launch_zero_dm_outliers_u16(uint16_t* data_tc, long nsamp, int nchans, const float* bp, int nbits,
                                       hipStream_t s)
{
    return launch_outliers<uint16_t>(data_tc, nsamp, nchans, bp, nbits, s);
}

}  // namespace aa_hip
