// Dedispersion of samples wider than 8 bits (16-bit filterbanks, and 32-bit
// ones the reader scales to u16): the rows stay u16 and each output is the
// reference's float accumulation in channel order,
//   acc = 0; for c = 0 .. nchans-1: acc += (float)L[c][t + delay(c, d)];
//   out[d][t] = (acc / nchans) / inBin
// (src/aa_device_dedispersion_kernel.cu:422-450).  Once a sum passes 2^24
// the float adds round, so the order is part of the result and no integer
// or pre-summation trick applies: one float add per channel-add.
//
// dedisperse_u16_kernel: a 256-thread work-group owns 16 DM trials x 512
// samples.  Channels are staged 8 at a time in LDS (double buffered): each
// channel's window starts at its delay for the tile's first DM and spans the
// tile plus the delay spread of the tile's 16 DMs (at most kWRowMax samples,
// checked on the host).  Windows are loaded as aligned dwords (the odd start
// is carried as a parity), coalesced across lanes.  Wave w sums DMs
// 4w .. 4w+3, lane L samples L + 64 j (j < 8): each ds_read_u16 is a
// stride-1 read across the wave.  Per channel-add: one LDS read, one
// conversion, one add (VALU-bound; ceiling about a third of the FP32 add
// rate).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "ddtr_engine.hpp"
#include "ddtr_kernels.h"

namespace aa_hip {
namespace {

constexpr int kWT = 512;          // samples per tile
constexpr int kWDms = 16;         // DM trials per tile
constexpr int kWCh = 8;           // channels per LDS chunk
constexpr int kWRowMax = 1024;    // staged samples per channel window (tile + spread up to 510)
constexpr int kWRowStride = kWRowMax + 2;
constexpr int kWLoads = (kWCh * (kWRowMax / 2 + 1) + 255) / 256;   // dwords staged per thread per chunk

__device__ __forceinline__ int wide_delay(float dmshift, float shift_temp)
{
    return (int)__fmul_rn(dmshift, shift_temp);
}

struct wide_args {
    const uint16_t* L;
    long pitch;         // samples between rows
    long row_len;       // valid samples per row
    int nchans;
    const float* dmshifts;
    int ndms, t_processed;
    float mstartdm, mdmstep;
    float fnchans, fbin, fscale;
    float* out;
    long out_pitch;
    int win;            // staged window: kWT + largest delay spread of a tile, even
};

// Aligned: rows start on 4-byte boundaries (L aligned, pitch even), so the
// window dwords are single loads; otherwise each dword is two u16 loads.
template <bool Aligned>
__global__ __launch_bounds__(256) void dedisperse_u16_kernel(wide_args a)
{
    __shared__ __align__(16) uint16_t rows[2][kWCh][kWRowStride];
    __shared__ int offs[2][kWCh][kWDms];   // delay(c, d0 + k) - delay(c, d0) + parity of the window start
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long t0 = (long)blockIdx.x * kWT;
    const int d0 = blockIdx.y * kWDms;
    const int dlast = min(d0 + kWDms - 1, a.ndms - 1);
    const int ndw = a.win / 2 + 1;   // dwords staged per channel window
    const int nchunks = (a.nchans + kWCh - 1) / kWCh;

    float acc[4][8];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[k][j] = 0.0f;

    const float st0 = __fmaf_rn((float)d0, a.mdmstep, a.mstartdm);
    // Loads of chunk `ch0`'s windows into registers (kWCh windows of ndw
    // dwords, at most kWLoads per thread).
    uint32_t v[kWLoads];
    auto load = [&](int ch0) {
#pragma unroll
        for (int i = 0; i < kWLoads; ++i) {
            const int item = tid + 256 * i;
            const int ch = item / ndw, w = item - ch * ndw;
            const int c = ch0 + ch;
            uint32_t x = 0u;
            if (ch < kWCh && c < a.nchans) {
                const long base = t0 + wide_delay(a.dmshifts[c], st0);
                const long pos = (base & ~1L) + 2L * w;
                const uint16_t* row = a.L + (size_t)c * (size_t)a.pitch;
                if (Aligned && pos + 1 < a.row_len) x = *reinterpret_cast<const uint32_t*>(row + pos);
                else if (pos + 1 < a.row_len) x = (uint32_t)row[pos] | (uint32_t)row[pos + 1] << 16;
                else if (pos < a.row_len) x = row[pos];
            }
            v[i] = x;
        }
    };
    auto store = [&](int buf, int ch0) {
#pragma unroll
        for (int i = 0; i < kWLoads; ++i) {
            const int item = tid + 256 * i;
            const int ch = item / ndw, w = item - ch * ndw;
            if (ch < kWCh) *reinterpret_cast<uint32_t*>(&rows[buf][ch][2 * w]) = v[i];
        }
        if (tid < kWCh * kWDms) {
            const int ch = tid / kWDms, k = tid - ch * kWDms;
            const int c = ch0 + ch;
            int o = 0;
            if (c < a.nchans) {
                const float ds = a.dmshifts[c];
                const int b = wide_delay(ds, st0);
                const int d = min(d0 + k, dlast);
                o = wide_delay(ds, __fmaf_rn((float)d, a.mdmstep, a.mstartdm)) - b + (int)((t0 + b) & 1L);
            }
            offs[buf][ch][k] = o;
        }
    };

    load(0);
    store(0, 0);
    __syncthreads();
    for (int q = 0; q < nchunks; ++q) {
        const int buf = q & 1;
        if (q + 1 < nchunks) load((q + 1) * kWCh);
        const int nch = min(kWCh, a.nchans - q * kWCh);
        for (int ch = 0; ch < nch; ++ch) {   // channel order: the float sum rounds
            const uint16_t* row = rows[buf][ch] + lane;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint16_t* p = row + offs[buf][ch][4 * wave + k];
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[k][j] = __fadd_rn(acc[k][j], (float)p[64 * j]);
            }
        }
        if (q + 1 < nchunks) store(buf ^ 1, (q + 1) * kWCh);
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int d = d0 + 4 * wave + k;
        if (d >= a.ndms) continue;
        float* orow = a.out + (size_t)d * (size_t)a.out_pitch;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long t = t0 + lane + 64 * j;
            if (t < a.t_processed)
                orow[t] = a.fscale != 0.0f ? acc[k][j] * a.fscale : __fdiv_rn(__fdiv_rn(acc[k][j], a.fnchans), a.fbin);
        }
    }
}

// [t][c] u16 -> [c][t] u16 with a row pitch (samples), 64 x 64 tiles.
__global__ __launch_bounds__(256) void transpose_u16_pitch_kernel(const uint16_t* __restrict__ in, long nsamp,
                                                                  int nchans, uint16_t* __restrict__ out,
                                                                  long out_pitch)
{
    __shared__ uint16_t tile[64][66];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const long t0 = (long)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int r = ty + 4 * i;
        const long t = t0 + r;
        const int c = c0 + tx;
        tile[r][tx] = (t < nsamp && c < nchans) ? in[(size_t)t * (size_t)nchans + (size_t)c] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int cc = ty + 4 * i;
        const int c = c0 + cc;
        const long t = t0 + tx;
        if (c < nchans && t < nsamp) out[(size_t)c * (size_t)out_pitch + (size_t)t] = tile[tx][cc];
    }
}

// One x2 level of u16 rows: the reference's float mean (exact for u16
// operands) truncated back to an integer sample, i.e. floor((a + b) / 2)
// (src/aa_device_binning_kernel.cu:33-45, src/aa_device_corner_turn_kernel.cu:91-98).
__global__ __launch_bounds__(256) void bin2_u16_kernel(const uint16_t* __restrict__ in, long in_pitch, long n_out,
                                                       uint16_t* __restrict__ out, long out_pitch)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (t >= n_out) return;
    const uint16_t* row = in + (size_t)c * (size_t)in_pitch;
    out[(size_t)c * (size_t)out_pitch + (size_t)t] = (uint16_t)(((uint32_t)row[2 * t] + (uint32_t)row[2 * t + 1]) >> 1);
}

inline unsigned blocks(long n, long b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int wide_window(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep)
{
    if (ndms <= 0) return kWT + 2;
    int spread = 0;
    for (int d0 = 0; d0 < ndms; d0 += kWDms) {
        const int d1 = std::min(d0 + kWDms - 1, ndms - 1);
        for (int c = 0; c < nchans; ++c)
            spread = std::max(spread, host_delay(h_dmshifts[c], d1, mstartdm, mdmstep) -
                                          host_delay(h_dmshifts[c], d0, mstartdm, mdmstep));
    }
    int win = kWT + spread + 2;   // + 1 for the odd start, rounded up to even
    win += win & 1;
    return win <= kWRowMax ? win : -1;
}

hipError_t launch_dedisperse_wide_u16(const uint16_t* L, size_t pitch, long row_len, int nchans,
                                      const float* d_dmshifts, int ndms, int t_processed, float mstartdm,
                                      float mdmstep, int inBin, float* out, size_t out_pitch, int win, hipStream_t s)
{
    if (!L || !d_dmshifts || !out || nchans <= 0 || ndms < 0 || t_processed < 0 || inBin <= 0 ||
        (size_t)row_len > pitch || out_pitch < (size_t)t_processed || win < 2 || win > kWRowMax || (win & 1) ||
        ((uintptr_t)L & 1))
        return hipErrorInvalidValue;
    if (ndms == 0 || t_processed == 0) return hipSuccess;
    wide_args a{};
    a.L = L;
    a.pitch = (long)pitch;
    a.row_len = row_len;
    a.nchans = nchans;
    a.dmshifts = d_dmshifts;
    a.ndms = ndms;
    a.t_processed = t_processed;
    a.mstartdm = mstartdm;
    a.mdmstep = mdmstep;
    a.fnchans = (float)nchans;
    a.fbin = (float)inBin;
    const auto pow2 = [](long v) { return v > 0 && (v & (v - 1)) == 0; };
    a.fscale = pow2(nchans) && pow2(inBin) ? (float)(1.0 / ((double)nchans * (double)inBin)) : 0.0f;
    a.out = out;
    a.out_pitch = (long)out_pitch;
    a.win = win;
    const dim3 grid(blocks(t_processed, kWT), blocks(ndms, kWDms));
    if (((uintptr_t)L & 3) == 0 && (pitch & 1) == 0)
        hipLaunchKernelGGL(dedisperse_u16_kernel<true>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(dedisperse_u16_kernel<false>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_transpose_u16_pitch(const uint16_t* in_tc, long nsamp, int nchans, uint16_t* out_ct,
                                      size_t out_pitch, hipStream_t s)
{
    if (!in_tc || !out_ct || nsamp < 0 || nchans <= 0 || (size_t)nsamp > out_pitch) return hipErrorInvalidValue;
    if (nsamp == 0) return hipSuccess;
    hipLaunchKernelGGL(transpose_u16_pitch_kernel, dim3(blocks(nsamp, 64), blocks(nchans, 64)), dim3(256), 0, s,
                       in_tc, nsamp, nchans, out_ct, (long)out_pitch);
    return hipGetLastError();
}

hipError_t launch_bin2_u16(const uint16_t* in, size_t in_pitch, int nchans, long n_out, uint16_t* out,
                           size_t out_pitch, hipStream_t s)
{
    if (!in || !out || nchans <= 0 || n_out < 0 || in_pitch < 2 * (size_t)n_out || out_pitch < (size_t)n_out)
        return hipErrorInvalidValue;
    if (n_out == 0) return hipSuccess;
    hipLaunchKernelGGL(bin2_u16_kernel, dim3(blocks(n_out, 256), nchans), dim3(256), 0, s, in, (long)in_pitch, n_out,
                       out, (long)out_pitch);
    return hipGetLastError();
}

}  // namespace aa_hip
