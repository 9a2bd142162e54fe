// DDTR kernels for gfx950 (MI355X): corner turn, x2 binning, direct
// dedispersion and synthetic-data generation.  The LDS-tiled dedisperser
// lives in ddtr_lds.hip.
//
// Delay semantics (shared with every dedispersion kernel in this library):
//   shift_temp = fma((float)d, mdmstep, mstartdm)            one rounding
//   delay      = trunc(dmshift[c] * shift_temp)              one rounding
// which is the reference's failsafe kernel as its compiler emits it
// (src/aa_device_dedispersion_kernel.cu:430,440).  The output scaling is
// ((float)sum / nchans) / inBin, left to right (:448).  Sums of 8-bit
// samples are exact integers, so the result is independent of summation
// order.
#include "ddtr_kernels.h"

#include <algorithm>

namespace aa_hip {

namespace {

constexpr int kTile = 64;

__device__ __forceinline__ int ddtr_delay(float dmshift, float shift_temp)
{
    return (int)__fmul_rn(dmshift, shift_temp);  // v_cvt_i32_f32 truncates
}

// ---------------------------------------------------------------- corner turn
// 64x64 tile through LDS.  Loads are coalesced along channels (the fast axis
// of [t][c]); stores are coalesced along time (the fast axis of [c][t]).
template <typename Tin>
__global__ __launch_bounds__(256) void corner_turn_kernel(const Tin* __restrict__ in, int nsamp,
                                                          int nchans, uint8_t* __restrict__ out,
                                                          size_t out_pitch, int* __restrict__ err)
{
    __shared__ uint8_t tile[kTile][kTile + 4];  // [t][c]; +4 keeps column reads conflict-free
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;  // 0..3
    const long t0 = (long)blockIdx.x * kTile;
    const int c0 = blockIdx.y * kTile;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < kTile / 4; ++i) {
        const int r = ty + 4 * i;
        const long t = t0 + r;
        const int c = c0 + tx;
        uint8_t v = 0;
        if (t < nsamp && c < nchans) {
            const Tin x = in[(size_t)t * (size_t)nchans + (size_t)c];
            if (sizeof(Tin) > 1 && x > 255) bad = true;
            v = (uint8_t)x;
        }
        tile[r][tx] = v;
    }
    if (sizeof(Tin) > 1 && bad) atomicOr(err, 1);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kTile / 4; ++i) {
        const int r = ty + 4 * i;
        const int c = c0 + r;
        const long t = t0 + tx;
        if (c < nchans && t < nsamp) out[(size_t)c * out_pitch + (size_t)t] = tile[tx][r];
    }
}

__global__ __launch_bounds__(256) void transpose_u16_kernel(const uint16_t* __restrict__ in, int rows,
                                                            int cols, uint16_t* __restrict__ out)
{
    __shared__ uint16_t tile[kTile][kTile + 2];
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;
    const long r0 = (long)blockIdx.x * kTile;
    const int c0 = blockIdx.y * kTile;
#pragma unroll
    for (int i = 0; i < kTile / 4; ++i) {
        const int rr = ty + 4 * i;
        const long r = r0 + rr;
        const int c = c0 + tx;
        tile[rr][tx] = (r < rows && c < cols) ? in[(size_t)r * cols + c] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kTile / 4; ++i) {
        const int cc = ty + 4 * i;
        const int c = c0 + cc;
        const long r = r0 + tx;
        if (c < cols && r < rows) out[(size_t)c * rows + (size_t)r] = tile[tx][cc];
    }
}

// -------------------------------------------------------------------- binning
// floor((a+b)/2) equals the reference's (a+b)/2.0f followed by the
// truncating conversion back to an integer sample.
__global__ __launch_bounds__(256) void bin2_u8_vec_kernel(const uint8_t* __restrict__ in, size_t in_pitch,
                                                          int n_out, uint8_t* __restrict__ out,
                                                          size_t out_pitch)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;  // quad of outputs
    const int c = blockIdx.y;
    const long t = 4 * q;
    if (t >= n_out) return;
    const uint8_t* row = in + (size_t)c * in_pitch;
    uint8_t* orow = out + (size_t)c * out_pitch;
    if (t + 3 < n_out) {
        const uint2 w = *reinterpret_cast<const uint2*>(row + 2 * t);
        // even bytes + odd bytes of each 16-bit pair, as 16-bit fields
        const uint32_t lo_e = w.x & 0x00FF00FFu, lo_o = (w.x >> 8) & 0x00FF00FFu;
        const uint32_t hi_e = w.y & 0x00FF00FFu, hi_o = (w.y >> 8) & 0x00FF00FFu;
        const uint32_t lo = ((lo_e + lo_o) >> 1) & 0x00FF00FFu;  // outputs t, t+1 in bytes 0 and 2
        const uint32_t hi = ((hi_e + hi_o) >> 1) & 0x00FF00FFu;  // outputs t+2, t+3
        const uint32_t packed = (lo & 0xFFu) | ((lo >> 8) & 0xFF00u) | ((hi & 0xFFu) << 16) |
                                ((hi << 8) & 0xFF000000u);
        *reinterpret_cast<uint32_t*>(orow + t) = packed;
    } else {
        for (long k = t; k < n_out; ++k)
            orow[k] = (uint8_t)(((unsigned)row[2 * k] + (unsigned)row[2 * k + 1]) >> 1);
    }
}

__global__ __launch_bounds__(256) void bin2_u8_scalar_kernel(const uint8_t* __restrict__ in,
                                                             size_t in_pitch, int n_out,
                                                             uint8_t* __restrict__ out,
                                                             size_t out_pitch)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (t >= n_out) return;
    const uint8_t* row = in + (size_t)c * in_pitch;
    out[(size_t)c * out_pitch + t] = (uint8_t)(((unsigned)row[2 * t] + (unsigned)row[2 * t + 1]) >> 1);
}

// Reference bin_gpu on u16 rows (src/aa_bin_gpu.cpp:8-44): a float temp is
// zeroed, bin writes temp[c*(nsamp/2)+t] for c < 32*(nchans/32) and
// t < 8*((nsamp+1)/16) (its launch grid), then swap truncates temp[j] back
// into data[j] for j = c*nsamp + t, c < 8*(nchans/8), t < 32*(nsamp/32).
__global__ void bin_ref_u16_kernel(const uint16_t* __restrict__ data, float* __restrict__ temp,
                                   int nchans, int nsamp)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int out_nsamp = nsamp / 2;
    const long t_cov = 8L * ((nsamp + 1) / 16);
    const int c_cov = 32 * (nchans / 32);
    if (c >= c_cov || t >= t_cov) return;
    const uint16_t* row = data + (size_t)c * nsamp;
    temp[(size_t)c * out_nsamp + t] = ((float)row[2 * t] + (float)row[2 * t + 1]) / 2.0f;
}

__global__ void swap_ref_u16_kernel(uint16_t* __restrict__ data, const float* __restrict__ temp,
                                    int nchans, int nsamp)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (c >= 8 * (nchans / 8) || t >= 32L * (nsamp / 32)) return;
    const size_t j = (size_t)c * nsamp + t;
    data[j] = (uint16_t)temp[j];
}

// ---------------------------------------------------------- direct dedisperse
// One thread per (DM, output sample).  Used as the universally valid
// variant and as the shim behind the u16 drop-in dedisperse().
template <typename T>
__global__ __launch_bounds__(256) void dedisperse_direct_kernel(
    const T* __restrict__ L, size_t pitch, int nchans, const float* __restrict__ dmshifts,
    int ndms, int t_processed, float mstartdm, float mdmstep, float fnchans, float fbin,
    float* __restrict__ out, size_t out_pitch)
{
    const long t = (long)blockIdx.x * 64 + (threadIdx.x & 63);
    const int d = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (t >= t_processed || d >= ndms) return;
    const float shift_temp = __fmaf_rn((float)d, mdmstep, mstartdm);
    const T* p = L + t;
    uint32_t sum = 0;
    for (int c = 0; c < nchans; ++c) {
        const int delay = ddtr_delay(dmshifts[c], shift_temp);
        sum += (uint32_t)p[(size_t)c * pitch + (size_t)delay];
    }
    out[(size_t)d * out_pitch + (size_t)t] = __fdiv_rn(__fdiv_rn((float)sum, fnchans), fbin);
}

// ----------------------------------------------------------- synthetic input
__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void synth_noise_u8_kernel(uint64_t key, int64_t n, double mean,
                                                             double sigma, uint8_t* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double u1 = 1.0 - (double)(splitmix64(key + 2 * (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(splitmix64(key + 2 * (uint64_t)i + 1) >> 11) * (1.0 / 9007199254740992.0);
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        double v = rint(mean + sigma * z);
        v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
        out[i] = (uint8_t)v;
    }
}

__global__ __launch_bounds__(256) void inject_pulse_u8_kernel(uint8_t* __restrict__ data, int nsamp,
                                                              int nchans, const float* __restrict__ dmshifts,
                                                              float dm_over_tsamp, int t0, int width,
                                                              int amplitude)
{
    const int c = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (c >= nchans || i >= width) return;
    const int shift = (int)floorf(__fmul_rn(dmshifts[c], dm_over_tsamp));
    const long t = (long)t0 + shift + i;
    if (t < 0 || t >= nsamp) return;
    uint8_t* p = data + (size_t)t * nchans + c;
    const int v = (int)*p + amplitude;
    *p = (uint8_t)(v > 255 ? 255 : v);
}

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Zero-DM filter on [t][c] samples, in place (src/aa_device_zero_dm_kernel.cu:
// 37-88): per time sample, x' = (x - mean over channels) + bandpass[c], in
// float as written there, clamped to [0, maxval] and truncated.  One wave per
// time sample; the channel sum is an exact integer, which is what the
// reference's float accumulator holds while sums stay below 2^24.  HBM-bound:
// each sample is read twice (sum, then filter; the second read mostly hits
// L2) and written once.
// ---------------------------------------------------------------------------
// u16 rows holding 8-bit values -> u8 rows (the reference's dedisperse()
// input layout to the LDS kernels'); a value above 255 sets *err.
__global__ __launch_bounds__(256) void narrow_u16_rows_kernel(const uint16_t* __restrict__ in, long in_pitch,
                                                              long n, uint8_t* __restrict__ out, size_t out_pitch,
                                                              int* __restrict__ err)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const long c = blockIdx.y;
    if (t >= n) return;
    const uint16_t v = in[c * in_pitch + t];
    out[c * (long)out_pitch + t] = (uint8_t)v;
    if (v > 255) atomicOr(err, 1);
}

__device__ __forceinline__ uint32_t zero_dm_wave_sum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
    return v;
}

__device__ __forceinline__ uint32_t zero_dm_one(uint32_t x, float mean, float bp, float maxval)
{
    float r = __fadd_rn(__fsub_rn((float)x, mean), bp);
    r = r < 0.0f ? 0.0f : (r > maxval ? maxval : r);
    return (uint32_t)r;
}

// u8 rows of a multiple of 16 channels, at most 1024 NQ: the row is read
// once into registers (NQ 16-byte quads per lane), summed, filtered and
// written back.
template <int NQ>
__global__ __launch_bounds__(256) void zero_dm_u8_reg_kernel(uint8_t* data, int nsamp, int nchans,
                                                             const float* __restrict__ bp, float maxval)
{
    const int lane = threadIdx.x & 63;
    const long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= nsamp) return;   // the whole wave leaves; no barrier here
    uint4* row = reinterpret_cast<uint4*>(data + t * (long)nchans);
    const int nv = nchans >> 4;
    uint4 v[NQ];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int q = lane + 64 * i;
        v[i] = q < nv ? row[q] : make_uint4(0, 0, 0, 0);
        sum = __builtin_amdgcn_sad_u8(v[i].x, 0u, sum);
        sum = __builtin_amdgcn_sad_u8(v[i].y, 0u, sum);
        sum = __builtin_amdgcn_sad_u8(v[i].z, 0u, sum);
        sum = __builtin_amdgcn_sad_u8(v[i].w, 0u, sum);
    }
    sum = zero_dm_wave_sum(sum);
    const float mean = __fdiv_rn((float)sum, (float)nchans);
    const float4* b4 = reinterpret_cast<const float4*>(bp);
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int q = lane + 64 * i;
        if (q >= nv) continue;
        uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 b = b4[4 * q + k];
            const uint32_t x = w[k];
            w[k] = zero_dm_one(x & 0xffu, mean, b.x, maxval) | zero_dm_one((x >> 8) & 0xffu, mean, b.y, maxval) << 8 |
                   zero_dm_one((x >> 16) & 0xffu, mean, b.z, maxval) << 16 |
                   zero_dm_one(x >> 24, mean, b.w, maxval) << 24;
        }
        row[q] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// u8 rows of a multiple of 4 channels: one dword (4 channels) per lane step.
__global__ __launch_bounds__(256) void zero_dm_u8x4_kernel(uint8_t* data, int nsamp, int nchans,
                                                           const float* __restrict__ bp, float maxval)
{
    const int lane = threadIdx.x & 63;
    const long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= nsamp) return;   // the whole wave leaves; no barrier here
    uint32_t* row = reinterpret_cast<uint32_t*>(data + t * (long)nchans);
    const int nq = nchans >> 2;
    uint32_t sum = 0;
    for (int q = lane; q < nq; q += 64) sum = __builtin_amdgcn_sad_u8(row[q], 0u, sum);
    sum = zero_dm_wave_sum(sum);
    const float mean = __fdiv_rn((float)sum, (float)nchans);
    for (int q = lane; q < nq; q += 64) {
        const uint32_t x = row[q];
        const float4 b = reinterpret_cast<const float4*>(bp)[q];
        row[q] = zero_dm_one(x & 0xffu, mean, b.x, maxval) | zero_dm_one((x >> 8) & 0xffu, mean, b.y, maxval) << 8 |
                 zero_dm_one((x >> 16) & 0xffu, mean, b.z, maxval) << 16 |
                 zero_dm_one(x >> 24, mean, b.w, maxval) << 24;
    }
}

template <class T>
__global__ __launch_bounds__(256) void zero_dm_kernel(T* data, int nsamp, int nchans, const float* __restrict__ bp,
                                                      float maxval)
{
    const int lane = threadIdx.x & 63;
    const long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= nsamp) return;
    T* row = data + t * (long)nchans;
    uint32_t sum = 0;
    for (int c = lane; c < nchans; c += 64) sum += row[c];
    sum = zero_dm_wave_sum(sum);
    const float mean = __fdiv_rn((float)sum, (float)nchans);
    for (int c = lane; c < nchans; c += 64) row[c] = (T)zero_dm_one(row[c], mean, bp[c], maxval);
}

}  // namespace

// ------------------------------------------------------------------ launchers
hipError_t launch_corner_turn_u8(const uint8_t* in_tc, int nsamp, int nchans, uint8_t* out_ct,
                                 size_t out_pitch, hipStream_t s)
{
    if (nsamp < 0 || nchans <= 0 || (size_t)nsamp > out_pitch) return hipErrorInvalidValue;
    if (nsamp == 0) return hipSuccess;
    dim3 grid(ceil_div(nsamp, kTile), ceil_div(nchans, kTile));
    hipLaunchKernelGGL(corner_turn_kernel<uint8_t>, grid, dim3(256), 0, s, in_tc, nsamp, nchans,
                       out_ct, out_pitch, (int*)nullptr);
    return hipGetLastError();
}

hipError_t launch_corner_turn_u16(const uint16_t* in_tc, int nsamp, int nchans, uint8_t* out_ct,
                                  size_t out_pitch, int* d_err, hipStream_t s)
{
    if (nsamp < 0 || nchans <= 0 || (size_t)nsamp > out_pitch || !d_err) return hipErrorInvalidValue;
    if (nsamp == 0) return hipSuccess;
    dim3 grid(ceil_div(nsamp, kTile), ceil_div(nchans, kTile));
    hipLaunchKernelGGL(corner_turn_kernel<uint16_t>, grid, dim3(256), 0, s, in_tc, nsamp, nchans,
                       out_ct, out_pitch, d_err);
    return hipGetLastError();
}

hipError_t launch_transpose_u16(const uint16_t* in, int rows, int cols, uint16_t* out, hipStream_t s)
{
    if (rows < 0 || cols <= 0) return hipErrorInvalidValue;
    if (rows == 0) return hipSuccess;
    dim3 grid(ceil_div(rows, kTile), ceil_div(cols, kTile));
    hipLaunchKernelGGL(transpose_u16_kernel, grid, dim3(256), 0, s, in, rows, cols, out);
    return hipGetLastError();
}

hipError_t launch_narrow_u16(const uint16_t* in, long in_pitch, int rows, long n, uint8_t* out, size_t out_pitch,
                             int* d_err, hipStream_t s)
{
    if (rows <= 0 || n < 0 || in_pitch < n || (long)out_pitch < n || !in || !out || !d_err)
        return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(narrow_u16_rows_kernel, dim3(ceil_div(n, 256), rows), dim3(256), 0, s, in, in_pitch, n, out,
                       out_pitch, d_err);
    return hipGetLastError();
}

static float zero_dm_max(int nbits) { return (float)((1u << (nbits > 16 ? 16 : nbits)) - 1u); }

hipError_t launch_zero_dm_u8(uint8_t* data_tc, int nsamp, int nchans, const float* bp, int nbits, hipStream_t s)
{
    if (nsamp < 0 || nchans <= 0 || !data_tc || !bp || nbits < 1 || nbits > 16) return hipErrorInvalidValue;
    if (nsamp == 0) return hipSuccess;
    const dim3 grid((unsigned)ceil_div((long)nsamp, 4L));
    const float mx = zero_dm_max(nbits);
    if (nchans % 16 == 0 && nchans <= 8192 && (uintptr_t)data_tc % 16 == 0 && (uintptr_t)bp % 16 == 0) {
        if (nchans <= 1024) hipLaunchKernelGGL(zero_dm_u8_reg_kernel<1>, grid, dim3(256), 0, s, data_tc, nsamp, nchans, bp, mx);
        else if (nchans <= 2048) hipLaunchKernelGGL(zero_dm_u8_reg_kernel<2>, grid, dim3(256), 0, s, data_tc, nsamp, nchans, bp, mx);
        else if (nchans <= 4096) hipLaunchKernelGGL(zero_dm_u8_reg_kernel<4>, grid, dim3(256), 0, s, data_tc, nsamp, nchans, bp, mx);
        else hipLaunchKernelGGL(zero_dm_u8_reg_kernel<8>, grid, dim3(256), 0, s, data_tc, nsamp, nchans, bp, mx);
    } else if (nchans % 4 == 0 && (uintptr_t)data_tc % 4 == 0 && (uintptr_t)bp % 16 == 0)
        hipLaunchKernelGGL(zero_dm_u8x4_kernel, grid, dim3(256), 0, s, data_tc, nsamp, nchans, bp, mx);
    else
        hipLaunchKernelGGL(zero_dm_kernel<uint8_t>, grid, dim3(256), 0, s, data_tc, nsamp, nchans, bp,
                           zero_dm_max(nbits));
    return hipGetLastError();
}

hipError_t launch_zero_dm_u16(uint16_t* data_tc, int nsamp, int nchans, const float* bp, int nbits, hipStream_t s)
{
    if (nsamp < 0 || nchans <= 0 || !data_tc || !bp || nbits < 1 || nbits > 16) return hipErrorInvalidValue;
    if (nsamp == 0) return hipSuccess;
    const dim3 grid((unsigned)ceil_div((long)nsamp, 4L));
    hipLaunchKernelGGL(zero_dm_kernel<uint16_t>, grid, dim3(256), 0, s, data_tc, nsamp, nchans, bp,
                       zero_dm_max(nbits));
    return hipGetLastError();
}

hipError_t launch_bin2_u8(const uint8_t* in, size_t in_pitch, int nchans, int n_out, uint8_t* out,
                          size_t out_pitch, hipStream_t s)
{
    if (nchans <= 0 || n_out < 0 || in_pitch < 2 * (size_t)n_out || out_pitch < (size_t)n_out)
        return hipErrorInvalidValue;
    if (n_out == 0) return hipSuccess;
    const bool vec = ((uintptr_t)in % 8 == 0) && (in_pitch % 8 == 0) && ((uintptr_t)out % 4 == 0) &&
                     (out_pitch % 4 == 0);
    if (vec) {
        dim3 grid(ceil_div(ceil_div(n_out, 4), 256), nchans);
        hipLaunchKernelGGL(bin2_u8_vec_kernel, grid, dim3(256), 0, s, in, in_pitch, n_out, out, out_pitch);
    } else {
        dim3 grid(ceil_div(n_out, 256), nchans);
        hipLaunchKernelGGL(bin2_u8_scalar_kernel, grid, dim3(256), 0, s, in, in_pitch, n_out, out,
                           out_pitch);
    }
    return hipGetLastError();
}

hipError_t launch_bin2_u16_packed(uint16_t* data, int nchans, int nsamp, float* temp, hipStream_t s)
{
    // temp must hold nchans*nsamp floats (the reference passes its float
    // output buffer here).
    if (nchans <= 0 || nsamp < 2) return hipErrorInvalidValue;
    const size_t n = (size_t)nchans * (size_t)nsamp;
    hipError_t e = hipMemsetAsync(temp, 0, n * sizeof(float), s);
    if (e != hipSuccess) return e;
    dim3 g1(ceil_div(std::max(1L, 8L * ((nsamp + 1) / 16)), 256), nchans);
    hipLaunchKernelGGL(bin_ref_u16_kernel, g1, dim3(256), 0, s, data, temp, nchans, nsamp);
    dim3 g2(ceil_div(nsamp, 256), nchans);
    hipLaunchKernelGGL(swap_ref_u16_kernel, g2, dim3(256), 0, s, data, temp, nchans, nsamp);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipMemsetAsync(temp, 0, n * sizeof(float), s);
}

namespace {
template <typename T>
hipError_t launch_direct(const T* L, size_t pitch, int nchans, const float* d_dmshifts, int ndms,
                         int t_processed, float mstartdm, float mdmstep, int inBin, float* out,
                         size_t out_pitch, hipStream_t s)
{
    if (!L || !d_dmshifts || !out || nchans <= 0 || ndms < 0 || t_processed < 0 || inBin <= 0 ||
        out_pitch < (size_t)t_processed)
        return hipErrorInvalidValue;
    if (ndms == 0 || t_processed == 0) return hipSuccess;
    dim3 grid(ceil_div(t_processed, 64), ceil_div(ndms, 4));
    hipLaunchKernelGGL(dedisperse_direct_kernel<T>, grid, dim3(256), 0, s, L, pitch, nchans,
                       d_dmshifts, ndms, t_processed, mstartdm, mdmstep, (float)nchans,
                       (float)inBin, out, out_pitch);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_dedisperse_direct_u8(const uint8_t* L, size_t pitch, int nchans,
                                       const float* d_dmshifts, int ndms, int t_processed,
                                       float mstartdm, float mdmstep, int inBin, float* out,
                                       size_t out_pitch, hipStream_t s)
{
    return launch_direct(L, pitch, nchans, d_dmshifts, ndms, t_processed, mstartdm, mdmstep, inBin,
                         out, out_pitch, s);
}

hipError_t launch_dedisperse_direct_u16(const uint16_t* L, size_t pitch, int nchans,
                                        const float* d_dmshifts, int ndms, int t_processed,
                                        float mstartdm, float mdmstep, int inBin, float* out,
                                        size_t out_pitch, hipStream_t s)
{
    return launch_direct(L, pitch, nchans, d_dmshifts, ndms, t_processed, mstartdm, mdmstep, inBin,
                         out, out_pitch, s);
}

hipError_t launch_synth_noise_u8(uint64_t seed, int64_t n, double mean, double sigma, uint8_t* out_tc,
                                 hipStream_t s)
{
    if (n < 0 || !out_tc) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    // SplitMix64 of the seed keys the stream, as in the CPU generator.
    uint64_t key = seed + 0x9E3779B97F4A7C15ull;
    key = (key ^ (key >> 30)) * 0xBF58476D1CE4E5B9ull;
    key = (key ^ (key >> 27)) * 0x94D049BB133111EBull;
    key = key ^ (key >> 31);
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(synth_noise_u8_kernel, dim3(blocks), dim3(256), 0, s, key, n, mean, sigma, out_tc);
    return hipGetLastError();
}

hipError_t launch_inject_pulse_u8(uint8_t* data_tc, int nsamp, int nchans, const float* d_dmshifts,
                                  float dm_over_tsamp, int t0, int width, int amplitude, hipStream_t s)
{
    if (!data_tc || !d_dmshifts || nchans <= 0 || width < 0 || nsamp < 0) return hipErrorInvalidValue;
    if (width == 0 || nsamp == 0) return hipSuccess;
    dim3 grid(ceil_div(width, 256), nchans);
    hipLaunchKernelGGL(inject_pulse_u8_kernel, grid, dim3(256), 0, s, data_tc, nsamp, nchans,
                       d_dmshifts, dm_over_tsamp, t0, width, amplitude);
    return hipGetLastError();
}

}  // namespace aa_hip
