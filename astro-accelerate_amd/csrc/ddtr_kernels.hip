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

// Wide corner turn fused with the x2 binning cascade.  A work-group of
// kCtT threads turns a tile of kCtT samples x 128 channels:
//  * load: lane (cq = lane & 7, tg = lane >> 3) reads samples 8 tg .. 8 tg + 7
//    of channels 16 cq .. 16 cq + 15 with eight 16-byte loads (one 128-byte
//    line of each [t][c] row per 8 lanes);
//  * in registers: each 4 x 4 byte block (4 samples x 4 channels) is
//    transposed with 8 v_perm, so a channel's 8 samples are two dwords;
//  * LDS [c][t] (row stride kCtRow bytes): one 8-byte write per channel;
//  * store: lane (tq = lane % (kCtT / 16), cr = lane / (kCtT / 16)) writes 16
//    samples of channels cr + 16 j (j < 8) to level 0 with 16-byte stores
//    (kCtT contiguous bytes of a [c][t] row per row of lanes), and their
//    floor averages to the next NLEV levels (8, 4, 2 bytes per lane), each
//    level k holding L_k[t] = floor((L_{k-1}[2t] + L_{k-1}[2t+1]) / 2) for
//    t < nsamp >> k.  1024-sample tiles give every level's row segment at
//    least one whole 128-byte line (level 3: 128 bytes).
// Work-group g turns channel group g % nGroups of time tile g / nGroups:
// the work-groups in flight read whole [t][c] rows (every channel group of
// a run of time tiles), not one 128-byte line of each of many rows
// (2.51 ms against 3.47 ms with the time tile varying fastest, 4096
// channels x 1,048,544 samples).
// HBM traffic: every input byte read once, every level byte written once.
constexpr int kCtT = 1024, kCtC = 128, kCtRow = kCtT + 8;

struct ct_levels {
    uint8_t* ptr[4];
    size_t pitch[4];
};

__device__ __forceinline__ uint32_t perm_b32(uint32_t hi, uint32_t lo, uint32_t sel)
{
    return __builtin_amdgcn_perm(hi, lo, sel);
}

// floor averages of adjacent bytes of a and b -> 4 bytes (a's two, then b's two)
__device__ __forceinline__ uint32_t halve_pairs(uint32_t a, uint32_t b)
{
    const uint32_t sa = (((a & 0x00FF00FFu) + ((a >> 8) & 0x00FF00FFu)) >> 1) & 0x00FF00FFu;
    const uint32_t sb = (((b & 0x00FF00FFu) + ((b >> 8) & 0x00FF00FFu)) >> 1) & 0x00FF00FFu;
    return perm_b32(sb, sa, 0x06040200u);
}

// 16 bytes of channel data from a load of Tin samples; u16 values above 255
// set *bad.
template <typename Tin>
__device__ __forceinline__ uint4 ct_load16(const Tin* p, bool& bad);

template <>
__device__ __forceinline__ uint4 ct_load16<uint8_t>(const uint8_t* p, bool&)
{
    return *reinterpret_cast<const uint4*>(p);
}

template <>
__device__ __forceinline__ uint4 ct_load16<uint16_t>(const uint16_t* p, bool& bad)
{
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    bad |= ((a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w) & 0xFF00FF00u) != 0u;
    return make_uint4(perm_b32(a.y, a.x, 0x06040200u), perm_b32(a.w, a.z, 0x06040200u),
                      perm_b32(b.y, b.x, 0x06040200u), perm_b32(b.w, b.z, 0x06040200u));
}

// Tile g of the turn of one chunk (nsamp x nchans): channel group
// g % ngroups of time tile g / ngroups.
template <typename Tin, int NLEV>
__device__ __forceinline__ void ct_wide_tile(const Tin* __restrict__ in, long nsamp, int nchans, const ct_levels& lv,
                                             int* __restrict__ err, long g)
{
    __shared__ __align__(16) uint8_t tile[kCtC * kCtRow];
    const int tid = threadIdx.x;
    const int ngroups = (nchans + kCtC - 1) / kCtC;
    const long t0 = (g / ngroups) * kCtT;
    const int c0 = (int)(g % ngroups) * kCtC;
    {
        const int cq = tid & 7, tg = tid >> 3;
        const int c = c0 + 16 * cq;
        const long tb = t0 + 8 * tg;
        bool bad = false;
        uint4 v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = make_uint4(0u, 0u, 0u, 0u);
            if (c < nchans && tb + k < nsamp) v[k] = ct_load16<Tin>(in + (size_t)(tb + k) * (size_t)nchans + c, bad);
        }
        if (sizeof(Tin) > 1 && bad) atomicOr(err, 1);
        // word q of v[k] holds channels 4q..4q+3 of sample k
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t a[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = q == 0 ? v[k].x : q == 1 ? v[k].y : q == 2 ? v[k].z : v[k].w;
            uint32_t lo[4], hi[4];   // channel 4q + j: samples 0..3 in lo[j], 4..7 in hi[j]
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t* A = a + 4 * h;
                const uint32_t p01l = perm_b32(A[1], A[0], 0x05010400u), p01h = perm_b32(A[1], A[0], 0x07030602u);
                const uint32_t p23l = perm_b32(A[3], A[2], 0x05010400u), p23h = perm_b32(A[3], A[2], 0x07030602u);
                uint32_t* B = h == 0 ? lo : hi;
                B[0] = perm_b32(p23l, p01l, 0x05040100u);
                B[1] = perm_b32(p23l, p01l, 0x07060302u);
                B[2] = perm_b32(p23h, p01h, 0x05040100u);
                B[3] = perm_b32(p23h, p01h, 0x07060302u);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                *reinterpret_cast<uint2*>(tile + (16 * cq + 4 * q + j) * kCtRow + 8 * tg) = make_uint2(lo[j], hi[j]);
        }
    }
    __syncthreads();
    const int tq = tid % (kCtT / 16), cr = tid / (kCtT / 16);
    const long t = t0 + 16 * tq;
#pragma unroll 2
    for (int j = 0; j < 8; ++j) {
        const int cl = cr + 16 * j;
        const int c = c0 + cl;
        if (c >= nchans || t >= nsamp) continue;
        const uint8_t* src = tile + cl * kCtRow + 16 * tq;
        const uint2 w01 = *reinterpret_cast<const uint2*>(src);
        const uint2 w23 = *reinterpret_cast<const uint2*>(src + 8);
        uint32_t w[4] = {w01.x, w01.y, w23.x, w23.y};
        uint8_t* o0 = lv.ptr[0] + (size_t)c * lv.pitch[0] + (size_t)t;
        if (t + 16 <= nsamp) {
            *reinterpret_cast<uint4*>(o0) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
            for (int i = 0; i < (int)(nsamp - t); ++i) o0[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
        }
        // level k: 16 >> k outputs at t >> k, valid below nsamp >> k
#pragma unroll
        for (int k = 1; k <= NLEV; ++k) {
            if (k == 1) {
                w[0] = halve_pairs(w[0], w[1]);
                w[1] = halve_pairs(w[2], w[3]);
            } else {
                w[0] = halve_pairs(w[0], w[1]);   // k = 3: w[1] is 0, the result's low half
                w[1] = 0u;
            }
            const long tk = t >> k;
            const long nk = nsamp >> k;
            uint8_t* ok = lv.ptr[k] + (size_t)c * lv.pitch[k] + (size_t)tk;
            if (tk + (16 >> k) <= nk) {
                if (k == 1) *reinterpret_cast<uint2*>(ok) = make_uint2(w[0], w[1]);
                else if (k == 2) *reinterpret_cast<uint32_t*>(ok) = w[0];
                else *reinterpret_cast<uint16_t*>(ok) = (uint16_t)w[0];
            } else {
                for (int i = 0; i < (int)(nk - tk); ++i) ok[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
            }
        }
    }
}

template <typename Tin, int NLEV>
__global__ __launch_bounds__(kCtT) void corner_turn_wide_kernel(const Tin* __restrict__ in, long nsamp, int nchans,
                                                                ct_levels lv, int* __restrict__ err)
{
    ct_wide_tile<Tin, NLEV>(in, nsamp, nchans, lv, err, blockIdx.x);
}

// The turns of many u8 chunks (one per job, all writing NLEV binned levels)
// in one launch: work-group g turns tile g - tile_end[j - 1] of job j, the
// first job whose tile_end exceeds g.
template <int NLEV>
__global__ __launch_bounds__(kCtT) void corner_turn_wide_batch_kernel(const ct_job* __restrict__ jobs, int njobs)
{
    const long g = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (jobs[mid].tile_end > g) hi = mid;
        else lo = mid + 1;
    }
    const ct_job& J = jobs[lo];
    ct_levels lv;
#pragma unroll
    for (int k = 0; k <= kCtMaxLevels; ++k) {
        lv.ptr[k] = J.ptr[k];
        lv.pitch[k] = J.pitch[k];
    }
    ct_wide_tile<uint8_t, NLEV>(J.in, J.nsamp, J.nchans, lv, nullptr, g - (lo ? jobs[lo - 1].tile_end : 0));
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
// variant and as the shim behind the u16 drop-in dedisperse().  u8 sums are
// exact integers; u16 samples are summed as the reference's cache kernel
// does, one float accumulator in channel order (it rounds once the sum
// passes 2^24).
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
    float total;
    if constexpr (sizeof(T) == 1) {
        uint32_t sum = 0;
        for (int c = 0; c < nchans; ++c) {
            const int delay = ddtr_delay(dmshifts[c], shift_temp);
            sum += (uint32_t)p[(size_t)c * pitch + (size_t)delay];
        }
        total = (float)sum;
    } else {
        total = 0.0f;
        for (int c = 0; c < nchans; ++c) {
            const int delay = ddtr_delay(dmshifts[c], shift_temp);
            total = __fadd_rn(total, (float)p[(size_t)c * pitch + (size_t)delay]);
        }
    }
    out[(size_t)d * out_pitch + (size_t)t] = __fdiv_rn(__fdiv_rn(total, fnchans), fbin);
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
template <typename Tin>
static void launch_ct_wide(const Tin* in, long nsamp, int nchans, const ct_levels& lv, int nlev, int* d_err,
                           hipStream_t s)
{
    const dim3 grid((unsigned)((long)ceil_div(nsamp, kCtT) * ceil_div(nchans, kCtC)));   // one per 128 KiB of input
    switch (nlev) {
    case 0: hipLaunchKernelGGL((corner_turn_wide_kernel<Tin, 0>), grid, dim3(kCtT), 0, s, in, nsamp, nchans, lv, d_err); break;
    case 1: hipLaunchKernelGGL((corner_turn_wide_kernel<Tin, 1>), grid, dim3(kCtT), 0, s, in, nsamp, nchans, lv, d_err); break;
    case 2: hipLaunchKernelGGL((corner_turn_wide_kernel<Tin, 2>), grid, dim3(kCtT), 0, s, in, nsamp, nchans, lv, d_err); break;
    default: hipLaunchKernelGGL((corner_turn_wide_kernel<Tin, 3>), grid, dim3(kCtT), 0, s, in, nsamp, nchans, lv, d_err); break;
    }
}

static bool ct_wide_fits(const void* in, int nchans, uint8_t* const* ptrs, const size_t* pitches, int nlev)
{
    if (nchans % 16 != 0 || (uintptr_t)in % 16 != 0) return false;
    for (int k = 0; k <= nlev; ++k)
        if ((uintptr_t)ptrs[k] % 16 != 0 || pitches[k] % 16 != 0) return false;
    return true;
}

bool corner_turn_levels_wide(const void* in_tc, int nchans, uint8_t* const* ptrs, const size_t* pitches, int nlev)
{
    return in_tc && nchans > 0 && nlev >= 0 && nlev <= kCtMaxLevels && ct_wide_fits(in_tc, nchans, ptrs, pitches, nlev);
}

long corner_turn_tiles(long nsamp, int nchans)
{
    return (long)ceil_div(nsamp, kCtT) * ceil_div(nchans, kCtC);
}

hipError_t launch_corner_turn_batch(const ct_job* d_jobs, int njobs, long ntiles, int nlev, hipStream_t s)
{
    if (!d_jobs || njobs <= 0 || ntiles <= 0 || ntiles > 0x7fffffffL || nlev < 0 || nlev > kCtMaxLevels)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)ntiles);
    switch (nlev) {
    case 0: hipLaunchKernelGGL((corner_turn_wide_batch_kernel<0>), grid, dim3(kCtT), 0, s, d_jobs, njobs); break;
    case 1: hipLaunchKernelGGL((corner_turn_wide_batch_kernel<1>), grid, dim3(kCtT), 0, s, d_jobs, njobs); break;
    case 2: hipLaunchKernelGGL((corner_turn_wide_batch_kernel<2>), grid, dim3(kCtT), 0, s, d_jobs, njobs); break;
    default: hipLaunchKernelGGL((corner_turn_wide_batch_kernel<3>), grid, dim3(kCtT), 0, s, d_jobs, njobs); break;
    }
    return hipGetLastError();
}

hipError_t launch_corner_turn_levels(const void* in_tc, int sample_bytes, long nsamp, int nchans,
                                     uint8_t* const* ptrs, const size_t* pitches, int nlev, int* d_err,
                                     hipStream_t s, int* fused)
{
    *fused = 0;
    if (nsamp < 0 || nchans <= 0 || nlev < 0 || nlev > kCtMaxLevels || !in_tc || !ptrs || !pitches ||
        (sample_bytes != 1 && sample_bytes != 2) || (sample_bytes == 2 && !d_err) || nsamp > 0x7fffffffL)
        return hipErrorInvalidValue;
    for (int k = 0; k <= nlev; ++k)
        if (!ptrs[k] || (size_t)(nsamp >> k) > pitches[k]) return hipErrorInvalidValue;
    if (nsamp == 0) {
        *fused = nlev;
        return hipSuccess;
    }
    if (!ct_wide_fits(in_tc, nchans, ptrs, pitches, nlev)) {
        const dim3 grid(ceil_div(nsamp, kTile), ceil_div(nchans, kTile));
        if (sample_bytes == 1)
            hipLaunchKernelGGL(corner_turn_kernel<uint8_t>, grid, dim3(256), 0, s, (const uint8_t*)in_tc, (int)nsamp,
                               nchans, ptrs[0], pitches[0], (int*)nullptr);
        else
            hipLaunchKernelGGL(corner_turn_kernel<uint16_t>, grid, dim3(256), 0, s, (const uint16_t*)in_tc,
                               (int)nsamp, nchans, ptrs[0], pitches[0], d_err);
        return hipGetLastError();
    }
    ct_levels lv{};
    for (int k = 0; k <= nlev; ++k) {
        lv.ptr[k] = ptrs[k];
        lv.pitch[k] = pitches[k];
    }
    if (sample_bytes == 1) launch_ct_wide((const uint8_t*)in_tc, nsamp, nchans, lv, nlev, d_err, s);
    else launch_ct_wide((const uint16_t*)in_tc, nsamp, nchans, lv, nlev, d_err, s);
    *fused = nlev;
    return hipGetLastError();
}

hipError_t launch_corner_turn_u8(const uint8_t* in_tc, int nsamp, int nchans, uint8_t* out_ct,
                                 size_t out_pitch, hipStream_t s)
{
    if (nsamp < 0 || nchans <= 0 || (size_t)nsamp > out_pitch) return hipErrorInvalidValue;
    int fused = 0;
    return launch_corner_turn_levels(in_tc, 1, nsamp, nchans, &out_ct, &out_pitch, 0, nullptr, s, &fused);
}

hipError_t launch_corner_turn_u16(const uint16_t* in_tc, int nsamp, int nchans, uint8_t* out_ct,
                                  size_t out_pitch, int* d_err, hipStream_t s)
{
    if (nsamp < 0 || nchans <= 0 || (size_t)nsamp > out_pitch || !d_err) return hipErrorInvalidValue;
    int fused = 0;
    return launch_corner_turn_levels(in_tc, 2, nsamp, nchans, &out_ct, &out_pitch, 0, d_err, s, &fused);
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
