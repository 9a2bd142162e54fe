// Column normalisation of a [rows][cols] array in place: the reference's
// normalization_transposed_kernel (src/aa_device_zero_dm_kernel.cu:100-164),
// which both DDTR normalisation options run:
//   - input normalisation on each [t][c] chunk before its corner turn, with
//     cols = nchans and factor (2^nbits - 1)/2 (src/aa_zero_dm.cpp:49-60,
//     called at include/aa_permitted_pipelines_generic.hpp:731-735);
//   - output normalisation on each [dm][t] plane, with cols = t_processed
//     and factor 127.5 (src/aa_zero_dm.cpp:62-74, called at
//     include/aa_permitted_pipelines_generic.hpp:823-831).
// Per column: 32 float partial sums, partial y adding rows y, y+32, ... in
// row order; then p[y] += p[y+i] for i = 16, 8, 4, 2, 1; s = p[0]/rows -
// factor; every value becomes x - s (integers: clamped to [0, 2^nbits - 1]
// and truncated).  The float sums round, so the order is kept exactly:
//   - normalize_columns_kernel keeps it literally (64 columns per
//     workgroup, wave w owns partials w and w + 16).  It serves the float
//     planes (few rows, many columns) and any integer data.
//   - For integer data whose partial sums cannot pass 2^24
//     (ceil(rows/32) * type max <= 2^24, e.g. u8 chunks up to 2,105,376
//     rows), every float partial is an exact integer, so the tall input
//     chunks are summed in integers over row segments spread across the
//     chip (atomic adds into 32 x cols counters), and only the tree runs in
//     float.  Same result, many more workgroups.
// HBM-bound: the data are read twice (sum, then apply) and written once.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "ddtr_kernels.h"

namespace aa_hip {
namespace {

__device__ __forceinline__ float norm_value(float x, float s, float maxval)
{
    float r = __fsub_rn(x, s);
    return r < 0.0f ? 0.0f : (r > maxval ? maxval : r);
}

// The 16-8-4-2-1 tree over p[0..15] (p[y] already holds partials y + y+16)
// and the column offset s.
__device__ __forceinline__ float norm_offset(float (&p)[16], long rows, float factor)
{
#pragma unroll
    for (int i = 8; i > 0; i >>= 1)
#pragma unroll
        for (int y = 0; y < i; ++y) p[y] = __fadd_rn(p[y], p[y + i]);
    return __fsub_rn(__fdiv_rn(p[0], (float)rows), factor);
}

template <class T>
__global__ __launch_bounds__(1024) void normalize_columns_kernel(T* data, long rows, long cols, float factor,
                                                                 float maxval)
{
    __shared__ float part[16][64];
    __shared__ float offs[64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long c = (long)blockIdx.x * 64 + lane;
    const bool on = c < cols;
    float a0 = 0.0f, a1 = 0.0f;
    if (on) {
#pragma unroll 4
        for (long r = w; r < rows; r += 32) {
            a0 = __fadd_rn(a0, (float)data[r * cols + c]);
            if (r + 16 < rows) a1 = __fadd_rn(a1, (float)data[(r + 16) * cols + c]);
        }
    }
    part[w][lane] = __fadd_rn(a0, a1);
    __syncthreads();
    if (w == 0) {
        float p[16];
#pragma unroll
        for (int y = 0; y < 16; ++y) p[y] = part[y][lane];
        offs[lane] = norm_offset(p, rows, factor);
    }
    __syncthreads();
    if (!on) return;
    const float s = offs[lane];
    for (long r = w; r < rows; r += 16) {
        T* v = data + r * cols + c;
        if constexpr (std::is_same<T, float>::value)
            *v = __fsub_rn(*v, s);
        else
            *v = (T)(uint32_t)norm_value((float)*v, s, maxval);
    }
}

// Float planes of a multiple of 4 columns: the same order, 4 columns per
// lane (16-byte accesses), 256 columns per workgroup.
__global__ __launch_bounds__(1024) void normalize_columns_f32x4_kernel(float* data, long rows, long cols,
                                                                       float factor)
{
    __shared__ float4 part[16][64];
    __shared__ float4 offs[64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long nq = cols >> 2;
    const long q = (long)blockIdx.x * 64 + lane;
    const bool on = q < nq;
    float4* d = reinterpret_cast<float4*>(data);
    float4 a0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), a1 = a0;
    if (on) {
#pragma unroll 4
        for (long r = w; r < rows; r += 32) {
            const float4 x = d[r * nq + q];
            a0.x = __fadd_rn(a0.x, x.x);
            a0.y = __fadd_rn(a0.y, x.y);
            a0.z = __fadd_rn(a0.z, x.z);
            a0.w = __fadd_rn(a0.w, x.w);
            if (r + 16 < rows) {
                const float4 y = d[(r + 16) * nq + q];
                a1.x = __fadd_rn(a1.x, y.x);
                a1.y = __fadd_rn(a1.y, y.y);
                a1.z = __fadd_rn(a1.z, y.z);
                a1.w = __fadd_rn(a1.w, y.w);
            }
        }
    }
    part[w][lane] = make_float4(__fadd_rn(a0.x, a1.x), __fadd_rn(a0.y, a1.y), __fadd_rn(a0.z, a1.z),
                                __fadd_rn(a0.w, a1.w));
    __syncthreads();
    if (w < 4) {   // wave k finishes component k of every lane's 4 columns
        float p[16];
#pragma unroll
        for (int y = 0; y < 16; ++y) {
            const float4 v = part[y][lane];
            p[y] = w == 0 ? v.x : w == 1 ? v.y : w == 2 ? v.z : v.w;
        }
        const float sv = norm_offset(p, rows, factor);
        float* o = reinterpret_cast<float*>(&offs[lane]);
        o[w] = sv;
    }
    __syncthreads();
    if (!on) return;
    const float4 s = offs[lane];
    for (long r = w; r < rows; r += 16) {
        float4 x = d[r * nq + q];
        x.x = __fsub_rn(x.x, s.x);
        x.y = __fsub_rn(x.y, s.y);
        x.z = __fsub_rn(x.z, s.z);
        x.w = __fsub_rn(x.w, s.w);
        d[r * nq + q] = x;
    }
}

__global__ void norm_zero_kernel(uint32_t* p, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Integer partial sums of one row segment: lane = 4 columns (one dword of
// u8), wave w = partials w and w + 16.  seg_rows is a multiple of 32, so a
// row's partial index is its global row index mod 32.
__global__ __launch_bounds__(1024) void norm_partial_u8x4_kernel(const uint8_t* data, long rows, long cols,
                                                                 long seg_rows, uint32_t* partial)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long nq = cols >> 2;
    const long q = (long)blockIdx.x * 64 + lane;
    if (q >= nq) return;   // no barrier in this kernel
    const long r0 = (long)blockIdx.y * seg_rows;
    const long r1 = rows < r0 + seg_rows ? rows : r0 + seg_rows;
    const uint32_t* d = reinterpret_cast<const uint32_t*>(data);
    uint32_t a[2][4] = {{0u, 0u, 0u, 0u}, {0u, 0u, 0u, 0u}};
#pragma unroll 4
    for (long r = r0 + w; r < r1; r += 16) {
        const uint32_t x = d[r * nq + q];
        const int k = (int)((r >> 4) & 1);   // partial w (k = 0) or w + 16 (k = 1)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t v = (x >> (8 * b)) & 0xffu;
            a[0][b] += k ? 0u : v;
            a[1][b] += k ? v : 0u;
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (a[k][b]) atomicAdd(&partial[(long)(w + 16 * k) * cols + 4 * q + b], a[k][b]);
}

// Same, one column per lane (any column count, u8 or u16).
template <class T>
__global__ __launch_bounds__(1024) void norm_partial_kernel(const T* data, long rows, long cols, long seg_rows,
                                                            uint32_t* partial)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long c = (long)blockIdx.x * 64 + lane;
    if (c >= cols) return;   // no barrier in this kernel
    const long r0 = (long)blockIdx.y * seg_rows;
    const long r1 = rows < r0 + seg_rows ? rows : r0 + seg_rows;
    uint32_t a0 = 0u, a1 = 0u;
#pragma unroll 4
    for (long r = r0 + w; r < r1; r += 32) {
        a0 += data[r * cols + c];
        if (r + 16 < r1) a1 += data[(r + 16) * cols + c];
    }
    if (a0) atomicAdd(&partial[(long)w * cols + c], a0);
    if (a1) atomicAdd(&partial[(long)(w + 16) * cols + c], a1);
}

// Exact integer partials -> float (exact) -> the reference's tree -> s[c].
__global__ void norm_offsets_kernel(const uint32_t* partial, long rows, long cols, float factor, float* offs)
{
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    float p[16];
#pragma unroll
    for (int y = 0; y < 16; ++y)
        p[y] = __fadd_rn((float)partial[(long)y * cols + c], (float)partial[(long)(y + 16) * cols + c]);
    offs[c] = norm_offset(p, rows, factor);
}

// Apply: x' = clamp(x - s[c]); blockIdx.x over column groups, rows strided
// over blockIdx.y.  u8 rows of a multiple of 16 columns: 16 columns per
// lane (one 16-byte access), four rows in flight.
__device__ __forceinline__ uint32_t norm_word(uint32_t x, const float4& s, float maxval)
{
    return (uint32_t)norm_value((float)(x & 0xffu), s.x, maxval) |
           (uint32_t)norm_value((float)((x >> 8) & 0xffu), s.y, maxval) << 8 |
           (uint32_t)norm_value((float)((x >> 16) & 0xffu), s.z, maxval) << 16 |
           (uint32_t)norm_value((float)(x >> 24), s.w, maxval) << 24;
}

__global__ __launch_bounds__(256) void norm_apply_u8x16_kernel(uint8_t* data, long rows, long cols,
                                                               const float* __restrict__ offs, float maxval)
{
    const long nv = cols >> 4;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    uint4* d = reinterpret_cast<uint4*>(data);
    const float4* o = reinterpret_cast<const float4*>(offs) + 4 * v;
    const float4 s0 = o[0], s1 = o[1], s2 = o[2], s3 = o[3];
    const long step = gridDim.y;
    long r = blockIdx.y;
    for (; r + 3 * step < rows; r += 4 * step) {
        uint4 x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = d[(r + k * step) * nv + v];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            d[(r + k * step) * nv + v] = make_uint4(norm_word(x[k].x, s0, maxval), norm_word(x[k].y, s1, maxval),
                                                    norm_word(x[k].z, s2, maxval), norm_word(x[k].w, s3, maxval));
    }
    for (; r < rows; r += step) {
        const uint4 x = d[r * nv + v];
        d[r * nv + v] = make_uint4(norm_word(x.x, s0, maxval), norm_word(x.y, s1, maxval),
                                   norm_word(x.z, s2, maxval), norm_word(x.w, s3, maxval));
    }
}

__global__ void norm_apply_u8x4_kernel(uint8_t* data, long rows, long cols, const float* __restrict__ offs,
                                       float maxval)
{
    const long nq = cols >> 2;
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    uint32_t* d = reinterpret_cast<uint32_t*>(data);
    const float4 s = reinterpret_cast<const float4*>(offs)[q];
    for (long r = blockIdx.y; r < rows; r += gridDim.y) d[r * nq + q] = norm_word(d[r * nq + q], s, maxval);
}

template <class T>
__global__ void norm_apply_kernel(T* data, long rows, long cols, const float* __restrict__ offs, float maxval)
{
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    const float s = offs[c];
    for (long r = blockIdx.y; r < rows; r += gridDim.y)
        data[r * cols + c] = (T)(uint32_t)norm_value((float)data[r * cols + c], s, maxval);
}

inline unsigned blocks_of(long n, long b) { return (unsigned)((n + b - 1) / b); }

template <class T>
hipError_t normalize_int(T* data, long rows, long cols, float factor, int nbits, void* scratch, hipStream_t s)
{
    if (!data || rows < 0 || cols <= 0 || nbits < 1 || nbits > 16) return hipErrorInvalidValue;
    if (rows == 0) return hipSuccess;
    const float maxval = (float)((1u << nbits) - 1u);
    const long type_max = std::is_same<T, uint8_t>::value ? 255 : 65535;
    const bool exact_int = scratch && ((rows + 31) / 32) * type_max <= (1L << 24);
    if (!exact_int) {
        hipLaunchKernelGGL(normalize_columns_kernel<T>, dim3(blocks_of(cols, 64)), dim3(1024), 0, s, data, rows,
                           cols, factor, maxval);
        return hipGetLastError();
    }
    uint32_t* partial = static_cast<uint32_t*>(scratch);
    float* offs = reinterpret_cast<float*>(partial + 32 * cols);
    hipLaunchKernelGGL(norm_zero_kernel, dim3(blocks_of(32 * cols, 256)), dim3(256), 0, s, partial, 32 * cols);
    const bool x4 = std::is_same<T, uint8_t>::value && cols % 4 == 0 && (uintptr_t)data % 4 == 0;
    const long colblocks = x4 ? blocks_of(cols / 4, 64) : blocks_of(cols, 64);
    // About 2048 workgroups of 16 waves, row segments a multiple of 32 rows.
    long nseg = (2048 + colblocks - 1) / colblocks;
    const long max_seg = (rows + 511) / 512;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    long seg_rows = (rows + nseg - 1) / nseg;
    seg_rows = (seg_rows + 31) / 32 * 32;
    nseg = (rows + seg_rows - 1) / seg_rows;
    const dim3 grid((unsigned)colblocks, (unsigned)nseg);
    if (x4)
        hipLaunchKernelGGL(norm_partial_u8x4_kernel, grid, dim3(1024), 0, s, (const uint8_t*)data, rows, cols,
                           seg_rows, partial);
    else
        hipLaunchKernelGGL(norm_partial_kernel<T>, grid, dim3(1024), 0, s, (const T*)data, rows, cols, seg_rows,
                           partial);
    hipLaunchKernelGGL(norm_offsets_kernel, dim3(blocks_of(cols, 256)), dim3(256), 0, s, (const uint32_t*)partial,
                       rows, cols, factor, offs);
    const unsigned ry = (unsigned)(rows < 2048 ? rows : 2048);
    if (x4 && cols % 16 == 0 && (uintptr_t)data % 16 == 0 && (uintptr_t)offs % 16 == 0)
        hipLaunchKernelGGL(norm_apply_u8x16_kernel, dim3(blocks_of(cols / 16, 256), ry), dim3(256), 0, s,
                           (uint8_t*)data, rows, cols, (const float*)offs, maxval);
    else if (x4 && (uintptr_t)offs % 16 == 0)
        hipLaunchKernelGGL(norm_apply_u8x4_kernel, dim3(blocks_of(cols / 4, 256), ry), dim3(256), 0, s,
                           (uint8_t*)data, rows, cols, (const float*)offs, maxval);
    else
        hipLaunchKernelGGL(norm_apply_kernel<T>, dim3(blocks_of(cols, 256), ry), dim3(256), 0, s, data, rows, cols,
                           (const float*)offs, maxval);
    return hipGetLastError();
}

}  // namespace

size_t normalize_columns_scratch_bytes(long cols) { return (size_t)(cols > 0 ? cols : 0) * 33u * sizeof(float); }

hipError_t launch_normalize_columns_u8(uint8_t* data, long rows, long cols, float factor, int nbits, void* scratch,
                                       hipStream_t s)
{
    if (nbits > 8) return hipErrorInvalidValue;
    return normalize_int<uint8_t>(data, rows, cols, factor, nbits, scratch, s);
}

hipError_t launch_normalize_columns_u16(uint16_t* data, long rows, long cols, float factor, int nbits,
                                        void* scratch, hipStream_t s)
{
    return normalize_int<uint16_t>(data, rows, cols, factor, nbits, scratch, s);
}

hipError_t launch_normalize_columns_f32(float* data, long rows, long cols, float factor, hipStream_t s)
{
    if (!data || rows < 0 || cols <= 0) return hipErrorInvalidValue;
    if (rows == 0) return hipSuccess;
    if (cols % 4 == 0 && (uintptr_t)data % 16 == 0)
        hipLaunchKernelGGL(normalize_columns_f32x4_kernel, dim3(blocks_of(cols / 4, 64)), dim3(1024), 0, s, data,
                           rows, cols, factor);
    else
        hipLaunchKernelGGL(normalize_columns_kernel<float>, dim3(blocks_of(cols, 64)), dim3(1024), 0, s, data,
                           rows, cols, factor, 0.0f);
    return hipGetLastError();
}

}  // namespace aa_hip
