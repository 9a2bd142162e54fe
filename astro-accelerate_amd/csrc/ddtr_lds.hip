// LDS-tiled dedispersion kernel for gfx950 (MI355X).
//
// Work-group tile: kDms DM trials (one per wave) x kT output samples (64
// lanes x kR samples each).  Channels are streamed kC at a time: the rows of
// a channel chunk (u8 samples, the window the tile's DMs need) are copied
// into LDS with global_load_lds_dwordx4 (straight from global memory to LDS,
// no registers), double-buffered, so the copy of chunk k+1 is in flight while
// chunk k is summed.  One barrier per chunk.
//
// Inner loop, per wave (one DM), per channel pair: each lane reads its
// 32-sample run from the two LDS rows with three aligned ds_read_b128 each,
// then v_perm_b32 with scalar selectors does the byte realignment and the
// u8 -> 16-bit unpack in one instruction, and v_add3_u32 adds both channels
// into 16 packed accumulators (two samples per dword, 16-bit fields).  The
// packed fields are flushed into 32-bit accumulators every 256 channels
// (256 * 255 < 2^16).  The misalignment of a lane run is the same for all
// lanes of the wave (lane runs are multiples of 16 samples apart), so it is
// resolved by a scalar branch over the 16 (dword shift a, dword shift b)
// combinations; the byte shift inside a dword goes into the selectors.
//
// Every sum is an exact integer; the output is ((float)sum / nchans) / inBin
// as in the reference's failsafe kernel (src/aa_device_dedispersion_kernel.cu:448),
// with delays trunc(dmshift[c] * fma(d, mdmstep, mstartdm)) (:430,440).
#include "ddtr_lds.h"
#include "ddtr_pair_common.h"

#include <cstring>

#include <algorithm>
#include <vector>

namespace aa_hip {

namespace {

constexpr int kR = 32;              // samples per lane
constexpr int kT = 64 * kR;         // samples per tile

// Channel-chunk shape: KC channels per LDS chunk, staged rows of up to RQ
// uint4.  A staged row holds the even quads of the row's window in its first
// half and the odd quads in its second half (kHalf apart).  A lane run
// starts at quad E + 2*lane; every lane has the parity of E, so each of the
// run's three reads is a stride-1 access across the wave (no LDS bank
// conflicts), where a plain layout would be stride 2.  Both shapes fill the
// same LDS buffers: 32 channels x 152 quads, or 16 channels x 304 quads for
// plans whose tiles need a longer window (larger delay spread).
template <int KC, int RQ>
struct lds_shape {
    static constexpr int kC = KC;
    static constexpr int kRowQuads = RQ;
    static constexpr int kHalf = RQ / 2;
    static constexpr int kFlushChunks = 256 / KC;  // 16-bit fields hold 256 * 255
};
using shape_c32 = lds_shape<32, kLdsRowQuadsMax>;
using shape_c16 = lds_shape<16, 2 * kLdsRowQuadsMax>;
constexpr int kBufQuads = 32 * kLdsRowQuadsMax;
constexpr int kOffsMax = kDms * 32;
static_assert(shape_c16::kC * shape_c16::kRowQuads == kBufQuads, "shapes share the LDS buffers");

// Two separate LDS arrays (not one [2][...] array), so that the compiler can
// tell the buffer being summed from the buffer being filled by the
// direct-to-LDS copies and does not wait for those copies before each read.
__shared__ uint4 g_rows0[kBufQuads];
__shared__ uint4 g_rows1[kBufQuads];
__shared__ int g_offs0[kOffsMax];
__shared__ int g_offs1[kOffsMax];

__device__ __forceinline__ int dev_delay(float dmshift, int d, float mstartdm, float mdmstep)
{
    const float shift_temp = __fmaf_rn((float)d, mdmstep, mstartdm);
    return (int)__fmul_rn(dmshift, shift_temp);
}

// One channel pair: wa / wb are the 12 dwords read from each channel's row
// (aligned to 16 bytes), QA / QB the dword shift of the run start, sel*0 /
// sel*1 the byte selectors of even / odd sample pairs.
template <int QA, int QB>
__device__ __forceinline__ void accumulate_pair(uint32_t (&acc)[kR / 2], const uint32_t (&wa)[12],
                                                const uint32_t (&wb)[12], uint32_t sela0, uint32_t sela1,
                                                uint32_t selb0, uint32_t selb1)
{
#pragma unroll
    for (int j = 0; j < kR / 2; ++j) {
        const int ja = QA + (j >> 1);
        const int jb = QB + (j >> 1);
        const uint32_t xa = __builtin_amdgcn_perm(wa[ja + 1], wa[ja], (j & 1) ? sela1 : sela0);
        const uint32_t xb = __builtin_amdgcn_perm(wb[jb + 1], wb[jb], (j & 1) ? selb1 : selb0);
        // one v_add3_u32 (two channels, two samples), written as a single
        // instruction so the add cannot be split and moved out of the case
        asm("v_add3_u32 %0, %0, %1, %2" : "+v"(acc[j]) : "v"(xa), "v"(xb));
    }
}

template <int QA>
__device__ __forceinline__ void accumulate_one(uint32_t (&acc)[kR / 2], const uint32_t (&wa)[12], uint32_t sela0,
                                               uint32_t sela1)
{
#pragma unroll
    for (int j = 0; j < kR / 2; ++j) {
        const int ja = QA + (j >> 1);
        const uint32_t xa = __builtin_amdgcn_perm(wa[ja + 1], wa[ja], (j & 1) ? sela1 : sela0);
        asm("v_add_u32 %0, %0, %1" : "+v"(acc[j]) : "v"(xa));
    }
}

// Reads the 3 quads of a lane run from a split row: `row` is the row base,
// E the run's first quad for lane 0 (uniform).  Quad 2k is at row[k], quad
// 2k+1 at row[kHalf + k].
template <class S>
__device__ __forceinline__ void load_run(const uint4* row, int E, int lane, uint32_t (&w)[12])
{
    constexpr int kHalf = S::kHalf;
    // Offsets as integer arithmetic on the parity p (a select between two
    // pointers would hide the LDS array from the compiler's alias analysis
    // and make it wait for the in-flight copies before these reads).
    const int p = E & 1;
    const int m = (E >> 1) + lane;
    const u32x4* lo = reinterpret_cast<const u32x4*>(row + m + p * kHalf);              // quads E, E+2
    const u32x4* mid = reinterpret_cast<const u32x4*>(row + m + kHalf + p * (1 - kHalf));  // quad E+1
    u32x4 v0 = lo[0], v1 = mid[0], v2 = lo[1];
    // Every variant uses a different subset of the 12 dwords; marking all of
    // them as used keeps each read a single ds_read_b128 (the compiler would
    // otherwise split a partly used quad into narrower LDS reads).
    asm("" : "+v"(v0), "+v"(v1), "+v"(v2));
    w[0] = v0.x; w[1] = v0.y; w[2] = v0.z; w[3] = v0.w;
    w[4] = v1.x; w[5] = v1.y; w[6] = v1.z; w[7] = v1.w;
    w[8] = v2.x; w[9] = v2.y; w[10] = v2.z; w[11] = v2.w;
}

// Selector for the bytes at offsets r, r+1 of an 8-byte window, placed in
// bytes 0 and 2 of the result (selector 0x0c yields 0x00 in bytes 1 and 3).
__device__ __forceinline__ uint32_t pair_selector(int r)
{
    return 0x0c000c00u | ((uint32_t)(r + 1) << 16) | (uint32_t)r;
}

// Direct global -> LDS copy of 16 bytes per lane (global_load_lds_dwordx4):
// lane l's 16 bytes land at LDS byte address lds_base + 16*l.  Issued from
// inline asm so that the compiler does not track it: its wait-count pass
// would otherwise stall LDS reads of the other buffer on these copies.  The
// kernel waits for them explicitly (vmcnt(0)) before each barrier; M0 is
// saved and restored around the copy.
__device__ __forceinline__ void lds_copy16(const void* src, uint32_t lds_base)
{
    uint32_t saved;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %2\n\t"
        "global_load_lds_dwordx4 %1, off\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(saved)
        : "v"(src), "s"(lds_base)
        : "memory");
}

__device__ __forceinline__ uint32_t lds_addr(const uint4* p)
{
    return (uint32_t)(uintptr_t)p;
}

struct tile_ctx {
    const uint8_t* __restrict__ L;
    const float* __restrict__ dmshifts;
    int d0, lane, wave, tid, Wq;
    long t0;
};

// Issues the copies of chunk `chunk` into `rows`: wave w copies rows w and
// w + 16 of the chunk, 64 uint4 per instruction.  The row base comes from
// a scalar load of the row's dmshift, so no vector load precedes the copies
// (vmcnt is in order: waiting on such a load would wait on the copies too).
template <class S>
__device__ __forceinline__ void stage_chunk(const lds_params& p, const tile_ctx& t, int chunk, uint4* rows)
{
    constexpr int kC = S::kC, kHalf = S::kHalf;
    const int c0 = chunk * kC;
    const int nrows = (p.nchans - c0) < kC ? (p.nchans - c0) : kC;
    for (int row = t.wave; row < nrows; row += kDms) {
        const int c = c0 + row;
        const float ds = t.dmshifts[c];
        const int b16 = __builtin_amdgcn_readfirstlane(dev_delay(ds, t.d0, p.mstartdm, p.mdmstep) & ~15);
        const uint8_t* src = t.L + (size_t)c * p.pitch + (size_t)t.t0 + (size_t)b16;
        uint4* dst = rows + row * S::kRowQuads;
        const int n_even = (t.Wq + 1) >> 1, n_odd = t.Wq >> 1;
        for (int j = 0; j < n_even; j += 64) {
            if (j + t.lane < n_even)
                lds_copy16(src + 32 * (size_t)(j + t.lane), __builtin_amdgcn_readfirstlane(lds_addr(dst + j)));
        }
        for (int j = 0; j < n_odd; j += 64) {
            if (j + t.lane < n_odd)
                lds_copy16(src + 16 + 32 * (size_t)(j + t.lane),
                           __builtin_amdgcn_readfirstlane(lds_addr(dst + kHalf + j)));
        }
    }
}

// Offset e = delay(c, d0 + wd) - floor16(delay(c, d0)) of (wd, ci) = (tid / kC, tid % kC).
template <class S>
__device__ __forceinline__ int chunk_offset(const lds_params& p, const tile_ctx& t, int chunk)
{
    constexpr int kC = S::kC;
    int e = 0;
    if (t.tid < kDms * kC) {
        const int wd = t.tid / kC, ci = t.tid % kC;
        const int c = chunk * kC + ci;
        const int dd = t.d0 + wd;
        if (c < p.nchans && dd < p.ndms) {
            const float ds = t.dmshifts[c];
            e = dev_delay(ds, dd, p.mstartdm, p.mdmstep) - (dev_delay(ds, t.d0, p.mstartdm, p.mdmstep) & ~15);
        }
    }
    return e;
}

// All channel pairs of the chunk whose two runs have dword shifts (QA, QB):
// bit 2p of `m` set means pair (2p, 2p+1), whose offsets sit in lanes 2p and
// 2p+1 of e_v.  A plain loop, so the accumulators stay in their registers.
template <class S, int QA, int QB>
__device__ __forceinline__ void sum_group(uint64_t m, int e_v, const uint4* rows, int lane, uint32_t (&acc16)[kR / 2])
{
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        const int ea = __builtin_amdgcn_readlane(e_v, l);
        const int eb = __builtin_amdgcn_readlane(e_v, l + 1);
        uint32_t wa[12], wb[12];
        load_run<S>(rows + l * S::kRowQuads, ea >> 4, lane, wa);
        load_run<S>(rows + (l + 1) * S::kRowQuads, eb >> 4, lane, wb);
        const uint32_t sa0 = pair_selector(ea & 3), sa1 = sa0 + 0x00020002u;
        const uint32_t sb0 = pair_selector(eb & 3), sb1 = sb0 + 0x00020002u;
        accumulate_pair<QA, QB>(acc16, wa, wb, sa0, sa1, sb0, sb1);
    }
}

template <class S, int V>
__device__ __forceinline__ void sum_groups(int variant, bool voter, int e_v, const uint4* rows, int lane,
                                           uint32_t (&acc16)[kR / 2])
{
    const uint64_t m = __builtin_amdgcn_ballot_w64(voter && variant == V);
    sum_group<S, V / 4, V % 4>(m, e_v, rows, lane, acc16);
    if constexpr (V + 1 < 16) sum_groups<S, V + 1>(variant, voter, e_v, rows, lane, acc16);
}

// Sums one staged chunk into the wave's packed accumulators.  Lane l holds
// the offset of chunk channel l; pairs are grouped by their (dword shift a,
// dword shift b) variant with one ballot per variant.
template <class S>
__device__ __forceinline__ void sum_chunk(const lds_params& p, const tile_ctx& t, int chunk, const uint4* rows,
                                          const int* offs, uint32_t (&acc16)[kR / 2])
{
    constexpr int kC = S::kC;
    const int nrows = (p.nchans - chunk * kC) < kC ? (p.nchans - chunk * kC) : kC;
    const int e_v = t.lane < nrows ? offs[t.wave * kC + t.lane] : 0;
    const int q = (e_v >> 2) & 3;
    const int q_next = __shfl_down(q, 1, 64);
    const bool voter = ((t.lane & 1) == 0) && (t.lane + 1 < nrows);
    sum_groups<S, 0>(q * 4 + q_next, voter, e_v, rows, t.lane, acc16);
    if (nrows & 1) {  // odd channel count in the last chunk: its last channel alone
        const int ea = __builtin_amdgcn_readlane(e_v, nrows - 1);
        uint32_t wa[12];
        load_run<S>(rows + (nrows - 1) * S::kRowQuads, ea >> 4, t.lane, wa);
        const uint32_t sa0 = pair_selector(ea & 3), sa1 = sa0 + 0x00020002u;
        switch ((ea >> 2) & 3) {
            case 0: accumulate_one<0>(acc16, wa, sa0, sa1); break;
            case 1: accumulate_one<1>(acc16, wa, sa0, sa1); break;
            case 2: accumulate_one<2>(acc16, wa, sa0, sa1); break;
            default: accumulate_one<3>(acc16, wa, sa0, sa1); break;
        }
    }
}

__device__ __forceinline__ void flush(uint32_t (&acc16)[kR / 2], uint32_t (&acc32)[kR])
{
#pragma unroll
    for (int j = 0; j < kR / 2; ++j) {
        acc32[2 * j] += acc16[j] & 0xFFFFu;
        acc32[2 * j + 1] += acc16[j] >> 16;
        acc16[j] = 0;
    }
}

// One pipeline step: sum chunk `chunk` from (rows_cur, offs_cur) while the
// copy of chunk + 1 into rows_next is in flight; publish its offsets, wait
// for the copies, barrier.
template <class S>
__device__ __forceinline__ void pipeline_step(const lds_params& p, const tile_ctx& t, int chunk, int nchunks,
                                              bool active, const uint4* rows_cur, const int* offs_cur,
                                              uint4* rows_next, int* offs_next, uint32_t (&acc16)[kR / 2],
                                              uint32_t (&acc32)[kR])
{
    const bool more = chunk + 1 < nchunks;
    if (more) stage_chunk<S>(p, t, chunk + 1, rows_next);
    if (active) {
        sum_chunk<S>(p, t, chunk, rows_cur, offs_cur, acc16);
        if ((chunk + 1) % S::kFlushChunks == 0 || !more) flush(acc16, acc32);
    }
    if (more) {
        const int e = chunk_offset<S>(p, t, chunk + 1);
        if (t.tid < kDms * S::kC) offs_next[t.tid] = e;
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's copies have landed
    __syncthreads();
}

// One tile: DM trials d0 .. d0 + kDms - 1 (one per wave) x samples
// t0 .. t0 + kT - 1 of the range described by p.  Every wave of the
// work-group takes part in every barrier; waves past the last DM trial only
// stage rows.  Ends with all LDS reads done (last barrier), so the caller
// may start another tile at once.
template <class S>
__device__ __forceinline__ void process_tile(const lds_params& p, int d0, long t0)
{
    tile_ctx t;
    t.L = p.L;
    t.dmshifts = p.dmshifts;
    t.d0 = d0;
    t.t0 = t0;
    t.tid = threadIdx.x;
    t.lane = t.tid & 63;
    t.wave = __builtin_amdgcn_readfirstlane(t.tid >> 6);
    t.Wq = p.row_quads;
    const int d = t.d0 + t.wave;
    const bool active = d < p.ndms;
    const int nchunks = (p.nchans + S::kC - 1) / S::kC;

    uint32_t acc16[kR / 2];
    uint32_t acc32[kR];
#pragma unroll
    for (int j = 0; j < kR / 2; ++j) acc16[j] = 0;
#pragma unroll
    for (int j = 0; j < kR; ++j) acc32[j] = 0;

    stage_chunk<S>(p, t, 0, g_rows0);
    {
        const int e = chunk_offset<S>(p, t, 0);
        if (t.tid < kDms * S::kC) g_offs0[t.tid] = e;
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    for (int chunk = 0; chunk < nchunks; chunk += 2) {
        pipeline_step<S>(p, t, chunk, nchunks, active, g_rows0, g_offs0, g_rows1, g_offs1, acc16, acc32);
        if (chunk + 1 >= nchunks) break;
        pipeline_step<S>(p, t, chunk + 1, nchunks, active, g_rows1, g_offs1, g_rows0, g_offs0, acc16, acc32);
    }

    if (active) {
        const long tl = t.t0 + (long)kR * t.lane;   // a multiple of 4
        float* orow = p.out + (size_t)d * p.out_pitch;
        const bool aligned = (reinterpret_cast<uintptr_t>(orow) & 15u) == 0;
#pragma unroll
        for (int j = 0; j < kR; j += 4) store_out4(orow, tl + j, p.t_processed, aligned, out_values4(acc32 + j, p));
    }
}

// The XCD this work-group runs on (hardware register XCC_ID, bits 0-3), so
// that it starts from the queue whose tiles share its XCD's L2.
__device__ __forceinline__ int batch_xcd()
{
    return (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u);
}

// ---------------------------------------------------------------------------
// Quad pre-summation (chunk shape 3).  As the pair shape, with rows that sum
// the four channels of a group: in a 16-DM tile the three delays of channels
// c+1, c+2, c+3 relative to c take a few combinations (about 3.4), and one
// row per combination,
//     S[i] = L[c][base + i] + sum_m L[c+m][base + v_m + i],
// serves every DM whose combination it is, at offset e = delay(c, d) - base.
// A DM then adds 4 rows per 16-channel chunk instead of 8: half the LDS
// reads of the pair shape, for about twice its build work.  A group with
// more than four combinations is built as two pairs (or single channels).
// Rows are narrower (at most 192 quads: window offsets up to 256 samples),
// so the two buffers take 96 KiB.  Accumulator layout, plan records (host
// computed, one per DM tile and chunk) and the chunk pipeline are the pair
// shape's.
// ---------------------------------------------------------------------------
constexpr int kQIt = 3;                        // quads per lane per row, at most
constexpr int kQRQ = 64 * kQIt;                // 192 quads per row
constexpr int kQRowDw = 4 * kQRQ;              // 768 dwords per row
constexpr int kQRows = 16;                     // rows per buffer (one per wave)
constexpr int kQT = 1024;                      // samples per tile
constexpr int kQR = pair_geom<kQT>::R;         // 16 samples per lane
constexpr int kQSplit = pair_geom<kQT>::SPLIT;
constexpr int kQMaxOffset = kQRowDw - kQSplit; // e + 511 < kQRowDw
constexpr int kQGroups = kPairChunk / 4;       // 4-channel groups per chunk
constexpr int kQMaxGroupRows = 4;
static_assert(kQGroups * kQMaxGroupRows == kQRows && kQRows == kDms, "one row per wave");
// Four rows of up to 4 x 255 per chunk: 4080 per 16-bit field, flushed
// every kPFlushChunks chunks.
static_assert(kPFlushChunks * 4 * 4 * 255 <= 65535, "16-bit fields do not overflow");

struct quad_row {
    long long off[4];   // channel m: bytes from the level start to the dword floor of its window
    uint32_t sh;        // byte shift of channel m's window in bits 8m..8m+1 (channel 0: 0)
    uint32_t nch;       // channels summed: 4, 2 or 1 (unused channels repeat channel 0's window)
    int pad0, pad1;
};
struct quad_rec {
    uint16_t ent[kDms][kPMaxEnt];  // per DM slot: dword offset (row * kQRowDw + e) of each entry
    quad_row row[kQRows];
    int nrows, nent, pad0, pad1;
};
static_assert(sizeof(quad_row) == 48 && sizeof(quad_rec) % 16 == 0, "records are 16-byte aligned");
static_assert(kQRows * kQRowDw <= 65536, "entry offsets fit 16 bits");

__shared__ uint4 g_qrow[2 * kQRows * kQRQ];   // 2 x 48 KiB
__device__ __forceinline__ uint4* quad_buf(int b) { return g_qrow + b * (kQRows * kQRQ); }

// Registers of one row in flight: channel 0's low and high dwords (its
// window is dword aligned), and for channels 1-3 two dwords each around the
// low and the high sample.
template <int NI>
struct quad_load {
    uint32_t a[NI][2];
    uint32_t x[NI][3][4];
    uint32_t sh, nch;
};

// Loads of one row, issued whatever the row's use (no register merges on a
// branch; a wave without a row reads its zeroed record's window at the level
// start).  A lane over-reads at most 63 quads past the needed width, inside
// the 4096-byte slack every row has past its longest window.
template <int NI>
__device__ __forceinline__ void quad_issue_loads(const lds_params& p, long t0, const quad_row& R, int lane,
                                                 quad_load<NI>& B)
{
    const uint8_t* r[4] = {p.L + t0 + R.off[0], p.L + t0 + R.off[1], p.L + t0 + R.off[2], p.L + t0 + R.off[3]};
    B.sh = R.sh;
    B.nch = R.nch;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int q = lane + 64 * i;
        const uint32_t* pa = reinterpret_cast<const uint32_t*>(r[0] + 4 * q);
        B.a[i][0] = pa[0];
        B.a[i][1] = pa[kQSplit / 4];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const uint32_t* pm = reinterpret_cast<const uint32_t*>(r[m + 1] + 4 * q);
            B.x[i][m][0] = pm[0];
            B.x[i][m][1] = pm[1];
            B.x[i][m][2] = pm[kQSplit / 4];
            B.x[i][m][3] = pm[kQSplit / 4 + 1];
        }
    }
}

// Row sums into `rows` (row `wave`): dword j of quad q is the 16-bit sums of
// the channels' byte j of the low dword and of the high dword.  MASK: rows
// of fewer than four channels (the unused ones are masked off).
template <int NI, bool MASK>
__device__ __forceinline__ void quad_store_row(int wave, int lane, const quad_load<NI>& B, uint4* rows)
{
    uint4* r = rows + wave * kQRQ + lane;
    uint32_t mask[3];   // channel m + 1 summed
#pragma unroll
    for (int m = 0; m < 3; ++m) mask[m] = !MASK || B.nch > (uint32_t)(m + 1) ? ~0u : 0u;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        uint32_t lo[3], hi[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const uint32_t sh = (B.sh >> (8 * (m + 1))) & 3u;
            const uint32_t mk = mask[m];
            lo[m] = __builtin_amdgcn_alignbyte(B.x[i][m][1], B.x[i][m][0], sh) & mk;
            hi[m] = __builtin_amdgcn_alignbyte(B.x[i][m][3], B.x[i][m][2], sh) & mk;
        }
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t sel = 0x0c000c00u | ((uint32_t)(4 + j) << 16) | (uint32_t)j;
            o[j] = __builtin_amdgcn_perm(B.a[i][1], B.a[i][0], sel) + __builtin_amdgcn_perm(hi[0], lo[0], sel) +
                   __builtin_amdgcn_perm(hi[1], lo[1], sel) + __builtin_amdgcn_perm(hi[2], lo[2], sel);
        }
        r[64 * i] = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

template <int NI>
__device__ __forceinline__ void quad_store(int wave, int lane, const quad_load<NI>& B, uint4* rows)
{
    if (B.nch == 4) quad_store_row<NI, false>(wave, lane, B, rows);
    else quad_store_row<NI, true>(wave, lane, B, rows);
}

// All n entries of DM slot `wave` in chunk record rc, read from the record.
__device__ __forceinline__ void quad_sum_any(const quad_rec& rc, int wave, int n, const uint32_t* base,
                                             uint32_t (&acc)[kQR / 2])
{
    const uint32_t* ew = reinterpret_cast<const uint32_t*>(&rc.ent[wave][0]);
    for (int i = 0; i < n; ++i) {
        const uint32_t e = ew[i >> 1];
        pair_add1<kQR>(base, (int)((i & 1) ? (e >> 16) : (e & 0xffffu)), acc);
    }
}

__device__ __forceinline__ pair_ents quad_ents_of(const quad_rec& rc, int wave)
{
    pair_ents E;
    const uint32_t* ew = reinterpret_cast<const uint32_t*>(&rc.ent[wave][0]);
#pragma unroll
    for (int i = 0; i < kPMaxEnt / 2; ++i) E.e[i] = ew[i];
    E.n = rc.nent;
    return E;
}

// The pair shape's pipeline: step k stores chunk k+1's rows (loaded during
// step k-1), issues the loads of chunk k+2, sums chunk k, one barrier.
template <int NI>
__device__ __forceinline__ void process_quad_tile(const lds_params& p, int dt, long t0)
{
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool active = dt * kDms + wave < p.ndms;
    const int nchunks = (p.nchans + kPairChunk - 1) / kPairChunk;
    const quad_rec* recs = reinterpret_cast<const quad_rec*>(p.pair_recs) + (size_t)dt * nchunks;
    uint32_t acc16[kQR / 2];
    uint32_t acc32[kQR];
#pragma unroll
    for (int j = 0; j < kQR / 2; ++j) acc16[j] = 0;
#pragma unroll
    for (int j = 0; j < kQR; ++j) acc32[j] = 0;

    quad_load<NI> B;
    quad_issue_loads(p, t0, recs[0].row[wave], lane, B);
    if (wave < recs[0].nrows) quad_store(wave, lane, B, quad_buf(0));
    const int k1 = min(1, nchunks - 1);
    bool have_next = k1 == 1 && wave < recs[k1].nrows;   // B holds chunk k+1's row
    quad_issue_loads(p, t0, recs[k1].row[wave], lane, B);
    int kr = min(2, nchunks - 1);
    quad_row rn = recs[kr].row[wave];                     // row record of chunk k+2
    int nrows_n = recs[kr].nrows;
    pair_ents E = quad_ents_of(recs[0], wave);
    const uint32_t* lane_dw = reinterpret_cast<const uint32_t*>(quad_buf(0)) + lane;
    __syncthreads();
    for (int k = 0; k < nchunks; ++k) {
        const uint32_t* base = lane_dw + ((k & 1) ? kQRows * kQRowDw : 0);   // rows of chunk k
        if (have_next) quad_store(wave, lane, B, quad_buf((k + 1) & 1));
        have_next = k + 2 < nchunks && wave < nrows_n;
        quad_issue_loads(p, t0, rn, lane, B);
        if (active) {
            if (E.n == kQGroups) {   // the usual full chunk: four rows, one batch of reads
                pair_x4<kQR> X;
                pair_read4<kQR>(base, E.e[0], E.e[1], X);
                pair_acc4<kQR>(X, acc16);
            } else {
                quad_sum_any(recs[k], wave, E.n, base, acc16);
            }
            if ((k + 1) % kPFlushChunks == 0 || k + 1 == nchunks) pair_flush<kQR>(acc16, acc32);
        }
        kr = min(k + 3, nchunks - 1);
        rn = recs[kr].row[wave];
        nrows_n = recs[kr].nrows;
        E = quad_ents_of(recs[min(k + 1, nchunks - 1)], wave);
        __syncthreads();
    }

    if (active) {
        const int d = dt * kDms + wave;
        float* orow = p.out + (size_t)d * p.out_pitch;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < kQR / 2; ++j) {
                const long tt = t0 + h * kQSplit + lane + 64 * j;
                if (tt < p.t_processed)
                    orow[tt] = out_value(acc32[h * (kQR / 2) + j], p);
            }
    }
}

template <int NI>
__global__ __launch_bounds__(kThreads) void dedisperse_quad_kernel(lds_params p)
{
    const int ntiles = p.n_dm_tiles * p.n_t_tiles;
    const int per_xcd = (ntiles + 7) / 8;
    const int w = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (w >= ntiles) return;  // whole work-group leaves before any barrier
    process_quad_tile<NI>(p, w % p.n_dm_tiles, (long)(w / p.n_dm_tiles) * kQT);
}

// One range: a flat grid of tiles (DM-major within a time tile).
template <class S>
__global__ __launch_bounds__(kThreads) void dedisperse_lds_kernel(lds_params p)
{
    // XCD-aware tile order: work-groups are dispatched round-robin over the
    // 8 XCDs, so each XCD takes a contiguous block of tiles (consecutive DM
    // tiles of a time tile share most of their input window in its L2).
    const int ntiles = p.n_dm_tiles * p.n_t_tiles;
    const int per_xcd = (ntiles + 7) / 8;
    const int w = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (w >= ntiles) return;  // whole work-group leaves before any barrier
    process_tile<S>(p, (w % p.n_dm_tiles) * kDms, (long)(w / p.n_dm_tiles) * kT);
}

// Many ranges (of many beams) in one persistent launch.  Each of the 8
// queues (one per XCD) holds the x-th eighth of every job, job after job
// (lds_job), so the XCDs work through the jobs side by side, as a launch per
// range would; within an XCD, neighbouring tiles share input rows in its L2.
// A work-group takes tiles from its own XCD's queue and, once it is empty,
// from the others in turn.  Every work-group leaves when all 8 queues are
// empty.
__global__ __launch_bounds__(kThreads) void dedisperse_lds_batch_kernel(const lds_job* __restrict__ jobs, int njobs,
                                                                        const long* __restrict__ bounds,
                                                                        int* __restrict__ counters)
{
    __shared__ int s_q[2];   // queue, position
    const int xcd = batch_xcd();
    int q = 0;  // queues found empty so far (thread 0)
    for (;;) {
        if (threadIdx.x == 0) {
            int qx = -1, pos = 0;
            while (q < 8) {
                const int x = (xcd + q) & 7;
                const long i = bounds[x] + (long)atomicAdd(&counters[x], 1);
                if (i < bounds[x + 1]) {
                    qx = x;
                    pos = (int)(i - bounds[x]);   // the host keeps the tile count below 2^31
                    break;
                }
                ++q;
            }
            s_q[0] = qx;
            s_q[1] = pos;
        }
        __syncthreads();
        // wave-uniform (scalar) from here on: the tile's indices and addresses
        const int qx = __builtin_amdgcn_readfirstlane(s_q[0]);
        const long pos = __builtin_amdgcn_readfirstlane(s_q[1]);
        __syncthreads();
        if (qx < 0) break;
        long local;
        const lds_job& J = batch_job_of(jobs, njobs, qx, pos, &local);
        const int d0 = (int)(local % J.p.n_dm_tiles) * kDms;
        const long t0 = (local / J.p.n_dm_tiles) * kT;
        if (J.variant == 0) process_tile<shape_c32>(J.p, d0, t0);
        else process_tile<shape_c16>(J.p, d0, t0);
    }
}

__global__ void batch_reset_kernel(int* counters)
{
    if (threadIdx.x < 8) counters[threadIdx.x] = 0;
}

}  // namespace

template <int NDM>
static int pair_plan(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, size_t pitch,
                     int tile, std::vector<pair_rec_t<NDM>>* recs);
static int quad_plan(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, size_t pitch,
                     std::vector<quad_rec>* recs);

int lds_row_quads(int variant, const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep)
{
    if (variant < 0 || variant >= kLdsVariants) return -1;
    if (variant == 2) return pair_plan<kDms>(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 0, kWT2, nullptr);
    if (variant == 4) return pair_plan<kDms>(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 0, kWT4, nullptr);
    if (variant == 5) return pair_plan<kDms>(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 0, kWT3, nullptr);
    if (variant == 6) return pair_plan<kDms2>(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 0, kWT2, nullptr);
    if (variant == 3) return quad_plan(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 0, nullptr);
    // e_max over every tile and channel: delay(c, last DM of the tile) -
    // floor16(delay(c, first DM)); a lane run reads 48 bytes starting at
    // 16*floor(e/16) + 32*lane.
    long e_max = 0;
    for (int dt = 0; dt * kDms < ndms; ++dt) {
        const int da = dt * kDms;
        const int db = (da + kDms - 1 < ndms) ? da + kDms - 1 : ndms - 1;
        const float s0 = __builtin_fmaf((float)da, mdmstep, mstartdm);
        const float s1 = __builtin_fmaf((float)db, mdmstep, mstartdm);
        for (int c = 0; c < nchans; ++c) {
            const long a = (long)(int)(h_dmshifts[c] * s0);
            const long b = (long)(int)(h_dmshifts[c] * s1);
            const long e = b - (a & ~15L);
            if (e > e_max) e_max = e;
        }
    }
    const long quads = (e_max >> 4) + 2 * 63 + 3;  // last lane's third quad
    const long cap = variant == 0 ? shape_c32::kRowQuads : shape_c16::kRowQuads;
    return quads > cap ? -1 : (int)quads;
}

// Pair shape plan for tiles of `tile` samples (1024: shape 2, 2016: shape
// 4): quads per row (window offsets up to e_max need dwords up to
// e_max + tile/2 - 1), or -1 when some offset exceeds a row or a pair's
// relative delay is negative.  With recs, also the records
// of every (DM tile, chunk), DM-tile-major, for rows `pitch` bytes apart.
// Delays follow the kernels' rule (truncation of
// dmshift * fmaf(d, mdmstep, mstartdm) in float).
template <int NDM>
static int pair_plan(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, size_t pitch,
                     int tile, std::vector<pair_rec_t<NDM>>* recs)
{
    const int max_off = pair_max_offset(tile);
    auto make_row = [pitch](int a, int base, int b, int vb) {
        // channel a's window at base; channel b (or none: b < 0) at base + vb
        pair_row r{};
        r.offa = (long long)a * (long long)pitch + base;
        const int bb = b >= 0 ? b : a;
        const int wb = b >= 0 ? base + vb : base;
        r.offb = (long long)bb * (long long)pitch + (wb & ~3);
        r.sh = (uint32_t)(wb & 3);
        r.mb = b >= 0 ? ~0u : 0u;
        return r;
    };
    if (nchans <= 0 || ndms <= 0) return -1;
    const int nchunks = (nchans + kPairChunk - 1) / kPairChunk;
    const int ndt = (ndms + NDM - 1) / NDM;
    if (recs) recs->assign((size_t)ndt * nchunks, pair_rec_t<NDM>{});
    std::vector<int> dl((size_t)NDM * nchans), d0l((size_t)nchans);
    int e_max = 0;
    for (int dt = 0; dt < ndt; ++dt) {
        const int d0 = dt * NDM;
        const float s0 = __builtin_fmaf((float)d0, mdmstep, mstartdm);
        for (int c = 0; c < nchans; ++c) d0l[(size_t)c] = (int)(h_dmshifts[c] * s0);
        for (int w = 0; w < NDM; ++w) {
            const int d = std::min(d0 + w, ndms - 1);
            const float sd = __builtin_fmaf((float)d, mdmstep, mstartdm);
            for (int c = 0; c < nchans; ++c) dl[(size_t)w * nchans + c] = (int)(h_dmshifts[c] * sd);
        }
        auto D = [&](int w, int c) { return dl[(size_t)w * nchans + c]; };
        // Values taken by the relative delay of the pair starting at channel a.
        auto pair_values = [&](int a) {
            if (a + 1 >= nchans) return 1;
            int lo = D(0, a + 1) - D(0, a), hi = lo;
            for (int w = 1; w < NDM; ++w) {
                lo = std::min(lo, D(w, a + 1) - D(w, a));
                hi = std::max(hi, D(w, a + 1) - D(w, a));
            }
            return hi - lo + 1;
        };
        for (int k = 0; k < nchunks; ++k) {
            pair_rec_t<NDM>* R = recs ? &(*recs)[(size_t)dt * nchunks + k] : nullptr;
            int nrows = 0, npairs = 0, nsplit = 0;
            int split_row[kPairs], split_e[kPairs][NDM];
            // Rows left over when every pair of more than two values is
            // split: a pair of exactly three values takes one of them and
            // is built as three pair rows (one entry per DM instead of two)
            // while they last, in pair order.
            int spare = kPRows;
            for (int pp = 0; pp < kPairs && k * kPairChunk + 2 * pp < nchans; ++pp)
                spare -= std::min(2, pair_values(k * kPairChunk + 2 * pp));
            for (int pp = 0; pp < kPairs; ++pp) {
                const int a = k * kPairChunk + 2 * pp;
                if (a >= nchans) break;
                ++npairs;
                const bool hasb = a + 1 < nchans;
                const int basea = d0l[(size_t)a] & ~3;
                int ea[NDM], v[NDM], vmin = 0, vmax = 0;
                for (int w = 0; w < NDM; ++w) {
                    ea[w] = D(w, a) - basea;
                    if (ea[w] < 0 || ea[w] > max_off) return -1;
                    e_max = std::max(e_max, ea[w]);
                    v[w] = hasb ? D(w, a + 1) - D(w, a) : 0;
                    if (v[w] < 0) return -1;   // relative delay of a pair must not be negative
                    vmin = w ? std::min(vmin, v[w]) : v[w];
                    vmax = w ? std::max(vmax, v[w]) : v[w];
                }
                const int nv = vmax - vmin + 1;
                const bool three_rows = nv == 3 && spare > 0;
                if (three_rows) --spare;
                if (nv <= 2 || three_rows) {
                    if (R) {
                        for (int i = 0; i < nv; ++i)
                            R->row[nrows + i] = make_row(a, basea, hasb ? a + 1 : -1, vmin + i);
                        for (int w = 0; w < NDM; ++w)
                            R->ent[w][pp] = (uint16_t)((nrows + v[w] - vmin) * kPRowDw + ea[w]);
                    }
                    nrows += nv;
                } else {   // split: two single-channel rows
                    const int baseb = d0l[(size_t)a + 1] & ~3;
                    for (int w = 0; w < NDM; ++w) {
                        split_e[nsplit][w] = D(w, a + 1) - baseb;
                        if (split_e[nsplit][w] < 0 || split_e[nsplit][w] > max_off) return -1;
                        e_max = std::max(e_max, split_e[nsplit][w]);
                    }
                    if (R) {
                        R->row[nrows] = make_row(a, basea, -1, 0);
                        R->row[nrows + 1] = make_row(a + 1, baseb, -1, 0);
                        for (int w = 0; w < NDM; ++w) R->ent[w][pp] = (uint16_t)(nrows * kPRowDw + ea[w]);
                    }
                    split_row[nsplit++] = nrows + 1;
                    nrows += 2;
                }
            }
            if (R) {
                for (int s = 0; s < nsplit; ++s)
                    for (int w = 0; w < NDM; ++w)
                        R->ent[w][npairs + s] = (uint16_t)(split_row[s] * kPRowDw + split_e[s][w]);
                R->nrows = nrows;
                R->nent = npairs + nsplit;
                // Each DM's entries in order of residue (offset & 3): the sum
                // does not depend on their order, and the 16-byte-read kernel
                // adds two entries with one v_add3 per dword where their
                // residues agree.
                for (int w = 0; w < NDM; ++w)
                    std::stable_sort(R->ent[w], R->ent[w] + R->nent,
                                     [](uint16_t x, uint16_t y) { return (x & 3) < (y & 3); });
            }
        }
    }
    return (e_max + tile / 2 + 3) / 4;
}

// Quad shape plan (shape 3): as pair_plan, with 4-channel groups; a group
// whose relative delays take more than four combinations in a tile is built
// as two pairs (each as in the pair shape: one row per value of its relative
// delay when there are at most two, else two single-channel rows).  Entry
// slots: one per group in order, the further entries of split groups after
// them.
static int quad_plan(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, size_t pitch,
                     std::vector<quad_rec>* recs)
{
    if (nchans <= 0 || ndms <= 0) return -1;
    const int nchunks = (nchans + kPairChunk - 1) / kPairChunk;
    const int ndt = (ndms + kDms - 1) / kDms;
    if (recs) recs->assign((size_t)ndt * nchunks, quad_rec{});
    std::vector<int> dl((size_t)kDms * nchans), d0l((size_t)nchans);
    int e_max = 0;
    bool fits = true;
    auto chan_off = [pitch](int c, int start) { return (long long)c * (long long)pitch + (start & ~3); };
    for (int dt = 0; dt < ndt && fits; ++dt) {
        const int d0 = dt * kDms;
        const float s0 = __builtin_fmaf((float)d0, mdmstep, mstartdm);
        for (int c = 0; c < nchans; ++c) d0l[(size_t)c] = (int)(h_dmshifts[c] * s0);
        for (int w = 0; w < kDms; ++w) {
            const int d = std::min(d0 + w, ndms - 1);
            const float sd = __builtin_fmaf((float)d, mdmstep, mstartdm);
            for (int c = 0; c < nchans; ++c) dl[(size_t)w * nchans + c] = (int)(h_dmshifts[c] * sd);
        }
        auto D = [&](int w, int c) { return dl[(size_t)w * nchans + c]; };
        for (int k = 0; k < nchunks && fits; ++k) {
            quad_rec* R = recs ? &(*recs)[(size_t)dt * nchunks + k] : nullptr;
            int nrows = 0, ngroups = 0, nextra = 0;
            int extra_ent[kPMaxEnt][kDms];   // further entries of split groups, in order
            // One row: channels ch[0..nc) at windows base + rel[m] (rel[0] = 0).
            auto add_row = [&](const int* ch, int nc, int base, const int* rel) {
                if (R) {
                    quad_row q{};
                    q.nch = (uint32_t)nc;
                    for (int m = 0; m < 4; ++m) {
                        const int c = m < nc ? ch[m] : ch[0];
                        const int start = m < nc ? base + rel[m] : base;
                        q.off[m] = chan_off(c, start);
                        q.sh |= (uint32_t)(start & 3) << (8 * m);
                    }
                    R->row[nrows] = q;
                }
                return nrows++;
            };
            auto offset_ok = [&](int e) {
                if (e < 0 || e > kQMaxOffset) return false;
                e_max = std::max(e_max, e);
                return true;
            };
            // entries of a two-channel piece (first channel a, second b or -1):
            // rows as in the pair shape; ent_out[w][0..n) the DM's entries.
            auto add_pair = [&](int a, int b, int (*ent_out)[2], int* n_out) {
                const int basea = d0l[(size_t)a] & ~3;
                int v[kDms], vmin = 0, vmax = 0;
                for (int w = 0; w < kDms; ++w) {
                    if (!offset_ok(D(w, a) - basea)) return false;
                    v[w] = b >= 0 ? D(w, b) - D(w, a) : 0;
                    if (v[w] < 0) return false;
                    vmin = w ? std::min(vmin, v[w]) : v[w];
                    vmax = w ? std::max(vmax, v[w]) : v[w];
                }
                if (b < 0 || vmax - vmin <= 1) {
                    int first = -1;
                    for (int i = 0; i <= vmax - vmin; ++i) {
                        const int ch[2] = {a, b};
                        const int rel[2] = {0, vmin + i};
                        const int r = add_row(ch, b >= 0 ? 2 : 1, basea, rel);
                        if (first < 0) first = r;
                    }
                    for (int w = 0; w < kDms; ++w) ent_out[w][0] = (first + v[w] - vmin) * kQRowDw + D(w, a) - basea;
                    *n_out = 1;
                    return true;
                }
                const int baseb = d0l[(size_t)b] & ~3;
                for (int w = 0; w < kDms; ++w)
                    if (!offset_ok(D(w, b) - baseb)) return false;
                const int rel0[1] = {0};
                const int ra = add_row(&a, 1, basea, rel0);
                const int rb = add_row(&b, 1, baseb, rel0);
                for (int w = 0; w < kDms; ++w) {
                    ent_out[w][0] = ra * kQRowDw + D(w, a) - basea;
                    ent_out[w][1] = rb * kQRowDw + D(w, b) - baseb;
                }
                *n_out = 2;
                return true;
            };
            for (int g = 0; g < kQGroups && fits; ++g) {
                const int a = k * kPairChunk + 4 * g;
                if (a >= nchans) break;
                ++ngroups;
                const int nc = std::min(4, nchans - a);
                const int basea = d0l[(size_t)a] & ~3;
                // combinations of relative delays over the tile's DMs
                int combo[kDms], ncombo = 0, rels[kDms][4];
                for (int w = 0; w < kDms && fits; ++w) {
                    if (!offset_ok(D(w, a) - basea)) fits = false;
                    rels[w][0] = 0;
                    for (int m = 1; m < nc; ++m) {
                        rels[w][m] = D(w, a + m) - D(w, a);
                        if (rels[w][m] < 0) fits = false;   // relative delays must not be negative
                    }
                    combo[w] = -1;
                    for (int u = 0; u < w && combo[w] < 0; ++u) {
                        bool same = true;
                        for (int m = 1; m < nc; ++m) same = same && rels[u][m] == rels[w][m];
                        if (same) combo[w] = combo[u];
                    }
                    if (combo[w] < 0) combo[w] = ncombo++;
                }
                if (!fits) break;
                if (ncombo <= kQMaxGroupRows) {
                    int first = -1;
                    for (int cb = 0; cb < ncombo; ++cb) {
                        int w0 = 0;
                        while (combo[w0] != cb) ++w0;
                        const int ch[4] = {a, a + 1, a + 2, a + 3};
                        const int r = add_row(ch, nc, basea, rels[w0]);
                        if (first < 0) first = r;
                    }
                    if (R)
                        for (int w = 0; w < kDms; ++w)
                            R->ent[w][g] = (uint16_t)((first + combo[w]) * kQRowDw + D(w, a) - basea);
                } else {   // two pairs
                    int e1[kDms][2], e2[kDms][2], n1 = 0, n2 = 0;
                    if (!add_pair(a, nc > 1 ? a + 1 : -1, e1, &n1)) { fits = false; break; }
                    if (nc > 2 && !add_pair(a + 2, nc > 3 ? a + 3 : -1, e2, &n2)) { fits = false; break; }
                    for (int w = 0; w < kDms; ++w) {
                        if (R) R->ent[w][g] = (uint16_t)e1[w][0];
                        int x = nextra;
                        for (int i = 1; i < n1; ++i) extra_ent[x++][w] = e1[w][i];
                        for (int i = 0; i < n2; ++i) extra_ent[x++][w] = e2[w][i];
                    }
                    nextra += (n1 - 1) + n2;
                }
            }
            if (!fits) break;
            if (R) {
                for (int x = 0; x < nextra; ++x)
                    for (int w = 0; w < kDms; ++w) R->ent[w][ngroups + x] = (uint16_t)extra_ent[x][w];
                R->nrows = nrows;
                R->nent = ngroups + nextra;
            }
        }
    }
    if (!fits) return -1;
    return (e_max + kQSplit + 3) / 4;
}

int lds_plan_tables(int variant, const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep,
                    size_t pitch, std::vector<uint8_t>* bytes)
{
    if (variant == 3) {
        std::vector<quad_rec> recs;
        const int q = quad_plan(h_dmshifts, nchans, ndms, mstartdm, mdmstep, pitch, bytes ? &recs : nullptr);
        if (q > 0 && bytes) {
            bytes->resize(recs.size() * sizeof(quad_rec));
            std::memcpy(bytes->data(), recs.data(), bytes->size());
        }
        return q;
    }
    if (variant == 2 || variant == 4 || variant == 5) {
        std::vector<pair_rec> recs;
        const int q = pair_plan<kDms>(h_dmshifts, nchans, ndms, mstartdm, mdmstep, pitch,
                                      variant == 4 ? kWT4 : variant == 5 ? kWT3 : kWT2, bytes ? &recs : nullptr);
        if (q > 0 && bytes) {
            bytes->resize(recs.size() * sizeof(pair_rec));
            std::memcpy(bytes->data(), recs.data(), bytes->size());
        }
        return q;
    }
    if (variant == 6) {
        std::vector<pair_rec2> recs;
        const int q = pair_plan<kDms2>(h_dmshifts, nchans, ndms, mstartdm, mdmstep, pitch, kWT2,
                                       bytes ? &recs : nullptr);
        if (q > 0 && bytes) {
            bytes->resize(recs.size() * sizeof(pair_rec2));
            std::memcpy(bytes->data(), recs.data(), bytes->size());
        }
        return q;
    }
    return -1;
}

int lds_pair_tables(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, size_t pitch,
                    std::vector<uint8_t>* bytes)
{
    std::vector<pair_rec> recs;
    const int q = pair_plan<kDms>(h_dmshifts, nchans, ndms, mstartdm, mdmstep, pitch, kWT2, bytes ? &recs : nullptr);
    if (q > 0 && bytes) {
        bytes->resize(recs.size() * sizeof(pair_rec));
        std::memcpy(bytes->data(), recs.data(), bytes->size());
    }
    return q;
}

long lds_tile_samples(int variant)
{
    return variant == 2 || variant == 6 ? kWT2 : variant == 3 ? kQT : variant == 4 ? kWT4 : variant == 5 ? kWT3 : kT;
}
int lds_tile_dms(int variant) { return variant == 6 ? kDms2 : kDms; }

hipError_t launch_dedisperse_lds(const lds_params& p, int variant, hipStream_t s)
{
    const int cap = variant == 0   ? shape_c32::kRowQuads
                    : variant == 1 ? shape_c16::kRowQuads
                    : variant == 3 ? kQRQ
                                   : kPRQ;
    if (variant < 0 || variant >= kLdsVariants || p.row_quads <= 0 || p.row_quads > cap) return hipErrorInvalidValue;
    if (variant >= 2 && !p.pair_recs) return hipErrorInvalidValue;
    if (p.ndms <= 0 || p.t_processed <= 0) return hipSuccess;
    lds_params q = p;
    const long tile = lds_tile_samples(variant);
    q.n_dm_tiles = (p.ndms + lds_tile_dms(variant) - 1) / lds_tile_dms(variant);
    q.n_t_tiles = (int)((p.t_processed + tile - 1) / tile);
    const long ntiles = (long)q.n_dm_tiles * q.n_t_tiles;
    const long grid = 8 * ((ntiles + 7) / 8);
    if (grid > 0x7fffffffL) return hipErrorInvalidValue;
    if (variant == 0)
        hipLaunchKernelGGL(dedisperse_lds_kernel<shape_c32>, dim3((unsigned)grid), dim3(kThreads), 0, s, q);
    else if (variant == 1)
        hipLaunchKernelGGL(dedisperse_lds_kernel<shape_c16>, dim3((unsigned)grid), dim3(kThreads), 0, s, q);
    else if (variant == 3) {
        if ((p.row_quads + 63) / 64 <= 2)
            hipLaunchKernelGGL(dedisperse_quad_kernel<2>, dim3((unsigned)grid), dim3(kThreads), 0, s, q);
        else
            hipLaunchKernelGGL(dedisperse_quad_kernel<3>, dim3((unsigned)grid), dim3(kThreads), 0, s, q);
    } else {
        // quads per lane: the row width rounded up to whole 64-quad items
        // (a row is at least half a tile wide)
        const int ni = (p.row_quads + 63) / 64;
        if (variant == 6) return launch_dedisperse_wide2(q, ni, (unsigned)grid, s);
        return launch_dedisperse_wide(q, variant == 4 ? 4 : variant == 5 ? 3 : 2, ni, (unsigned)grid, s);
    }
    return hipGetLastError();
}

void lds_job_tiles(lds_params& p, int variant)
{
    const long tile = lds_tile_samples(variant);
    p.n_dm_tiles = (p.ndms + lds_tile_dms(variant) - 1) / lds_tile_dms(variant);
    p.n_t_tiles = (int)((p.t_processed + tile - 1) / tile);
}

hipError_t launch_dedisperse_lds_batch(const lds_job* d_jobs, int njobs, const long* d_bounds, int* d_counters,
                                       int grid, hipStream_t s)
{
    if (njobs <= 0) return hipSuccess;
    if (grid <= 0 || grid % 8 != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, d_counters);
    hipLaunchKernelGGL(dedisperse_lds_batch_kernel, dim3((unsigned)grid), dim3(kThreads), 0, s, d_jobs, njobs, d_bounds,
                       d_counters);
    return hipGetLastError();
}

}  // namespace aa_hip
