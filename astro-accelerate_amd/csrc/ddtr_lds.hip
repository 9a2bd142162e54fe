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

#include <algorithm>
#include <vector>

namespace aa_hip {

namespace {

constexpr int kR = 32;              // samples per lane
constexpr int kT = 64 * kR;         // samples per tile
constexpr int kDms = 16;            // DM trials per work-group (= waves)
constexpr int kThreads = 64 * kDms; // 1024

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
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

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
        const long tl = t.t0 + (long)kR * t.lane;
        float* orow = p.out + (size_t)d * p.out_pitch;
#pragma unroll
        for (int j = 0; j < kR; ++j) {
            const long tt = tl + j;
            if (tt < p.t_processed) orow[tt] = __fdiv_rn(__fdiv_rn((float)acc32[j], p.fnchans), p.fbin);
        }
    }
}

// ---------------------------------------------------------------------------
// Pair pre-summation (chunk shape 2).  For a 16-DM tile the relative delay
// v = delay(c+1, d) - delay(c, d) of two adjacent channels takes only a few
// values (about two: the truncations of nearly equal delays differ by k or
// k+1).  Per chunk of 16 channels (8 pairs) the work-group builds, for each
// pair and each of its values v, one row of 16-bit sums
//     S_v[j] = L[c][base + j] + L[c+1][base + v + j],  j < 2432,
// base = 8-sample floor of the pair's smallest delay in the tile.  A DM
// trial then adds the row of its own v at offset delay(c, d) - base: two
// channels per 16-bit lane, exactly the sum of the single-channel kernel.
// Rows are built from global memory (L2) through registers: the loads of
// chunk k+1 are issued before chunk k is summed, the sums and LDS stores
// after; one barrier per chunk.  A row is 304 float4 quads (8 samples each)
// in four phase planes (quad q at plane q % 4, slot q / 4), so the five
// ds_read_b128 of a lane run (quads 4*lane + e/8 + m) are stride-1 across
// the wave.  16 rows per buffer; the host admits a range only if no chunk of
// any tile needs more rows or a longer window (pair_plan_fits).
// ---------------------------------------------------------------------------
constexpr int kPairs = 8;                    // pairs per chunk
constexpr int kPairChunk = 2 * kPairs;       // 16 channels
constexpr int kPRQ = 2 * kLdsRowQuadsMax;    // 304 quads per row (2432 samples)
constexpr int kPPlane = kPRQ / 4;            // 76 slots per phase plane
constexpr int kPRows = kBufQuads / kPRQ;     // 16 rows per buffer
constexpr int kPFlushChunks = 256 / kPairChunk;   // 128 pair sums (<= 510 each) per 16-bit field
// lane 63 reads quads 252 + e/8 .. +4 of a row
constexpr int kPairMaxOffset = 8 * (kPRQ - 257) + 7;
static_assert(kPRows * kPRQ == kBufQuads, "pair rows fill the LDS buffers");
static_assert(kPRows == kDms, "one row per wave is built");

// Per-chunk tables, three chunks in flight (summed, being built, next).
// A pair whose relative delay takes more than two values in the tile is
// split into two single-channel rows (each DM then adds both), so a chunk
// never needs more than 16 rows.
struct pair_tables {
    // per (DM slot, pair): x = row1 | row2 << 8 (0xff = none), y = e1 | e2 << 16
    int2 ent[3][kDms * kPairs];
    int2 row[3][kPRows];          // per row: (channel in chunk, v; -1 = single channel)
    int rbase[3][kPRows];         // per row: first sample of its window
    int nrows[3];
};
__shared__ pair_tables g_ptab;

__device__ __forceinline__ pair_tables& ptab() { return g_ptab; }

// Tables of chunk `chunk` into slot `sl`; run by wave 0 (lanes: DM slot
// lane % 16, pairs lane / 16 and lane / 16 + 4).
__device__ __forceinline__ void pair_table(const lds_params& p, int d0, int chunk, int sl, int lane)
{
    pair_tables& T = ptab();
    const int c0 = chunk * kPairChunk;
    const int w = lane & 15, g = lane >> 4;
    const int d = min(d0 + w, p.ndms - 1);
    int v[2], ea[2], eb[2], vmin[2], nv[2], nr[2], basea[2], baseb[2];
    bool have[2], hasb[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int a = c0 + 2 * (g + 4 * h);
        v[h] = ea[h] = eb[h] = basea[h] = baseb[h] = 0;
        have[h] = a < p.nchans;
        hasb[h] = a + 1 < p.nchans;
        if (have[h]) {
            const float dsa = p.dmshifts[a];
            const int da = dev_delay(dsa, d, p.mstartdm, p.mdmstep);
            basea[h] = dev_delay(dsa, d0, p.mstartdm, p.mdmstep) & ~7;
            ea[h] = da - basea[h];
            if (hasb[h]) {
                const float dsb = p.dmshifts[a + 1];
                const int db = dev_delay(dsb, d, p.mstartdm, p.mdmstep);
                baseb[h] = dev_delay(dsb, d0, p.mstartdm, p.mdmstep) & ~7;
                eb[h] = db - baseb[h];
                v[h] = db - da;
            }
        }
        int lo = v[h], hi = v[h];
#pragma unroll
        for (int m = 8; m > 0; m >>= 1) {
            lo = min(lo, __shfl_xor(lo, m, 16));
            hi = max(hi, __shfl_xor(hi, m, 16));
        }
        vmin[h] = lo;
        nv[h] = hi - lo + 1;
        nr[h] = !have[h] ? 0 : (nv[h] <= 2 ? nv[h] : 2);   // split pair: two single rows
    }
    // first row of each of the 8 pairs (pair g + 4h lives in lanes 16g..16g+15)
    int first[2];
    int total = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int f = 0;
#pragma unroll
        for (int gg = 0; gg < 4; ++gg) {
            if (gg == g) f = total;
            total += __builtin_amdgcn_readlane(nr[h], 16 * gg);
        }
        first[h] = f;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int pp = g + 4 * h;
        int r1 = 0xff, r2 = 0xff, e1 = ea[h], e2 = 0;
        if (nr[h] && nv[h] <= 2) {
            r1 = first[h] + (v[h] - vmin[h]);
            if (w < nv[h]) {
                T.row[sl][first[h] + w] = make_int2(2 * pp, vmin[h] + w);
                T.rbase[sl][first[h] + w] = basea[h];
            }
        } else if (nr[h]) {
            r1 = first[h];
            r2 = first[h] + 1;
            e2 = eb[h];
            if (w < 2) {
                T.row[sl][first[h] + w] = make_int2(2 * pp + w, -1);
                T.rbase[sl][first[h] + w] = w ? baseb[h] : basea[h];
            }
        }
        T.ent[sl][w * kPairs + pp] = make_int2(r1 | (r2 << 8), (e1 & 0xffff) | (e2 << 16));
    }
    if (lane == 0) T.nrows[sl] = total;
}

// Wave w builds row w of a chunk (when the chunk has that many rows).  Lane
// l writes whole quads (8 samples) q = 4 (l % 16 + 16 i) + l / 16: plane
// l / 16, slot l % 16 + 16 i, so 16 consecutive lanes write 16 consecutive
// slots of one plane (stride-1 ds_write_b128, no bank conflicts), and the
// wave's loads cover 512 contiguous bytes per i.  The row's channel,
// relative delay and window start are wave-uniform.  Items are loaded and
// stored in kPParts parts, each overlapped with a part of the summation, to
// keep the prefetch registers few.
constexpr int kPIt = (kPRQ + 63) / 64;         // quads per lane per row (5)
constexpr int kPParts = 2;
constexpr int kPHalf = (kPIt + kPParts - 1) / kPParts;
struct pair_build {
    uint32_t xa[kPHalf][2], xb[kPHalf][3];
};

__host__ __device__ constexpr int pair_part_items(int h)
{
    return kPIt - h * kPHalf < kPHalf ? kPIt - h * kPHalf : kPHalf;
}
__device__ __forceinline__ int pair_quad(int lane, int i) { return 4 * ((lane & 15) + 16 * i) + (lane >> 4); }

// Global loads of part H of row `wave` of chunk `chunk` (table slot sl).
template <int H>
__device__ __forceinline__ void pair_build_load(const lds_params& p, long t0, int chunk, int sl, int wave, int lane,
                                                pair_build& B)
{
    const pair_tables& T = ptab();
    const int nrows = __builtin_amdgcn_readfirstlane(T.nrows[sl]);
    if (wave >= nrows) return;
    const int2 rv = T.row[sl][wave];
    const int ch = __builtin_amdgcn_readfirstlane(rv.x);
    const int v = __builtin_amdgcn_readfirstlane(rv.y);
    const int base = __builtin_amdgcn_readfirstlane(T.rbase[sl][wave]);
    const int c0 = chunk * kPairChunk;
    const uint8_t* ra = p.L + (size_t)(c0 + ch) * p.pitch + t0 + base;
    const bool hasb = v >= 0 && c0 + ch + 1 < p.nchans;
    const uint8_t* rb = p.L + (size_t)(c0 + ch + 1) * p.pitch + t0 + ((base + (hasb ? v : 0)) & ~3);
#pragma unroll
    for (int i = 0; i < pair_part_items(H); ++i) {
        const int q = pair_quad(lane, H * kPHalf + i);
        uint32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0, b2 = 0;
        if (q < kPRQ) {
            const uint2 av = *reinterpret_cast<const uint2*>(ra + 8 * q);
            a0 = av.x;
            a1 = av.y;
            if (hasb) {
                const uint32_t* bp = reinterpret_cast<const uint32_t*>(rb + 8 * q);
                b0 = bp[0];
                b1 = bp[1];
                b2 = bp[2];
            }
        }
        B.xa[i][0] = a0;
        B.xa[i][1] = a1;
        B.xb[i][0] = b0;
        B.xb[i][1] = b1;
        B.xb[i][2] = b2;
    }
}

// Pair sums of part H of row `wave`, stored into `rows`.
template <int H>
__device__ __forceinline__ void pair_build_store(int sl, int wave, int lane, const pair_build& B, uint4* rows)
{
    const pair_tables& T = ptab();
    const int nrows = __builtin_amdgcn_readfirstlane(T.nrows[sl]);
    if (wave >= nrows) return;
    const uint32_t sh = (uint32_t)(__builtin_amdgcn_readfirstlane(T.row[sl][wave].y) & 3);  // -1 (single): xb = 0
    uint4* r = rows + wave * kPRQ + (lane >> 4) * kPPlane + (lane & 15);
#pragma unroll
    for (int i = 0; i < pair_part_items(H); ++i) {
        const int ii = H * kPHalf + i;
        if (pair_quad(lane, ii) < kPRQ) {
            const uint32_t x0 = __builtin_amdgcn_alignbyte(B.xb[i][1], B.xb[i][0], sh);
            const uint32_t x1 = __builtin_amdgcn_alignbyte(B.xb[i][2], B.xb[i][1], sh);
            uint4 o;
            o.x = __builtin_amdgcn_perm(0u, B.xa[i][0], 0x0c010c00u) + __builtin_amdgcn_perm(0u, x0, 0x0c010c00u);
            o.y = __builtin_amdgcn_perm(0u, B.xa[i][0], 0x0c030c02u) + __builtin_amdgcn_perm(0u, x0, 0x0c030c02u);
            o.z = __builtin_amdgcn_perm(0u, B.xa[i][1], 0x0c010c00u) + __builtin_amdgcn_perm(0u, x1, 0x0c010c00u);
            o.w = __builtin_amdgcn_perm(0u, B.xa[i][1], 0x0c030c02u) + __builtin_amdgcn_perm(0u, x1, 0x0c030c02u);
            r[16 * ii] = o;   // slot (lane % 16) + 16 ii of plane lane / 16
        }
    }
}

// acc16[j] += the 16 dwords (32 samples) of a lane run whose five quads
// are q[m], SD dwords and H half-dwords into the first quad.
template <int SD, int H>
__device__ __forceinline__ void pair_accumulate(const uint4* const (&q)[5], uint32_t (&acc)[kR / 2])
{
    constexpr int nq = (SD + 16 + H + 3) / 4;
    uint32_t w[20];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        u32x4 t = {0u, 0u, 0u, 0u};
        if (m < nq) t = *reinterpret_cast<const u32x4*>(q[m]);
        w[4 * m] = t.x;
        w[4 * m + 1] = t.y;
        w[4 * m + 2] = t.z;
        w[4 * m + 3] = t.w;
    }
#pragma unroll
    for (int j = 0; j < kR / 2; ++j) {
        const uint32_t x = H ? __builtin_amdgcn_alignbyte(w[SD + j + 1], w[SD + j], 2u) : w[SD + j];
        asm("v_add_u32 %0, %0, %1" : "+v"(acc[j]) : "v"(x));
    }
}

// All rows of the wave's entries (lane l < 16: entry l) whose sub-quad
// offset is SUB: a plain loop per offset class, so the accumulators stay in
// their registers (no switch joins).
template <int SUB>
__device__ __forceinline__ void pair_sum_class(uint64_t m, int row_v, int e_v, int lane, const uint4* rows,
                                               uint32_t (&acc)[kR / 2])
{
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        const int row = __builtin_amdgcn_readlane(row_v, l);
        const int e8 = __builtin_amdgcn_readlane(e_v, l) >> 3;
        // quad 4*lane + e8 + k of the row: plane (e8 + k) & 3, slot lane + (e8 + k) / 4
        const uint4* rb = rows + row * kPRQ + lane;
        const uint4* q[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) q[k] = rb + ((e8 + k) & 3) * kPPlane + ((e8 + k) >> 2);
        pair_accumulate<(SUB >> 1), (SUB & 1)>(q, acc);
    }
}

template <int SUB>
__device__ __forceinline__ void pair_sum_classes(bool valid, int sub, int row_v, int e_v, int lane,
                                                 const uint4* rows, uint32_t (&acc)[kR / 2])
{
    pair_sum_class<SUB>(__builtin_amdgcn_ballot_w64(valid && sub == SUB), row_v, e_v, lane, rows, acc);
    if constexpr (SUB + 1 < 8) pair_sum_classes<SUB + 1>(valid, sub, row_v, e_v, lane, rows, acc);
}

// Pairs [P0, P0 + kPairs / kPParts) of chunk slot sl: lanes 0..3 (or the
// part's pairs) hold the first row of each pair, lanes 8.. the second row
// of split pairs.
template <int P0>
__device__ __forceinline__ void pair_sum_half(int sl, int wave, int lane, const uint4* rows, uint32_t (&acc)[kR / 2])
{
    constexpr int NP = kPairs / kPParts;
    const pair_tables& T = ptab();
    const int pp = P0 + (lane & (NP - 1));
    const int2 en = T.ent[sl][wave * kPairs + pp];
    const bool second = (lane & 8) != 0;
    const int row = second ? ((en.x >> 8) & 0xff) : (en.x & 0xff);
    // host-checked to lie in [0, kPairMaxOffset]; clamped so that a row read
    // can never leave the buffer
    const int e = min(max(second ? ((en.y >> 16) & 0xffff) : (en.y & 0xffff), 0), kPairMaxOffset);
    const bool valid = (lane & 7) < NP && lane < 16 && row < kPRows;
    pair_sum_classes<0>(valid, e & 7, row, e, lane, rows, acc);
}

// Part H: load part H of chunk k+1's build, sum pairs of part H of chunk k,
// store the build part.
template <int H>
__device__ __forceinline__ void pair_part(const lds_params& p, long t0, int k, bool more, bool active, int sn,
                                          int wave, int lane, const uint4* rows_cur, uint4* rows_next,
                                          uint32_t (&acc16)[kR / 2])
{
    pair_build B;
    if (more) pair_build_load<H>(p, t0, k + 1, sn, wave, lane, B);
    if (active) pair_sum_half<H * (kPairs / kPParts)>(k % 3, wave, lane, rows_cur, acc16);
    if (more) pair_build_store<H>(sn, wave, lane, B, rows_next);
}

__device__ __forceinline__ void pair_step(const lds_params& p, long t0, int d0, int k, int nchunks, bool active,
                                          int wave, int lane, int tid, const uint4* rows_cur, uint4* rows_next,
                                          uint32_t (&acc16)[kR / 2], uint32_t (&acc32)[kR])
{
    const bool more = k + 1 < nchunks;
    const int sn = (k + 1) % 3;
    pair_part<0>(p, t0, k, more, active, sn, wave, lane, rows_cur, rows_next, acc16);
    pair_part<1>(p, t0, k, more, active, sn, wave, lane, rows_cur, rows_next, acc16);
    if (active && ((k + 1) % kPFlushChunks == 0 || !more)) flush(acc16, acc32);
    if (wave == 0 && k + 2 < nchunks) pair_table(p, d0, k + 2, (k + 2) % 3, lane);
    __syncthreads();
}

__device__ __forceinline__ void process_pair_tile(const lds_params& p, int d0, long t0)
{
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool active = d0 + wave < p.ndms;
    const int nchunks = (p.nchans + kPairChunk - 1) / kPairChunk;
    uint32_t acc16[kR / 2];
    uint32_t acc32[kR];
#pragma unroll
    for (int j = 0; j < kR / 2; ++j) acc16[j] = 0;
#pragma unroll
    for (int j = 0; j < kR; ++j) acc32[j] = 0;

    if (wave == 0) {
        pair_table(p, d0, 0, 0, lane);
        if (nchunks > 1) pair_table(p, d0, 1, 1, lane);
    }
    __syncthreads();
    {
        pair_build B;
        pair_build_load<0>(p, t0, 0, 0, wave, lane, B);
        pair_build_store<0>(0, wave, lane, B, g_rows0);
        pair_build_load<1>(p, t0, 0, 0, wave, lane, B);
        pair_build_store<1>(0, wave, lane, B, g_rows0);
    }
    __syncthreads();
    for (int k = 0; k < nchunks; k += 2) {
        pair_step(p, t0, d0, k, nchunks, active, wave, lane, tid, g_rows0, g_rows1, acc16, acc32);
        if (k + 1 >= nchunks) break;
        pair_step(p, t0, d0, k + 1, nchunks, active, wave, lane, tid, g_rows1, g_rows0, acc16, acc32);
    }

    if (active) {
        const int d = d0 + wave;
        const long tl = t0 + (long)kR * lane;
        float* orow = p.out + (size_t)d * p.out_pitch;
#pragma unroll
        for (int j = 0; j < kR; ++j) {
            const long tt = tl + j;
            if (tt < p.t_processed) orow[tt] = __fdiv_rn(__fdiv_rn((float)acc32[j], p.fnchans), p.fbin);
        }
    }
}

__global__ __launch_bounds__(kThreads) void dedisperse_pair_kernel(lds_params p)
{
    const int ntiles = p.n_dm_tiles * p.n_t_tiles;
    const int per_xcd = (ntiles + 7) / 8;
    const int w = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (w >= ntiles) return;  // whole work-group leaves before any barrier
    process_pair_tile(p, (w % p.n_dm_tiles) * kDms, (long)(w / p.n_dm_tiles) * kT);
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

// Many ranges (of many beams) in one persistent launch.  The tiles of all
// jobs form one list, split into 8 contiguous queues of about equal work, one
// per XCD; a work-group takes tiles from its own XCD's queue (neighbouring
// tiles share input rows in that XCD's L2) and, once it is empty, from the
// others in turn.  Every work-group leaves when all 8 queues are empty.
__global__ __launch_bounds__(kThreads) void dedisperse_lds_batch_kernel(const lds_job* __restrict__ jobs, int njobs,
                                                                        const long* __restrict__ bounds,
                                                                        int* __restrict__ counters)
{
    __shared__ long s_item;
    const int xcd = (int)(blockIdx.x % 8);
    int q = 0;  // queues found empty so far (thread 0)
    for (;;) {
        if (threadIdx.x == 0) {
            long item = -1;
            while (q < 8) {
                const int x = (xcd + q) & 7;
                const long i = bounds[x] + (long)atomicAdd(&counters[x], 1);
                if (i < bounds[x + 1]) {
                    item = i;
                    break;
                }
                ++q;
            }
            s_item = item;
        }
        __syncthreads();
        const long item = s_item;
        __syncthreads();
        if (item < 0) break;
        int lo = 0, hi = njobs - 1;  // last job with tile_begin <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].tile_begin <= item) lo = mid;
            else hi = mid - 1;
        }
        const lds_job& J = jobs[lo];
        const long local = item - J.tile_begin;
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

static bool pair_plan_fits(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep);

int lds_row_quads(int variant, const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep)
{
    if (variant < 0 || variant >= kLdsVariants) return -1;
    if (variant == 2) return pair_plan_fits(h_dmshifts, nchans, ndms, mstartdm, mdmstep) ? kPRQ : -1;
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

// Pair shape: every window offset of every tile fits a row (the same
// arithmetic as pair_table; rows per chunk never exceed 16 by construction).
static bool pair_plan_fits(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep)
{
    std::vector<int> dl((size_t)kDms * nchans), d0l((size_t)nchans);
    for (int d0 = 0; d0 < ndms; d0 += kDms) {
        const float s0 = __builtin_fmaf((float)d0, mdmstep, mstartdm);
        for (int c = 0; c < nchans; ++c) d0l[(size_t)c] = (int)(h_dmshifts[c] * s0);
        for (int w = 0; w < kDms; ++w) {
            const int d = std::min(d0 + w, ndms - 1);
            const float sd = __builtin_fmaf((float)d, mdmstep, mstartdm);
            for (int c = 0; c < nchans; ++c) dl[(size_t)w * nchans + c] = (int)(h_dmshifts[c] * sd);
        }
        for (int c = 0; c < nchans; ++c) {
            const int base = d0l[(size_t)c] & ~7;
            for (int w = 0; w < kDms; ++w) {
                const int e = dl[(size_t)w * nchans + c] - base;
                if (e < 0 || e > kPairMaxOffset) return false;
                if ((c & 1) == 0 && c + 1 < nchans && dl[(size_t)w * nchans + c + 1] < dl[(size_t)w * nchans + c])
                    return false;   // relative delay of a pair must not be negative
            }
        }
    }
    return true;
}

long lds_tile_samples() { return kT; }
int lds_tile_dms() { return kDms; }

hipError_t launch_dedisperse_lds(const lds_params& p, int variant, hipStream_t s)
{
    const int cap = variant == 0 ? shape_c32::kRowQuads : variant == 1 ? shape_c16::kRowQuads : kPRQ;
    if (variant < 0 || variant >= kLdsVariants || p.row_quads <= 0 || p.row_quads > cap) return hipErrorInvalidValue;
    if (p.ndms <= 0 || p.t_processed <= 0) return hipSuccess;
    lds_params q = p;
    q.n_dm_tiles = (p.ndms + kDms - 1) / kDms;
    q.n_t_tiles = (int)((p.t_processed + kT - 1) / kT);
    const long ntiles = (long)q.n_dm_tiles * q.n_t_tiles;
    const long grid = 8 * ((ntiles + 7) / 8);
    if (grid > 0x7fffffffL) return hipErrorInvalidValue;
    if (variant == 0)
        hipLaunchKernelGGL(dedisperse_lds_kernel<shape_c32>, dim3((unsigned)grid), dim3(kThreads), 0, s, q);
    else if (variant == 1)
        hipLaunchKernelGGL(dedisperse_lds_kernel<shape_c16>, dim3((unsigned)grid), dim3(kThreads), 0, s, q);
    else
        hipLaunchKernelGGL(dedisperse_pair_kernel, dim3((unsigned)grid), dim3(kThreads), 0, s, q);
    return hipGetLastError();
}

void lds_job_tiles(lds_params& p)
{
    p.n_dm_tiles = (p.ndms + kDms - 1) / kDms;
    p.n_t_tiles = (int)((p.t_processed + kT - 1) / kT);
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
