// Pair pre-summation rows (chunk shapes 2 and 4): geometry, plan records,
// the row build and the accumulator helpers shared by the LDS kernels
// (ddtr_lds.hip: byte-row and quad kernels, host planning) and the
// 16-byte-read pair kernels (ddtr_pair_wide.hip).  Internal header.
#pragma once

#include "ddtr_lds.h"

namespace aa_hip {

namespace {

constexpr int kDms = 16;            // DM trials per work-group (= waves)
constexpr int kThreads = 64 * kDms; // 1024
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Pair pre-summation (chunk shapes 2 and 4).  For a 16-DM tile the relative
// delay v = delay(c+1, d) - delay(c, d) of two adjacent channels takes only a
// few values (about two: the truncations of nearly equal delays differ by k
// or k+1).  Per chunk of 16 channels (8 pairs) the work-group builds, for each
// pair and each of its values v, one row of 16-bit sums
//     S_v[i] = L[c][base + i] + L[c+1][base + v + i],
// base = the pair's delay at the tile's first DM rounded down to 4 samples.
// A DM trial then adds the row of its own v at offset e = delay(c, d) - base:
// two channels per 16-bit field, exactly the sum of the single-channel
// kernel.
//
// Tiles are 16 DMs x T samples; each wave sums one DM.  Transposed rows: a
// row stores dword k = S[k] | S[k + T/2] << 16, so a DM adding its row at
// offset e reads dwords e + k whatever e is, and each 16-bit half of an
// accumulator holds one output sample (t0 + k low, t0 + T/2 + k high).  How
// a lane reads those dwords (16-byte reads, realigned across lanes) is
// described with dedisperse_wide_kernel below.
//
// Which rows a chunk needs and where each DM reads them depend only on the
// plan, not on the data: the host computes them once per range (pair_plan:
// one pair_rec per DM tile and chunk, read by every time tile through L2).
// Rows are built from global memory through registers; one barrier per
// chunk.
// ---------------------------------------------------------------------------
// Tile geometry for T samples: R = T / 64 samples per lane, halves of a
// dword T / 2 samples apart (the quad shape's 1024-sample tiles; the pair
// shapes use SPLIT only).
template <int T>
struct pair_geom {
    static constexpr int R = T / 64;
    static constexpr int SPLIT = T / 2;
};
constexpr int kPairs = 8;                      // pairs per chunk
constexpr int kPairChunk = 2 * kPairs;         // 16 channels
constexpr int kPIt = 5;                        // quads per lane per row
constexpr int kPRQ = 64 * kPIt;                // 320 quads per row
constexpr int kPRowDw = 4 * kPRQ;              // 1280 dwords per row
constexpr int kPRows = 16;                     // rows per buffer (one per wave)
constexpr int kPMaxEnt = 2 * kPairs;           // row entries per DM per chunk
constexpr int kPFlushChunks = 256 / kPairChunk;   // 128 pair sums (<= 510 each) per 16-bit field
// Tiles of the 16-byte-read kernel: 2 x J blocks of 252 dwords, J = 4
// (shape 4, 2016 samples), 3 (shape 5, 1512) or 2 (shape 2, 1008).
constexpr int kWBlk = 252;                     // dwords per block: 63 lanes x 4
constexpr int kWT4 = 2 * 4 * kWBlk;            // 2016 samples per shape-4 tile
constexpr int kWT2 = 2 * 2 * kWBlk;            // 1008 samples per shape-2 tile
constexpr int kWT3 = 2 * 3 * kWBlk;            // 1512 samples per shape-5 tile
// Largest window offset e of a tile of T samples: reads of a DM end at dword
// e + T/2 - 1, or for the 16-byte-read tiles at e + T/2 + 3 (lane 63's last
// quad), of an 1280-dword row.
__host__ __device__ constexpr int pair_max_offset(int T)
{
    return T % (2 * kWBlk) == 0 ? kPRowDw - (T / 2 + 4) : kPRowDw - T / 2;
}
static_assert(kPRows == kDms, "one row per wave is built");

// Plan record of one (DM tile, chunk).  A pair whose relative delay takes
// more than two values in the tile is split into two single-channel rows
// (each DM then adds both), so a chunk never needs more than 16 rows or a DM
// more than 16 entries; a pair of three values is instead built as three
// pair rows while the chunk's 16 rows allow.  Entries 0 .. npairs-1 belong
// to the pairs in order, the second rows of split pairs follow.
struct pair_row {
    long long offa, offb;   // bytes from the level start to the two channels' window starts (offb: dword floor)
    uint32_t sh, mb;        // byte shift of b's window in its dwords; ~0 pair row, 0 single row (b masked off)
    int pad0, pad1;
};
template <int NDM>
struct pair_rec_t {
    uint16_t ent[NDM][kPMaxEnt];   // per DM slot: dword offset (row * kPRowDw + e) of each entry
    pair_row row[kPRows];
    int nrows, nent, pad0, pad1;
};
using pair_rec = pair_rec_t<kDms>;
// Shape 6 tiles hold 32 DM trials (two per wave) over the same 16 rows.
constexpr int kDms2 = 2 * kDms;
using pair_rec2 = pair_rec_t<kDms2>;
static_assert(sizeof(pair_row) == 32 && sizeof(pair_rec) % 16 == 0 && sizeof(pair_rec2) % 16 == 0,
              "records are 16-byte aligned");
static_assert(kPRows * kPRowDw <= 65536, "entry offsets fit 16 bits");

// Two row buffers of 80 KiB (the whole LDS) in one array: buffer b starts
// at quad b * kPRows * kPRQ.
__shared__ uint4 g_prow[2 * kPRows * kPRQ];
__device__ __forceinline__ uint4* pair_buf(int b) { return g_prow + b * (kPRows * kPRQ); }

// Registers of one row in flight: quads q = lane + 64 i, i < NI.
template <int NI>
struct pair_build {
    uint32_t al[NI], ah[NI], bl[NI], bh[NI];
    uint32_t mb;   // ~0: pair row; 0: single-channel row (b's bytes select zero)
};

// Loads of one row (R: its record), issued whatever the row's use (no
// register merges on a branch): a wave without a row in the chunk reads the
// zeroed record's window at the level start.  Channel b's window starts at
// any byte: its dwords are read unaligned (one global_load_dword each; the
// memory path handles the misalignment).  A lane over-reads at most 63
// quads (252 bytes) past the window's needed width, inside the 4096-byte
// slack every row has past its longest window.
template <int NI, int T>
__device__ __forceinline__ void pair_load_item(int i, const uint8_t* ra, const uint8_t* rb, int lane,
                                               pair_build<NI>& B)
{
    const int q = lane + 64 * i;
    const uint32_t* pa = reinterpret_cast<const uint32_t*>(ra + 4 * q);
    B.al[i] = pa[0];
    B.ah[i] = pa[pair_geom<T>::SPLIT / 4];
    __builtin_memcpy(&B.bl[i], rb + 4 * q, 4);
    __builtin_memcpy(&B.bh[i], rb + 4 * q + pair_geom<T>::SPLIT, 4);
}

template <int NI, int T>
__device__ __forceinline__ void pair_build_load(const lds_params& p, long t0, const pair_row& R, int lane,
                                                pair_build<NI>& B)
{
    const uint8_t* ra = p.L + t0 + R.offa;
    const uint8_t* rb = p.L + t0 + R.offb + R.sh;
    B.mb = R.mb;
#pragma unroll
    for (int i = 0; i < NI; ++i) pair_load_item<NI, T>(i, ra, rb, lane, B);
}

// Pair sums of row `wave` into `rows`: dword m of quad q is
// (a[4q+m] + b[4q+m]) | (a[T/2+4q+m] + b[T/2+4q+m]) << 16.  Whole items are
// stored (rows are 320 quads wide).  Byte m of the low word goes to bits
// 0..7, byte m of the high word to bits 16..23; channel b's selectors pick
// zero bytes (0x0c) in a single-channel row.
struct pair_sel {
    uint32_t s[4];
};

__device__ __forceinline__ pair_sel pair_sel_of(uint32_t mb)
{
    const uint32_t z = 0x0c0c0c0cu;
    return pair_sel{{mb ? 0x0c040c00u : z, mb ? 0x0c050c01u : z, mb ? 0x0c060c02u : z, mb ? 0x0c070c03u : z}};
}

template <int NI>
__device__ __forceinline__ void pair_store_item(int i, const pair_build<NI>& B, const pair_sel& S, uint4* r)
{
    uint4 o;
    o.x = __builtin_amdgcn_perm(B.ah[i], B.al[i], 0x0c040c00u) + __builtin_amdgcn_perm(B.bh[i], B.bl[i], S.s[0]);
    o.y = __builtin_amdgcn_perm(B.ah[i], B.al[i], 0x0c050c01u) + __builtin_amdgcn_perm(B.bh[i], B.bl[i], S.s[1]);
    o.z = __builtin_amdgcn_perm(B.ah[i], B.al[i], 0x0c060c02u) + __builtin_amdgcn_perm(B.bh[i], B.bl[i], S.s[2]);
    o.w = __builtin_amdgcn_perm(B.ah[i], B.al[i], 0x0c070c03u) + __builtin_amdgcn_perm(B.bh[i], B.bl[i], S.s[3]);
    r[64 * i] = o;
}

template <int NI>
__device__ __forceinline__ void pair_build_store(int wave, int lane, const pair_build<NI>& B, uint4* rows)
{
    uint4* r = rows + wave * kPRQ + lane;
    const pair_sel S = pair_sel_of(B.mb);
#pragma unroll
    for (int i = 0; i < NI; ++i) pair_store_item<NI>(i, B, S, r);
}

// acc16[j] += row dword k = e + L + 64 j of the entry at dword offset o0.
template <int R>
__device__ __forceinline__ void pair_add1(const uint32_t* base, int o0, uint32_t (&acc)[R / 2])
{
    const uint32_t* r0 = base + o0;
#pragma unroll
    for (int j = 0; j < R / 2; ++j) acc[j] += r0[64 * j];
}

// Four rows' dwords of a lane (entries e01, e23), read ahead of their adds.
template <int R>
struct pair_x4 {
    uint32_t x[4][R / 2];
};

template <int R>
__device__ __forceinline__ void pair_read4(const uint32_t* base, uint32_t e01, uint32_t e23, pair_x4<R>& X)
{
    const uint32_t* r[4] = {base + (e01 & 0xffffu), base + (e01 >> 16), base + (e23 & 0xffffu), base + (e23 >> 16)};
#pragma unroll
    for (int j = 0; j < R / 2; ++j)
#pragma unroll
        for (int m = 0; m < 4; ++m) X.x[m][j] = r[m][64 * j];
}

template <int R>
__device__ __forceinline__ void pair_acc4(const pair_x4<R>& X, uint32_t (&acc)[R / 2])
{
#pragma unroll
    for (int j = 0; j < R / 2; ++j) acc[j] += X.x[0][j] + X.x[1][j];
#pragma unroll
    for (int j = 0; j < R / 2; ++j) acc[j] += X.x[2][j] + X.x[3][j];
}

// Entries of one DM in one chunk (16-bit dword offsets, two per dword),
// held in scalar registers.
struct pair_ents {
    uint32_t e[kPMaxEnt / 2];
    int n;
};

template <int NDM>
__device__ __forceinline__ pair_ents pair_ents_of(const pair_rec_t<NDM>& rc, int slot)
{
    pair_ents E;
    const uint32_t* ew = reinterpret_cast<const uint32_t*>(&rc.ent[slot][0]);
#pragma unroll
    for (int i = 0; i < kPMaxEnt / 2; ++i) E.e[i] = ew[i];
    E.n = rc.nent;
    return E;
}

template <int R>
__device__ __forceinline__ void pair_flush(uint32_t (&acc16)[R / 2], uint32_t (&acc32)[R])
{
#pragma unroll
    for (int j = 0; j < R / 2; ++j) {
        acc32[j] += acc16[j] & 0xFFFFu;
        acc32[R / 2 + j] += acc16[j] >> 16;
        acc16[j] = 0;
    }
}

// Job and tile of position `pos` of batch queue x (see lds_job): the first
// job whose part in queue x ends after pos; *local = the tile's index in
// that job.
__device__ __forceinline__ const lds_job& batch_job_of(const lds_job* __restrict__ jobs, int njobs, int x, long pos,
                                                       long* local)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (jobs[mid].qend[x] > pos) hi = mid;
        else lo = mid + 1;
    }
    const lds_job& J = jobs[lo];
    const long n = (long)J.p.n_dm_tiles * J.p.n_t_tiles;
    const long b = x * n / 8, e = (x + 1) * n / 8;
    *local = b + pos - (J.qend[x] - (e - b));
    return J;
}

}  // namespace

// Launches the 16-byte-read pair kernel with J blocks per lane (4: shape 4,
// 3: shape 5, 2: shape 2) and NI = ni quads per lane of a row, on `grid`
// work-groups.
hipError_t launch_dedisperse_wide(const lds_params& q, int J, int ni, unsigned grid, hipStream_t s);

// Launches shape 6: 1008-sample tiles of 32 DM trials, two per wave (DM
// slots w and w + 16 of wave w), over the 16 rows a chunk builds; NI = ni
// quads per lane of a row.
hipError_t launch_dedisperse_wide2(const lds_params& q, int ni, unsigned grid, hipStream_t s);

}  // namespace aa_hip
