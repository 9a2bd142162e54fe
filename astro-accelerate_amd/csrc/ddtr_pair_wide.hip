// 16-byte-read pair pre-summation kernels (chunk shapes 4 and 2).  Built
// with uniform branches left unstructurized (see the Makefile): each row
// entry branches on the scalar unit to the add code of its residue.
#include "ddtr_pair_common.h"

namespace aa_hip {

namespace {

// ---------------------------------------------------------------------------
// Pair pre-summation with 16-byte LDS reads (chunk shapes 4 and 2).  Rows
// and plan records are those above; this is how a DM reads a row.
// ds_read_b128 moves about twice the bytes per clock of the dword forms on
// this part, but only from 16-byte-aligned addresses, and a DM's row offset
// e is arbitrary.  So lane L reads the aligned quads at A + 252 j + 4 L
// (A = e rounded down to a multiple of 4 dwords, r = e - A, j < J) and needs
// the dwords e + 252 j + 4 L + q (q < 4): the last 4 - r of its own quad and
// the first r of lane L+1's, which a DPP wave shift (lane L reads lane L+1)
// brings over inside the add.  Lane 63 reads the quad after lane 62's (the
// first of block j+1) and owns no samples, so a tile is 2 x J x 252 samples
// (J = 4: 2016, shape 4; J = 3: 1512, shape 5; J = 2: 1008, shape 2): dword k = 252 j + 4 L + q of
// lane L < 63 holds samples t0 + k (low half) and t0 + 252 J + k (high
// half).  r is the same for the whole wave (one DM per wave): each entry
// branches, on the scalar unit, to the add code of its r.  With J <= 3 the
// full chunks add two entries at a time and keep the neighbour's dwords in
// registers of their own until the flush (wide_acc2n) instead of shifting
// them per entry.
// ---------------------------------------------------------------------------

// Lane L's x of lane L + 1 (lane 63 gets 0): DPP wave_shl:1.
__device__ __forceinline__ uint32_t wide_next(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x130, 0xf, 0xf, true);
}

// The lane's four quads of the row entry at dword offset e (from the lane's
// base: row buffer + 4 * lane): D[4j + m] = dword (e & ~3) + 252 j + 4 L + m.
template <int J>
__device__ __forceinline__ void wide_read(const uint32_t* lane_q, uint32_t e, uint32_t (&D)[4 * J])
{
    const uint32_t* p = lane_q + (e & ~3u);
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(p + kWBlk * j);
        D[4 * j] = v.x; D[4 * j + 1] = v.y; D[4 * j + 2] = v.z; D[4 * j + 3] = v.w;
    }
}

// acc += lane L+1's x, as one v_add_u32 with the DPP shift on its source:
// the empty asm keeps the compiler from merging this add with the previous
// one into a v_add3 (which cannot take a DPP operand).
__device__ __forceinline__ void wide_add_next(uint32_t& acc, uint32_t x)
{
    asm("" : "+v"(acc));
    acc += wide_next(x);
}

// acc[4j + q] += dword e + 252 j + 4 L + q of two entries (RA, RB: their e & 3):
// one v_add3 where both dwords are the lane's own, else one add each.
template <int RA, int RB, int J>
__device__ __forceinline__ void wide_acc2(const uint32_t (&Da)[4 * J], const uint32_t (&Db)[4 * J],
                                          uint32_t (&acc)[4 * J])
{
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int x = 4 * j + q;
            if (q + RA < 4 && q + RB < 4) {
                acc[x] += Da[x + RA] + Db[x + RB];
            } else if (q + RA < 4) {
                acc[x] += Da[x + RA];
                wide_add_next(acc[x], Db[x + RB - 4]);
            } else if (q + RB < 4) {
                acc[x] += Db[x + RB];
                wide_add_next(acc[x], Da[x + RA - 4]);
            } else {
                wide_add_next(acc[x], Da[x + RA - 4]);
                wide_add_next(acc[x], Db[x + RB - 4]);
            }
        }
}

// The lane's own dwords first, then the shifted ones: a DPP add reading an
// accumulator written by the instruction just before it needs wait states.
template <int RA, int J>
__device__ __forceinline__ void wide_acc1(const uint32_t (&Da)[4 * J], uint32_t (&acc)[4 * J])
{
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int q = 0; q < 4 - RA; ++q) acc[4 * j + q] += Da[4 * j + q + RA];
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int q = 4 - RA; q < 4; ++q) wide_add_next(acc[4 * j + q], Da[4 * j + q + RA - 4]);
}

// Two entries packed in one dword (low, high 16 bits).
template <int J>
__device__ __forceinline__ void wide_add2(const uint32_t* lane_q, uint32_t e2, uint32_t (&acc)[4 * J])
{
    uint32_t Da[4 * J], Db[4 * J];
    wide_read<J>(lane_q, e2 & 0xffffu, Da);
    wide_read<J>(lane_q, e2 >> 16, Db);
    switch (((e2 & 3u) << 2) | ((e2 >> 16) & 3u)) {
    case 0: wide_acc2<0, 0, J>(Da, Db, acc); break;
    case 1: wide_acc2<0, 1, J>(Da, Db, acc); break;
    case 2: wide_acc2<0, 2, J>(Da, Db, acc); break;
    case 3: wide_acc2<0, 3, J>(Da, Db, acc); break;
    case 4: wide_acc2<1, 0, J>(Da, Db, acc); break;
    case 5: wide_acc2<1, 1, J>(Da, Db, acc); break;
    case 6: wide_acc2<1, 2, J>(Da, Db, acc); break;
    case 7: wide_acc2<1, 3, J>(Da, Db, acc); break;
    case 8: wide_acc2<2, 0, J>(Da, Db, acc); break;
    case 9: wide_acc2<2, 1, J>(Da, Db, acc); break;
    case 10: wide_acc2<2, 2, J>(Da, Db, acc); break;
    case 11: wide_acc2<2, 3, J>(Da, Db, acc); break;
    case 12: wide_acc2<3, 0, J>(Da, Db, acc); break;
    case 13: wide_acc2<3, 1, J>(Da, Db, acc); break;
    case 14: wide_acc2<3, 2, J>(Da, Db, acc); break;
    default: wide_acc2<3, 3, J>(Da, Db, acc); break;
    }
}

template <int J>
__device__ __forceinline__ void wide_add1(const uint32_t* lane_q, uint32_t e, uint32_t (&acc)[4 * J])
{
    uint32_t Da[4 * J];
    wide_read<J>(lane_q, e, Da);
    switch (e & 3u) {
    case 0: wide_acc1<0, J>(Da, acc); break;
    case 1: wide_acc1<1, J>(Da, acc); break;
    case 2: wide_acc1<2, J>(Da, acc); break;
    default: wide_acc1<3, J>(Da, acc); break;
    }
}

// Two entries at once with every add on the lane's own dwords.  Dword
// q + r - 4 of a lane's quad (q >= 4 - r) belongs to the left neighbour's
// accumulator q; instead of shifting it over with a DPP add per entry, the
// lane adds it into accN[3j + q - 1], which holds, for its left neighbour,
// the not-yet-shifted part of accumulator q, and wide_take_next moves accN
// over once per flush.  Per quad of two entries: one v_add3 per target both
// entries add into, one add per target only one does: 4 + |RA - RB|
// instructions for 8 dword adds (8 with a DPP add per shifted dword).  The
// plan tables sort each DM's entries by residue, so most pairs agree.
template <int RA, int RB, int J>
__device__ __forceinline__ void wide_acc2n(const uint32_t (&Da)[4 * J], const uint32_t (&Db)[4 * J],
                                           uint32_t (&acc)[4 * J], uint32_t (&accN)[3 * J])
{
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int x = 4 * j + q, n = 3 * j + q - 1;
            if (q + RA < 4 && q + RB < 4) {
                acc[x] += Da[x + RA] + Db[x + RB];
            } else if (q + RA < 4) {
                acc[x] += Da[x + RA];
                accN[n] += Db[x + RB - 4];
            } else if (q + RB < 4) {
                acc[x] += Db[x + RB];
                accN[n] += Da[x + RA - 4];
            } else {
                accN[n] += Da[x + RA - 4] + Db[x + RB - 4];
            }
        }
}

// wide_acc2n for run-time residues (wave-uniform): ten bodies, the operands
// swapped when ra > rb.
template <int J>
__device__ __forceinline__ void wide_acc2n_r(const uint32_t (&Da)[4 * J], const uint32_t (&Db)[4 * J], uint32_t ra,
                                             uint32_t rb, uint32_t (&acc)[4 * J], uint32_t (&accN)[3 * J])
{
    switch ((ra << 2) | rb) {
    case 0: wide_acc2n<0, 0, J>(Da, Db, acc, accN); break;
    case 1: wide_acc2n<0, 1, J>(Da, Db, acc, accN); break;
    case 2: wide_acc2n<0, 2, J>(Da, Db, acc, accN); break;
    case 3: wide_acc2n<0, 3, J>(Da, Db, acc, accN); break;
    case 4: wide_acc2n<0, 1, J>(Db, Da, acc, accN); break;
    case 5: wide_acc2n<1, 1, J>(Da, Db, acc, accN); break;
    case 6: wide_acc2n<1, 2, J>(Da, Db, acc, accN); break;
    case 7: wide_acc2n<1, 3, J>(Da, Db, acc, accN); break;
    case 8: wide_acc2n<0, 2, J>(Db, Da, acc, accN); break;
    case 9: wide_acc2n<1, 2, J>(Db, Da, acc, accN); break;
    case 10: wide_acc2n<2, 2, J>(Da, Db, acc, accN); break;
    case 11: wide_acc2n<2, 3, J>(Da, Db, acc, accN); break;
    case 12: wide_acc2n<0, 3, J>(Db, Da, acc, accN); break;
    case 13: wide_acc2n<1, 3, J>(Db, Da, acc, accN); break;
    case 14: wide_acc2n<2, 3, J>(Db, Da, acc, accN); break;
    default: wide_acc2n<3, 3, J>(Da, Db, acc, accN); break;
    }
}

// acc += the entry whose quads are in D (r: its e & 3), shifted dwords by
// DPP adds.
template <int J>
__device__ __forceinline__ void wide_acc_r(const uint32_t (&D)[4 * J], uint32_t r, uint32_t (&acc)[4 * J])
{
    switch (r) {
    case 0: wide_acc1<0, J>(D, acc); break;
    case 1: wide_acc1<1, J>(D, acc); break;
    case 2: wide_acc1<2, J>(D, acc); break;
    default: wide_acc1<3, J>(D, acc); break;
    }
}

// Whether the kernel of J blocks per lane keeps accN: 3J more registers,
// which at J = 4 exceed the 128 a lane has at 16 waves per CU (the kernel
// then spills), so J = 4 keeps the DPP add per shifted dword.
template <int J>
constexpr bool wide_uses_accn = J <= 3;

// acc[4j + q] += lane L+1's accN[3j + q - 1] (q = 1..3), then accN = 0.
template <int J>
__device__ __forceinline__ void wide_take_next(uint32_t (&acc)[4 * J], uint32_t (&accN)[3 * J])
{
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            wide_add_next(acc[4 * j + q], accN[3 * j + q - 1]);
            accN[3 * j + q - 1] = 0;
        }
}

// The 8 entries of a full chunk, two at a time, with the row build spread
// over them: at slot s (two per entry pair: after the read of the pair's
// second entry, then after the read of the next pair's first), item
// kWItemAt[s] (if any) of this wave's row of the next chunk is stored (if
// `store`) and the same item of the row after that loaded into the freed
// registers.  The build's VALU work and LDS writes thus fill the waits on
// the entries' reads, and the global loads still have about a chunk to
// arrive.
constexpr int kWItemAt[kPairs] = {-1, 0, -1, 1, -1, 2, 3, 4};
template <int NI, int J>
__device__ __forceinline__ void wide_sum8_build(const uint32_t* lane_q, const uint32_t (&e)[kPMaxEnt / 2],
                                                uint32_t (&D0)[4 * J], uint32_t (&acc)[4 * J],
                                                uint32_t (&accN)[3 * J], pair_build<NI>& B, bool store, uint4* r,
                                                const uint8_t* ra, const uint8_t* rb, uint32_t mb_next, int lane)
{
    const pair_sel S = pair_sel_of(B.mb);
    uint32_t D1[4 * J];
    auto item = [&](int s) {
        const int it = kWItemAt[s];
        if (it >= 0 && it < NI) {
            if (store) pair_store_item<NI>(it, B, S, r);
            pair_load_item<NI, 2 * kWBlk * J>(it, ra, rb, lane, B);
        }
    };
#pragma unroll
    for (int i = 0; i < kPairs / 2; ++i) {
        wide_read<J>(lane_q, e[i] >> 16, D1);
        if constexpr (wide_uses_accn<J>) {
            item(2 * i);
            wide_acc2n_r<J>(D0, D1, e[i] & 3u, (e[i] >> 16) & 3u, acc, accN);
        } else {
            wide_acc_r<J>(D0, e[i] & 3u, acc);
            item(2 * i);
        }
        if (i + 1 < kPairs / 2) wide_read<J>(lane_q, e[i + 1] & 0xffffu, D0);
        if constexpr (!wide_uses_accn<J>) wide_acc_r<J>(D1, (e[i] >> 16) & 3u, acc);
        item(2 * i + 1);
    }
    B.mb = mb_next;
}

// All n entries of DM slot `wave` in chunk record rc, read from the record
// itself (a register array indexed at run time would live in scratch memory).
template <int J, int NDM>
__device__ __forceinline__ void wide_sum(const pair_rec_t<NDM>& rc, int slot, int n, const uint32_t* lane_q,
                                         uint32_t (&acc)[4 * J])
{
    const uint32_t* ew = reinterpret_cast<const uint32_t*>(&rc.ent[slot][0]);
    int i = 0;
    for (; i + 1 < n; i += 2) wide_add2<J>(lane_q, ew[i >> 1], acc);
    if (i < n) wide_add1<J>(lane_q, ew[i >> 1] & 0xffffu, acc);
}

// Software pipeline over the chunks: step k sums chunk k, stores chunk k+1's
// row from the registers loaded during step k-1, issues the loads of chunk
// k+2, and ends with the chunk's one barrier; a row's loads thus have about
// a whole step to arrive.  NI: quads per lane of a row (the range's row
// width rounded up to 64 quads); J: 252-dword blocks per lane read.
template <int NI, int J>
__device__ __forceinline__ void process_wide_tile(const lds_params& p, int dt, long t0)
{
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool active = dt * kDms + wave < p.ndms;
    const int nchunks = (p.nchans + kPairChunk - 1) / kPairChunk;
    const pair_rec* recs = reinterpret_cast<const pair_rec*>(p.pair_recs) + (size_t)dt * nchunks;
    constexpr int A = 4 * J, SPLIT = kWBlk * J, T = 2 * SPLIT;
    uint32_t acc16[A];
    uint32_t accN[3 * J];   // parts of the left neighbour's acc16 (wide_acc2n)
    uint32_t acc32[2 * A];
#pragma unroll
    for (int j = 0; j < A; ++j) acc16[j] = 0;
#pragma unroll
    for (int j = 0; j < 3 * J; ++j) accN[j] = 0;
#pragma unroll
    for (int j = 0; j < 2 * A; ++j) acc32[j] = 0;

    pair_build<NI> B;
    pair_build_load<NI, T>(p, t0, recs[0].row[wave], lane, B);
    if (wave < recs[0].nrows) pair_build_store(wave, lane, B, pair_buf(0));
    const int k1 = min(1, nchunks - 1);
    bool have_next = k1 == 1 && wave < recs[k1].nrows;   // B holds chunk k+1's row
    pair_build_load<NI, T>(p, t0, recs[k1].row[wave], lane, B);
    int kr = min(2, nchunks - 1);
    pair_row rn = recs[kr].row[wave];                     // row record of chunk k+2
    int nrows_n = recs[kr].nrows;
    pair_ents E = pair_ents_of(recs[0], wave);
    __syncthreads();
    const uint32_t* lane_q = reinterpret_cast<const uint32_t*>(pair_buf(0)) + 4 * lane;
    for (int k = 0; k < nchunks; ++k) {
        const uint32_t* base = lane_q + ((k & 1) ? kPRows * kPRowDw : 0);   // rows of chunk k
        uint4* rows_next = pair_buf((k + 1) & 1);
        // the first entry's reads go ahead of the row stores in the LDS queue
        uint32_t D0[A];
        wide_read<J>(base, E.e[0] & 0xffffu, D0);
        const bool store = have_next;
        have_next = k + 2 < nchunks && wave < nrows_n;
        if (active && E.n == kPairs) {   // the usual full chunk: straight-line code
            wide_sum8_build<NI, J>(base, E.e, D0, acc16, accN, B, store, rows_next + wave * kPRQ + lane,
                                p.L + t0 + rn.offa, p.L + t0 + rn.offb + rn.sh, rn.mb, lane);
        } else {
            if (store) pair_build_store(wave, lane, B, rows_next);
            pair_build_load<NI, T>(p, t0, rn, lane, B);
            if (active) wide_sum<J>(recs[k], wave, E.n, base, acc16);
        }
        if (active && ((k + 1) % kPFlushChunks == 0 || k + 1 == nchunks)) {
            if constexpr (wide_uses_accn<J>) wide_take_next<J>(acc16, accN);
#pragma unroll
            for (int j = 0; j < A; ++j) {
                acc32[j] += acc16[j] & 0xFFFFu;
                acc32[A + j] += acc16[j] >> 16;
                acc16[j] = 0;
            }
        }
        kr = min(k + 3, nchunks - 1);
        rn = recs[kr].row[wave];
        nrows_n = recs[kr].nrows;
        E = pair_ents_of(recs[min(k + 1, nchunks - 1)], wave);
        __syncthreads();
    }

    if (active && lane < 63) {
        const int d = dt * kDms + wave;
        float* orow = p.out + (size_t)d * p.out_pitch;
        const bool aligned = (reinterpret_cast<uintptr_t>(orow) & 15u) == 0;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int b = 0; b < J; ++b) {
                store_out4(orow, t0 + h * SPLIT + kWBlk * b + 4 * lane, p.t_processed, aligned,
                           out_values4(acc32 + h * A + 4 * b, p));
            }
    }
}

template <int NI, int J>
__global__ __launch_bounds__(kThreads) void dedisperse_wide_kernel(lds_params p)
{
    const int ntiles = p.n_dm_tiles * p.n_t_tiles;
    const int per_xcd = (ntiles + 7) / 8;
    const int w = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (w >= ntiles) return;  // whole work-group leaves before any barrier
    process_wide_tile<NI, J>(p, w % p.n_dm_tiles, (long)(w / p.n_dm_tiles) * (2 * kWBlk * J));
}

// Many pair-shape ranges (of many beams) sharing J and NI in one launch.
// The 8 queues are those of the byte-row batch kernel: queue x holds the
// x-th eighth of every job, job after job, so the XCDs work through the jobs
// side by side as launches per range would.  Work-group g runs on XCD g % 8
// (work-groups are dispatched to the XCDs round robin) and processes
// position g / 8 of queue g % 8: the hardware dispatcher hands each XCD its
// queue's tiles in order, so neighbouring DM tiles of one time tile run
// together and share their input rows in that XCD's L2, as in a launch per
// range.  One tile per work-group: a loop over tiles in each work-group
// (a persistent grid) keeps its loop state beside the tile's and, at
// J = 4, raises the kernel from 56 SGPRs / 121 VGPRs to 106 / 128 with
// scratch spills.  The job table is read through the constant address
// space, so its fields load into scalar registers as kernel arguments do.
template <int NI, int J>
__global__ __launch_bounds__(kThreads) void dedisperse_wide_batch_kernel(const lds_job* __restrict__ jobs, int njobs,
                                                                         const long* __restrict__ bounds)
{
    typedef const __attribute__((address_space(4))) lds_job const_job;
    const int x = (int)(blockIdx.x & 7u);
    const long pos = (long)(blockIdx.x >> 3);
    if (pos >= bounds[x + 1] - bounds[x]) return;   // whole work-group leaves before any barrier
    long local;
    const lds_job& j0 = batch_job_of(jobs, njobs, x, pos, &local);
    const const_job& jb = ((const const_job*)jobs)[&j0 - jobs];
    process_wide_tile<NI, J>((const lds_params&)jb.p, (int)(local % jb.p.n_dm_tiles),
                             (local / jb.p.n_dm_tiles) * (2L * kWBlk * J));
}


// ---------------------------------------------------------------------------
// Shape 6: 1008-sample tiles (J = 2) of 32 DM trials, two per wave (DM
// slots w and w + 16 of wave w).  A chunk still builds at most 16 rows (a
// pair's relative delay takes at most two values over the 32 trials, else
// three pair rows or two single-channel rows), one per wave, so the
// build and the row stores are shared by twice the trials of shape 2, and
// a chunk holds as many adds per wave as shape 4's.
// ---------------------------------------------------------------------------

// Entry i (< 8) of a DM's packed entries.
__device__ __forceinline__ uint32_t wide2_entry(const uint32_t (&e)[kPMaxEnt / 2], int i)
{
    return (i & 1) ? (e[i >> 1] >> 16) : (e[i >> 1] & 0xffffu);
}

// The 8 entries of each of the two DMs of a full chunk, DM a's entry i
// summed while DM b's is read and the other way round, with the row build
// spread over them: item kW2ItemAt[s] after slot s (even: DM a's entry s/2,
// odd: DM b's).
constexpr int kW2ItemAt[2 * kPairs] = {-1, -1, -1, 0, -1, -1, -1, 1, -1, -1, -1, 2, -1, 3, -1, 4};
template <int NI>
__device__ __forceinline__ void wide2_sum8_build(const uint32_t* lane_q, const uint32_t (&ea)[kPMaxEnt / 2],
                                                 const uint32_t (&eb)[kPMaxEnt / 2], uint32_t (&Da)[8],
                                                 uint32_t (&acca)[8], uint32_t (&accb)[8], pair_build<NI>& B,
                                                 bool store, uint4* r, const uint8_t* ra, const uint8_t* rb,
                                                 uint32_t mb_next, int lane)
{
    constexpr int J = 2;
    const pair_sel S = pair_sel_of(B.mb);
    uint32_t Db[4 * J];
    auto item = [&](int s) {
        const int it = kW2ItemAt[s];
        if (it >= 0 && it < NI) {
            if (store) pair_store_item<NI>(it, B, S, r);
            pair_load_item<NI, 2 * kWBlk * J>(it, ra, rb, lane, B);
        }
    };
#pragma unroll
    for (int i = 0; i < kPairs; ++i) {
        wide_read<J>(lane_q, wide2_entry(eb, i), Db);
        wide_acc_r<J>(Da, wide2_entry(ea, i) & 3u, acca);
        item(2 * i);
        if (i + 1 < kPairs) wide_read<J>(lane_q, wide2_entry(ea, i + 1), Da);
        wide_acc_r<J>(Db, wide2_entry(eb, i) & 3u, accb);
        item(2 * i + 1);
    }
    B.mb = mb_next;
}

template <int NI>
__device__ __forceinline__ void process_wide2_tile(const lds_params& p, int dt, long t0)
{
    constexpr int J = 2, A = 4 * J, SPLIT = kWBlk * J, T = 2 * SPLIT;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int da = dt * kDms2 + wave, db = da + kDms;
    const bool act_a = da < p.ndms, act_b = db < p.ndms;
    const int nchunks = (p.nchans + kPairChunk - 1) / kPairChunk;
    const pair_rec2* recs = reinterpret_cast<const pair_rec2*>(p.pair_recs) + (size_t)dt * nchunks;
    uint32_t acc16a[A], acc16b[A];
    uint32_t acc32a[2 * A], acc32b[2 * A];
#pragma unroll
    for (int j = 0; j < A; ++j) acc16a[j] = acc16b[j] = 0;
#pragma unroll
    for (int j = 0; j < 2 * A; ++j) acc32a[j] = acc32b[j] = 0;

    pair_build<NI> B;
    pair_build_load<NI, T>(p, t0, recs[0].row[wave], lane, B);
    if (wave < recs[0].nrows) pair_build_store(wave, lane, B, pair_buf(0));
    const int k1 = min(1, nchunks - 1);
    bool have_next = k1 == 1 && wave < recs[k1].nrows;   // B holds chunk k+1's row
    pair_build_load<NI, T>(p, t0, recs[k1].row[wave], lane, B);
    int kr = min(2, nchunks - 1);
    pair_row rn = recs[kr].row[wave];                     // row record of chunk k+2
    int nrows_n = recs[kr].nrows;
    pair_ents Ea = pair_ents_of(recs[0], wave), Eb = pair_ents_of(recs[0], wave + kDms);
    __syncthreads();
    const uint32_t* lane_q = reinterpret_cast<const uint32_t*>(pair_buf(0)) + 4 * lane;
    for (int k = 0; k < nchunks; ++k) {
        const uint32_t* base = lane_q + ((k & 1) ? kPRows * kPRowDw : 0);   // rows of chunk k
        uint4* rows_next = pair_buf((k + 1) & 1);
        uint32_t Da[A];
        wide_read<J>(base, Ea.e[0] & 0xffffu, Da);
        const bool store = have_next;
        have_next = k + 2 < nchunks && wave < nrows_n;
        if (act_a && act_b && Ea.n == kPairs) {   // the usual full chunk: straight-line code
            wide2_sum8_build<NI>(base, Ea.e, Eb.e, Da, acc16a, acc16b, B, store, rows_next + wave * kPRQ + lane,
                                 p.L + t0 + rn.offa, p.L + t0 + rn.offb + rn.sh, rn.mb, lane);
        } else {
            if (store) pair_build_store(wave, lane, B, rows_next);
            pair_build_load<NI, T>(p, t0, rn, lane, B);
            if (act_a) wide_sum<J>(recs[k], wave, Ea.n, base, acc16a);
            if (act_b) wide_sum<J>(recs[k], wave + kDms, Eb.n, base, acc16b);
        }
        if ((k + 1) % kPFlushChunks == 0 || k + 1 == nchunks) {
#pragma unroll
            for (int j = 0; j < A; ++j) {
                acc32a[j] += acc16a[j] & 0xFFFFu;
                acc32a[A + j] += acc16a[j] >> 16;
                acc16a[j] = 0;
                acc32b[j] += acc16b[j] & 0xFFFFu;
                acc32b[A + j] += acc16b[j] >> 16;
                acc16b[j] = 0;
            }
        }
        kr = min(k + 3, nchunks - 1);
        rn = recs[kr].row[wave];
        nrows_n = recs[kr].nrows;
        const int kn = min(k + 1, nchunks - 1);
        Ea = pair_ents_of(recs[kn], wave);
        Eb = pair_ents_of(recs[kn], wave + kDms);
        __syncthreads();
    }

    if (lane < 63) {
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            if (!(m == 0 ? act_a : act_b)) continue;
            float* orow = p.out + (size_t)(m == 0 ? da : db) * p.out_pitch;
            const bool aligned = (reinterpret_cast<uintptr_t>(orow) & 15u) == 0;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int b = 0; b < J; ++b) {
                    store_out4(orow, t0 + h * SPLIT + kWBlk * b + 4 * lane, p.t_processed, aligned,
                               out_values4((m == 0 ? acc32a : acc32b) + h * A + 4 * b, p));
                }
        }
    }
}

template <int NI>
__global__ __launch_bounds__(kThreads) void dedisperse_wide2_kernel(lds_params p)
{
    const int ntiles = p.n_dm_tiles * p.n_t_tiles;
    const int per_xcd = (ntiles + 7) / 8;
    const int w = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (w >= ntiles) return;  // whole work-group leaves before any barrier
    process_wide2_tile<NI>(p, w % p.n_dm_tiles, (long)(w / p.n_dm_tiles) * kWT2);
}

// Shape-6 ranges of a batch: the queues and the job lookup of
// dedisperse_wide_batch_kernel.
template <int NI>
__global__ __launch_bounds__(kThreads) void dedisperse_wide2_batch_kernel(const lds_job* __restrict__ jobs, int njobs,
                                                                          const long* __restrict__ bounds)
{
    typedef const __attribute__((address_space(4))) lds_job const_job;
    const int x = (int)(blockIdx.x & 7u);
    const long pos = (long)(blockIdx.x >> 3);
    if (pos >= bounds[x + 1] - bounds[x]) return;   // whole work-group leaves before any barrier
    long local;
    const lds_job& j0 = batch_job_of(jobs, njobs, x, pos, &local);
    const const_job& jb = ((const const_job*)jobs)[&j0 - jobs];
    process_wide2_tile<NI>((const lds_params&)jb.p, (int)(local % jb.p.n_dm_tiles),
                           (local / jb.p.n_dm_tiles) * (long)kWT2);
}

}  // namespace

hipError_t launch_dedisperse_wide_batch(const lds_job* d_jobs, int njobs, const long* d_bounds, long max_queue,
                                        int J, int ni, hipStream_t s)
{
    if (njobs <= 0 || max_queue <= 0) return hipSuccess;
    if (!d_jobs || !d_bounds || max_queue > 0x0fffffffL) return hipErrorInvalidValue;
    const long grid = 8 * max_queue;
    const dim3 g((unsigned)grid), b(kThreads);
#define AA_WIDE_BATCH(NI_, J_) \
    hipLaunchKernelGGL((dedisperse_wide_batch_kernel<NI_, J_>), g, b, 0, s, d_jobs, njobs, d_bounds)
    if (J == 4) {
        if (ni <= 4) AA_WIDE_BATCH(4, 4);
        else AA_WIDE_BATCH(5, 4);
    } else if (J == 3) {
        if (ni <= 3) AA_WIDE_BATCH(3, 3);
        else if (ni == 4) AA_WIDE_BATCH(4, 3);
        else AA_WIDE_BATCH(5, 3);
    } else if (J == 2) {
        if (ni <= 2) AA_WIDE_BATCH(2, 2);
        else if (ni == 3) AA_WIDE_BATCH(3, 2);
        else if (ni == 4) AA_WIDE_BATCH(4, 2);
        else AA_WIDE_BATCH(5, 2);
    } else {
        return hipErrorInvalidValue;
    }
#undef AA_WIDE_BATCH
    return hipGetLastError();
}

hipError_t launch_dedisperse_wide(const lds_params& q, int J, int ni, unsigned grid, hipStream_t s)
{
    const dim3 g(grid), b(kThreads);
    if (J == 4) {   // 2016-sample tiles: rows of at least 252 quads
        if (ni <= 4) hipLaunchKernelGGL((dedisperse_wide_kernel<4, 4>), g, b, 0, s, q);
        else hipLaunchKernelGGL((dedisperse_wide_kernel<5, 4>), g, b, 0, s, q);
    } else if (J == 3) {   // 1512-sample tiles: rows of at least 189 quads
        if (ni <= 3) hipLaunchKernelGGL((dedisperse_wide_kernel<3, 3>), g, b, 0, s, q);
        else if (ni == 4) hipLaunchKernelGGL((dedisperse_wide_kernel<4, 3>), g, b, 0, s, q);
        else hipLaunchKernelGGL((dedisperse_wide_kernel<5, 3>), g, b, 0, s, q);
    } else if (ni <= 2) {   // 1008-sample tiles: rows of at least 126 quads
        hipLaunchKernelGGL((dedisperse_wide_kernel<2, 2>), g, b, 0, s, q);
    } else if (ni == 3) {
        hipLaunchKernelGGL((dedisperse_wide_kernel<3, 2>), g, b, 0, s, q);
    } else if (ni == 4) {
        hipLaunchKernelGGL((dedisperse_wide_kernel<4, 2>), g, b, 0, s, q);
    } else {
        hipLaunchKernelGGL((dedisperse_wide_kernel<5, 2>), g, b, 0, s, q);
    }
    return hipGetLastError();
}

hipError_t launch_dedisperse_wide2_batch(const lds_job* d_jobs, int njobs, const long* d_bounds, long max_queue,
                                         int ni, hipStream_t s)
{
    if (njobs <= 0 || max_queue <= 0) return hipSuccess;
    if (!d_jobs || !d_bounds || max_queue > 0x0fffffffL) return hipErrorInvalidValue;
    const dim3 g((unsigned)(8 * max_queue)), b(kThreads);
    if (ni <= 2) hipLaunchKernelGGL(dedisperse_wide2_batch_kernel<2>, g, b, 0, s, d_jobs, njobs, d_bounds);
    else if (ni == 3) hipLaunchKernelGGL(dedisperse_wide2_batch_kernel<3>, g, b, 0, s, d_jobs, njobs, d_bounds);
    else if (ni == 4) hipLaunchKernelGGL(dedisperse_wide2_batch_kernel<4>, g, b, 0, s, d_jobs, njobs, d_bounds);
    else hipLaunchKernelGGL(dedisperse_wide2_batch_kernel<5>, g, b, 0, s, d_jobs, njobs, d_bounds);
    return hipGetLastError();
}

hipError_t launch_dedisperse_wide2(const lds_params& q, int ni, unsigned grid, hipStream_t s)
{
    const dim3 g(grid), b(kThreads);
    if (ni <= 2) hipLaunchKernelGGL(dedisperse_wide2_kernel<2>, g, b, 0, s, q);
    else if (ni == 3) hipLaunchKernelGGL(dedisperse_wide2_kernel<3>, g, b, 0, s, q);
    else if (ni == 4) hipLaunchKernelGGL(dedisperse_wide2_kernel<4>, g, b, 0, s, q);
    else hipLaunchKernelGGL(dedisperse_wide2_kernel<5>, g, b, 0, s, q);
    return hipGetLastError();
}

}  // namespace aa_hip
