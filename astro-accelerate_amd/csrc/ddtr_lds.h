// LDS-tiled dedispersion kernel (see ddtr_lds.hip for the algorithm).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace aa_hip {

// Largest staged row (uint4 units) of the 32-channel chunk shape: two
// 32-channel buffers plus the offset tables fill 160 KiB of LDS.  The
// 16-channel shape holds rows twice as long in the same buffers.
constexpr int kLdsRowQuadsMax = 152;
// Chunk shapes: 0 = 32 channels x 152 quads, 1 = 16 channels x 304 quads,
// 2 = 16 channels pre-summed in pairs (rows of 16-bit pair sums, 1008-sample
// tiles, window offsets up to 772 samples), 3 = 16 channels pre-summed in
// groups of four, 4 = shape 2 on 2016-sample tiles (offsets up to 268),
// 5 = shape 2 on 1512-sample tiles (offsets up to 520), 6 = shape 2 with
// 32 DM trials per tile (two per wave) sharing the chunk's rows.  Shapes 2,
// 4, 5 and 6 read rows with 16-byte LDS reads.
constexpr int kLdsVariants = 7;

struct lds_params {
    const uint8_t* L;        // nchans rows of u8 samples
    size_t pitch;            // bytes between rows (multiple of 16)
    int nchans;
    const float* dmshifts;   // device, nchans entries
    int ndms;
    int t_processed;
    float mstartdm, mdmstep;
    float fnchans, fbin;
    float fscale;            // 1 / (nchans * inBin) when both are powers of two (the two
                             // divisions are then exact and equal this product), else 0
    float* out;
    size_t out_pitch;        // floats between output rows
    int row_quads;           // staged row width in uint4 (from lds_row_quads)
    const void* pair_recs;   // chunk shapes 2-4: device plan tables (from lds_plan_tables)
    int n_dm_tiles, n_t_tiles;  // filled by the launcher / lds_job_tiles
};

// The output value (float)sum / nchans / inBin, each division rounded as the
// reference's; one multiplication when both divisors are powers of two.
__device__ __forceinline__ float out_value(uint32_t sum, const lds_params& p)
{
    return p.fscale != 0.0f ? (float)sum * p.fscale : __fdiv_rn(__fdiv_rn((float)sum, p.fnchans), p.fbin);
}

// Output of one lane: samples tt .. tt+3 of a DM row (tt a multiple of 4),
// one 16-byte store where the row start is 16-byte aligned and all four lie
// inside the plane's t_end samples, else one store per sample inside it.
typedef float f32x4_out __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_out4(float* orow, long tt, long t_end, bool aligned, f32x4_out v)
{
    if (aligned && tt + 3 < t_end) {
        *reinterpret_cast<f32x4_out*>(orow + tt) = v;
    } else {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            if (tt + m < t_end) orow[tt + m] = v[m];
    }
}

// The four sums s[0..3] as output values.
__device__ __forceinline__ f32x4_out out_values4(const uint32_t* s, const lds_params& p)
{
    return f32x4_out{out_value(s[0], p), out_value(s[1], p), out_value(s[2], p), out_value(s[3], p)};
}

// One range of a batched launch.  The launch has 8 queues (one per XCD);
// queue x holds, job after job, the x-th eighth of every job's tiles (tiles
// [x n / 8, (x+1) n / 8) of a job of n = n_dm_tiles * n_t_tiles tiles), so
// the XCDs work through the jobs side by side.  qend[x]: end of this job's
// part in queue x (queue positions from 0).
struct lds_job {
    lds_params p;
    int variant;             // chunk shape
    int pad;
    long qend[8];
};

// Staged row width (uint4) this geometry needs with chunk shape `variant`,
// or -1 when it exceeds that shape's rows.  Host only; evaluates the delay
// rule exactly as the kernel does.
int lds_row_quads(int variant, const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep);

// Chunk shape 2 (pair pre-summation) plan tables of one range over rows
// `pitch` bytes apart: fills `bytes` (when not null) with the records the
// kernel reads, to be copied to the device as they are, and returns the row
// width in quads, or -1 when the range does not fit the shape.  Host only.
int lds_pair_tables(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, size_t pitch,
                    std::vector<uint8_t>* bytes);

// Plan tables of chunk shape 2, 3 or 4 (as lds_pair_tables; -1 for other
// shapes, which need none).
int lds_plan_tables(int variant, const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep,
                    size_t pitch, std::vector<uint8_t>* bytes);

// Tile geometry: samples per tile of chunk shape `variant` and DM trials
// per tile (16, or 32 in shape 6).  Rows must hold the tile's samples
// readable past the last processed sample plus the maximum delay (the
// engine pads by 4 KiB).
long lds_tile_samples(int variant);
int lds_tile_dms(int variant);

hipError_t launch_dedisperse_lds(const lds_params& p, int variant, hipStream_t s);

// Fills p.n_dm_tiles and p.n_t_tiles for chunk shape `variant`.
void lds_job_tiles(lds_params& p, int variant);

// Persistent launch over jobs[0..njobs) (device array, in queue order);
// queue x is items d_bounds[x] .. d_bounds[x+1]-1 of the launch (d_bounds[x]
// = the lengths of queues 0..x-1 added), d_counters[0..7] is scratch (reset
// by the launch itself, with a kernel, so the sequence may be captured in a
// graph).  grid: multiple of 8.
hipError_t launch_dedisperse_lds_batch(const lds_job* d_jobs, int njobs, const long* d_bounds, int* d_counters,
                                       int grid, hipStream_t s);

// One launch over pair-shape jobs (chunk shapes 2, 4, 5) that share the
// tile width J (2, 3, 4 blocks: shapes 2, 5, 4) and the row width NI = ni
// quads per lane (the instantiation a launch per range would use: ni of at
// least J); jobs in queue order and d_bounds as for
// launch_dedisperse_lds_batch; max_queue: the longest queue (tiles).  Each
// work-group takes one tile of its XCD's queue.
hipError_t launch_dedisperse_wide_batch(const lds_job* d_jobs, int njobs, const long* d_bounds, long max_queue,
                                        int J, int ni, hipStream_t s);

// The same over shape-6 jobs (32 DM trials per tile) sharing NI = ni.
hipError_t launch_dedisperse_wide2_batch(const lds_job* d_jobs, int njobs, const long* d_bounds, long max_queue,
                                         int ni, hipStream_t s);


}  // namespace aa_hip
