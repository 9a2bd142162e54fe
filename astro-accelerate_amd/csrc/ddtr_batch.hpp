// Batched dedispersion of many beams (each its own plan: channel count,
// length, DM ranges), optionally replayed from a captured hipGraph.  No
// counterpart in the reference, whose pipeline runs one beam per process
// (include/aa_pipeline_api.hpp:113-117).
//
// Per call: every beam's chunk is corner-turned into its engine's level 0 and
// the binned levels the corner turn writes, all beams that take the wide
// corner turn in one launch per number of binned levels; deeper levels are
// binned as the ranges need (the engines' own kernels).  The (beam, range)
// jobs then run in launches over shared tile queues (8, one per XCD), one
// launch per kernel instantiation: the two byte-row chunk shapes together
// in a persistent launch (dedisperse_lds_batch_kernel), and the pair
// pre-summation shapes grouped by tile width and row width, one tile per
// work-group (dedisperse_wide_batch_kernel<NI, J>).  Ranges
// on the quad shape or on the direct kernel (no LDS shape fits) run as
// launches of their own after them (captured in the same graph).  Results
// are bit-identical to running each engine's execute_chunk.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ddtr_engine.hpp"
#include "ddtr_lds.h"

namespace aa_hip {

class ddtr_batch {
public:
    static int create(int nbeams, ddtr_engine *const *beams, ddtr_batch **out);
    ~ddtr_batch();

    // inputs[b]: beam b's [t][c] u8 chunk; t_processed, d_out, out_pitch are
    // flattened beam-major over each beam's ranges.  With use_graph, the
    // launch sequence is captured once per distinct set of buffers and
    // replayed afterwards.
    int execute(const void *const *inputs, const int *t_processed, float *const *d_out, const size_t *out_pitch,
                bool use_graph, hipStream_t s);

    // jobs (ranges) and tiles in the persistent launches, ranges launched
    // on their own, persistent launches
    int njobs() const {
        size_t n = 0;
        for (const auto &f : m_fam) n += f.jobs.size();
        return (int)n;
    }
    long ntiles() const {
        long n = 0;
        for (const auto &f : m_fam) n += f.ntiles;
        return n;
    }
    int ndirect() const { return (int)m_direct.size(); }
    // corner turns of the last call: batched launches, beams turned alone
    int nct_batched() const {
        int n = 0;
        for (const auto &f : m_ct) n += f.jobs.empty() ? 0 : 1;
        return n;
    }
    int nct_own() const { return m_ct_direct; }
    int nlaunches() const {
        int n = 0;
        for (const auto &f : m_fam) n += f.jobs.empty() ? 0 : 1;
        return n;
    }
    int nranges_total() const { return m_nranges_total; }

private:
    ddtr_batch() = default;
    // upload: copy the corner-turn job tables to the device first (the eager
    // pass; a captured pass reuses them)
    int enqueue(const void *const *inputs, const int *t_processed, float *const *d_out, const size_t *out_pitch,
                hipStream_t s, bool upload);
    hipError_t launch_families(hipStream_t s);
    void release_graph();

    int m_device = 0;
    int m_grid = 0;
    std::vector<ddtr_engine *> m_beams;
    int m_nranges_total = 0;
    // One batched launch's jobs.  Family 0: the byte-row shapes (chunk
    // shapes 0 and 1, dedisperse_lds_batch_kernel); family 1 + 4 (J - 2) +
    // (NI - 2): the pair shapes of tile width J (2: shape 2, 3: shape 5, 4:
    // shape 4) and NI quads per lane of a row (dedisperse_wide_batch_kernel).
    struct family {
        std::vector<lds_job> jobs;
        std::vector<long> bounds;  // 9 queue bounds (host copy)
        long ntiles = 0;
        lds_job *d_jobs = nullptr;
        size_t jobs_cap = 0;
        long *d_bounds = nullptr;  // 9 queue bounds
        int *d_counters = nullptr; // 8 queue heads (family 0)
    };
    // family 0: byte rows (one persistent launch); 1..12: pair shapes by
    // (J, NI); 13..16: shape 6 by NI
    static constexpr int kFamilies = 1 + 4 * 4;
    static int family_of(int variant, int row_quads);   // -1: launched on its own
    family m_fam[kFamilies];
    std::vector<range_launch> m_direct;
    // Batched corner turns by number of binned levels (0..kCtMaxLevels).
    struct ct_family {
        std::vector<ct_job> jobs;
        ct_job *d_jobs = nullptr;
        size_t cap = 0;
    };
    ct_family m_ct[kCtMaxLevels + 1];
    int m_ct_direct = 0;   // beams turned by launches of their own
    hipStream_t m_capture_stream = nullptr;  // for callers on the legacy default stream
    hipGraph_t m_graph = nullptr;
    hipGraphExec_t m_exec = nullptr;
    std::vector<uintptr_t> m_key;  // buffers and sizes the graph was captured for
};

}  // namespace aa_hip
