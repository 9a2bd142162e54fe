// Batched dedispersion of many beams (each its own plan: channel count,
// length, DM ranges), optionally replayed from a captured hipGraph.  No
// counterpart in the reference, whose pipeline runs one beam per process
// (include/aa_pipeline_api.hpp:113-117).
//
// Per call: every beam's chunk is corner-turned into its engine's level 0 and
// binned as its ranges need (the engines' own kernels).  The (beam, range)
// jobs on the two byte-row chunk shapes run in one persistent launch over a
// shared tile list (dedisperse_lds_batch_kernel); ranges on the pair or quad
// pre-summation shapes, or on the direct kernel (no LDS shape fits), run as
// launches of their own after it (captured in the same graph).  Results are
// bit-identical to running each engine's execute_chunk.
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
    // on their own
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
    int nranges_total() const { return m_nranges_total; }

private:
    ddtr_batch() = default;
    int enqueue(const void *const *inputs, const int *t_processed, float *const *d_out, const size_t *out_pitch,
                hipStream_t s);
    hipError_t launch_families(hipStream_t s);
    void release_graph();

    int m_device = 0;
    int m_grid = 0;
    std::vector<ddtr_engine *> m_beams;
    int m_nranges_total = 0;
    // The persistent launch's jobs (byte-row shapes).
    struct family {
        std::vector<lds_job> jobs;
        std::vector<long> bounds;  // 9 queue bounds (host copy)
        long ntiles = 0;
        lds_job *d_jobs = nullptr;
        size_t jobs_cap = 0;
        long *d_bounds = nullptr;  // 9 queue bounds
        int *d_counters = nullptr; // 8 queue heads
    };
    static constexpr int kFamilies = 1;
    family m_fam[kFamilies];
    std::vector<range_launch> m_direct;
    hipStream_t m_capture_stream = nullptr;  // for callers on the legacy default stream
    hipGraph_t m_graph = nullptr;
    hipGraphExec_t m_exec = nullptr;
    std::vector<uintptr_t> m_key;  // buffers and sizes the graph was captured for
};

}  // namespace aa_hip
