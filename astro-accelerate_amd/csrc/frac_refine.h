// Batched DM refinement of candidate windows (see frac_refine.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace aa_hip {

constexpr int kRefineMaxSamples = 4096;    // per window (1024 threads x 4)
constexpr int kRefineMaxChannels = 4096;   // per window (delay table in LDS)

struct refine_cand {
    int64_t x_off;      // window at x + x_off, [C][T] float32
    int64_t ds_off;     // delay per unit DM at ds + ds_off, C entries
    int64_t power_off;  // channel powers at power + power_off, C entries (loss 1)
    int C, T;
};

struct refine_args {
    const float* x;
    const float* ds;
    const refine_cand* cands;   // [ncand]
    const int* order;           // processing order (largest windows first)
    int ncand;
    float* u;                   // [ncand] residual DM, updated in place
    float* step;                // [ncand] Rprop step, updated in place
    double* prev_grad;          // [ncand] previous dL/du, updated in place
    double* loss;               // [ncand] L at the final u
    double* grad;               // [ncand] dL/du at the final u
    int nsteps;
    float step_max, step_min;
    int scan_n;                 // coarse scan points before the steps (0 = none)
    float scan_lo, scan_step;   // scan offsets relative to the starting u
    int* counter;               // queue head (reset by the launch)
    int loss_kind;              // 0: E = sum_t y^2; 1: E minus the expected noise energy
    double* power;              // loss 1: per-candidate channel powers (written by the launch)
};

// grid: resident work-groups (the kernel is persistent).
hipError_t launch_refine(const refine_args& a, int grid, hipStream_t s);
int refine_blocks_per_cu();

}  // namespace aa_hip
