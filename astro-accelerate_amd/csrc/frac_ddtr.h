// Fractional-delay dedispersion of candidate windows (FP32), its adjoint, and
// the kernels' launch descriptions.  See frac_ddtr.hip for the algorithm.
//
// For candidate b, trial k, channel c and output sample t:
//   u_bk  = trials[k] + dm[b], tau = ds[c] * u_bk  (evaluated in double)
//   i0    = floor(tau), f = (float)(tau - i0)
//   y[b][k][t] = sum_c (1 - f) * x[b][c][t + i0] + f * x[b][c][t + i0 + 1]
// with x[b][c][j] = 0 outside 0 <= j < T.  ds[c] is the dispersion delay per
// unit DM in samples (the reference's dmshift[c] / tsamp), dm[b] the
// candidate's continuous DM offset and trials[k] the trial DM offsets.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace aa_hip {

constexpr int kFracTile = 1024;        // output samples per work-group tile (64 lanes x 16)
constexpr int kFracRowQuads = 320;     // staged row length in float4
constexpr int kFracChunkRows = 12;     // rows per LDS chunk (two chunks resident)
// Trials (forward, DM adjoint) or channels (input adjoint) per wave at most:
// two keep the accumulators, the window and the staging registers within
// 128 VGPRs (1024-thread work-groups) without spills.
constexpr int kFracMaxPerWave = 2;
// Largest window offset (in samples from the staged row start) a lane run may
// have: lane 63 reads quads (16*63 + e)/4 .. +4, all below kFracRowQuads.
constexpr int kFracMaxOffset = 4 * (kFracRowQuads - 257) + 3;

struct frac_args {
    const float* x;        // [B][C][T] input windows
    const float* gy;       // [B][K][T] adjoint of y
    float* y;              // [B][K][T] output profiles
    float* gx;             // [B][C][T] adjoint of x
    double* gdm_part;      // partial d(loss)/d(dm) per work-group
    double* gdm;           // [B] d(loss)/d(dm)
    const float* ds;       // [C] delay per unit DM, samples
    const float* dm;       // [B]
    const float* trials;   // [K]
    int B, C, T, K;
    int n_ttiles;          // ceil(T / kFracTile)
    int n_groups;          // trial groups (forward, dm adjoint) or channel groups (x adjoint)
    int group_q;           // trials (or channels) per wave: 1 or 2
};

// Trials per wave (1 or 2) such that a group of 16*q consecutive trials
// keeps every staged window within a row; 0 if even q = 1 does not fit.
// Host only: ds_absmax = max |ds[c]|, trials as given.
int frac_trials_per_wave(const float* h_trials, int K, float ds_absmax);

hipError_t launch_frac_forward(const frac_args& a, hipStream_t s);
hipError_t launch_frac_grad_dm(const frac_args& a, hipStream_t s);   // writes gdm_part then gdm
hipError_t launch_frac_grad_x(const frac_args& a, hipStream_t s);

// Work-groups (hence gdm_part entries) of launch_frac_grad_dm.
size_t frac_grad_dm_parts(const frac_args& a);

}  // namespace aa_hip
