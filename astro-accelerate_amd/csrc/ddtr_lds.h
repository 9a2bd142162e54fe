// LDS-tiled dedispersion kernel (see ddtr_lds.hip for the algorithm).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace aa_hip {

// Largest staged row (uint4 units) that fits two 32-channel buffers plus the
// offset tables in 160 KiB of LDS.
constexpr int kLdsRowQuadsMax = 152;

struct lds_params {
    const uint8_t* L;        // nchans rows of u8 samples
    size_t pitch;            // bytes between rows (multiple of 16)
    int nchans;
    const float* dmshifts;   // device, nchans entries
    int ndms;
    int t_processed;
    float mstartdm, mdmstep;
    float fnchans, fbin;
    float* out;
    size_t out_pitch;        // floats between output rows
    int row_quads;           // staged row width in uint4 (from lds_row_quads)
    int n_dm_tiles, n_t_tiles;  // filled by the launcher
};

// Staged row width (uint4) this geometry needs, or -1 when it exceeds LDS.
// Host only; evaluates the delay rule exactly as the kernel does.
int lds_row_quads(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep);

// Samples per time tile; rows must hold this many readable bytes past the
// last processed sample plus the maximum delay (the engine pads by 4 KiB).
long lds_tile_samples();

hipError_t launch_dedisperse_lds(const lds_params& p, hipStream_t s);

}  // namespace aa_hip
