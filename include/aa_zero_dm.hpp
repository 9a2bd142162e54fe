// Drop-in replacement for zero_dm (include/aa_zero_dm.hpp:22-23,
// src/aa_zero_dm.cpp:9-37): the zero-DM filter on the reference's u16
// [t][c] device input, in place, on the legacy default stream, synchronous.
// Per time sample x' = (x - mean over channels) + normalization_factor[c]
// in float, clamped to [0, 2^nbits - 1], truncated.  The 4-argument form
// uses (2^nbits - 1) / 2 for every channel.
#ifndef ASTRO_ACCELERATE_AA_ZERO_DM_HPP
#define ASTRO_ACCELERATE_AA_ZERO_DM_HPP

namespace astroaccelerate {

void zero_dm(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits,
             float *normalization_factor);
void zero_dm(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits);

}  // namespace astroaccelerate
#endif
