// Drop-in replacements for the reference's include/aa_zero_dm_outliers.hpp
// (src/aa_zero_dm_outliers.cpp:6-67): zero-DM with outlier rejection on the
// reference's u16 [t][c] device input, in place, on the legacy default
// stream, synchronous.  Per time sample, the mean over the channels strictly
// inside mean +- 3 sigma (iterated up to 5 times, the reference's streaming
// statistics); x' = (x - mean) + normalization_factor[c] in float, clamped to
// [0, 2^nbits - 1], truncated.  The 4-argument form uses (2^nbits - 1) / 2
// for every channel.
#ifndef ASTRO_ACCELERATE_AA_ZERO_DM_OUTLIERS_HPP
#define ASTRO_ACCELERATE_AA_ZERO_DM_OUTLIERS_HPP

namespace astroaccelerate {

void zero_dm_outliers(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits,
                      float *normalization_factor);
void zero_dm_outliers(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits);

}  // namespace astroaccelerate
#endif
