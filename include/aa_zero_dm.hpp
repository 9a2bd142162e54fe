// Drop-in replacements for the reference's include/aa_zero_dm.hpp.
// zero_dm (include/aa_zero_dm.hpp:22-23,
// src/aa_zero_dm.cpp:9-37): the zero-DM filter on the reference's u16
// [t][c] device input, in place, on the legacy default stream, synchronous.
// Per time sample x' = (x - mean over channels) + normalization_factor[c]
// in float, clamped to [0, 2^nbits - 1], truncated.  The 4-argument form
// uses (2^nbits - 1) / 2 for every channel.
#ifndef ASTRO_ACCELERATE_AA_ZERO_DM_HPP
#define ASTRO_ACCELERATE_AA_ZERO_DM_HPP

#include <cstddef>

namespace astroaccelerate {

void zero_dm(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits,
             float *normalization_factor);
void zero_dm(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits);

// Input DDTR normalisation (include/aa_zero_dm.hpp:26, src/aa_zero_dm.cpp:
// 49-60): each column of a u16 array of `nChannels` rows of `nTimesamples`
// values moved so that its mean over the rows is (2^nbits - 1)/2, clamped
// to [0, 2^nbits - 1], truncated.  The reference's runner calls it with
// (d_input, nchans, nTimesamples, nbits), which normalises each channel of
// the [t][c] chunk over time.
void zero_dm_normalization_dm(unsigned short *d_input, size_t nTimesamples, size_t nChannels, int nbits);
// Output DDTR normalisation (include/aa_zero_dm.hpp:25, src/aa_zero_dm.cpp:
// 62-74): each time sample of a [nDMs][nTimesamples] float plane moved so
// that its mean over DMs is 127.5.
void post_DDTR_normalization(float *d_input, size_t nTimesamples, size_t nDMs);

}  // namespace astroaccelerate
#endif
