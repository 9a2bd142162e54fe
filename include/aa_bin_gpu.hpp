// Drop-in replacement for bin_gpu (src/aa_bin_gpu.cpp:8-44): one x2
// floor-average of u16 rows of nsamp samples, packed into rows of nsamp/2,
// in place, using d_output (nchans*nsamp floats) as scratch.
#ifndef ASTRO_ACCELERATE_AA_BIN_GPU_HPP
#define ASTRO_ACCELERATE_AA_BIN_GPU_HPP

namespace astroaccelerate {

void bin_gpu(unsigned short *const d_input, float *const d_output, const int nchans, const int nsamp);

}  // namespace astroaccelerate
#endif
