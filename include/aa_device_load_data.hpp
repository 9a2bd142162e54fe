// Drop-in replacements for include/aa_device_load_data.hpp
// (src/aa_device_load_data.cu:39-61): H2D copy of one chunk and the
// per-range dedispersion constants.
#ifndef ASTRO_ACCELERATE_AA_DEVICE_LOAD_DATA_HPP
#define ASTRO_ACCELERATE_AA_DEVICE_LOAD_DATA_HPP

namespace astroaccelerate {

void load_chunk_data(unsigned short *device_pointer, unsigned short const *const host_pointer,
                     int t_processed, int maxshift, int nchans, float *dmshifts,
                     float *d_dm_shifts, int nbits);
void set_dedispersion_constants(int t_processed, int maxshift);

}  // namespace astroaccelerate
#endif
