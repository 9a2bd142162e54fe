// Drop-in replacements for the reference's device-level DDTR functions, with
// their exact signatures, so that a runner written against the reference
// (include/aa_permitted_pipelines_generic.hpp:693,741,786,799,812) links
// against this library unchanged.  They operate on the reference's device
// layout (u16 samples, [c][t] after the corner turn) on the legacy default
// stream, and keep the reference's per-process "current constants" (set by
// load_chunk_data / set_dedispersion_constants) in host memory instead of
// __constant__ symbols.
#ifndef ASTRO_ACCELERATE_AA_DEDISPERSE_HPP
#define ASTRO_ACCELERATE_AA_DEDISPERSE_HPP

namespace astroaccelerate {

// include/aa_dedisperse.hpp:20.  Returns 0 on success, 1 on a device error.
// failsafe and d_dm_shifts are accepted and ignored: the gfx950 kernels are
// valid for every geometry and read the dmshift table uploaded by
// load_chunk_data.  nbits must be 8 (or 16 with 8-bit values).
int dedisperse(int i, int t_processed, int *inBin, float *dmshifts, unsigned short *d_input,
               float *d_output, float *d_dm_shifts, int nchans, float *tsamp, float *dm_low,
               float *dm_step, int const *const ndms, int nbits, int failsafe);

}  // namespace astroaccelerate
#endif
