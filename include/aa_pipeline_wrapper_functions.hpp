// C-style entry points of the dedispersion pipeline: the reference's
// include/aa_pipeline_wrapper_functions.hpp (src/aa_pipeline_wrapper_functions.cpp:6-15),
// dedispersion only.  Each runs aa_pipeline_generic on u16 [t][c] samples
// (8-bit values, as aa_sigproc_input::read_signal() provides them).
#ifndef ASTRO_ACCELERATE_AA_PIPELINE_WRAPPER_FUNCTIONS_HPP
#define ASTRO_ACCELERATE_AA_PIPELINE_WRAPPER_FUNCTIONS_HPP

#include <vector>

#include "aa_ddtr_plan.hpp"
#include "aa_filterbank_metadata.hpp"
#include "aa_pipeline.hpp"
#include "aa_pipeline_generic.hpp"

namespace astroaccelerate {

// Replaces src/aa_pipeline_wrapper_functions.cpp:6-9 (raw pointer input).
void dedisperse_telescope_data(const aa_filterbank_metadata &filterbank_data,
                               const aa_pipeline::pipeline_option &pipeline_options,
                               std::vector<aa_ddtr_plan::dm> dm_ranges, unsigned short const *const input_data);

// Replaces src/aa_pipeline_wrapper_functions.cpp:12-15 (std::vector input).
void dedisperse_telescope_data(const aa_filterbank_metadata &filterbank_data,
                               const aa_pipeline::pipeline_option &pipeline_options,
                               std::vector<aa_ddtr_plan::dm> dm_ranges, const std::vector<unsigned short> &input_data);

}  // namespace astroaccelerate
#endif
