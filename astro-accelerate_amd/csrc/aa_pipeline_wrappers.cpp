// dedisperse_telescope_data: the dedispersion-only pipeline over u16 input
// (src/aa_pipeline_wrapper_functions.cpp:6-15 in the reference).
#include "aa_pipeline_wrapper_functions.hpp"

namespace astroaccelerate {

void dedisperse_telescope_data(const aa_filterbank_metadata &filterbank_data,
                               const aa_pipeline::pipeline_option &pipeline_options,
                               std::vector<aa_ddtr_plan::dm> dm_ranges, unsigned short const *const input_data) {
    const std::vector<aa_pipeline::component> selected_components = {aa_pipeline::component::dedispersion};
    aa_pipeline_generic<unsigned short>(selected_components, pipeline_options, filterbank_data, dm_ranges,
                                        input_data);
}

void dedisperse_telescope_data(const aa_filterbank_metadata &filterbank_data,
                               const aa_pipeline::pipeline_option &pipeline_options,
                               std::vector<aa_ddtr_plan::dm> dm_ranges,
                               const std::vector<unsigned short> &input_data) {
    dedisperse_telescope_data(filterbank_data, pipeline_options, dm_ranges, input_data.data());
}

}  // namespace astroaccelerate
