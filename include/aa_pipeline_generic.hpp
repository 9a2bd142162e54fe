// Boilerplate runner over aa_pipeline_api<T>: the reference's
// aa_pipeline_generic (include/aa_pipeline_generic.hpp:22-150), restricted
// to the dedispersion component.  It builds the pipeline from the selected
// components, binds an aa_ddtr_plan of dm_ranges, and runs every (time
// chunk, range) on device 0.  The analysis and periodicity parameters are
// kept for signature compatibility; those components are outside this
// library, so a pipeline that selects them is not ready and does not run
// (logged, as any pipeline that cannot start).
#ifndef ASTRO_ACCELERATE_AA_PIPELINE_GENERIC_HPP
#define ASTRO_ACCELERATE_AA_PIPELINE_GENERIC_HPP

#include <vector>

#include "aa_ddtr_plan.hpp"
#include "aa_device_info.hpp"
#include "aa_filterbank_metadata.hpp"
#include "aa_log.hpp"
#include "aa_pipeline.hpp"
#include "aa_pipeline_api.hpp"

namespace astroaccelerate {

template <typename T>
void aa_pipeline_generic(const std::vector<aa_pipeline::component> &selected_components,
                         const aa_pipeline::pipeline_option &pipeline_options,
                         const aa_filterbank_metadata &filterbank_data, std::vector<aa_ddtr_plan::dm> dm_ranges,
                         T const *const input_data, const float &analysis_sigma_cutoff = 0.0,
                         const float &analysis_sigma_constant = 0.0,
                         const float &analysis_max_boxcar_width_in_sec = 0.0,
                         const bool &analysis_enable_msd_baseline_noise_algorithm = false,
                         const float &periodicity_sigma_cutoff = 0.0,
                         const float &periodicity_sigma_outlier_rejection_threshold = 0.0,
                         const int &periodicity_nHarmonics = 0, const bool &periodicity_candidate_algorithm = false,
                         const bool &periodicity_enable_outlier_rejection = false) {
    (void)analysis_sigma_cutoff;
    (void)analysis_sigma_constant;
    (void)analysis_max_boxcar_width_in_sec;
    (void)analysis_enable_msd_baseline_noise_algorithm;
    (void)periodicity_sigma_cutoff;
    (void)periodicity_sigma_outlier_rejection_threshold;
    (void)periodicity_nHarmonics;
    (void)periodicity_candidate_algorithm;
    (void)periodicity_enable_outlier_rejection;
    aa_pipeline::pipeline the_pipeline(selected_components.begin(), selected_components.end());
    aa_device_info selected_device(0);
    aa_pipeline_api<T> pipeline_manager(the_pipeline, pipeline_options, filterbank_data, input_data, selected_device);
    aa_ddtr_plan ddtr_plan;
    ddtr_plan.add_dm(dm_ranges);
    if (pipeline_manager.bind(ddtr_plan))
        LOG(log_level::notice, "ddtr_plan bound successfully.");
    else
        LOG(log_level::error, "Could not bind ddtr_plan.");
    if (pipeline_manager.ready())
        LOG(log_level::notice, "Pipeline is ready.");
    else
        LOG(log_level::notice, "Pipeline is not ready.");
    if (pipeline_manager.run())
        LOG(log_level::notice, "The pipeline finished successfully.");
    else
        LOG(log_level::notice, "The pipeline could not start or had errors.");
}

}  // namespace astroaccelerate
#endif
