// Pipeline component and option names, as in the reference
// (include/aa_pipeline.hpp:30-62).  This library runs the dedispersion
// component; requests for the other components are rejected at ready().
#ifndef ASTRO_ACCELERATE_AA_PIPELINE_HPP
#define ASTRO_ACCELERATE_AA_PIPELINE_HPP

#include <set>
#include <string>

namespace astroaccelerate {
namespace aa_pipeline {

enum class component : int { empty = 0, dedispersion, analysis, periodicity, fdas, jerk };

enum class component_option : int {
    empty = 0,
    zero_dm,
    zero_dm_with_outliers,
    input_DDTR_normalization,
    output_DDTR_normalization,
    set_bandpass_average,
    old_rfi,
    copy_ddtr_data_to_host,
    output_dmt,
    msd_baseline_noise,
    dered,
    candidate_algorithm,
    candidate_filtering,
    output_ffdot_plan,
    output_fdas_list,
    fdas_custom_fft,
    fdas_inbin,
    fdas_norm,
    timelog_export_to_file
};

typedef std::set<component> pipeline;
typedef std::set<component_option> pipeline_option;

inline const std::string component_name(const component &c) {
    switch (c) {
        case component::empty: return "empty";
        case component::dedispersion: return "De-dispersion";
        case component::analysis: return "Single pulse detection";
        case component::periodicity: return "Periodicity search";
        case component::fdas: return "Fourier domain acceleration search";
        case component::jerk: return "JERK search";
    }
    return "unknown";
}

}  // namespace aa_pipeline
}  // namespace astroaccelerate
#endif
