// Runner status codes, as in the reference (include/aa_pipeline_runner.hpp:25-30).
#ifndef ASTRO_ACCELERATE_AA_PIPELINE_RUNNER_HPP
#define ASTRO_ACCELERATE_AA_PIPELINE_RUNNER_HPP

namespace astroaccelerate {
class aa_pipeline_runner {
public:
    enum class status : int { error = -1, finished = 0, has_more = 1, finished_component = 2 };
};
}  // namespace astroaccelerate
#endif
