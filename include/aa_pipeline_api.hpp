// Pipeline front end for the dedispersion path: the reference's
// aa_pipeline_api<T> (include/aa_pipeline_api.hpp:91-1196) restricted to
// the dedispersion component.  T is the host sample type: unsigned short
// (the reference's widened 8-bit samples) or unsigned char.
//
// Sequencing follows the reference's generic runner
// (include/aa_permitted_pipelines_generic.hpp:606-913): each run(status)
// call processes one (time chunk, DM range); the range's DM-time plane is
// copied into output_buffer()[range][dm][inc/inBin + t] when the option
// copy_ddtr_data_to_host is set.  The device work runs through the
// library's HIP engine.
#ifndef ASTRO_ACCELERATE_AA_PIPELINE_API_HPP
#define ASTRO_ACCELERATE_AA_PIPELINE_API_HPP

#include <cstddef>
#include <memory>
#include <type_traits>

#include "aa_ddtr_hip.h"
#include "aa_ddtr_plan.hpp"
#include "aa_ddtr_strategy.hpp"
#include "aa_device_info.hpp"
#include "aa_filterbank_metadata.hpp"
#include "aa_pipeline.hpp"
#include "aa_pipeline_runner.hpp"

namespace astroaccelerate {

// Non-template implementation inside libaa_ddtr_hip.so.
class aa_ddtr_pipeline_impl;

class aa_ddtr_pipeline {
public:
    aa_ddtr_pipeline(const aa_pipeline::pipeline &requested_pipeline,
                     const aa_pipeline::pipeline_option &pipeline_options,
                     const aa_filterbank_metadata &filterbank_metadata, const void *input_data,
                     int input_bytes_per_sample, aa_device_info &card_info);
    ~aa_ddtr_pipeline();
    aa_ddtr_pipeline(const aa_ddtr_pipeline &) = delete;
    aa_ddtr_pipeline &operator=(const aa_ddtr_pipeline &) = delete;

    bool bind(const aa_ddtr_plan &plan);
    aa_ddtr_strategy ddtr_strategy();
    size_t get_nRanges();
    const int *get_ndms_array();
    int dm_low(const int range);
    int total_computed_samples();
    int samples_in_current(int i, int j);
    aa_filterbank_metadata metadata();
    bool ready();
    bool run();
    bool run(aa_pipeline_runner::status &status_code);
    int get_current_range();
    int get_current_tchunk();
    long int get_current_inc();
    bool cleanup();
    float ***output_buffer();
    // Hands every (time chunk, range) plane to fn on the device, where the
    // reference's runner calls analysis_GPU
    // (include/aa_permitted_pipelines_generic.hpp:848-889); NULL clears it.
    void set_ddtr_consumer(aa_hip_ddtr_consumer fn, void *user);

private:
    std::unique_ptr<aa_ddtr_pipeline_impl> m_impl;
};

template <typename T>
class aa_pipeline_api {
    static_assert(std::is_same<T, unsigned short>::value || std::is_same<T, unsigned char>::value,
                  "aa_pipeline_api supports unsigned short or unsigned char samples");

public:
    aa_pipeline_api(const aa_pipeline::pipeline &requested_pipeline,
                    const aa_pipeline::pipeline_option &pipeline_options,
                    const aa_filterbank_metadata &filterbank_metadata, T const *const input_data,
                    aa_device_info &card_info)
        : m_p(requested_pipeline, pipeline_options, filterbank_metadata, input_data, (int)sizeof(T),
              card_info) {}

    bool bind(aa_ddtr_plan plan) { return m_p.bind(plan); }
    aa_ddtr_strategy ddtr_strategy() { return m_p.ddtr_strategy(); }
    size_t get_nRanges() { return m_p.get_nRanges(); }
    const int *get_ndms_array() { return m_p.get_ndms_array(); }
    int dm_low(const int range) { return m_p.dm_low(range); }
    int total_computed_samples() { return m_p.total_computed_samples(); }
    int samples_in_current(int i, int j) { return m_p.samples_in_current(i, j); }
    aa_filterbank_metadata metadata() { return m_p.metadata(); }
    bool ready() { return m_p.ready(); }
    bool run() { return m_p.run(); }
    bool run(aa_pipeline_runner::status &status_code) { return m_p.run(status_code); }
    int get_current_range() { return m_p.get_current_range(); }
    int get_current_tchunk() { return m_p.get_current_tchunk(); }
    long int get_current_inc() { return m_p.get_current_inc(); }
    bool cleanup() { return m_p.cleanup(); }
    float ***output_buffer() { return m_p.output_buffer(); }
    // No reference counterpart: the device plane of every (time chunk, range)
    // goes to fn (see aa_hip_ddtr_consumer), the single-pulse search's entry.
    void set_ddtr_consumer(aa_hip_ddtr_consumer fn, void *user) { m_p.set_ddtr_consumer(fn, user); }
    // Not implemented in the reference either (include/aa_pipeline_api.hpp:1188-1195).
    bool handoff(aa_pipeline_api &) { return true; }

private:
    aa_ddtr_pipeline m_p;
};

}  // namespace astroaccelerate
#endif
