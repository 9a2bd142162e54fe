// Dedispersion strategy: same interface and arithmetic as the reference
// (include/aa_ddtr_strategy.hpp:33-228, src/aa_ddtr_strategy.cpp:27-428):
// dmshift table, DM grid rounded to multiples of 64 trials, redefined range
// starts, maxshift, t_processed per range and time chunk, memory requests.
#ifndef ASTRO_ACCELERATE_AA_DDTR_STRATEGY_HPP
#define ASTRO_ACCELERATE_AA_DDTR_STRATEGY_HPP

#include <cstddef>
#include <string>
#include <vector>

#include "aa_ddtr_plan.hpp"
#include "aa_device_info.hpp"
#include "aa_filterbank_metadata.hpp"

namespace astroaccelerate {

// Tile constants that the reference bakes into the plan granularity
// (include/aa_params.hpp:95-99); kept so plans match it exactly.
constexpr int SDIVINT = 8;
constexpr int SNUMREG = 14;
constexpr int SDIVINDM = 64;
constexpr int UNROLLS = 4;

class aa_strategy {
public:
    virtual ~aa_strategy() {}
    virtual std::string name() const = 0;
    virtual bool setup() = 0;
};

class aa_ddtr_strategy : public aa_strategy {
public:
    aa_ddtr_strategy();
    aa_ddtr_strategy(const aa_ddtr_plan &plan, const aa_filterbank_metadata &metadata,
                     const size_t &free_memory, const bool &enable_analysis,
                     aa_device_info *selected_device);
    virtual ~aa_ddtr_strategy() {}

    std::string name() const override { return "ddtr_strategy"; }
    bool setup() override;

    int maxshift() const { return m_maxshift; }
    const std::vector<std::vector<int>> &t_processed() const { return m_t_processed; }
    size_t nProcessedTimesamples() {
        size_t sum = 0;
        if (!m_t_processed.empty())
            for (int v : m_t_processed[0]) sum += (size_t)v;
        return sum;
    }
    std::vector<float> dmshifts() const { return m_dmshifts; }
    size_t bandpass_normalization_size() const { return m_bandpass_normalization.size(); }
    const float *bandpass_normalization_pointer() const { return m_bandpass_normalization.data(); }
    std::vector<float> bandpass_normalization() const { return m_bandpass_normalization; }
    const aa_ddtr_plan::dm dm(const size_t &i) const { return str_dm.at(i); }
    size_t get_nRanges() const { return str_dm.size(); }
    size_t ndms_size() const { return m_ndms.size(); }
    int ndms(const size_t &i) const { return m_ndms.at(i); }
    int max_ndms() const { return m_max_ndms; }
    const int *ndms_data() const { return m_ndms.data(); }
    // Truncates to int like the reference (include/aa_ddtr_strategy.hpp:126-128);
    // use dm(range).low for the exact float.
    int dm_low(const int range) const { return (int)str_dm.at(range).low; }
    int num_tchunks() const { return m_num_tchunks; }
    float power() const { return m_power; }
    int enable_msd_baseline_noise() const { return m_enable_msd_baseline_noise ? 1 : 0; }
    bool ready() const { return m_ready; }
    const aa_filterbank_metadata metadata() const { return m_metadata; }
    bool configured_for_analysis() const { return m_configured_for_analysis; }
    int strategy_case() const { return m_case; }  // 1..4, src/aa_ddtr_strategy.cpp:155-160
    static bool print_info(const aa_ddtr_strategy &strategy);

private:
    bool strategy(const aa_ddtr_plan &plan, const size_t &free_memory, const bool &enable_analysis);

    bool m_ready = false;
    bool m_strategy_already_calculated = false;
    aa_device_info *m_selected_device = nullptr;
    bool m_configured_for_analysis = false;
    bool is_setup = false;
    aa_filterbank_metadata m_metadata;
    std::vector<int> m_ndms;
    std::vector<float> m_dmshifts;
    std::vector<aa_ddtr_plan::dm> str_dm;
    int m_maxshift = 0;
    int m_num_tchunks = 0;
    int m_total_ndms = 0;
    float m_max_dm = 0.0f;
    int m_maxshift_high = 0;
    int m_max_ndms = 0;
    float m_power = 0.0f;
    std::vector<std::vector<int>> m_t_processed;
    bool m_enable_msd_baseline_noise = false;
    std::vector<float> m_bandpass_normalization;
    int m_case = 0;
};

}  // namespace astroaccelerate
#endif
