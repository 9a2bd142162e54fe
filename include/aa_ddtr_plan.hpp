// User dedispersion plan: same interface as the reference
// (include/aa_ddtr_plan.hpp:17-158).  A list of DM ranges {low, high, step,
// inBin, outBin}, the dispersion power (default 2.0) and an optional
// bandpass vector.
#ifndef ASTRO_ACCELERATE_AA_DDTR_PLAN_HPP
#define ASTRO_ACCELERATE_AA_DDTR_PLAN_HPP

#include <cstddef>
#include <vector>

namespace astroaccelerate {

class aa_ddtr_plan {
public:
    aa_ddtr_plan() = default;

    struct dm {
        float low;
        float high;
        float step;
        int inBin;
        int outBin;  // kept for interface compatibility; unused by dedispersion
    };

    bool add_dm(const float &low, const float &high, const float &step, const int &inBin,
                const int &outBin) {
        m_ranges.push_back(dm{low, high, step, inBin, outBin});
        return true;
    }
    bool add_dm(aa_ddtr_plan::dm &range) {
        m_ranges.push_back(range);
        return true;
    }
    bool add_dm(std::vector<aa_ddtr_plan::dm> &ranges) {
        m_ranges.insert(m_ranges.end(), ranges.begin(), ranges.end());
        return true;
    }

    size_t range() const { return m_ranges.size(); }

    // Out-of-range indices yield an all-zero range, as in the reference.
    const dm user_dm(const size_t &i) const {
        if (i < m_ranges.size()) return m_ranges[i];
        return dm{0.0f, 0.0f, 0.0f, 0, 0};
    }

    bool set_power(const float &power) {
        m_power = power;
        return true;
    }
    float power() const { return m_power; }

    bool set_enable_msd_baseline_noise(const bool flag) {
        m_msd_baseline_noise = flag;
        return true;
    }
    bool enable_msd_baseline_noise() const { return m_msd_baseline_noise; }

    bool bind_bandpass_normalization(const float *values, int count) {
        if (count <= 0) return false;
        m_bandpass.assign(values, values + count);
        return true;
    }
    bool bind_bandpass_normalization_vector(std::vector<float> &values) {
        m_bandpass = values;
        return true;
    }
    size_t bandpass_normalization_size() const { return m_bandpass.size(); }
    const float *bandpass_data_pointer() const { return m_bandpass.empty() ? nullptr : m_bandpass.data(); }

private:
    std::vector<dm> m_ranges;
    float m_power = 2.0f;
    bool m_msd_baseline_noise = false;
    std::vector<float> m_bandpass;
};

}  // namespace astroaccelerate
#endif
