// Dedispersion strategy (host only).  Restates src/aa_ddtr_strategy.cpp:27-428
// so that plans, DM grids and chunking agree with the reference bit for bit:
// every float/double/int conversion happens where the reference performs it.
// Built with -ffp-contract=off so no product below is fused into an add.
#include <cmath>
#include <cstdio>
#include <cstdint>

#include "aa_ddtr_strategy.hpp"
#include "aa_log.hpp"

namespace astroaccelerate {

namespace {
constexpr unsigned long long kMaxOutputSize = 17179869184ULL;  // src/aa_ddtr_strategy.cpp:7
constexpr size_t kReservedBytes = 104857600;                   // :77

// ceil((top * dmshift_max) / tsamp) with float operands (:113,125,136).
int dispersive_extent(float top_dm, float dmshift_max, float tsamp) {
    return (int)std::ceil((top_dm * dmshift_max) / tsamp);
}
}  // namespace

aa_ddtr_strategy::aa_ddtr_strategy() = default;

aa_ddtr_strategy::aa_ddtr_strategy(const aa_ddtr_plan &plan, const aa_filterbank_metadata &metadata,
                                   const size_t &free_memory, const bool &enable_analysis,
                                   aa_device_info *selected_device)
    : m_selected_device(selected_device), m_configured_for_analysis(enable_analysis),
      m_metadata(metadata), m_power(plan.power()),
      m_enable_msd_baseline_noise(plan.enable_msd_baseline_noise()) {
    strategy(plan, free_memory, enable_analysis);
}

bool aa_ddtr_strategy::strategy(const aa_ddtr_plan &plan, const size_t &free_memory,
                                const bool &enable_analysis) {
    if (m_strategy_already_calculated) return true;

    const int nchans = m_metadata.nchans();
    const int nsamp = m_metadata.nsamples();
    const float fch1 = (float)m_metadata.fch1();
    const float foff = (float)m_metadata.foff();
    const float tsamp = (float)m_metadata.tsamp();
    const int nbits = m_metadata.nbits();

    // Bandpass used by zero-DM filtering (:48-60).
    if (plan.bandpass_normalization_size() == (size_t)nchans && nchans > 0) {
        m_bandpass_normalization.assign(plan.bandpass_data_pointer(), plan.bandpass_data_pointer() + nchans);
    } else {
        m_bandpass_normalization.assign((size_t)std::max(nchans, 0), (float)(((1 << nbits) - 1.0) / 2.0));
    }

    const size_t nranges = plan.range();
    if (nranges == 0 || nchans <= 0) {
        LOG(log_level::error, "ddtr strategy: the plan has no DM ranges (or nchans <= 0).");
        return false;
    }

    // dmshift table (:81-92): the channel frequency is a float multiply and a
    // float add; the power-2 branch takes reciprocal squares in double.
    m_dmshifts.resize((size_t)nchans);
    for (int c = 0; c < nchans; ++c) {
        const float f = fch1 + foff * (float)c;
        double inv;
        if (m_power != 2.0f)
            inv = (1.0 / (double)std::pow(f, m_power)) - (1.0 / (double)std::pow(fch1, m_power));
        else
            inv = (1.0 / std::pow((double)f, (double)m_power)) - (1.0 / std::pow((double)fch1, (double)m_power));
        m_dmshifts[(size_t)c] = (float)((double)4148.741601f * inv);
    }

    // DM trials per range, rounded up to whole blocks of 64 (:94-101).
    m_ndms.resize(nranges);
    m_max_ndms = 0;
    m_total_ndms = 0;
    for (size_t i = 0; i < nranges; ++i) {
        const aa_ddtr_plan::dm u = plan.user_dm(i);
        const int blocks = (((int)((u.high - u.low) / u.step)) + SDIVINDM) / SDIVINDM;
        float whole;
        std::modf((float)blocks, &whole);
        m_ndms[i] = (int)whole * SDIVINDM;
        m_max_ndms = std::max(m_max_ndms, m_ndms[i]);
        m_total_ndms += m_ndms[i];
    }

    // Ranges redefined so that each starts where the previous one ends (:104-110).
    str_dm.assign(nranges, aa_ddtr_plan::dm{0, 0, 0, 0, 0});
    str_dm[0].low = plan.user_dm(0).low;
    str_dm[0].high = str_dm[0].low + ((float)m_ndms[0] * plan.user_dm(0).step);
    str_dm[0].step = plan.user_dm(0).step;
    const float modulo = (float)(SDIVINT * 2 * SNUMREG);
    auto binned_maxshift = [&](size_t r) {
        const float top = str_dm[r].low + str_dm[r].step * (float)m_ndms[r];
        int m = dispersive_extent(top, m_dmshifts[(size_t)nchans - 1], tsamp);
        const int ib = plan.user_dm(r).inBin;
        m = (int)(((float)m / (float)ib + modulo) / modulo);
        return m * (SDIVINT * 2 * SNUMREG) * ib;
    };
    m_maxshift_high = 0;
    for (size_t i = 1; i < nranges; ++i) {
        str_dm[i].low = str_dm[i - 1].high;
        str_dm[i].high = str_dm[i].low + (float)m_ndms[i] * plan.user_dm(i).step;
        str_dm[i].step = plan.user_dm(i).step;
        if (plan.user_dm(i - 1).inBin > 1) m_maxshift_high = std::max(m_maxshift_high, binned_maxshift(i - 1));
    }
    if (plan.user_dm(nranges - 1).inBin > 1)
        m_maxshift_high = std::max(m_maxshift_high, binned_maxshift(nranges - 1));
    if (m_maxshift_high == 0) {
        const size_t r = nranges - 1;
        const float top = str_dm[r].low + str_dm[r].step * (float)m_ndms[r];
        m_maxshift_high = dispersive_extent(top, m_dmshifts[(size_t)nchans - 1], tsamp);
    }
    m_max_dm = std::ceil(str_dm[nranges - 1].high);
    m_maxshift = m_maxshift_high + (SNUMREG * 2 * SDIVINT);
    for (size_t i = 0; i < nranges; ++i) {  // :420-423
        str_dm[i].inBin = plan.user_dm(i).inBin;
        str_dm[i].outBin = plan.user_dm(i).outBin;
    }
    if (m_maxshift >= nsamp) {
        LOG(log_level::error, "Your maximum DM trial exceeds the number of samples you have. Reduce your maximum DM trial.");
        return false;
    }

    // Time chunking (:155-389).  Granularity 224 samples (448 for 4-bit).
    const size_t gpu_memory = free_memory - kReservedBytes;
    const double spdt_fraction = 3.0 / 4.0;
    const int gran = (nbits == 4) ? (SDIVINT * 4 * SNUMREG) : (SDIVINT * 2 * SNUMREG);
    const int inbin_last = plan.user_dm(nranges - 1).inBin;
    const size_t spdt = enable_analysis ? (size_t)(sizeof(float) * (size_t)m_max_ndms * spdt_fraction) : 0;

    unsigned int max_tsamps;
    bool fits;
    if (nchans < m_max_ndms) {
        max_tsamps = (unsigned int)(gpu_memory / (sizeof(unsigned short) * (size_t)nchans +
                                                  sizeof(float) * (size_t)m_max_ndms + spdt));
        if (((size_t)max_tsamps) * sizeof(float) * (size_t)m_max_ndms > kMaxOutputSize)
            max_tsamps = (unsigned int)(kMaxOutputSize / (sizeof(float) * (size_t)m_max_ndms));
        if ((unsigned int)m_maxshift > max_tsamps) {
            LOG(log_level::error, "The selected GPU does not have enough memory for this number of dispersion trials.");
            return false;
        }
        fits = (unsigned int)nsamp < max_tsamps;
        m_case = fits ? 1 : 3;
    } else {
        max_tsamps = (unsigned int)(gpu_memory / ((size_t)nchans * (sizeof(float) + sizeof(unsigned short)) + spdt));
        if (((size_t)max_tsamps) * sizeof(float) * (size_t)nchans > kMaxOutputSize)
            max_tsamps = (unsigned int)(kMaxOutputSize / (sizeof(float) * (size_t)nchans));
        if (m_maxshift > (int)max_tsamps) {
            LOG(log_level::error, "The selected GPU does not have enough memory for this number of dispersion trials.");
            return false;
        }
        fits = (unsigned int)nsamp < max_tsamps;
        m_case = fits ? 2 : 4;
    }

    // Per-range share of a chunk of `local` raw samples: a whole number of
    // granules of the range's binned time axis.
    auto share = [&](float local, int ib) { return (int)std::floor((local / (float)ib) / (float)gran) * gran; };

    m_t_processed.assign(nranges, std::vector<int>());
    if (m_case == 1) {
        unsigned long local = (unsigned long)std::floor(((double)(nsamp - m_maxshift) / (double)inbin_last) / (double)gran);
        local = local * (unsigned long)gran * (unsigned long)inbin_last;
        for (size_t i = 0; i < nranges; ++i) m_t_processed[i].assign(1, share((float)local, plan.user_dm(i).inBin));
        m_num_tchunks = 1;
    } else if (m_case == 2) {
        int local = (int)std::floor(((float)(nsamp - m_maxshift) / (float)inbin_last) / (float)gran);
        local = local * gran * inbin_last;
        for (size_t i = 0; i < nranges; ++i) m_t_processed[i].assign(1, share((float)local, plan.user_dm(i).inBin));
        m_num_tchunks = 1;
    } else {
        const int samp_block_size = (int)(max_tsamps - (unsigned int)m_maxshift);
        int local = (int)std::floor(((float)samp_block_size / (float)inbin_last) / (float)gran);
        local = local * gran * inbin_last;
        if (local <= 0) {
            LOG(log_level::error, "ddtr strategy: no whole chunk fits in device memory.");
            return false;
        }
        const int num_blocks = (int)(((float)(nsamp - m_maxshift)) / ((float)local));
        int remainder = nsamp - (num_blocks * local) - m_maxshift;
        // The reference divides the floored bin count by a float and
        // truncates the quotient on assignment (:238, :354).
        remainder = (int)((float)((int)std::floor((float)remainder / (float)inbin_last)) / (float)gran);
        remainder = remainder * gran * inbin_last;
        const int rem_block = remainder > 0 ? 1 : 0;
        for (size_t i = 0; i < nranges; ++i) {
            const int ib = plan.user_dm(i).inBin;
            m_t_processed[i].assign((size_t)(num_blocks + rem_block), share((float)local, ib));
            if (rem_block) m_t_processed[i][(size_t)num_blocks] = share((float)remainder, ib);
        }
        m_num_tchunks = num_blocks + rem_block;
    }
    if (m_num_tchunks <= 0) {
        LOG(log_level::error, "ddtr strategy: zero time chunks.");
        return false;
    }

    // Memory requests (:401-415); only the output request decides readiness.
    const size_t time_samps = (size_t)((long)m_t_processed[0][0] + m_maxshift);
    if (m_selected_device) {
        m_selected_device->request_memory(time_samps * (size_t)nchans * sizeof(unsigned short));
        const size_t out_rows = (size_t)(nchans < m_max_ndms ? m_max_ndms : nchans);
        if (!m_selected_device->request_memory(time_samps * out_rows * sizeof(float))) {
            LOG(log_level::error, "ddtr strategy: could not request output memory.");
            return false;
        }
    }

    m_ready = true;
    m_strategy_already_calculated = true;
    return true;
}

bool aa_ddtr_strategy::setup() {
    is_setup = true;
    return is_setup && ready();
}

bool aa_ddtr_strategy::print_info(const aa_ddtr_strategy &s) {
    LOG(log_level::dev_debug, "DDTR STRATEGY INFORMATION:");
    LOG(log_level::dev_debug, "ddtr dm ranges:\t\t" + std::to_string(s.get_nRanges()));
    for (size_t i = 0; i < s.get_nRanges(); ++i) {
        const aa_ddtr_plan::dm r = s.dm(i);
        LOG(log_level::dev_debug, "     dm (low,high,step,inBin,outBin) " + std::to_string(r.low) + "," +
                                      std::to_string(r.high) + "," + std::to_string(r.step) + "," +
                                      std::to_string(r.inBin) + "," + std::to_string(r.outBin) +
                                      "  ndms " + std::to_string(s.ndms(i)));
    }
    LOG(log_level::dev_debug, "ddtr maxshift:\t\t" + std::to_string(s.maxshift()));
    LOG(log_level::dev_debug, "ddtr num_tchunks:\t" + std::to_string(s.num_tchunks()));
    LOG(log_level::dev_debug, "ddtr max_ndms:\t\t" + std::to_string(s.max_ndms()));
    return true;
}

}  // namespace astroaccelerate
