// Dedispersion pipeline behind aa_pipeline_api<T> and aa_py_pipeline_api.
// Sequencing follows the reference's generic runner
// (include/aa_permitted_pipelines_generic.hpp:606-913): one (time chunk, DM
// range) per run(status) call; range 0 of a chunk uploads t_processed[0] +
// maxshift samples starting at sample `inc` and corner-turns them; every
// range dedisperses on its binning level and, with copy_ddtr_data_to_host,
// lands in output_buffer()[range][dm][inc/inBin + t]
// (src/aa_device_save_data.cu:41-50).  When every chunk is done the next call
// reports finished_component and frees device memory, and the call after
// that returns false with status finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "aa_ddtr_hip.h"
#include "aa_log.hpp"
#include "aa_pipeline_api.hpp"
#include "ddtr_engine.hpp"
#include "ddtr_kernels.h"

namespace aa_hip {

// set_bandpass_average (src/aa_host_cpu_msd.cpp:7-39, called at
// include/aa_permitted_pipelines_generic.hpp:533-551 with the [t][c] host
// input): out[i] = (float)(sum / nsamp) over input[i*nsamp .. i*nsamp +
// nsamp) for i < nchans, a double sum of integers (exact, so integer sums
// give the same float).  The runs are nsamp consecutive samples of the [t][c]
// buffer, as the reference indexes it, not channel i.
void bandpass_average(const void *in_tc, int bytes_per_sample, size_t nchans, size_t nsamp, float *out) {
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned k = 0; k < nt; ++k) {
        pool.emplace_back([=] {
            for (size_t i = k; i < nchans; i += nt) {
                uint64_t sum = 0;
                if (bytes_per_sample == 1) {
                    const uint8_t *p = (const uint8_t *)in_tc + i * nsamp;
                    for (size_t j = 0; j < nsamp; ++j) sum += p[j];
                } else {
                    const uint16_t *p = (const uint16_t *)in_tc + i * nsamp;
                    for (size_t j = 0; j < nsamp; ++j) sum += p[j];
                }
                out[i] = (float)((double)sum / (double)nsamp);
            }
        });
    }
    for (auto &t : pool) t.join();
}

}  // namespace aa_hip

namespace astroaccelerate {

class aa_ddtr_pipeline_impl {
public:
    aa_ddtr_pipeline_impl(const aa_pipeline::pipeline &components, const aa_pipeline::pipeline_option &options,
                          const aa_filterbank_metadata &metadata, const void *input, int bytes_per_sample,
                          aa_device_info &card)
        : m_components(components), m_options(options), m_metadata(metadata), m_input(input),
          m_bytes(bytes_per_sample), m_card(card) {}

    ~aa_ddtr_pipeline_impl() {
        release_device();
        free_host_output();
    }

    bool bind(const aa_ddtr_plan &plan) {
        m_ready = false;
        if (m_components.find(aa_pipeline::component::dedispersion) == m_components.end()) {
            LOG(log_level::error, "bind: the pipeline does not contain dedispersion.");
            return false;
        }
        if (m_bound) {
            LOG(log_level::error, "bind: a dedispersion plan is already bound.");
            return false;
        }
        const bool analysis = m_components.find(aa_pipeline::component::analysis) != m_components.end();
        aa_ddtr_strategy s(plan, m_metadata, m_card.free_memory(), analysis, &m_card);
        if (!s.ready()) return false;
        m_plan = plan;
        m_strategy = s;
        m_bound = true;
        return true;
    }

    aa_ddtr_strategy ddtr_strategy() { return m_bound ? m_strategy : aa_ddtr_strategy(); }

    void set_ddtr_consumer(aa_hip_ddtr_consumer fn, void *user) {
        m_consumer = fn;
        m_consumer_user = fn ? user : nullptr;
    }

    bool has(aa_pipeline::component_option o) const { return m_options.find(o) != m_options.end(); }

    bool ready() {
        if (m_ready) return true;
        for (auto c : m_components) {
            if (c != aa_pipeline::component::dedispersion) {
                LOG(log_level::error, aa_pipeline::component_name(c) +
                                          " is outside this library (dedispersion only).");
                return false;
            }
        }
        if (!m_bound) {
            LOG(log_level::error, "De-dispersion plan is not ok.");
            return false;
        }
        using opt = aa_pipeline::component_option;
        for (auto o : m_options) {
            // Options that change the dedispersed values are refused; options
            // that only concern output files or other components are ignored.
            const bool changes_values = o == opt::old_rfi;
            if (changes_values) {
                LOG(log_level::error, "pipeline option " + std::to_string((int)o) +
                                          " (old RFI) is not supported by this library.");
                aa_hip::set_last_error("unsupported pipeline option");
                return false;
            }
            if (o != opt::copy_ddtr_data_to_host && o != opt::empty && o != opt::zero_dm &&
                o != opt::zero_dm_with_outliers &&
                o != opt::input_DDTR_normalization && o != opt::output_DDTR_normalization &&
                o != opt::set_bandpass_average)
                LOG(log_level::notice, "pipeline option " + std::to_string((int)o) + " has no effect here (ignored).");
        }
        const int nb = m_metadata.nbits();
        const bool narrow = nb == 1 || nb == 2 || nb == 4 || nb == 8;
        const bool wide = nb == 16 || nb == 32;   // 32-bit files arrive scaled to u16 (the reader's rule)
        if (!narrow && !wide) {
            LOG(log_level::error, "filterbank data of 1, 2, 4, 8, 16 or 32 bits is supported (nbits=" +
                                      std::to_string(nb) + ").");
            aa_hip::set_last_error("unsupported nbits");
            return false;
        }
        if (wide && m_bytes != 2) {
            LOG(log_level::error, std::to_string(nb) + "-bit data needs a 16-bit (unsigned short) input buffer.");
            aa_hip::set_last_error("unsupported nbits for an 8-bit input buffer");
            return false;
        }
        if (!m_input) {
            LOG(log_level::error, "no input data.");
            return false;
        }
        if (!m_strategy.setup()) return false;
        m_ready = true;
        return true;
    }

    bool run() {
        aa_pipeline_runner::status st;
        while (run(st)) {
        }
        return st != aa_pipeline_runner::status::error;
    }

    bool run(aa_pipeline_runner::status &status) {
        if (!m_ready) {
            LOG(log_level::error, "Pipeline could not start/resume because the pipeline is not ready.");
            status = aa_pipeline_runner::status::finished;
            return false;
        }
        const int nranges = (int)m_strategy.get_nRanges();
        if (m_chunk < m_strategy.num_tchunks() && !setup()) {
            LOG(log_level::error, "Pipeline could not start/resume because the runner could not be set up.");
            status = aa_pipeline_runner::status::error;
            return false;
        }
        if (m_chunk >= m_strategy.num_tchunks()) {
            if (!m_notified_finish) {
                release_device();
                m_notified_finish = true;
                status = aa_pipeline_runner::status::finished_component;
                return true;
            }
            status = aa_pipeline_runner::status::finished;
            return false;
        }
        if (!step(m_chunk, m_range)) {
            release_device();
            status = aa_pipeline_runner::status::error;
            return false;
        }
        ++m_range;
        if (m_range >= nranges) {
            m_inc += m_strategy.t_processed()[0][(size_t)m_chunk];
            ++m_chunk;
            m_range = 0;
        }
        status = aa_pipeline_runner::status::has_more;
        return true;
    }

    size_t get_nRanges() { return m_strategy.get_nRanges(); }
    const int *get_ndms_array() { return m_strategy.ndms_data(); }
    int dm_low(int range) { return m_strategy.dm_low(range); }
    int total_computed_samples() {
        int total = 0;
        if (!m_strategy.t_processed().empty())
            for (int v : m_strategy.t_processed()[0]) total += v;
        return total;
    }
    int samples_in_current(int i, int j) { return m_strategy.t_processed()[(size_t)i][(size_t)j]; }
    aa_filterbank_metadata metadata() { return m_metadata; }
    int current_range() const { return m_range; }
    int current_chunk() const { return m_chunk; }
    long current_inc() const { return m_inc; }

    bool cleanup() {
        release_device();
        free_host_output();
        return true;
    }

    float ***output_buffer() {
        if (m_options.find(aa_pipeline::component_option::copy_ddtr_data_to_host) == m_options.end())
            LOG(log_level::error, "Could not get data from DDTR. Enable option copy_ddtr_data_to_host.");
        return m_host_out;
    }

private:
    // Allocates device and host buffers on first use (the reference's
    // runner setup, include/aa_permitted_pipelines_generic.hpp:1090-1106).
    bool setup() {
        if (m_engine) return true;
        if (m_finished_setup_failed) return false;
        const int nranges = (int)m_strategy.get_nRanges();
        const int nchunks = m_strategy.num_tchunks();
        const int nchans = m_metadata.nchans();
        std::vector<float> low, step;
        std::vector<int> ndms, inbin;
        int max_tp0 = 0;
        for (int c = 0; c < nchunks; ++c) max_tp0 = std::max(max_tp0, m_strategy.t_processed()[0][(size_t)c]);
        for (int r = 0; r < nranges; ++r) {
            low.push_back(m_strategy.dm((size_t)r).low);
            step.push_back(m_strategy.dm((size_t)r).step);
            ndms.push_back(m_strategy.ndms((size_t)r));
            inbin.push_back(m_strategy.dm((size_t)r).inBin);
        }
        const std::vector<float> dmshifts = m_strategy.dmshifts();
        aa_hip::ddtr_engine *e = nullptr;
        // Samples wider than 8 bits keep u16 levels and float channel sums.
        m_wide = m_metadata.nbits() > 8;
        int st = aa_hip::ddtr_engine::create(m_card.currently_selected_card_id(), nchans, dmshifts.data(), nranges,
                                             low.data(), step.data(), ndms.data(), inbin.data(),
                                             m_metadata.tsamp(), m_strategy.maxshift(), max_tp0, &e,
                                             m_wide ? 16 : 8);
        if (st != AA_HIP_SUCCESS) {
            LOG(log_level::error, std::string("pipeline setup: ") + aa_hip::last_error());
            m_finished_setup_failed = true;
            return false;
        }
        m_engine.reset(e);
        (void)hipSetDevice(m_engine->device());
        const size_t in_bytes = ((size_t)max_tp0 + (size_t)m_strategy.maxshift()) * (size_t)nchans * (size_t)m_bytes;
        hipError_t he = hipMalloc(&m_d_input, in_bytes);
        size_t max_out = 0;
        for (int r = 0; r < nranges; ++r) {
            int tp = 0;
            for (int c = 0; c < nchunks; ++c) tp = std::max(tp, m_strategy.t_processed()[(size_t)r][(size_t)c]);
            max_out = std::max(max_out, (size_t)tp * (size_t)ndms[(size_t)r]);
        }
        if (he == hipSuccess) he = hipMalloc(&m_d_output, std::max<size_t>(max_out, 1) * sizeof(float));
        m_zero_dm = has(aa_pipeline::component_option::zero_dm);
        // zero_dm takes precedence when both are set, as in the reference's
        // runner (include/aa_permitted_pipelines_generic.hpp:706-719).
        m_zero_dm_outliers = !m_zero_dm && has(aa_pipeline::component_option::zero_dm_with_outliers);
        if (he == hipSuccess && (m_zero_dm || m_zero_dm_outliers)) {   // the strategy's bandpass normalisation, one float per channel
            std::vector<float> bp = m_strategy.bandpass_normalization();
            // set_bandpass_average replaces it before the first chunk
            // (include/aa_permitted_pipelines_generic.hpp:533-551).
            if (has(aa_pipeline::component_option::set_bandpass_average)) {
                bp.assign((size_t)nchans, 0.0f);
                aa_hip::bandpass_average(m_input, m_bytes, (size_t)nchans, (size_t)m_metadata.nsamples(), bp.data());
            }
            if (bp.size() != (size_t)nchans) {
                LOG(log_level::error, "zero-DM: the bandpass normalisation does not have nchans values");
                aa_hip::set_last_error("zero-DM: bandpass normalisation size");
                m_finished_setup_failed = true;
                release_device();
                return false;
            }
            he = hipMalloc(&m_d_bandpass, bp.size() * sizeof(float));
            if (he == hipSuccess)
                he = hipMemcpy(m_d_bandpass, bp.data(), bp.size() * sizeof(float), hipMemcpyHostToDevice);
        }
        m_in_norm = has(aa_pipeline::component_option::input_DDTR_normalization);
        m_out_norm = has(aa_pipeline::component_option::output_DDTR_normalization);
        if (he == hipSuccess && m_in_norm)
            he = hipMalloc(&m_d_norm_scratch, aa_hip::normalize_columns_scratch_bytes(nchans));
        if (he == hipSuccess) he = hipStreamCreate(&m_stream);
        if (he == hipSuccess) he = hipStreamCreate(&m_copy_stream);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&m_ev_ingested, hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&m_ev_copied, hipEventDisableTiming);
        if (he != hipSuccess) {
            LOG(log_level::error, std::string("pipeline setup: ") + hipGetErrorString(he));
            m_finished_setup_failed = true;
            release_device();
            return false;
        }
        if (!m_host_out &&
            m_options.find(aa_pipeline::component_option::copy_ddtr_data_to_host) != m_options.end() &&
            !allocate_host_output()) {
            LOG(log_level::error, "pipeline setup: pinned host allocation of the DM-time output failed");
            aa_hip::set_last_error("pipeline setup: pinned host allocation of the DM-time output failed");
            free_host_output();
            m_finished_setup_failed = true;
            release_device();
            return false;
        }
        // With host copies and more than one range, a second output buffer
        // lets each range's copy overlap the next range's kernels, when the
        // memory budget left after the strategy's booking holds it.
        if (m_host_out && m_strategy.get_nRanges() > 1) {
            const size_t extra = std::max<size_t>(max_out, 1) * sizeof(float);
            if (extra <= m_card.free_memory() && hipMalloc(&m_d_output2, extra) == hipSuccess) {
                hipError_t oe = hipStreamCreate(&m_d2h_stream);
                if (oe == hipSuccess) oe = hipEventCreateWithFlags(&m_ev_range, hipEventDisableTiming);
                if (oe == hipSuccess && m_card.request_memory(extra)) {
                    m_overlap = true;
                } else {
                    if (m_d2h_stream) (void)hipStreamDestroy(m_d2h_stream);
                    if (m_ev_range) (void)hipEventDestroy(m_ev_range);
                    (void)hipFree(m_d_output2);
                    m_d2h_stream = nullptr;
                    m_ev_range = nullptr;
                    m_d_output2 = nullptr;
                }
            } else {
                m_d_output2 = nullptr;
                (void)hipGetLastError();
            }
        }
        return true;
    }

    // Host output as float***[range][dm] row pointers into one block per
    // range, so a whole range is one strided device-to-host copy.  The
    // blocks are pinned host memory (the copies then run at link speed);
    // every column of a block is written by the chunk that computes it, as
    // the reference's malloc'd rows are.
    bool allocate_host_output() {
        free_host_output();
        const int nranges = (int)m_strategy.get_nRanges();
        m_host_out = (float ***)std::calloc((size_t)nranges, sizeof(float **));
        m_host_blocks.assign((size_t)nranges, nullptr);
        for (int r = 0; r < nranges; ++r) {
            size_t total = 0;
            for (int v : m_strategy.t_processed()[(size_t)r]) total += (size_t)v;
            const int nd = m_strategy.ndms((size_t)r);
            void *blk = nullptr;
            if (hipHostMalloc(&blk, std::max<size_t>(total * (size_t)nd, 1) * sizeof(float), hipHostMallocDefault) !=
                hipSuccess)
                blk = nullptr;
            m_host_blocks[(size_t)r] = (float *)blk;
            m_host_out[r] = (float **)std::calloc((size_t)std::max(nd, 1), sizeof(float *));
            m_host_total.push_back(total);
            if (!blk || !m_host_out[r]) return false;
            for (int d = 0; d < nd; ++d) m_host_out[r][d] = m_host_blocks[(size_t)r] + (size_t)d * total;
        }
        return true;
    }

    void free_host_output() {
        if (m_host_out) {
            for (size_t r = 0; r < m_host_blocks.size(); ++r) {
                std::free(m_host_out[r]);
                if (m_host_blocks[r]) (void)hipHostFree(m_host_blocks[r]);
            }
            std::free(m_host_out);
        }
        m_host_out = nullptr;
        m_host_blocks.clear();
        m_host_total.clear();
    }

    void release_device() {
        if (m_engine) (void)hipSetDevice(m_engine->device());
        if (m_copy_stream) {
            (void)hipStreamSynchronize(m_copy_stream);
            (void)hipStreamDestroy(m_copy_stream);
        }
        if (m_stream) {
            (void)hipStreamSynchronize(m_stream);
            (void)hipStreamDestroy(m_stream);
        }
        if (m_d2h_stream) {
            (void)hipStreamSynchronize(m_d2h_stream);
            (void)hipStreamDestroy(m_d2h_stream);
        }
        if (m_ev_range) (void)hipEventDestroy(m_ev_range);
        if (m_d_output2) (void)hipFree(m_d_output2);
        m_d2h_stream = nullptr;
        m_ev_range = nullptr;
        m_d_output2 = nullptr;
        m_overlap = false;
        m_spec_chunk = m_spec_range = -1;
        if (m_ev_ingested) (void)hipEventDestroy(m_ev_ingested);
        if (m_ev_copied) (void)hipEventDestroy(m_ev_copied);
        if (m_d_input) (void)hipFree(m_d_input);
        if (m_d_output) (void)hipFree(m_d_output);
        if (m_d_bandpass) (void)hipFree(m_d_bandpass);
        m_d_bandpass = nullptr;
        if (m_d_norm_scratch) (void)hipFree(m_d_norm_scratch);
        m_d_norm_scratch = nullptr;
        m_copy_stream = nullptr;
        m_ev_ingested = nullptr;
        m_ev_copied = nullptr;
        m_prefetched_chunk = -1;
        m_stream = nullptr;
        m_d_input = nullptr;
        m_d_output = nullptr;
        m_engine.reset();
    }

    // Input of time chunk c: t_processed[0][c] + maxshift samples of every
    // channel ([t][c]).
    size_t chunk_input_bytes(int c) const {
        const size_t rows = (size_t)m_strategy.t_processed()[0][(size_t)c] + (size_t)m_strategy.maxshift();
        return rows * (size_t)m_metadata.nchans() * (size_t)m_bytes;
    }

    bool step(int chunk, int range) {
        const int nchans = m_metadata.nchans();
        const int tp = m_strategy.t_processed()[(size_t)range][(size_t)chunk];
        (void)hipSetDevice(m_engine->device());
        // A range launched ahead by the previous call (never range 0) only
        // needs its copy.
        if (!(m_spec_chunk == chunk && m_spec_range == range)) {
            if (range == 0) {
                const int tp0 = m_strategy.t_processed()[0][(size_t)chunk];
                hipError_t he = hipSuccess;
                if (m_prefetched_chunk == chunk) {   // copied during the previous chunk's ranges
                    he = hipStreamWaitEvent(m_stream, m_ev_copied, 0);
                } else {
                    const size_t bytes = chunk_input_bytes(chunk);
                    const char *src = (const char *)m_input + (size_t)m_inc * (size_t)nchans * (size_t)m_bytes;
                    he = hipMemcpyAsync(m_d_input, src, bytes, hipMemcpyHostToDevice, m_stream);
                }
                if (he != hipSuccess) {
                    LOG(log_level::error, std::string("H2D copy: ") + hipGetErrorString(he));
                    return false;
                }
                if (m_zero_dm) {   // the reference filters the [t][c] chunk before its corner turn
                    const int rows = tp0 + m_strategy.maxshift();
                    he = m_bytes == 1
                             ? aa_hip::launch_zero_dm_u8((uint8_t *)m_d_input, rows, nchans, m_d_bandpass,
                                                         m_metadata.nbits(), m_stream)
                             : aa_hip::launch_zero_dm_u16((uint16_t *)m_d_input, rows, nchans, m_d_bandpass,
                                                          sample_nbits(), m_stream);
                    if (he != hipSuccess) {
                        LOG(log_level::error, std::string("zero-DM: ") + hipGetErrorString(he));
                        return false;
                    }
                } else if (m_zero_dm_outliers) {
                    const long rows = (long)tp0 + m_strategy.maxshift();
                    he = m_bytes == 1 ? aa_hip::launch_zero_dm_outliers_u8((uint8_t *)m_d_input, rows, nchans,
                                                                           m_d_bandpass, m_metadata.nbits(), m_stream)
                                      : aa_hip::launch_zero_dm_outliers_u16((uint16_t *)m_d_input, rows, nchans,
                                                                            m_d_bandpass, sample_nbits(),
                                                                            m_stream);
                    if (he != hipSuccess) {
                        LOG(log_level::error, std::string("zero-DM with outliers: ") + hipGetErrorString(he));
                        return false;
                    }
                }
                if (m_in_norm) {   // after zero-DM, before the corner turn (same rows)
                    const long rows = (long)tp0 + m_strategy.maxshift();
                    const int nb = sample_nbits();
                    const float factor = (float)((std::pow(2.0, nb) - 1.0) / 2.0);
                    he = m_bytes == 1 ? aa_hip::launch_normalize_columns_u8((uint8_t *)m_d_input, rows, nchans, factor,
                                                                            nb, m_d_norm_scratch, m_stream)
                                      : aa_hip::launch_normalize_columns_u16((uint16_t *)m_d_input, rows, nchans,
                                                                             factor, nb, m_d_norm_scratch, m_stream);
                    if (he != hipSuccess) {
                        LOG(log_level::error, std::string("input DDTR normalisation: ") + hipGetErrorString(he));
                        return false;
                    }
                }
                int st = m_engine->ingest(m_d_input, m_bytes == 1 ? 8 : 16, tp0, m_stream);
                if (st != AA_HIP_SUCCESS) {
                    LOG(log_level::error, std::string("corner turn: ") + aa_hip::last_error());
                    return false;
                }
                // The corner turn has taken the chunk out of m_d_input: the next
                // chunk's copy starts behind it on the copy stream and overlaps
                // this chunk's ranges (and their output copies).
                if (chunk + 1 < m_strategy.num_tchunks()) {
                    const long inc_next = m_inc + tp0;
                    const char *src = (const char *)m_input + (size_t)inc_next * (size_t)nchans * (size_t)m_bytes;
                    he = hipEventRecord(m_ev_ingested, m_stream);
                    if (he == hipSuccess) he = hipStreamWaitEvent(m_copy_stream, m_ev_ingested, 0);
                    if (he == hipSuccess)
                        he = hipMemcpyAsync(m_d_input, src, chunk_input_bytes(chunk + 1), hipMemcpyHostToDevice,
                                            m_copy_stream);
                    if (he == hipSuccess) he = hipEventRecord(m_ev_copied, m_copy_stream);
                    if (he != hipSuccess) {
                        LOG(log_level::error, std::string("H2D prefetch: ") + hipGetErrorString(he));
                        return false;
                    }
                    m_prefetched_chunk = chunk + 1;
                }
                if (m_bytes == 2 && !m_wide) {
                    st = m_engine->take_input_overflow(m_stream);
                    if (st != 0) {
                        LOG(log_level::error, st > 0 ? std::string("input sample above 255 in 8-bit data")
                                                     : std::string(aa_hip::last_error()));
                        aa_hip::set_last_error("input sample above 255 in 8-bit data");
                        return false;
                    }
                }
            }
            if (!launch_range(chunk, range)) return false;
        }
        m_spec_chunk = m_spec_range = -1;
        if (m_host_out) {
            const int ib = m_strategy.dm((size_t)range).inBin;
            const size_t col = (size_t)(m_inc / ib);
            const size_t total = m_host_total[(size_t)range];
            const int nd = m_strategy.ndms((size_t)range);
            if (col + (size_t)tp > total) {
                LOG(log_level::error, "output column offset exceeds the range's total samples.");
                return false;
            }
            hipStream_t cs = m_stream;
            hipError_t he = hipSuccess;
            if (m_overlap) {   // the copy waits for this range's kernels on its own stream
                cs = m_d2h_stream;
                he = hipEventRecord(m_ev_range, m_stream);
                if (he == hipSuccess) he = hipStreamWaitEvent(cs, m_ev_range, 0);
            }
            if (he == hipSuccess)
                he = hipMemcpy2DAsync(m_host_blocks[(size_t)range] + col, total * sizeof(float), out_buffer(range),
                                      (size_t)tp * sizeof(float), (size_t)tp * sizeof(float), (size_t)nd,
                                      hipMemcpyDeviceToHost, cs);
            if (he != hipSuccess) {
                LOG(log_level::error, std::string("D2H copy: ") + hipGetErrorString(he));
                return false;
            }
            // The next range of this chunk starts computing into the other
            // buffer while this range's copy runs; run() still returns only
            // once this range's host data are complete.
            if (m_overlap && range + 1 < (int)m_strategy.get_nRanges()) {
                if (!launch_range(chunk, range + 1)) return false;
                m_spec_chunk = chunk;
                m_spec_range = range + 1;
            }
            he = hipStreamSynchronize(cs);
            if (he != hipSuccess) {
                LOG(log_level::error, std::string("device error: ") + hipGetErrorString(he));
                return false;
            }
            return true;
        }
        hipError_t he = hipStreamSynchronize(m_stream);
        if (he != hipSuccess) {
            LOG(log_level::error, std::string("device error: ") + hipGetErrorString(he));
            return false;
        }
        return true;
    }

    // Output buffer of a range: ranges alternate between two buffers when
    // the copies overlap the next range's kernels.
    float *out_buffer(int range) const { return m_overlap && (range & 1) ? m_d_output2 : m_d_output; }

    // Dedispersion of one (chunk, range) on m_stream into out_buffer(range),
    // then the output normalisation when asked for.
    bool launch_range(int chunk, int range) {
        const int tp = m_strategy.t_processed()[(size_t)range][(size_t)chunk];
        float *out = out_buffer(range);
        int st = m_engine->run_range(range, tp, out, (size_t)tp, m_stream);
        if (st != AA_HIP_SUCCESS) {
            LOG(log_level::error, std::string("dedispersion: ") + aa_hip::last_error());
            return false;
        }
        if (m_out_norm) {   // each time sample's mean over the range's DMs moved to 127.5
            const hipError_t he =
                aa_hip::launch_normalize_columns_f32(out, m_strategy.ndms((size_t)range), tp, 127.5f, m_stream);
            if (he != hipSuccess) {
                LOG(log_level::error, std::string("output DDTR normalisation: ") + hipGetErrorString(he));
                return false;
            }
        }
        if (m_consumer) {   // where the reference's runner calls analysis_GPU on the plane
            const auto &rd = m_strategy.dm((size_t)range);
            aa_hip_ddtr_plane pl;
            pl.range = range;
            pl.time_chunk = chunk;
            pl.d_plane = out;
            pl.ndms = m_strategy.ndms((size_t)range);
            pl.nsamp = tp;
            pl.pitch = (size_t)tp;
            pl.dm_low = rd.low;
            pl.dm_step = rd.step;
            pl.inBin = rd.inBin;
            pl.tsamp = (double)((float)m_metadata.tsamp() * (float)rd.inBin);
            pl.inc = m_inc;
            pl.device = m_engine->device();
            pl.stream = (void *)m_stream;
            const int rc = m_consumer(&pl, m_consumer_user);
            (void)hipSetDevice(m_engine->device());
            if (rc != 0) {
                LOG(log_level::error, "DDTR plane consumer failed (" + std::to_string(rc) + ") on range " +
                                          std::to_string(range) + ", chunk " + std::to_string(chunk));
                aa_hip::set_last_error("DDTR plane consumer returned an error");
                return false;
            }
        }
        return true;
    }

    aa_pipeline::pipeline m_components;
    aa_pipeline::pipeline_option m_options;
    aa_filterbank_metadata m_metadata;
    const void *m_input;
    int m_bytes;
    bool m_wide = false;   // nbits > 8: u16 levels, float channel sums
    // Sample width of the data in device memory: 32-bit files are scaled to
    // u16 by the reader, so the conditioning kernels see 16-bit samples.
    int sample_nbits() const { return std::min(m_metadata.nbits(), 16); }
    aa_device_info &m_card;
    aa_ddtr_plan m_plan;
    aa_ddtr_strategy m_strategy;
    bool m_bound = false;
    bool m_ready = false;
    bool m_finished_setup_failed = false;
    bool m_notified_finish = false;
    int m_chunk = 0;
    int m_range = 0;
    long m_inc = 0;
    std::unique_ptr<aa_hip::ddtr_engine> m_engine;
    void *m_d_input = nullptr;
    float *m_d_output = nullptr;
    hipStream_t m_stream = nullptr;
    hipStream_t m_copy_stream = nullptr;   // next chunk's input copy
    hipEvent_t m_ev_ingested = nullptr;    // corner turn of the current chunk done
    hipEvent_t m_ev_copied = nullptr;      // next chunk's input on the device
    int m_prefetched_chunk = -1;
    bool m_zero_dm = false;                // zero-DM filter before each chunk's corner turn
    float *m_d_bandpass = nullptr;         // its per-channel normalisation
    bool m_overlap = false;                // D2H copies overlap the next range's kernels
    float *m_d_output2 = nullptr;          // the second output buffer (odd ranges) then
    hipStream_t m_d2h_stream = nullptr;
    hipEvent_t m_ev_range = nullptr;       // a range's kernels done
    int m_spec_chunk = -1, m_spec_range = -1;   // (chunk, range) already launched
    bool m_zero_dm_outliers = false;       // zero-DM with outlier rejection instead
    bool m_in_norm = false;                // input DDTR normalisation of each chunk
    bool m_out_norm = false;               // output DDTR normalisation of each plane
    void *m_d_norm_scratch = nullptr;      // input normalisation's column sums
    aa_hip_ddtr_consumer m_consumer = nullptr;   // hand-off of each plane (analysis_GPU's place)
    void *m_consumer_user = nullptr;
    float ***m_host_out = nullptr;
    std::vector<float *> m_host_blocks;
    std::vector<size_t> m_host_total;
};

// ------------------------------------------------------- aa_ddtr_pipeline
aa_ddtr_pipeline::aa_ddtr_pipeline(const aa_pipeline::pipeline &requested_pipeline,
                                   const aa_pipeline::pipeline_option &pipeline_options,
                                   const aa_filterbank_metadata &filterbank_metadata, const void *input_data,
                                   int input_bytes_per_sample, aa_device_info &card_info)
    : m_impl(new aa_ddtr_pipeline_impl(requested_pipeline, pipeline_options, filterbank_metadata, input_data,
                                       input_bytes_per_sample, card_info)) {}
aa_ddtr_pipeline::~aa_ddtr_pipeline() = default;
bool aa_ddtr_pipeline::bind(const aa_ddtr_plan &plan) { return m_impl->bind(plan); }
aa_ddtr_strategy aa_ddtr_pipeline::ddtr_strategy() { return m_impl->ddtr_strategy(); }
size_t aa_ddtr_pipeline::get_nRanges() { return m_impl->get_nRanges(); }
const int *aa_ddtr_pipeline::get_ndms_array() { return m_impl->get_ndms_array(); }
int aa_ddtr_pipeline::dm_low(const int range) { return m_impl->dm_low(range); }
int aa_ddtr_pipeline::total_computed_samples() { return m_impl->total_computed_samples(); }
int aa_ddtr_pipeline::samples_in_current(int i, int j) { return m_impl->samples_in_current(i, j); }
aa_filterbank_metadata aa_ddtr_pipeline::metadata() { return m_impl->metadata(); }
bool aa_ddtr_pipeline::ready() { return m_impl->ready(); }
bool aa_ddtr_pipeline::run() { return m_impl->run(); }
bool aa_ddtr_pipeline::run(aa_pipeline_runner::status &status_code) { return m_impl->run(status_code); }
int aa_ddtr_pipeline::get_current_range() { return m_impl->current_range(); }
int aa_ddtr_pipeline::get_current_tchunk() { return m_impl->current_chunk(); }
long int aa_ddtr_pipeline::get_current_inc() { return m_impl->current_inc(); }
bool aa_ddtr_pipeline::cleanup() { return m_impl->cleanup(); }
float ***aa_ddtr_pipeline::output_buffer() { return m_impl->output_buffer(); }
void aa_ddtr_pipeline::set_ddtr_consumer(aa_hip_ddtr_consumer fn, void *user) { m_impl->set_ddtr_consumer(fn, user); }

}  // namespace astroaccelerate
