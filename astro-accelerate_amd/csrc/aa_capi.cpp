// extern "C" surface of libaa_ddtr_hip.so (declared in include/aa_ddtr_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "aa_ddtr_hip.h"
#include "aa_ddtr_plan.hpp"
#include "aa_ddtr_strategy.hpp"
#include "aa_device_info.hpp"
#include "aa_log.hpp"
#include "aa_pipeline_api.hpp"
#include "ddtr_batch.hpp"
#include "ddtr_engine.hpp"

using namespace astroaccelerate;

struct aa_hip_ddtr {
    std::unique_ptr<aa_hip::ddtr_engine> engine;
};

struct aa_hip_ddtr_batch {
    std::unique_ptr<aa_hip::ddtr_batch> batch;
};

namespace {

// A Python-side pipeline owns its device description (the reference builds
// one on the stack and keeps a reference to it past its lifetime).
struct py_pipeline {
    aa_device_info device;
    aa_ddtr_pipeline pipe;
    py_pipeline(const aa_pipeline::pipeline &c, const aa_pipeline::pipeline_option &o,
                const aa_filterbank_metadata &m, const void *input, int bytes, int card, size_t cap = 0)
        : device(card, cap), pipe(c, o, m, input, bytes, device) {}
};

aa_pipeline::pipeline components_of(const pipeline_components_struct &p) {
    aa_pipeline::pipeline out;
    if (p.dedispersion) out.insert(aa_pipeline::component::dedispersion);
    if (p.analysis) out.insert(aa_pipeline::component::analysis);
    if (p.periodicity) out.insert(aa_pipeline::component::periodicity);
    if (p.fdas) out.insert(aa_pipeline::component::fdas);
    return out;
}

aa_pipeline::pipeline_option options_of(const pipeline_component_option_struct &o) {
    using opt = aa_pipeline::component_option;
    aa_pipeline::pipeline_option out;
    if (o.zero_dm) out.insert(opt::zero_dm);
    if (o.zero_dm_with_outliers) out.insert(opt::zero_dm_with_outliers);
    if (o.input_DDTR_normalization) out.insert(opt::input_DDTR_normalization);
    if (o.output_DDTR_normalization) out.insert(opt::output_DDTR_normalization);
    if (o.set_bandpass_average) out.insert(opt::set_bandpass_average);
    if (o.old_rfi) out.insert(opt::old_rfi);
    if (o.copy_ddtr_data_to_host) out.insert(opt::copy_ddtr_data_to_host);
    if (o.msd_baseline_noise) out.insert(opt::msd_baseline_noise);
    if (o.output_dmt) out.insert(opt::output_dmt);
    if (o.output_ffdot_plan) out.insert(opt::output_ffdot_plan);
    if (o.output_fdas_list) out.insert(opt::output_fdas_list);
    if (o.candidate_algorithm) out.insert(opt::candidate_algorithm);
    if (o.fdas_custom_fft) out.insert(opt::fdas_custom_fft);
    if (o.fdas_inbin) out.insert(opt::fdas_inbin);
    if (o.fdas_norm) out.insert(opt::fdas_norm);
    return out;
}

aa_filterbank_metadata metadata_of(const aa_py_filterbank_metadata_struct &m) {
    return aa_filterbank_metadata(m.m_tstart, m.m_tsamp, m.m_nbits, m.m_nsamples, m.m_fch1, m.m_foff, m.m_nchans);
}

int bad(const char *msg) {
    aa_hip::set_last_error(msg);
    return AA_HIP_ERROR_INVALID_ARGUMENT;
}

int from_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return AA_HIP_SUCCESS;
    aa_hip::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? AA_HIP_ERROR_OUT_OF_MEMORY : AA_HIP_ERROR_DEVICE;
}

}  // namespace

template <class T>
static int input_norm(T *d, long nsamp, int nchans, int nbits, int max_bits, void *stream) {
    if (nsamp < 0 || nchans <= 0 || nbits < 1 || nbits > max_bits) return bad("invalid geometry or nbits");
    if (nsamp == 0) return AA_HIP_SUCCESS;
    if (!d) return bad("NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    void *scratch = nullptr;
    hipError_t e = hipMalloc(&scratch, aa_hip::normalize_columns_scratch_bytes(nchans));
    if (e != hipSuccess) return from_hip(e, "input DDTR normalisation scratch");
    const float factor = (float)((std::pow(2.0, nbits) - 1.0) / 2.0);
    if constexpr (sizeof(T) == 1)
        e = aa_hip::launch_normalize_columns_u8(d, nsamp, nchans, factor, nbits, scratch, s);
    else
        e = aa_hip::launch_normalize_columns_u16(d, nsamp, nchans, factor, nbits, scratch, s);
    const hipError_t w = hipStreamSynchronize(s);   // the scratch is freed below
    const hipError_t f = hipFree(scratch);
    return from_hip(e != hipSuccess ? e : w != hipSuccess ? w : f, "input DDTR normalisation");
}

extern "C" {

const char *aa_hip_last_error(void) { return aa_hip::last_error(); }
const char *aa_hip_version(void) { return "aa_ddtr_hip 0.1.0 (gfx950)"; }

// ------------------------------------------------------------- plan (FFI)
void *aa_py_ddtr_plan(void) { return new (std::nothrow) aa_ddtr_plan(); }
void aa_py_ddtr_plan_delete(void const *obj) { delete static_cast<aa_ddtr_plan const *>(obj); }
bool aa_py_ddtr_plan_add_dm(void *obj, float low, float high, float step, int inBin, int outBin) {
    return obj && static_cast<aa_ddtr_plan *>(obj)->add_dm(low, high, step, inBin, outBin);
}
size_t aa_py_ddtr_plan_range(void const *obj) { return obj ? static_cast<aa_ddtr_plan const *>(obj)->range() : 0; }
bool aa_py_ddtr_plan_set_power(void *obj, float power) {
    return obj && static_cast<aa_ddtr_plan *>(obj)->set_power(power);
}
float aa_py_ddtr_plan_power(void const *obj) { return obj ? static_cast<aa_ddtr_plan const *>(obj)->power() : 0.0f; }
bool aa_py_ddtr_plan_set_enable_msd_baseline(void *obj, bool flag) {
    return obj && static_cast<aa_ddtr_plan *>(obj)->set_enable_msd_baseline_noise(flag);
}
bool aa_py_ddtr_plan_enable_msd_baseline_noise(void const *obj) {
    return obj && static_cast<aa_ddtr_plan const *>(obj)->enable_msd_baseline_noise();
}
bool aa_py_ddtr_plan_bind_bandpass_normalization(void *obj, float const *data, int data_size) {
    return obj && data && static_cast<aa_ddtr_plan *>(obj)->bind_bandpass_normalization(data, data_size);
}

// --------------------------------------------------------- metadata (FFI)
aa_py_filterbank_metadata_struct *aa_py_filterbank_metadata(double tstart, double tsamp, int nbits, int nsamples,
                                                            double fch1, double foff, int nchans, double src_raj) {
    auto *m = new (std::nothrow) aa_py_filterbank_metadata_struct;
    if (!m) return nullptr;
    m->m_tstart = tstart;
    m->m_tsamp = tsamp;
    m->m_fch1 = fch1;
    m->m_foff = foff;
    m->m_nbits = nbits;
    m->m_nsamples = nsamples;
    m->m_nchans = nchans;
    m->m_src_raj = src_raj;
    return m;
}
void aa_py_filterbank_metadata_delete(aa_py_filterbank_metadata_struct const *obj) { delete obj; }
double aa_py_filterbank_metadata_tstart(aa_py_filterbank_metadata_struct const *obj) { return obj ? obj->m_tstart : 0.0; }

// --------------------------------------------------------- pipeline (FFI)
void *aa_py_pipeline_api(pipeline_components_struct pipeline, pipeline_component_option_struct options,
                         aa_py_filterbank_metadata_struct metadata, unsigned short const *input_data,
                         int card_number) {
    return new (std::nothrow) py_pipeline(components_of(pipeline), options_of(options), metadata_of(metadata),
                                          input_data, 2, card_number);
}
void *aa_py_pipeline_api_u8(pipeline_components_struct pipeline, pipeline_component_option_struct options,
                            aa_py_filterbank_metadata_struct metadata, unsigned char const *input_data,
                            int card_number) {
    return new (std::nothrow) py_pipeline(components_of(pipeline), options_of(options), metadata_of(metadata),
                                          input_data, 1, card_number);
}
void *aa_hip_py_pipeline_api_u8_capped(pipeline_components_struct pipeline, pipeline_component_option_struct options,
                                       aa_py_filterbank_metadata_struct metadata, unsigned char const *input_data,
                                       int card_number, size_t max_device_bytes) {
    return new (std::nothrow) py_pipeline(components_of(pipeline), options_of(options), metadata_of(metadata),
                                          input_data, 1, card_number, max_device_bytes);
}
void aa_py_pipeline_api_delete(void const *obj) { delete static_cast<py_pipeline const *>(obj); }
bool aa_py_pipeline_api_bind_ddtr_plan(void *obj, void const *plan) {
    if (!obj || !plan) return false;
    return static_cast<py_pipeline *>(obj)->pipe.bind(*static_cast<aa_ddtr_plan const *>(plan));
}
void *aa_py_pipeline_api_ddtr_strategy(void *obj) {
    if (!obj) return nullptr;
    return new (std::nothrow) aa_ddtr_strategy(static_cast<py_pipeline *>(obj)->pipe.ddtr_strategy());
}
void aa_py_ddtr_strategy_delete(void const *obj) { delete static_cast<aa_ddtr_strategy const *>(obj); }
size_t aa_py_get_ddtr_nRanges(void *obj) { return obj ? static_cast<py_pipeline *>(obj)->pipe.get_nRanges() : 0; }
const int *aa_py_get_ndms_array(void *obj) {
    return obj ? static_cast<py_pipeline *>(obj)->pipe.get_ndms_array() : nullptr;
}
int aa_py_get_t_proc(void *obj, int i, int j) { return obj ? static_cast<py_pipeline *>(obj)->pipe.samples_in_current(i, j) : 0; }
int aa_py_dm_low(void *obj, int range) { return obj ? static_cast<py_pipeline *>(obj)->pipe.dm_low(range) : 0; }
int aa_py_total_computed_samples(void *obj) {
    return obj ? static_cast<py_pipeline *>(obj)->pipe.total_computed_samples() : 0;
}
bool aa_py_pipeline_api_run(void *obj, int *status_code_int) {
    if (!obj) {
        if (status_code_int) *status_code_int = (int)aa_pipeline_runner::status::error;
        return false;
    }
    py_pipeline *p = static_cast<py_pipeline *>(obj);
    aa_pipeline_runner::status st = aa_pipeline_runner::status::error;
    bool more = false;
    if (p->pipe.ready()) {
        more = p->pipe.run(st);
        if (st == aa_pipeline_runner::status::error) more = false;
    }
    if (status_code_int) *status_code_int = (int)st;
    return more;
}
float ***aa_py_buffer(void *obj) { return obj ? static_cast<py_pipeline *>(obj)->pipe.output_buffer() : nullptr; }
int aa_py_current_range(void *obj) { return obj ? static_cast<py_pipeline *>(obj)->pipe.get_current_range() : 0; }
int aa_py_current_time_chunk(void *obj) { return obj ? static_cast<py_pipeline *>(obj)->pipe.get_current_tchunk() : 0; }
long aa_py_current_inc(void *obj) { return obj ? static_cast<py_pipeline *>(obj)->pipe.get_current_inc() : 0; }
bool aa_py_cleanup(void *obj) { return obj && static_cast<py_pipeline *>(obj)->pipe.cleanup(); }
bool aa_hip_pipeline_set_ddtr_consumer(void *pipeline, aa_hip_ddtr_consumer fn, void *user) {
    if (!pipeline) return false;
    static_cast<py_pipeline *>(pipeline)->pipe.set_ddtr_consumer(fn, user);
    return true;
}

// --------------------------------------------------------- strategy (host)
void *aa_py_ddtr_strategy_create(void const *plan, aa_py_filterbank_metadata_struct metadata, size_t free_memory,
                                 bool enable_analysis) {
    if (!plan) return nullptr;
    aa_device_info budget = aa_device_info::memory_only(free_memory);
    (void)budget.free_memory();
    return new (std::nothrow) aa_ddtr_strategy(*static_cast<aa_ddtr_plan const *>(plan), metadata_of(metadata),
                                               free_memory, enable_analysis, &budget);
}
bool aa_py_ddtr_strategy_ready(void const *obj) { return obj && static_cast<aa_ddtr_strategy const *>(obj)->ready(); }
int aa_py_ddtr_strategy_maxshift(void const *obj) { return obj ? static_cast<aa_ddtr_strategy const *>(obj)->maxshift() : 0; }
int aa_py_ddtr_strategy_num_tchunks(void const *obj) {
    return obj ? static_cast<aa_ddtr_strategy const *>(obj)->num_tchunks() : 0;
}
int aa_py_ddtr_strategy_max_ndms(void const *obj) { return obj ? static_cast<aa_ddtr_strategy const *>(obj)->max_ndms() : 0; }
size_t aa_py_ddtr_strategy_nranges(void const *obj) {
    return obj ? static_cast<aa_ddtr_strategy const *>(obj)->get_nRanges() : 0;
}
int aa_py_ddtr_strategy_ndms(void const *obj, int range) {
    auto *s = static_cast<aa_ddtr_strategy const *>(obj);
    if (!s || range < 0 || (size_t)range >= s->ndms_size()) return -1;
    return s->ndms((size_t)range);
}
int aa_py_ddtr_strategy_t_processed(void const *obj, int range, int chunk) {
    auto *s = static_cast<aa_ddtr_strategy const *>(obj);
    if (!s || range < 0 || (size_t)range >= s->t_processed().size()) return -1;
    const auto &row = s->t_processed()[(size_t)range];
    if (chunk < 0 || (size_t)chunk >= row.size()) return -1;
    return row[(size_t)chunk];
}
bool aa_py_ddtr_strategy_dm(void const *obj, int range, float *low, float *high, float *step, int *inBin,
                            int *outBin) {
    auto *s = static_cast<aa_ddtr_strategy const *>(obj);
    if (!s || range < 0 || (size_t)range >= s->get_nRanges()) return false;
    const aa_ddtr_plan::dm d = s->dm((size_t)range);
    if (low) *low = d.low;
    if (high) *high = d.high;
    if (step) *step = d.step;
    if (inBin) *inBin = d.inBin;
    if (outBin) *outBin = d.outBin;
    return true;
}
bool aa_py_ddtr_strategy_dmshifts(void const *obj, float *out, int nchans) {
    auto *s = static_cast<aa_ddtr_strategy const *>(obj);
    if (!s || !out) return false;
    const std::vector<float> v = s->dmshifts();
    if ((size_t)nchans > v.size()) return false;
    std::memcpy(out, v.data(), sizeof(float) * (size_t)nchans);
    return true;
}

// --------------------------------------------------------- device layer
static int ddtr_create(int device, int nchans, const float *dmshifts, int nranges, const float *dm_low,
                       const float *dm_step, const int *ndms, const int *inBin, double tsamp, int maxshift,
                       int max_t_processed0, aa_hip_ddtr **out, int sample_bits) {
    if (!out) return bad("out is NULL");
    *out = nullptr;
    aa_hip::ddtr_engine *e = nullptr;
    const int st = aa_hip::ddtr_engine::create(device, nchans, dmshifts, nranges, dm_low, dm_step, ndms, inBin,
                                               tsamp, maxshift, max_t_processed0, &e, sample_bits);
    if (st != AA_HIP_SUCCESS) return st;
    aa_hip_ddtr *h = new (std::nothrow) aa_hip_ddtr;
    if (!h) {
        delete e;
        return AA_HIP_ERROR_OUT_OF_MEMORY;
    }
    h->engine.reset(e);
    *out = h;
    return AA_HIP_SUCCESS;
}

int aa_hip_ddtr_create(int device, int nchans, const float *dmshifts, int nranges, const float *dm_low,
                       const float *dm_step, const int *ndms, const int *inBin, double tsamp, int maxshift,
                       int max_t_processed0, aa_hip_ddtr **out) {
    return ddtr_create(device, nchans, dmshifts, nranges, dm_low, dm_step, ndms, inBin, tsamp, maxshift,
                       max_t_processed0, out, 8);
}

int aa_hip_ddtr_create_u16(int device, int nchans, const float *dmshifts, int nranges, const float *dm_low,
                           const float *dm_step, const int *ndms, const int *inBin, double tsamp, int maxshift,
                           int max_t_processed0, aa_hip_ddtr **out) {
    return ddtr_create(device, nchans, dmshifts, nranges, dm_low, dm_step, ndms, inBin, tsamp, maxshift,
                       max_t_processed0, out, 16);
}

int aa_hip_ddtr_destroy(aa_hip_ddtr *h) {
    delete h;
    return AA_HIP_SUCCESS;
}

int aa_hip_ddtr_set_kernel(aa_hip_ddtr *h, int kernel) {
    if (!h) return bad("NULL handle");
    return h->engine->set_kernel(kernel);
}

size_t aa_hip_ddtr_device_bytes(const aa_hip_ddtr *h) { return h ? h->engine->device_bytes() : 0; }

int aa_hip_ddtr_execute_chunk(aa_hip_ddtr *h, const void *d_input_tc, int input_bits, const int *t_processed,
                              float *const *d_out, const size_t *out_pitch, void *stream) {
    if (!h) return bad("NULL handle");
    return h->engine->execute_chunk(d_input_tc, input_bits, t_processed, d_out, out_pitch, (hipStream_t)stream);
}

int aa_hip_ddtr_ingest(aa_hip_ddtr *h, const void *d_input_tc, int input_bits, int t_processed0, void *stream) {
    if (!h) return bad("NULL handle");
    return h->engine->ingest(d_input_tc, input_bits, t_processed0, (hipStream_t)stream);
}

int aa_hip_ddtr_run_range(aa_hip_ddtr *h, int range, int t_processed, float *d_out, size_t out_pitch, void *stream) {
    if (!h) return bad("NULL handle");
    return h->engine->run_range(range, t_processed, d_out, out_pitch, (hipStream_t)stream);
}

int aa_hip_ddtr_set_range_variant(aa_hip_ddtr *h, int range, int variant) {
    if (!h || !h->engine) return bad("NULL handle");
    return h->engine->set_range_variant(range, variant);
}

int aa_hip_ddtr_autotune_range(aa_hip_ddtr *h, int range, int t_processed, float *d_out, size_t out_pitch,
                               int reps, void *stream, int *best_variant, float *variant_ms) {
    if (!h || !h->engine) return bad("NULL handle");
    return h->engine->autotune_range(range, t_processed, d_out, out_pitch, reps, (hipStream_t)stream, best_variant,
                                     variant_ms);
}

int aa_hip_ddtr_batch_create(int nbeams, aa_hip_ddtr *const *beams, aa_hip_ddtr_batch **out) {
    if (!out) return bad("NULL output handle");
    *out = nullptr;
    if (nbeams <= 0 || !beams) return bad("no beams");
    std::vector<aa_hip::ddtr_engine *> engines((size_t)nbeams);
    for (int b = 0; b < nbeams; ++b) {
        if (!beams[b] || !beams[b]->engine) return bad("NULL beam handle");
        engines[(size_t)b] = beams[b]->engine.get();
    }
    aa_hip::ddtr_batch *p = nullptr;
    const int st = aa_hip::ddtr_batch::create(nbeams, engines.data(), &p);
    if (st != AA_HIP_SUCCESS) return st;
    auto *h = new (std::nothrow) aa_hip_ddtr_batch();
    if (!h) {
        delete p;
        return AA_HIP_ERROR_OUT_OF_MEMORY;
    }
    h->batch.reset(p);
    *out = h;
    return AA_HIP_SUCCESS;
}

int aa_hip_ddtr_batch_destroy(aa_hip_ddtr_batch *h) {
    delete h;
    return AA_HIP_SUCCESS;
}

int aa_hip_ddtr_batch_execute(aa_hip_ddtr_batch *h, const void *const *d_inputs_tc, const int *t_processed,
                              float *const *d_out, const size_t *out_pitch, int use_graph, void *stream) {
    if (!h || !h->batch) return bad("NULL handle");
    return h->batch->execute(d_inputs_tc, t_processed, d_out, out_pitch, use_graph != 0, (hipStream_t)stream);
}

int aa_hip_ddtr_batch_stats(const aa_hip_ddtr_batch *h, int *njobs, long *ntiles, int *ndirect) {
    if (!h || !h->batch) return bad("NULL handle");
    if (njobs) *njobs = h->batch->njobs();
    if (ntiles) *ntiles = h->batch->ntiles();
    if (ndirect) *ndirect = h->batch->ndirect();
    return AA_HIP_SUCCESS;
}

int aa_hip_ddtr_batch_ct_stats(const aa_hip_ddtr_batch *h, int *batched, int *own) {
    if (!h || !h->batch) return bad("NULL handle");
    if (batched) *batched = h->batch->nct_batched();
    if (own) *own = h->batch->nct_own();
    return AA_HIP_SUCCESS;
}

int aa_hip_corner_turn_u8(const uint8_t *d_in_tc, int nsamp, int nchans, uint8_t *d_out_ct, size_t out_pitch,
                          void *stream) {
    if (!d_in_tc || !d_out_ct) return bad("NULL buffer");
    return from_hip(aa_hip::launch_corner_turn_u8(d_in_tc, nsamp, nchans, d_out_ct, out_pitch, (hipStream_t)stream),
                    "corner turn");
}

int aa_hip_corner_turn_u16(const uint16_t *d_in_tc, int nsamp, int nchans, uint8_t *d_out_ct, size_t out_pitch,
                           int *d_err, void *stream) {
    if (!d_in_tc || !d_out_ct || !d_err) return bad("NULL buffer");
    return from_hip(
        aa_hip::launch_corner_turn_u16(d_in_tc, nsamp, nchans, d_out_ct, out_pitch, d_err, (hipStream_t)stream),
        "corner turn");
}

int aa_hip_corner_turn_levels(const void *d_in_tc, int sample_bytes, int nsamp, int nchans,
                              uint8_t *const *d_levels, const size_t *pitches, int nlev, int *d_err, int *fused,
                              void *stream) {
    if (!d_in_tc || !d_levels || !pitches) return bad("NULL buffer");
    if (nlev < 0 || nlev > aa_hip::kCtMaxLevels) return bad("nlev must be 0..3");
    if (sample_bytes == 2 && !d_err) return bad("u16 input needs an error flag");
    int done = 0;
    const int st = from_hip(aa_hip::launch_corner_turn_levels(d_in_tc, sample_bytes, nsamp, nchans, d_levels, pitches,
                                                              nlev, d_err, (hipStream_t)stream, &done),
                            "corner turn");
    if (fused) *fused = done;
    return st;
}

int aa_hip_bin2_u8(const uint8_t *d_in, size_t in_pitch, int nchans, int n_out, uint8_t *d_out, size_t out_pitch,
                   void *stream) {
    if (!d_in || !d_out) return bad("NULL buffer");
    return from_hip(aa_hip::launch_bin2_u8(d_in, in_pitch, nchans, n_out, d_out, out_pitch, (hipStream_t)stream),
                    "binning");
}

int aa_hip_zero_dm_u8(uint8_t *d_data_tc, int nsamp, int nchans, const float *d_bandpass, int nbits, void *stream) {
    if (nsamp < 0 || nchans <= 0 || nbits < 1 || nbits > 8) return bad("invalid geometry or nbits");
    if (nsamp == 0) return AA_HIP_SUCCESS;
    if (!d_data_tc || !d_bandpass) return bad("NULL buffer");
    return from_hip(aa_hip::launch_zero_dm_u8(d_data_tc, nsamp, nchans, d_bandpass, nbits, (hipStream_t)stream),
                    "zero-DM");
}

int aa_hip_zero_dm_u16(uint16_t *d_data_tc, int nsamp, int nchans, const float *d_bandpass, int nbits,
                       void *stream) {
    if (nsamp < 0 || nchans <= 0 || nbits < 1 || nbits > 16) return bad("invalid geometry or nbits");
    if (nsamp == 0) return AA_HIP_SUCCESS;
    if (!d_data_tc || !d_bandpass) return bad("NULL buffer");
    return from_hip(aa_hip::launch_zero_dm_u16(d_data_tc, nsamp, nchans, d_bandpass, nbits, (hipStream_t)stream),
                    "zero-DM");
}

int aa_hip_zero_dm_outliers_u8(uint8_t *d_data_tc, long nsamp, int nchans, const float *d_bandpass, int nbits,
                               void *stream) {
    if (nsamp < 0 || nchans <= 0 || nbits < 1 || nbits > 8) return bad("invalid geometry or nbits");
    if (nsamp == 0) return AA_HIP_SUCCESS;
    if (!d_data_tc || !d_bandpass) return bad("NULL buffer");
    return from_hip(aa_hip::launch_zero_dm_outliers_u8(d_data_tc, nsamp, nchans, d_bandpass, nbits,
                                                       (hipStream_t)stream),
                    "zero-DM with outliers");
}

int aa_hip_zero_dm_outliers_u16(uint16_t *d_data_tc, long nsamp, int nchans, const float *d_bandpass, int nbits,
                                void *stream) {
    if (nsamp < 0 || nchans <= 0 || nbits < 1 || nbits > 16) return bad("invalid geometry or nbits");
    if (nsamp == 0) return AA_HIP_SUCCESS;
    if (!d_data_tc || !d_bandpass) return bad("NULL buffer");
    return from_hip(aa_hip::launch_zero_dm_outliers_u16(d_data_tc, nsamp, nchans, d_bandpass, nbits,
                                                        (hipStream_t)stream),
                    "zero-DM with outliers");
}

int aa_hip_input_ddtr_normalization_u8(uint8_t *d_data_tc, long nsamp, int nchans, int nbits, void *stream) {
    return input_norm(d_data_tc, nsamp, nchans, nbits, 8, stream);
}

int aa_hip_input_ddtr_normalization_u16(uint16_t *d_data_tc, long nsamp, int nchans, int nbits, void *stream) {
    return input_norm(d_data_tc, nsamp, nchans, nbits, 16, stream);
}

int aa_hip_output_ddtr_normalization(float *d_plane, int ndms, long nsamp, void *stream) {
    if (ndms < 0 || nsamp < 0) return bad("invalid geometry");
    if (ndms == 0 || nsamp == 0) return AA_HIP_SUCCESS;
    if (!d_plane) return bad("NULL buffer");
    return from_hip(aa_hip::launch_normalize_columns_f32(d_plane, ndms, nsamp, 127.5f, (hipStream_t)stream),
                    "output DDTR normalisation");
}

int aa_hip_bandpass_average(const void *h_in_tc, int bytes_per_sample, size_t nchans, size_t nsamp,
                            float *h_bandpass) {
    if (bytes_per_sample != 1 && bytes_per_sample != 2) return bad("bytes_per_sample must be 1 or 2");
    if (nchans == 0) return AA_HIP_SUCCESS;
    if (nsamp == 0) return bad("nsamp must be positive");
    if (!h_in_tc || !h_bandpass) return bad("NULL buffer");
    aa_hip::bandpass_average(h_in_tc, bytes_per_sample, nchans, nsamp, h_bandpass);
    return AA_HIP_SUCCESS;
}

int aa_hip_dedisperse_u8(const uint8_t *d_in, size_t pitch, int row_len, int nchans, const float *h_dmshifts,
                         const float *d_dmshifts, int ndms, int t_processed, float mstartdm, float mdmstep,
                         int inBin, float *d_out, size_t out_pitch, int kernel, void *stream) {
    if (!d_in || !h_dmshifts || !d_dmshifts || !d_out) return bad("NULL buffer");
    if (nchans <= 0 || ndms < 0 || t_processed < 0 || inBin <= 0 || row_len < 0 || (size_t)row_len > pitch ||
        out_pitch < (size_t)t_processed)
        return bad("invalid geometry");
    if (kernel < AA_HIP_KERNEL_AUTO || kernel > AA_HIP_KERNEL_LDS_QUAD) return bad("unknown kernel variant");
    const int st = aa_hip::check_delays(h_dmshifts, nchans, ndms, mstartdm, mdmstep, t_processed, row_len);
    if (st != AA_HIP_SUCCESS) return st;
    aa_hip::range_launch rl;
    const int rs = aa_hip::resolve_launch(kernel, d_in, pitch, nchans, d_dmshifts, h_dmshifts, ndms, t_processed,
                                          mstartdm, mdmstep, inBin, d_out, out_pitch, aa_hip::lds_choice{-2, 0}, &rl);
    if (rs != AA_HIP_SUCCESS) return rs;
    if (rl.variant >= 2 && kernel == AA_HIP_KERNEL_AUTO) {
        // This handle-free call keeps no plan tables: left to the automatic
        // choice, it takes a byte-row shape when one fits, which needs none
        // and keeps the call asynchronous.
        aa_hip::lds_choice c = aa_hip::choose_lds(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 0);
        if (c.variant < 0) c = aa_hip::choose_lds(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 1);
        if (c.variant >= 0) {
            rl.variant = c.variant;
            rl.row_quads = c.row_quads;
        }
    }
    if (rl.variant < 2) return from_hip(aa_hip::launch_range(rl, (hipStream_t)stream), "dedisperse");
    // Chunk shapes 2 and 3 read plan tables; this handle-free call builds them,
    // launches, waits for the stream and frees them (handles keep theirs).
    std::vector<uint8_t> bytes;
    if (aa_hip::lds_plan_tables(rl.variant, h_dmshifts, nchans, ndms, mstartdm, mdmstep, pitch, &bytes) !=
        rl.row_quads)
        return bad("plan tables disagree with the chosen row width");
    void *t = nullptr;
    hipError_t e = hipMalloc(&t, bytes.size());
    if (e == hipSuccess) e = hipMemcpy(t, bytes.data(), bytes.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rl.pair_recs = t;
        e = aa_hip::launch_range(rl, (hipStream_t)stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (t) (void)hipFree(t);
    return from_hip(e, "dedisperse");
}

size_t aa_hip_spd_workspace_bytes(void) { return aa_hip::spd_workspace_bytes(); }

size_t aa_hip_spd_noalloc_scratch_bytes(void) { return aa_hip::spd_static_scratch_bytes(); }

int aa_hip_spd_first_pass_noalloc(const float *d_plane, size_t pitch, int ndms, int nsamp, int max_width,
                                  float *d_msd, float *d_snr, size_t snr_pitch, uint16_t *d_width,
                                  size_t width_pitch, void *stream) {
    if (!d_plane || !d_msd) return bad("NULL buffer");
    if (ndms <= 0 || nsamp <= 0 || pitch < (size_t)nsamp) return bad("invalid plane geometry");
    if (max_width < 1 || max_width > AA_HIP_SPD_MAX_WIDTH || max_width > nsamp)
        return bad("invalid max_width: the static workspace holds moments for widths 1..32 (and max_width <= nsamp)");
    if ((d_snr && snr_pitch < (size_t)nsamp) || (d_width && width_pitch < (size_t)nsamp)) return bad("invalid output pitch");
    if (aa_hip::spd_workspace_bytes() > aa_hip::spd_static_scratch_bytes())
        return bad("the first pass needs more workspace than the static scratch array holds");
    return from_hip(aa_hip::launch_spd_first_pass_static(d_plane, pitch, ndms, nsamp, max_width, d_msd, d_snr,
                                                         snr_pitch, d_width, width_pitch, (hipStream_t)stream),
                    "single-pulse first pass (no allocation)");
}

int aa_hip_spd_first_pass(const float *d_plane, size_t pitch, int ndms, int nsamp, int max_width, float *d_msd,
                          float *d_snr, size_t snr_pitch, uint16_t *d_width, size_t width_pitch, void *d_work,
                          size_t work_bytes, void *stream) {
    static_assert(AA_HIP_SPD_MAX_WIDTH == SPD_MAX_WIDTH, "one width limit");
    if (!d_plane || !d_msd || !d_work) return bad("NULL buffer");
    if (ndms <= 0 || nsamp <= 0 || pitch < (size_t)nsamp) return bad("invalid plane geometry");
    if (max_width < 1 || max_width > AA_HIP_SPD_MAX_WIDTH || max_width > nsamp) return bad("invalid max_width");
    if ((d_snr && snr_pitch < (size_t)nsamp) || (d_width && width_pitch < (size_t)nsamp)) return bad("invalid output pitch");
    if (work_bytes < aa_hip::spd_workspace_bytes()) return bad("workspace too small");
    if ((uintptr_t)d_work % 256 != 0) return bad("workspace not 256-byte aligned");
    return from_hip(aa_hip::launch_spd_first_pass(d_plane, pitch, ndms, nsamp, max_width, d_msd, d_snr, snr_pitch,
                                                  d_width, width_pitch, d_work, (hipStream_t)stream),
                    "single-pulse first pass");
}

int aa_hip_synth_noise_u8(uint64_t seed, int64_t n, double mean, double sigma, uint8_t *d_out_tc, void *stream) {
    if (!d_out_tc) return bad("NULL buffer");
    return from_hip(aa_hip::launch_synth_noise_u8(seed, n, mean, sigma, d_out_tc, (hipStream_t)stream), "synth");
}

int aa_hip_inject_pulse_u8(uint8_t *d_data_tc, int nsamp, int nchans, const float *d_dmshifts, double dm,
                           double tsamp, int t0, int width, int amplitude, void *stream) {
    if (!d_data_tc || !d_dmshifts) return bad("NULL buffer");
    const float dm_over_tsamp = (float)(dm / tsamp);
    return from_hip(aa_hip::launch_inject_pulse_u8(d_data_tc, nsamp, nchans, d_dmshifts, dm_over_tsamp, t0, width,
                                                   amplitude, (hipStream_t)stream),
                    "inject");
}

}  // extern "C"
