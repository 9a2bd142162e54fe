/*
 * C ABI of the MI355X (gfx950) dedispersion library libaa_ddtr_hip.so.
 *
 * Two layers, both plain C (no HIP, torch or C++ types in any signature):
 *
 *  1. aa_py_* -- the entry points the reference's Python FFI binds with
 *     ctypes (cmake/py_astro_accelerate.py.in), restricted to the
 *     dedispersion path.  Same names, argument meaning and return
 *     conventions; each declaration cites the reference definition it
 *     replaces.  Host buffers in, host float*** out.
 *
 *  2. aa_hip_* -- device-level entry points of the hot path: the pieces
 *     the reference's runner calls on device memory (corner turn, binning,
 *     dedisperse, chunk execution), re-cut as explicit-stream calls on
 *     caller-owned device buffers with no process-global state.  Streams are
 *     passed as `void*` holding a hipStream_t (NULL = legacy default stream).
 *
 * Return convention of the aa_hip_* layer: 0 (AA_HIP_SUCCESS) on success,
 * a negative aa_hip_status on error; aa_hip_last_error() returns a message
 * for the calling thread's most recent failure.
 */
#ifndef AA_DDTR_HIP_H
#define AA_DDTR_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
typedef enum {
    AA_HIP_SUCCESS = 0,
    AA_HIP_ERROR_INVALID_ARGUMENT = -1,
    AA_HIP_ERROR_DEVICE = -2,          /* a HIP runtime call failed */
    AA_HIP_ERROR_OUT_OF_MEMORY = -3,
    AA_HIP_ERROR_UNSUPPORTED = -4,     /* geometry/option outside this library's scope */
    AA_HIP_ERROR_INPUT_RANGE = -5,     /* a u16 input sample exceeded 255 */
    AA_HIP_ERROR_NO_DEVICE = -6
} aa_hip_status;

/* Kernel variants for aa_hip_dedisperse_u8 / aa_hip_ddtr_set_kernel. */
typedef enum {
    AA_HIP_KERNEL_AUTO = 0,     /* LDS kernel when a chunk shape fits (tuned shape if set), else direct */
    AA_HIP_KERNEL_DIRECT = 1,   /* one thread per (DM, sample), global gathers */
    AA_HIP_KERNEL_LDS = 2,      /* LDS-staged windows, packed integer sums, first shape that fits */
    AA_HIP_KERNEL_LDS_C32 = 3,  /* LDS kernel, 32-channel chunks (rows up to 2432 bytes) */
    AA_HIP_KERNEL_LDS_C16 = 4,  /* LDS kernel, 16-channel chunks (rows up to 4864 bytes) */
    AA_HIP_KERNEL_LDS_PAIR = 5, /* LDS kernel, 16-channel chunks pre-summed in channel pairs
                                   (the widest tile that fits: 2016, 1512 or 1008
                                   samples, read with 16-byte LDS reads) */
    AA_HIP_KERNEL_LDS_QUAD = 6  /* LDS kernel, 16-channel chunks pre-summed in groups of four */
} aa_hip_kernel;

/* Message for the calling thread's most recent aa_hip_* failure ("" if none). */
const char *aa_hip_last_error(void);
/* Library version string. */
const char *aa_hip_version(void);

/* ======================================================================
 * Layer 1: reference FFI (aa_py_*), dedispersion path only.
 * ====================================================================== */

/* Replaces src/aa_py_ddtr_plan.cpp:9-11 (aa_py_ddtr_plan). */
void *aa_py_ddtr_plan(void);
/* Replaces src/aa_py_ddtr_plan.cpp:13-15. */
void aa_py_ddtr_plan_delete(void const *obj);
/* Replaces src/aa_py_ddtr_plan.cpp:17-24. */
bool aa_py_ddtr_plan_add_dm(void *obj, float low, float high, float step, int inBin, int outBin);
/* Replaces src/aa_py_ddtr_plan.cpp:26-28. */
size_t aa_py_ddtr_plan_range(void const *obj);
/* Replaces src/aa_py_ddtr_plan.cpp:30-32. */
bool aa_py_ddtr_plan_set_power(void *obj, float power);
/* Replaces src/aa_py_ddtr_plan.cpp:34-36. */
float aa_py_ddtr_plan_power(void const *obj);
/* Replaces src/aa_py_ddtr_plan.cpp:38-40. */
bool aa_py_ddtr_plan_set_enable_msd_baseline(void *obj, bool flag);
/* Replaces src/aa_py_ddtr_plan.cpp:42-44. */
bool aa_py_ddtr_plan_enable_msd_baseline_noise(void const *obj);
/* Replaces src/aa_py_ddtr_plan.cpp:46-48. */
bool aa_py_ddtr_plan_bind_bandpass_normalization(void *obj, float const *data, int data_size);

/* Same field order as the reference's aa_py_filterbank_metadata_struct
 * (include/aa_py_filterbank_metadata.hpp:15-25). */
typedef struct {
    double m_tstart;
    double m_tsamp;
    double m_fch1;
    double m_foff;
    int m_nbits;
    int m_nsamples;
    int m_nchans;
    double m_src_raj;
} aa_py_filterbank_metadata_struct;

/* Replaces src/aa_py_filterbank_metadata.cpp:8-27. */
aa_py_filterbank_metadata_struct *aa_py_filterbank_metadata(double tstart, double tsamp, int nbits,
                                                            int nsamples, double fch1, double foff,
                                                            int nchans, double src_raj);
/* Replaces src/aa_py_filterbank_metadata.cpp:29-31. */
void aa_py_filterbank_metadata_delete(aa_py_filterbank_metadata_struct const *obj);
/* Replaces src/aa_py_filterbank_metadata.cpp:33-35. */
double aa_py_filterbank_metadata_tstart(aa_py_filterbank_metadata_struct const *obj);

/* Same layout as include/aa_py_pipeline_api.hpp:20-25. */
typedef struct {
    bool dedispersion;
    bool analysis;
    bool periodicity;
    bool fdas;
} pipeline_components_struct;

/* Same layout as include/aa_py_pipeline_api.hpp:32-48.  Honoured:
 * zero_dm, zero_dm_with_outliers, input_DDTR_normalization,
 * output_DDTR_normalization, set_bandpass_average and
 * copy_ddtr_data_to_host, each as the reference's runner applies it.
 * Refused: old_rfi (run() fails, "unsupported pipeline option" in
 * aa_hip_last_error()).  The others concern output files or components
 * outside dedispersion and are ignored (logged). */
typedef struct {
    bool zero_dm;
    bool zero_dm_with_outliers;
    bool input_DDTR_normalization;
    bool output_DDTR_normalization;
    bool set_bandpass_average;
    bool old_rfi;
    bool copy_ddtr_data_to_host;
    bool msd_baseline_noise;
    bool output_dmt;
    bool output_ffdot_plan;
    bool output_fdas_list;
    bool candidate_algorithm;
    bool fdas_custom_fft;
    bool fdas_inbin;
    bool fdas_norm;
} pipeline_component_option_struct;

/* Replaces src/aa_py_pipeline_api.cpp:9-44: a pipeline over a borrowed host
 * buffer of u16 samples in [t][c] order (8-bit values, as the reference's
 * SIGPROC reader widens them). */
void *aa_py_pipeline_api(pipeline_components_struct pipeline,
                         pipeline_component_option_struct options,
                         aa_py_filterbank_metadata_struct metadata,
                         unsigned short const *input_data, int card_number);
/* No reference counterpart: the same pipeline over u8 [t][c] host samples,
 * which halves host memory and PCIe traffic. */
void *aa_py_pipeline_api_u8(pipeline_components_struct pipeline,
                            pipeline_component_option_struct options,
                            aa_py_filterbank_metadata_struct metadata,
                            unsigned char const *input_data, int card_number);
/* No reference counterpart over the FFI (the C++ API takes the cap through
 * aa_device_info's constructor, include/aa_device_info.hpp:61-130): the u8
 * pipeline with the device memory it may plan for capped at
 * max_device_bytes (0: no cap), which sets the time-chunk size. */
void *aa_hip_py_pipeline_api_u8_capped(pipeline_components_struct pipeline,
                                       pipeline_component_option_struct options,
                                       aa_py_filterbank_metadata_struct metadata,
                                       unsigned char const *input_data, int card_number,
                                       size_t max_device_bytes);
/* Replaces src/aa_py_pipeline_api.cpp:46-48. */
void aa_py_pipeline_api_delete(void const *obj);
/* Replaces src/aa_py_pipeline_api.cpp:50-52. */
bool aa_py_pipeline_api_bind_ddtr_plan(void *obj, void const *plan);
/* Replaces src/aa_py_pipeline_api.cpp:54-56 (returns a new strategy copy,
 * released with aa_py_ddtr_strategy_delete). */
void *aa_py_pipeline_api_ddtr_strategy(void *obj);
/* Replaces src/aa_py_ddtr_strategy.cpp:6-8. */
void aa_py_ddtr_strategy_delete(void const *obj);
/* Replaces src/aa_py_pipeline_api.cpp:58-60. */
size_t aa_py_get_ddtr_nRanges(void *obj);
/* Replaces src/aa_py_pipeline_api.cpp:62-64. */
const int *aa_py_get_ndms_array(void *obj);
/* Replaces src/aa_py_pipeline_api.cpp:66-68. */
int aa_py_get_t_proc(void *obj, int i, int j);
/* Replaces src/aa_py_pipeline_api.cpp:70-72 (returns int, truncating a
 * fractional range start exactly as the reference does). */
int aa_py_dm_low(void *obj, int range);
/* Replaces src/aa_py_pipeline_api.cpp:74-76. */
int aa_py_total_computed_samples(void *obj);
/* Replaces src/aa_py_pipeline_api.cpp:92-103; the reference's `int&` is a
 * pointer at the ABI level. */
bool aa_py_pipeline_api_run(void *obj, int *status_code_int);
/* Replaces src/aa_py_pipeline_api.cpp:104-106: float***[range][dm][t]. */
float ***aa_py_buffer(void *obj);
/* Replace src/aa_py_pipeline_api.cpp:140-151. */
int aa_py_current_range(void *obj);
int aa_py_current_time_chunk(void *obj);
long aa_py_current_inc(void *obj);
/* Replaces src/aa_py_pipeline_api.cpp:152-154. */
bool aa_py_cleanup(void *obj);

/* No reference counterpart (the reference builds strategies only inside a
 * pipeline): strategy of a plan for explicit metadata and free device
 * memory, computed on the host only.  Accessors mirror
 * include/aa_ddtr_strategy.hpp:55-151. */
void *aa_py_ddtr_strategy_create(void const *plan, aa_py_filterbank_metadata_struct metadata,
                                 size_t free_memory, bool enable_analysis);
bool aa_py_ddtr_strategy_ready(void const *obj);
int aa_py_ddtr_strategy_maxshift(void const *obj);
int aa_py_ddtr_strategy_num_tchunks(void const *obj);
int aa_py_ddtr_strategy_max_ndms(void const *obj);
size_t aa_py_ddtr_strategy_nranges(void const *obj);
int aa_py_ddtr_strategy_ndms(void const *obj, int range);
int aa_py_ddtr_strategy_t_processed(void const *obj, int range, int chunk);
/* Strategy-redefined range: low/high/step (floats) and inBin/outBin. */
bool aa_py_ddtr_strategy_dm(void const *obj, int range, float *low, float *high, float *step,
                            int *inBin, int *outBin);
/* Copies nchans dmshift values into out. */
bool aa_py_ddtr_strategy_dmshifts(void const *obj, float *out, int nchans);

/* ======================================================================
 * Layer 2: device-level hot path (aa_hip_*).
 * ====================================================================== */

typedef struct aa_hip_ddtr aa_hip_ddtr;

/* Creates a dedispersion plan bound to `device`.  Replaces the state that
 * src/aa_device_load_data.cu:39-61 stores in __constant__ memory
 * (dm_shifts, i_nchans, i_nsamp, i_t_processed_s), made per handle.
 * dm_low/dm_step are the strategy-redefined ranges (aa_ddtr_strategy::dm),
 * tsamp is the metadata tsamp (narrowed to float as the reference does),
 * maxshift the strategy maxshift, max_t_processed0 the largest
 * t_processed[0][chunk] this handle will execute. */
int aa_hip_ddtr_create(int device, int nchans, const float *dmshifts, int nranges,
                       const float *dm_low, const float *dm_step, const int *ndms,
                       const int *inBin, double tsamp, int maxshift, int max_t_processed0,
                       aa_hip_ddtr **out);
/* As aa_hip_ddtr_create, for samples wider than 8 bits (16-bit filterbanks,
 * and 32-bit ones the reader scales to u16; src/aa_sigproc_input.cpp:400-433):
 * the cascade levels stay u16 and each output is the float channel sum in
 * channel order, as the reference's cache kernel forms it
 * (src/aa_device_dedispersion_kernel.cu:422-450; it rounds once a sum passes
 * 2^24).  Ingest takes input_bits 16 only; the kernel choice is the 16-bit
 * LDS kernel (AA_HIP_KERNEL_AUTO) or the direct kernel (DIRECT, and AUTO
 * where a tile's delay spread exceeds the LDS rows).  Not usable in
 * batches. */
int aa_hip_ddtr_create_u16(int device, int nchans, const float *dmshifts, int nranges,
                           const float *dm_low, const float *dm_step, const int *ndms,
                           const int *inBin, double tsamp, int maxshift, int max_t_processed0,
                           aa_hip_ddtr **out);
int aa_hip_ddtr_destroy(aa_hip_ddtr *h);
/* Selects the dedispersion kernel variant (aa_hip_kernel) for this handle. */
int aa_hip_ddtr_set_kernel(aa_hip_ddtr *h, int kernel);
/* Device bytes held by the handle (dmshift table, cascade levels). */
size_t aa_hip_ddtr_device_bytes(const aa_hip_ddtr *h);

/* One time chunk, end to end on the device: corner turn of
 * t_processed[0] + maxshift rows of d_input_tc (u8 when input_bits == 8,
 * u16 carrying 8-bit values when input_bits == 16), the x2 binning cascade,
 * and the dedispersion of every range into d_out[r] (ndms_r rows of
 * t_processed[r] floats, row pitch out_pitch[r] floats).  Replaces the
 * per-chunk DDTR block of include/aa_permitted_pipelines_generic.hpp:687-840
 * (load_chunk_data is the caller's H2D; D2H is the caller's). */
int aa_hip_ddtr_execute_chunk(aa_hip_ddtr *h, const void *d_input_tc, int input_bits,
                              const int *t_processed, float *const *d_out,
                              const size_t *out_pitch, void *stream);
/* The same chunk split in two steps, for callers that interleave work:
 * ingest = corner turn into the handle's level-0 buffer; run_range =
 * binning (when the range's bin exceeds the resident level) + one range. */
int aa_hip_ddtr_ingest(aa_hip_ddtr *h, const void *d_input_tc, int input_bits, int t_processed0,
                       void *stream);
int aa_hip_ddtr_run_range(aa_hip_ddtr *h, int range, int t_processed, float *d_out,
                          size_t out_pitch, void *stream);

/* Tile tuning (no reference counterpart: the reference picks its kernel by a
 * fixed LDS-size test, src/aa_dedisperse.cpp:90-131).  The LDS kernel has
 * seven channel-chunk shapes: 0 = 32 channels x 2432-byte rows, 1 = 16
 * channels x 4864-byte rows, 2 = 16 channels pre-added in channel pairs
 * (rows of 16-bit pair sums, one per distinct relative delay of the pair in
 * a 16-DM tile) on 1008-sample tiles (window offsets up to 772 samples),
 * 3 = 16 channels pre-added in groups of four, 4 = shape 2 on 2016-sample
 * tiles (offsets up to 268), 5 = shape 2 on 1512-sample tiles (offsets up
 * to 520), 6 = shape 2 with 32-DM tiles (two DM trials per wave over the
 * same rows); shapes 2, 4, 5 and 6 read rows with 16-byte LDS reads.
 * set_range_variant records a preference for range r (-1 = first that
 * fits, in the order 4, 5, 2, 3, 0, 1, 6), used whenever it fits.
 * autotune_range times every shape that fits range r on `stream` (reps
 * launches each, writing d_out), keeps the fastest as the preference and
 * reports it: *best_variant (-1 if none fits) and variant_ms[0..6] (-1 =
 * does not fit; the array holds 7 floats).
 * The range's level must be resident (after ingest). */
int aa_hip_ddtr_set_range_variant(aa_hip_ddtr *h, int range, int variant);
int aa_hip_ddtr_autotune_range(aa_hip_ddtr *h, int range, int t_processed, float *d_out,
                               size_t out_pitch, int reps, void *stream, int *best_variant,
                               float *variant_ms);

/* Many beams in one launch (no reference counterpart: the reference runs one
 * beam per process, include/aa_pipeline_api.hpp:113-117).  A batch borrows
 * the per-beam handles (their plans, levels and tuning); they must outlive
 * it and share one device.  execute: per beam b, d_inputs_tc[b] is its
 * [t][c] u8 chunk; t_processed, d_out and out_pitch are flattened
 * beam-major over each beam's ranges (as execute_chunk takes them per beam).
 * Every beam is corner-turned and binned, then all beams' ranges run in one
 * launch per kernel instantiation over shared tile queues: the two byte-row
 * LDS shapes in one persistent launch, the pair pre-summation shapes in one
 * launch per tile width and row width; ranges on the quad shape, or that no
 * chunk shape fits, run launches of their own after them (in the same
 * graph).  With use_graph != 0 the
 * sequence is captured into a hipGraph on the first call for a given set of
 * buffers and replayed by later calls with the same buffers.  Results are
 * bit-identical to aa_hip_ddtr_execute_chunk per beam.  Calls on one batch
 * (or on its beams) must not overlap. */
typedef struct aa_hip_ddtr_batch aa_hip_ddtr_batch;
int aa_hip_ddtr_batch_create(int nbeams, aa_hip_ddtr *const *beams, aa_hip_ddtr_batch **out);
int aa_hip_ddtr_batch_destroy(aa_hip_ddtr_batch *h);
int aa_hip_ddtr_batch_execute(aa_hip_ddtr_batch *h, const void *const *d_inputs_tc,
                              const int *t_processed, float *const *d_out,
                              const size_t *out_pitch, int use_graph, void *stream);
/* Jobs (ranges) in the batched launches, their tiles, and ranges launched
 * on their own (quad shape or direct kernel) in the last execute. */
int aa_hip_ddtr_batch_stats(const aa_hip_ddtr_batch *h, int *njobs, long *ntiles, int *ndirect);
/* Corner turns of the last execute: *batched = launches that turned many
 * beams at once (one per number of binned levels written), *own = beams
 * turned by a launch of their own (channel counts that are not a multiple
 * of 16, or unaligned buffers). */
int aa_hip_ddtr_batch_ct_stats(const aa_hip_ddtr_batch *h, int *batched, int *own);

/* Replaces corner_turn_SM_inplace (src/aa_corner_turn.cpp:118-121) for the
 * 8-bit path: [t][c] -> [c][t] into a separate u8 buffer with row pitch
 * out_pitch bytes (>= nsamp). */
int aa_hip_corner_turn_u8(const uint8_t *d_in_tc, int nsamp, int nchans, uint8_t *d_out_ct,
                          size_t out_pitch, void *stream);
/* As above for u16 input holding 8-bit samples.  A sample above 255 sets
 * *d_err (device int) to 1; the samples are narrowed regardless. */
int aa_hip_corner_turn_u16(const uint16_t *d_in_tc, int nsamp, int nchans, uint8_t *d_out_ct,
                           size_t out_pitch, int *d_err, void *stream);
/* Replaces corner_turn_SM_inplace (src/aa_corner_turn.cpp:118-121) followed
 * by the runner's bin_gpu cascade (include/aa_permitted_pipelines_generic.hpp:
 * 774-812, src/aa_bin_gpu.cpp:8-44) in one pass over the input: u8
 * (sample_bytes 1) or u16 holding 8-bit samples (sample_bytes 2; a value
 * above 255 sets *d_err) [t][c] -> level 0 u8 [c][t] at d_levels[0] and the
 * x2 levels k = 1..nlev (nlev <= 3) at d_levels[k]:
 * L_k[c][t] = floor((L_{k-1}[c][2t] + L_{k-1}[c][2t+1]) / 2), t < nsamp >> k.
 * Row k holds pitches[k] >= nsamp >> k bytes.  *fused (may be NULL) is set to
 * the number of binned levels written: nlev when nchans is a multiple of 16
 * and the input, rows and pitches are 16-byte aligned, else 0 (then only
 * level 0 is written and the levels come from aa_hip_bin2_u8). */
int aa_hip_corner_turn_levels(const void *d_in_tc, int sample_bytes, int nsamp, int nchans,
                              uint8_t *const *d_levels, const size_t *pitches, int nlev, int *d_err,
                              int *fused, void *stream);
/* Replaces bin_gpu (src/aa_bin_gpu.cpp:8-44) for the 8-bit path, one x2
 * level: out[c][t] = floor((in[c][2t] + in[c][2t+1]) / 2), t < n_out. */
int aa_hip_bin2_u8(const uint8_t *d_in, size_t in_pitch, int nchans, int n_out, uint8_t *d_out,
                   size_t out_pitch, void *stream);
/* Replaces dedisperse() (src/aa_dedisperse.cpp:5-163) + the failsafe kernel
 * (src/aa_device_dedispersion_kernel.cu:422-450) for one range on u8 rows:
 * out[d*out_pitch + t] = ((float)sum_c L[c*pitch + t + delay(c,d)] / nchans) / inBin,
 * delay(c,d) = trunc(dmshift[c] * fma(d, mdmstep, mstartdm)).  The dmshift
 * table is passed twice: h_dmshifts (host) is used to check, before the
 * launch, that every delay lies in [0, row_len - t_processed]; d_dmshifts
 * (device) is what the kernel reads.  row_len is the number of valid
 * samples per row (<= pitch); rows must also hold at least 4096 readable
 * bytes past row_len (the LDS kernel reads whole windows).  This call keeps
 * no state: with AA_HIP_KERNEL_AUTO it uses a byte-row LDS shape when one
 * fits (asynchronous on `stream`); the pre-summation kernels (PAIR, QUAD,
 * or AUTO when only they fit) build their plan tables for the call and
 * synchronise `stream` before returning. */
int aa_hip_dedisperse_u8(const uint8_t *d_in, size_t pitch, int row_len, int nchans,
                         const float *h_dmshifts, const float *d_dmshifts, int ndms, int t_processed,
                         float mstartdm, float mdmstep, int inBin, float *d_out, size_t out_pitch,
                         int kernel, void *stream);

/* ------------------------------------------------------------------------
 * Hand-off of DM-time planes to the single-pulse search and its first pass.
 *
 * The reference's runner passes each range's device plane d_DDTR_output,
 * [nDMs][nTimesamples] floats, to analysis_GPU right after dedispersing it
 * (include/aa_permitted_pipelines_generic.hpp:848-889; the plane is read at
 * src/aa_device_analysis.cu:126-133).  Here a pipeline hands every range's
 * plane to a consumer callback with its strides, and aa_hip_spd_first_pass
 * is the first stage such a consumer runs.
 * ---------------------------------------------------------------------- */
typedef struct {
    int range;              /* DM range */
    int time_chunk;         /* time chunk */
    float *d_plane;         /* device plane: d_plane[dm * pitch + t] */
    int ndms;               /* nDMs of the range */
    int nsamp;              /* nTimesamples: t_processed[range][time_chunk] */
    size_t pitch;           /* floats between DM rows */
    float dm_low, dm_step;  /* strategy-redefined DM grid of the range */
    int inBin;              /* binning of the range */
    double tsamp;           /* sampling time of the plane (metadata tsamp x inBin) */
    long inc;               /* first input sample of the chunk (column inc / inBin of the range) */
    int device;             /* the device holding the plane (current during the call) */
    void *stream;           /* hipStream_t the plane was written on */
} aa_hip_ddtr_plane;
/* Called once per (time chunk, range), after the range's dedispersion and
 * output normalisation are enqueued and before its host copy.  Work the
 * consumer enqueues on plane->stream is ordered before the copy and before
 * the buffer is reused; work on other streams must wait for plane->stream
 * and finish before the callback returns.  A non-zero return stops the
 * pipeline with an error (as a failed analysis_GPU does). */
typedef int (*aa_hip_ddtr_consumer)(const aa_hip_ddtr_plane *plane, void *user);
/* No reference counterpart over the FFI (the reference calls analysis_GPU
 * itself): sets (fn != NULL) or clears the consumer of a pipeline from
 * aa_py_pipeline_api / _u8 / aa_hip_py_pipeline_api_u8_capped. */
bool aa_hip_pipeline_set_ddtr_consumer(void *pipeline, aa_hip_ddtr_consumer fn, void *user);

/* First pass of the single-pulse search on one plane (d_plane[dm * pitch + t],
 * ndms x nsamp floats).  For widths w = 1..max_width (max_width <=
 * AA_HIP_SPD_MAX_WIDTH and <= nsamp) and every start sample t with t + w <=
 * nsamp, B_w(t) = sum of x[dm][t .. t+w-1]:
 *   d_msd[2(w-1)], d_msd[2(w-1)+1] = mean and population sd of all B_w over
 *     the plane (the per-width MSD that src/aa_device_MSD_plane_profile.cu
 *     estimates; exact here), float2 per width as the reference's d_MSD;
 *   d_snr[dm * snr_pitch + t] = max over the widths that fit at t of
 *     (B_w(t) - mean_w) / sd_w, and d_width[dm * width_pitch + t] = that w
 *     (the first boxcar plane of src/aa_device_SPS_long_kernel.cu:42-168;
 *     ties go to the smaller width, sd_w == 0 gives SNR 0).
 * d_snr and d_width may be NULL (MSD only).  d_work: device scratch of
 * aa_hip_spd_workspace_bytes() bytes, 256-byte aligned.  Asynchronous on
 * `stream`; the result does not depend on scheduling. */
#define AA_HIP_SPD_MAX_WIDTH 32
size_t aa_hip_spd_workspace_bytes(void);
int aa_hip_spd_first_pass(const float *d_plane, size_t pitch, int ndms, int nsamp, int max_width, float *d_msd,
                          float *d_snr, size_t snr_pitch, uint16_t *d_width, size_t width_pitch, void *d_work,
                          size_t work_bytes, void *stream);
/* No counterpart in the reference: aa_hip_spd_first_pass for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments without the
 * workspace, same results bit for bit.  The workspace is a file-scope device
 * array of aa_hip_spd_noalloc_scratch_bytes() bytes (786,960: the fixed
 * workspace of the pass for any plane size and max_width <= 32), one copy per
 * device, held until the process ends; its address is looked up on the first
 * call on each device (the current device).  max_width outside 1..32 is
 * rejected.  Calls on one device share the array: calls that overlap (two
 * streams, two host threads without ordering) race on it and must be
 * serialised by the caller.  Allocates nothing, asynchronous on `stream`. */
size_t aa_hip_spd_noalloc_scratch_bytes(void);
int aa_hip_spd_first_pass_noalloc(const float *d_plane, size_t pitch, int ndms, int nsamp, int max_width,
                                  float *d_msd, float *d_snr, size_t snr_pitch, uint16_t *d_width,
                                  size_t width_pitch, void *stream);

/* Synthetic input on the device (no reference counterpart; benchmarks and
 * tests).  Noise: N(mean, sigma^2) rounded to nearest and clipped to
 * [0,255], a SplitMix64 counter stream keyed by `seed` over the flat [t][c]
 * index (Box-Muller in double). */
int aa_hip_synth_noise_u8(uint64_t seed, int64_t n, double mean, double sigma, uint8_t *d_out_tc,
                          void *stream);
/* Adds `amplitude` (saturating at 255) to `width` samples starting at
 * t0 + floor(dmshift[c] * (float)(dm / tsamp)) of each channel c
 * (the injector convention of src/aa_host_fake_signal_generator.cpp:75-78). */
int aa_hip_inject_pulse_u8(uint8_t *d_data_tc, int nsamp, int nchans, const float *d_dmshifts,
                           double dm, double tsamp, int t0, int width, int amplitude, void *stream);

/* ------------------------------------------------------------------------
 * Fractional-delay (differentiable) dedispersion of FP32 candidate windows.
 * No counterpart in the reference (it has no autograd, fractional delays or
 * batching); it uses the reference's delay table, as samples per unit DM
 * (dmshift[c] / tsamp, src/aa_ddtr_strategy.cpp:81-92).
 *
 * x is [B][C][T] float (T = nsamp, a multiple of 4), dm[B] the candidates'
 * continuous DM offsets, trials[K] the trial DM offsets.  For each b, k, t:
 *   u = trials[k] + dm[b]; tau = delay_per_dm[c] * u (in double);
 *   y[b][k][t] = sum_c (1 - f) x[b][c][t + i0] + f x[b][c][t + i0 + 1],
 *   i0 = floor(tau), f = (float)(tau - i0), x = 0 outside [0, T).
 * The adjoint gives gx = (dy/dx)^T gy and gdm[b] = sum_{k,t} gy * dy/d(dm[b])
 * (accumulated in double).  Device buffers are caller-owned, 16-byte
 * aligned; calls on one handle must not overlap.
 * ---------------------------------------------------------------------- */
typedef struct aa_hip_frac aa_hip_frac;
/* dm_absmax bounds |dm[b]| for sizing the input adjoint's channel groups
 * (larger values stay correct, through a slower path). */
int aa_hip_frac_create(int device, int nchans, const float *delay_per_dm, int ntrials,
                       const float *trials, int nsamp, int max_batch, float dm_absmax,
                       aa_hip_frac **out);
int aa_hip_frac_destroy(aa_hip_frac *h);
/* Trials (forward, DM adjoint) and channels (input adjoint) per wave chosen
 * for this plan (1 or 2). */
int aa_hip_frac_grouping(const aa_hip_frac *h, int *trials_per_wave, int *chans_per_wave);
/* y: [nbatch][K][T] float. */
int aa_hip_frac_forward(aa_hip_frac *h, const float *d_x, int nbatch, const float *d_dm, float *d_y,
                        void *stream);
/* gy: [nbatch][K][T]; writes gx ([nbatch][C][T], may be NULL) and gdm
 * ([nbatch] double, may be NULL; needs d_x).  Recomputes the interpolation
 * from x: nothing from the forward pass is stored. */
int aa_hip_frac_adjoint(aa_hip_frac *h, const float *d_x, const float *d_gy, int nbatch,
                        const float *d_dm, float *d_gx, double *d_gdm, void *stream);

/* Batched DM refinement of candidate windows of uneven size (no reference
 * counterpart).  Candidate b's window is float [nchans[b]][nsamp[b]] at
 * d_x + x_offsets[b] (elements), its delay-per-unit-DM table nchans[b]
 * floats at d_delay_per_dm + ds_offsets[b]; up to 4096 x 4096.  For its
 * residual DM u[b] the loss L and dL/du are evaluated by one fused forward +
 * adjoint pass (the adjoint recomputes the interpolation).  The loss
 * (aa_hip_refine_set_loss; default AA_HIP_REFINE_LOSS_ENERGY) is
 *   ENERGY:        E = sum_t y(t)^2 of the dedispersed profile y (the
 *                  aa_hip_frac_forward transform with one trial at offset 0);
 *   EXCESS_ENERGY: E - N, N = sum_c v_c ((1-f_c)^2 (T-|i_c|)+ + f_c^2 (T-|i_c+1|)+)
 *                  the energy white noise of the window's channel powers
 *                  v_c = mean_t x[c][t]^2 would give at the same delays
 *                  (i_c, f_c: integer and fractional delay of channel c):
 *                  interpolation smooths noise least where every f_c is 0,
 *                  so E alone peaks at u = 0 on weak candidates.
 * With scan_n > 0, L is first evaluated (forward only)
 * at u + scan_lo + k * scan_step, k < scan_n, and the best point starts the
 * steps.  Then u[b] climbs L by nsteps sign-based adaptive steps:
 * step[b] grows x1.2 (up to step_max) while the sign of dL/du holds and
 * halves (down to step_min) when it flips; u += sign(dL/du) * step.  One
 * persistent launch for all candidates and steps.  u receives the best point
 * visited (largest L), loss and grad L and dL/du there; step and prev_grad
 * (initialise prev_grad to 0) carry the step state.  nsteps = 0 only
 * evaluates (after the scan, if any). */
typedef struct aa_hip_refine aa_hip_refine;
int aa_hip_refine_create(int device, int ncand, const int64_t *x_offsets, const int64_t *ds_offsets,
                         const int *nchans, const int *nsamp, aa_hip_refine **out);
int aa_hip_refine_destroy(aa_hip_refine *h);
#define AA_HIP_REFINE_LOSS_ENERGY 0
#define AA_HIP_REFINE_LOSS_EXCESS_ENERGY 1
int aa_hip_refine_set_loss(aa_hip_refine *h, int loss_kind);
int aa_hip_refine_run(aa_hip_refine *h, const float *d_x, const float *d_delay_per_dm, float *d_u,
                      float *d_step, double *d_prev_grad, double *d_loss, double *d_grad, int scan_n,
                      float scan_lo, float scan_step, int nsteps, float step_max, float step_min,
                      void *stream);

/* Zero-DM filter, in place on [t][c] samples before the corner turn.
 * Replaces zero_dm(d_input, nchans, nsamp, nbits, d_normalization_factor)
 * (include/aa_zero_dm.hpp:22, src/aa_zero_dm.cpp:28-37,
 * src/aa_device_zero_dm_kernel.cu:37-88): per time sample,
 * x' = (x - mean over channels) + d_bandpass[c] in float, clamped to
 * [0, 2^nbits - 1] and truncated.  d_bandpass: nchans device floats (the
 * strategy's bandpass normalisation, default (2^nbits - 1) / 2).  u8: nbits
 * 1-8; u16: 1-16. */
int aa_hip_zero_dm_u8(uint8_t *d_data_tc, int nsamp, int nchans, const float *d_bandpass, int nbits,
                      void *stream);
int aa_hip_zero_dm_u16(uint16_t *d_data_tc, int nsamp, int nchans, const float *d_bandpass, int nbits,
                       void *stream);

/* Zero-DM with outlier rejection, in place on [t][c] samples before the
 * corner turn.  Replaces zero_dm_outliers(d_input, nchans, nsamp, nbits,
 * d_normalization_factor) (include/aa_zero_dm_outliers.hpp:22,
 * src/aa_zero_dm_outliers.cpp:47-67,
 * src/aa_device_zero_dm_outliers_kernel.cu:190-233): per time sample, the
 * mean and standard deviation over the channels, then up to 5 times over
 * the channels strictly inside mean +- 3 sigma (stopping once the mean
 * drops by less than 0.001), computed with the reference's streaming
 * accumulators and merge order; x' = (x - mean) + d_bandpass[c] in float,
 * clamped to [0, 2^nbits - 1], truncated (0 where the mean is NaN, i.e. no
 * channel was inside the limits).  u8: nbits 1-8; u16: 1-16. */
int aa_hip_zero_dm_outliers_u8(uint8_t *d_data_tc, long nsamp, int nchans, const float *d_bandpass, int nbits,
                               void *stream);
int aa_hip_zero_dm_outliers_u16(uint16_t *d_data_tc, long nsamp, int nchans, const float *d_bandpass, int nbits,
                                void *stream);

/* Input DDTR normalisation, in place on a [t][c] chunk before its corner
 * turn.  Replaces zero_dm_normalization_dm(d_input, nchans, nsamp, nbits)
 * as the reference's runner calls it (include/aa_zero_dm.hpp:26,
 * src/aa_zero_dm.cpp:49-60, src/aa_device_zero_dm_kernel.cu:100-164,
 * include/aa_permitted_pipelines_generic.hpp:731-735): per channel,
 * s = (float sum over the chunk's nsamp samples, in the reference's order) /
 * nsamp - (2^nbits - 1)/2; x' = x - s in float, clamped to [0, 2^nbits - 1],
 * truncated.  u8: nbits 1-8; u16: 1-16.  Allocates 33 * nchans * 4 bytes
 * of device scratch and returns after the stream has finished the work. */
int aa_hip_input_ddtr_normalization_u8(uint8_t *d_data_tc, long nsamp, int nchans, int nbits, void *stream);
int aa_hip_input_ddtr_normalization_u16(uint16_t *d_data_tc, long nsamp, int nchans, int nbits, void *stream);

/* Output DDTR normalisation, in place on one range's dense [ndms][nsamp]
 * float plane.  Replaces post_DDTR_normalization(d_output, t_processed,
 * ndms) (include/aa_zero_dm.hpp:25, src/aa_zero_dm.cpp:62-74,
 * include/aa_permitted_pipelines_generic.hpp:823-831): per time sample,
 * s = (float sum over the DMs in the reference's order) / ndms - 127.5;
 * x' = x - s. */
int aa_hip_output_ddtr_normalization(float *d_plane, int ndms, long nsamp, void *stream);

/* The set_bandpass_average option's bandpass, on the host
 * (src/aa_host_cpu_msd.cpp:7-39, include/aa_permitted_pipelines_generic.hpp:
 * 533-551): h_bandpass[i] = mean of h_in_tc[i*nsamp .. i*nsamp + nsamp - 1]
 * for i < nchans, over the [t][c] input of nsamp x nchans samples of
 * bytes_per_sample 1 (u8) or 2 (u16), indexed as the reference indexes it. */
int aa_hip_bandpass_average(const void *h_in_tc, int bytes_per_sample, size_t nchans, size_t nsamp,
                            float *h_bandpass);

/* SIGPROC filterbank input.  Replaces aa_sigproc_input::read_metadata /
 * read_signal (include/aa_sigproc_input.hpp:31-47, src/aa_sigproc_input.cpp:52-80,
 * 85-541) for callers over the C ABI; the C++ class of the same name is in
 * include/aa_sigproc_input.hpp.  The header fields are those the reference
 * reads; nsamples comes from the data size.  aa_hip_sigproc_read_u8 keeps
 * 1/2/4/8-bit data as one byte per sample ([t][c], the dedispersion path's
 * input) instead of widening to u16, in pinned host memory when `pinned`
 * (needs a GPU; host-to-device copies from it run asynchronously);
 * aa_hip_sigproc_read_u16 widens as the reference does (any nbits; 32-bit
 * floats scaled to 0..65535).  A file is read once, through one of the two;
 * the buffer lives until aa_hip_sigproc_close. */
typedef struct aa_hip_sigproc aa_hip_sigproc;
typedef struct aa_hip_filterbank_header {
    int telescope_id, machine_id, data_type, barycentric, pulsarcentric;
    int nbits, nsamples, nchans, nifs;
    double az_start, za_start, src_raj, src_dej, tstart, tsamp;
    double fch1, foff, fchannel, refdm, period;
} aa_hip_filterbank_header;
int aa_hip_sigproc_open(const char *path, aa_hip_sigproc **out, aa_hip_filterbank_header *hdr);
int aa_hip_sigproc_read_u8(aa_hip_sigproc *h, int pinned, const uint8_t **data, size_t *bytes);
int aa_hip_sigproc_read_u16(aa_hip_sigproc *h, const uint16_t **data, size_t *count);
void aa_hip_sigproc_close(aa_hip_sigproc *h);

/* The reference's Python FFI of the SIGPROC reader: an aa_sigproc_input
 * object (include/aa_sigproc_input.hpp) behind a void pointer.
 * Replaces src/aa_py_sigproc_input.cpp:6-8 (NULL on a NULL path). */
void *aa_py_sigproc_input(char const *path);
/* Replaces src/aa_py_sigproc_input.cpp:10-12. */
void aa_py_sigproc_input_del(void const *obj);
/* Replaces src/aa_py_sigproc_input.cpp:14-25 (also fills m_src_raj, which
 * the reference leaves unset); all zero when the header cannot be read. */
aa_py_filterbank_metadata_struct aa_py_sigproc_input_read_metadata(void *obj);
/* Replaces src/aa_py_sigproc_input.cpp:27-29: reads the data widened to u16
 * ([t][c], any nbits); false if called before read_metadata or twice. */
bool aa_py_sigproc_input_read_signal(void *obj);
/* Replaces src/aa_py_sigproc_input.cpp:31-33: nsamples x nchans u16 samples
 * (NULL before read_signal); valid until aa_py_sigproc_input_del. */
unsigned short const *aa_py_sigproc_input_input_buffer(void const *obj);

#ifdef __cplusplus
}
#endif
#endif /* AA_DDTR_HIP_H */
