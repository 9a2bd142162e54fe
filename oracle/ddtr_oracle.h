/*
 * CPU oracle for AstroAccelerate's dedispersion transform (DDTR).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product library
 * (astro-accelerate_amd/) never links, loads or calls it.
 *
 * It restates, in plain C with OpenMP, the reference's
 *   - dmshift table           src/aa_ddtr_strategy.cpp:81-92
 *   - strategy (DM grid, maxshift, t_processed, chunking)
 *                             src/aa_ddtr_strategy.cpp:94-415
 *   - corner turn             src/aa_corner_turn.cpp:118-121 (a permutation)
 *   - x2 binning cascade      src/aa_device_binning_kernel.cu:33-45 +
 *                             src/aa_device_corner_turn_kernel.cu:91-98
 *   - failsafe direct sum     src/aa_device_dedispersion_kernel.cu:422-450
 *     with the argument derivation of src/aa_dedisperse.cpp:133-155
 *   - runner sequencing       include/aa_permitted_pipelines_generic.hpp:687-903
 *   - AA fake-signal injector src/aa_host_fake_signal_generator.cpp:13-99
 *
 * All float arithmetic is binary32, evaluated in the order the reference
 * evaluates it, with contraction disabled (-ffp-contract=off) except for the
 * explicit fmaf() that the reference's compiled kernel performs.
 */
#ifndef DDTR_ORACLE_H
#define DDTR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float low, high, step;
    int inBin, outBin;
} oracle_dm_range;

#define ORACLE_MAX_RANGES 64

typedef struct {
    int ok;                 /* 1 when the reference strategy would be ready() */
    int nranges;
    int nchans;
    int nsamp;
    int maxshift;
    int max_ndms;
    int total_ndms;
    int num_tchunks;
    int strategy_case;      /* 1..4 as in src/aa_ddtr_strategy.cpp:155-160 */
    int ndms[ORACLE_MAX_RANGES];
    oracle_dm_range dm[ORACLE_MAX_RANGES];   /* strategy-redefined ranges */
    /* t_processed[range][chunk], row-major with stride ORACLE_MAX_CHUNKS */
    int *t_processed;
    int max_chunks;
    float *dmshifts;        /* nchans entries, malloc'd */
    double tsamp;           /* metadata tsamp (narrowed to float where used) */
} oracle_strategy;

/* dmshift[c] for c < nchans (src/aa_ddtr_strategy.cpp:81-92). */
void oracle_dmshifts(int nchans, double fch1, double foff, float power, float *out);

/* Full strategy computation.  t_processed is malloc'd (free with
 * oracle_strategy_free).  free_memory is the byte count the reference would
 * have read from aa_device_info::free_memory(). */
int oracle_strategy_compute(const oracle_dm_range *user, int nranges,
                            int nchans, int nsamp, double fch1, double foff,
                            double tsamp, int nbits, float power,
                            size_t free_memory, int enable_analysis,
                            oracle_strategy *out);
void oracle_strategy_free(oracle_strategy *s);
int oracle_strategy_t_processed(const oracle_strategy *s, int range, int chunk);

/* delay(c,d) = trunc(dmshift[c] * fmaf((float)d, mdmstep, mstartdm)). */
int oracle_delay(float dmshift, int d, float mstartdm, float mdmstep);

/* Fill delays[d*nchans + c] for one range. */
void oracle_delay_table(const float *dmshifts, int nchans, int ndms,
                        float mstartdm, float mdmstep, int *delays);

/* mstartdm and mdmstep as the runner derives them
 * (include/aa_permitted_pipelines_generic.hpp:778, src/aa_dedisperse.cpp:137,150-153). */
void oracle_range_scalars(float tsamp_original, int inBin, float dm_low,
                          float dm_step, float *mstartdm, float *mdmstep);

/* [t][c] -> [c][t] for nsamp rows of the input.  out_pitch >= nsamp. */
void oracle_corner_turn_u8(const uint8_t *in_tc, int nsamp, int nchans,
                           uint8_t *out_ct, size_t out_pitch);
void oracle_corner_turn_u16(const uint16_t *in_tc, int nsamp, int nchans,
                            uint8_t *out_ct, size_t out_pitch);

/* Zero-DM filter on [t][c] samples in place (src/aa_device_zero_dm_kernel.cu:
 * 37-88, src/aa_zero_dm.cpp:28-37): per time sample, mean = (channel sum) /
 * nchans in float; x' = (x - mean) + bandpass[c] in float, clamped to
 * [0, 2^nbits - 1] and truncated. */
void oracle_zero_dm_u8(uint8_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits);
void oracle_zero_dm_u16(uint16_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits);
/* Zero-DM with outlier rejection on [t][c] samples in place
 * (src/aa_device_zero_dm_outliers_kernel.cu:190-233): per time sample the
 * mean over the channels inside mean +- 3 sigma (the reference's streaming
 * statistics, up to 5 rejection passes); x' = (x - mean) + bandpass[c],
 * clamped, truncated, NaN -> 0. */
void oracle_zero_dm_outliers_u8(uint8_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits);
void oracle_zero_dm_outliers_u16(uint16_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits);
/* Column normalisation of a [rows][cols] array in place
 * (src/aa_device_zero_dm_kernel.cu:100-164): per column, s = (float sum in
 * the reference's order) / rows - factor; x' = x - s, clamped to
 * [0, 2^nbits - 1] and truncated for integer data. */
void oracle_normalize_columns_u8(uint8_t *data, size_t rows, size_t cols, float factor, int nbits);
void oracle_normalize_columns_u16(uint16_t *data, size_t rows, size_t cols, float factor, int nbits);
void oracle_normalize_columns_f32(float *data, size_t rows, size_t cols, float factor);
/* set_bandpass_average (src/aa_host_cpu_msd.cpp:7-39): means of the first
 * nchans runs of nsamp consecutive samples of the [t][c] input. */
void oracle_bandpass_average_u16(const uint16_t *in_tc, size_t nchans, size_t nsamp, float *bandpass);
/* One x2 level: out[c][t] = floor((in[c][2t] + in[c][2t+1]) / 2), t < n_out. */
void oracle_bin_u8(const uint8_t *in, size_t in_pitch, int nchans, int n_out,
                   uint8_t *out, size_t out_pitch);

/* Direct sum over one range:
 * out[d*out_pitch + t] = ((float)sum_c L[c][t + delay(c,d)] / nchans) / inBin. */
void oracle_dedisperse_range(const uint8_t *L, size_t pitch, int row_len,
                             int nchans, const float *dmshifts, int ndms,
                             int t_processed, float mstartdm, float mdmstep,
                             int inBin, float *out, size_t out_pitch);

/* Only the DM rows listed in dm_idx (out row k = DM index dm_idx[k]). */
void oracle_dedisperse_rows(const uint8_t *L, size_t pitch, int nchans, const float *dmshifts,
                            const int *dm_idx, int nrows, int t_processed, float mstartdm,
                            float mdmstep, int inBin, float *out, size_t out_pitch);

/* Whole DDTR over all chunks and ranges of a strategy.
 * input_tc: nsamp x nchans samples in [t][c] order (u8).
 * out[r] points at an [ndms_r][total_t_r] float plane, total_t_r = sum over
 * chunks of t_processed[r][chunk]; chunk k of range r lands at column
 * inc_k / inBin_r as in src/aa_device_save_data.cu:44-49. */
int oracle_ddtr_run_u8(const oracle_strategy *s, const uint8_t *input_tc,
                       float **out);

/* Samples wider than 8 bits, kept as u16 rows: the transpose, one x2 level
 * (truncated float mean, exact for u16), and the float channel sum in
 * channel order, (acc / nchans) / inBin, as the reference's cache kernel
 * forms it; oracle_ddtr_run_u16 runs oracle_ddtr_run_u8's sequence on them. */
void oracle_transpose_u16(const uint16_t *in_tc, int nsamp, int nchans, uint16_t *out_ct, size_t out_pitch);
void oracle_bin_u16(const uint16_t *in, size_t in_pitch, int nchans, int n_out, uint16_t *out, size_t out_pitch);
void oracle_dedisperse_rows_u16(const uint16_t *L, size_t pitch, int nchans, const float *dmshifts,
                                const int *dm_idx, int nrows, int t_processed, float mstartdm, float mdmstep,
                                int inBin, float *out, size_t out_pitch);
int oracle_ddtr_run_u16(const oracle_strategy *s, const uint16_t *input_tc, float **out);

/* Synthetic data: SplitMix64 counter PRNG + Box-Muller (double), rounded to
 * nearest and clipped to [0,255].  value(t,c) depends only on (seed, t, c). */
void oracle_gen_noise_u8(uint64_t seed, int nsamp, int nchans, double mean,
                         double sigma, uint8_t *out_tc);

/* Add `amplitude` to `width` samples starting at t0 + shift[c] in every
 * channel, shift[c] = delay of DM `dm` at tsamp (AA injector convention,
 * src/aa_host_fake_signal_generator.cpp:75-78), saturating at 255. */
void oracle_inject_pulse_u8(uint8_t *data_tc, int nsamp, int nchans,
                            const float *dmshifts, double dm, double tsamp,
                            int t0, int width, int amplitude);

/* AA inverse-Gaussian mask (src/aa_host_fake_signal_generator.cpp:13-34).
 * Returns maximum_pos; mask has `length` entries. */
int oracle_aa_mask(float sigma, int length, float *mask);

/* AA fake signal (src/aa_host_fake_signal_generator.cpp:36-99), u16 [t][c],
 * zero background; out must hold nsamples*nchans entries. */
int oracle_aa_fake_signal_u16(const float *dmshifts, int nchans, int nsamples,
                              double tsamp, int maxshift, double dm_pos,
                              int signal_start, int width, float sigma,
                              int period_ms, uint16_t *out);

#ifdef __cplusplus
}
#endif
#endif
