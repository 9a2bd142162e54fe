/*
 * CPU oracle for AstroAccelerate DDTR -- TEST INFRASTRUCTURE ONLY.
 * See ddtr_oracle.h for scope and for who may use this file.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 * Every function cites the reference file:line it restates.
 */
#include "ddtr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Reference tile constants that are part of the *plan* semantics
 * (include/aa_params.hpp:95-99): SDIVINT 8, SNUMREG 14, SDIVINDM 64. */
#define SDIVINT 8
#define SNUMREG 14
#define SDIVINDM 64
#define MAX_OUTPUT_SIZE 17179869184ULL /* src/aa_ddtr_strategy.cpp:7 */

/* src/aa_ddtr_strategy.cpp:81-92.  fch1/foff are narrowed to float first
 * (src/aa_ddtr_strategy.cpp:42-43); fch1 + foff*c is a float multiply then a
 * float add; the power-2 branch evaluates the reciprocal powers in double. */
void oracle_dmshifts(int nchans, double fch1_d, double foff_d, float power, float *out)
{
    const float fch1 = (float)fch1_d;
    const float foff = (float)foff_d;
    for (int c = 0; c < nchans; c++) {
        float f = fch1 + foff * (float)c;
        if (power != 2.0f) {
            /* pow(float, float) resolves to the float overload in the reference */
            out[c] = (float)(4148.741601f * ((1.0 / (double)powf(f, power)) -
                                             (1.0 / (double)powf(fch1, power))));
        } else {
            out[c] = (float)(4148.741601f * ((1.0 / pow((double)f, (double)power)) -
                                             (1.0 / pow((double)fch1, (double)power))));
        }
    }
}

int oracle_strategy_t_processed(const oracle_strategy *s, int range, int chunk)
{
    return s->t_processed[range * s->max_chunks + chunk];
}

void oracle_strategy_free(oracle_strategy *s)
{
    free(s->t_processed);
    free(s->dmshifts);
    s->t_processed = NULL;
    s->dmshifts = NULL;
}

/* src/aa_ddtr_strategy.cpp:27-428, argument types kept as in the reference so
 * that every integer/float conversion happens at the same place. */
int oracle_strategy_compute(const oracle_dm_range *user, int nranges,
                            int nchans, int nsamp, double fch1, double foff,
                            double tsamp_d, int nbits, float power,
                            size_t free_memory, int enable_analysis,
                            oracle_strategy *s)
{
    memset(s, 0, sizeof(*s));
    s->tsamp = tsamp_d;
    s->nchans = nchans;
    s->nsamp = nsamp;
    s->nranges = nranges;
    if (nranges <= 0 || nranges > ORACLE_MAX_RANGES) return 0; /* :63-66 */

    const float tsamp = (float)tsamp_d;
    float *dmshifts = (float *)malloc(sizeof(float) * (size_t)nchans);
    oracle_dmshifts(nchans, fch1, foff, power, dmshifts);
    s->dmshifts = dmshifts;
    const size_t gpu_memory = free_memory - 104857600; /* :77 */
    const double SPDT_fraction = 3.0 / 4.0;           /* :79 */

    int max_ndms = 0, total_ndms = 0;
    for (int i = 0; i < nranges; i++) { /* :94-101 */
        int nblk = (((int)((user[i].high - user[i].low) / user[i].step)) + SDIVINDM) / SDIVINDM;
        float n;
        modff((float)nblk, &n);
        s->ndms[i] = (int)((int)n * SDIVINDM);
        if (max_ndms < s->ndms[i]) max_ndms = s->ndms[i];
        total_ndms += s->ndms[i];
    }
    s->max_ndms = max_ndms;
    s->total_ndms = total_ndms;

    /* :104-110 strategy-redefined ranges */
    s->dm[0].low = user[0].low;
    s->dm[0].high = s->dm[0].low + ((float)s->ndms[0] * user[0].step);
    s->dm[0].step = user[0].step;
    int maxshift = 0, maxshift_high = 0;
    const float modulo = (float)(SDIVINT * 2 * SNUMREG);
    for (int i = 1; i < nranges; i++) {
        s->dm[i].low = s->dm[i - 1].high;
        s->dm[i].high = s->dm[i].low + (float)s->ndms[i] * user[i].step;
        s->dm[i].step = user[i].step;
        if (user[i - 1].inBin > 1) { /* :112-121 */
            float top = s->dm[i - 1].low + s->dm[i - 1].step * (float)s->ndms[i - 1];
            maxshift = (int)ceilf((top * dmshifts[nchans - 1]) / tsamp);
            maxshift = (int)(((float)maxshift / (float)user[i - 1].inBin + modulo) / modulo);
            maxshift = maxshift * (SDIVINT * 2 * SNUMREG) * user[i - 1].inBin;
            if (maxshift > maxshift_high) maxshift_high = maxshift;
        }
    }
    if (user[nranges - 1].inBin > 1) { /* :124-133 */
        const int r = nranges - 1;
        float top = s->dm[r].low + s->dm[r].step * (float)s->ndms[r];
        maxshift = (int)ceilf((top * dmshifts[nchans - 1]) / tsamp);
        maxshift = (int)(((float)maxshift / (float)user[r].inBin + modulo) / modulo);
        maxshift = maxshift * (SDIVINT * 2 * SNUMREG) * user[r].inBin;
        if (maxshift > maxshift_high) maxshift_high = maxshift;
    }
    if (maxshift_high == 0) { /* :135-137 */
        const int r = nranges - 1;
        float top = s->dm[r].low + s->dm[r].step * (float)s->ndms[r];
        maxshift_high = (int)ceilf((top * dmshifts[nchans - 1]) / tsamp);
    }
    maxshift = maxshift_high + (SNUMREG * 2 * SDIVINT); /* :140 */
    s->maxshift = maxshift;
    for (int i = 0; i < nranges; i++) { /* :420-423 */
        s->dm[i].inBin = user[i].inBin;
        s->dm[i].outBin = user[i].outBin;
    }
    if (maxshift >= nsamp) return 0; /* :144-153 */

    const int gran = (nbits == 4) ? (SDIVINT * 4 * SNUMREG) : (SDIVINT * 2 * SNUMREG);
    const int inBin_last = user[nranges - 1].inBin;
    unsigned int max_tsamps;
    int num_blocks = 1, rem_block = 0;
    int local_int = 0, remainder = 0;
    unsigned long local_ul = 0;
    int kase;

    if (nchans < max_ndms) { /* :166-275 */
        size_t SPDT = enable_analysis ? (size_t)(sizeof(float) * (size_t)max_ndms * SPDT_fraction) : 0;
        max_tsamps = (unsigned int)(gpu_memory / (sizeof(unsigned short) * (size_t)nchans +
                                                  sizeof(float) * (size_t)max_ndms + SPDT));
        size_t out_per_chunk = ((size_t)max_tsamps) * sizeof(float) * ((size_t)max_ndms);
        if (out_per_chunk > MAX_OUTPUT_SIZE)
            max_tsamps = (unsigned int)(MAX_OUTPUT_SIZE / (sizeof(float) * (size_t)max_ndms));
        if ((unsigned int)maxshift > max_tsamps) return 0;
        if ((unsigned int)nsamp < max_tsamps) {
            kase = 1;
            local_ul = (unsigned long)floor(((double)(nsamp - maxshift) / (double)inBin_last) / (double)gran);
            local_ul = local_ul * (unsigned long)gran * (unsigned long)inBin_last;
        } else {
            kase = 3;
        }
    } else { /* :276-389 */
        size_t SPDT = enable_analysis ? (size_t)(sizeof(float) * (size_t)max_ndms * SPDT_fraction) : 0;
        max_tsamps = (unsigned int)(gpu_memory / ((size_t)nchans * (sizeof(float) + sizeof(unsigned short)) + SPDT));
        size_t out_per_chunk = ((size_t)max_tsamps) * sizeof(float) * ((size_t)nchans);
        if (out_per_chunk > MAX_OUTPUT_SIZE)
            max_tsamps = (unsigned int)(MAX_OUTPUT_SIZE / (sizeof(float) * (size_t)nchans));
        if (maxshift > (int)max_tsamps) return 0;
        if ((unsigned int)nsamp < max_tsamps) {
            kase = 2;
            local_int = (int)floorf(((float)(nsamp - maxshift) / (float)inBin_last) / (float)gran);
            local_int = local_int * gran * inBin_last;
        } else {
            kase = 4;
        }
    }
    s->strategy_case = kase;

    if (kase == 3 || kase == 4) { /* :209-274, :324-388 */
        int samp_block_size = (int)(max_tsamps - (unsigned int)maxshift);
        local_int = (int)floorf(((float)samp_block_size / (float)inBin_last) / (float)gran);
        local_int = local_int * gran * inBin_last;
        /* No whole granule fits: the reference divides by zero here (UB);
         * both this oracle and the library report "not ready". */
        if (local_int <= 0) return 0;
        num_blocks = (int)(((float)(nsamp - maxshift)) / ((float)local_int));
        remainder = nsamp - (num_blocks * local_int) - maxshift;
        /* (int)floor(x) / (float)gran is a float quotient truncated on assignment */
        remainder = (int)((float)((int)floorf((float)remainder / (float)inBin_last)) / (float)gran);
        remainder = remainder * gran * inBin_last;
        rem_block = remainder > 0 ? 1 : 0;
    }

    const int nchunks = (kase == 1 || kase == 2) ? 1 : (num_blocks + rem_block);
    if (nchunks <= 0) return 0;
    s->max_chunks = nchunks;
    s->num_tchunks = nchunks;
    s->t_processed = (int *)calloc((size_t)nranges * (size_t)nchunks, sizeof(int));
    for (int i = 0; i < nranges; i++) {
        const float ib = (float)user[i].inBin;
        if (kase == 1) {
            s->t_processed[i * nchunks] =
                (int)floorf(((float)local_ul / ib) / (float)gran) * gran;
        } else if (kase == 2) {
            s->t_processed[i * nchunks] =
                (int)floorf(((float)local_int / ib) / (float)gran) * gran;
        } else {
            for (int j = 0; j < num_blocks; j++)
                s->t_processed[i * nchunks + j] =
                    (int)floorf(((float)local_int / ib) / (float)gran) * gran;
            if (rem_block)
                s->t_processed[i * nchunks + num_blocks] =
                    (int)floorf(((float)remainder / ib) / (float)gran) * gran;
        }
    }

    /* :401-415 memory requests against the device's available memory; only
     * the output request decides readiness. */
    {
        size_t available = free_memory;
        long time_samps = (long)s->t_processed[0] + maxshift;
        size_t in_req = (size_t)time_samps * (size_t)nchans * sizeof(unsigned short);
        if (available > in_req) available -= in_req;
        size_t out_req = (size_t)time_samps * (size_t)(nchans < max_ndms ? max_ndms : nchans) * sizeof(float);
        if (!(available > out_req)) return 0;
    }
    s->ok = 1;
    return 1;
}

/* src/aa_device_dedispersion_kernel.cu:430,440: shift_temp = mstartdm + d*mdmstep
 * (contracted to one FMA by the CUDA compiler), then
 * __float2int_rz(dm_shifts[c] * shift_temp). */
int oracle_delay(float dmshift, int d, float mstartdm, float mdmstep)
{
    float shift_temp = fmaf((float)d, mdmstep, mstartdm);
    float prod = dmshift * shift_temp;
    return (int)prod; /* truncation toward zero */
}

void oracle_delay_table(const float *dmshifts, int nchans, int ndms,
                        float mstartdm, float mdmstep, int *delays)
{
    for (int d = 0; d < ndms; d++)
        for (int c = 0; c < nchans; c++)
            delays[(size_t)d * nchans + c] = oracle_delay(dmshifts[c], d, mstartdm, mdmstep);
}

/* include/aa_permitted_pipelines_generic.hpp:778 (tsamp = tsamp_original*inBin)
 * and src/aa_dedisperse.cpp:137,150-153 ((float)(startdm/(*tsamp))). */
void oracle_range_scalars(float tsamp_original, int inBin, float dm_low,
                          float dm_step, float *mstartdm, float *mdmstep)
{
    float tsamp = tsamp_original * (float)inBin;
    *mstartdm = dm_low / tsamp;
    *mdmstep = dm_step / tsamp;
}

/* src/aa_corner_turn.cpp:118-121: [t][c] -> [c][t]; u16 holds 8-bit values
 * (src/aa_sigproc_input.cpp:438-457). */
void oracle_corner_turn_u8(const uint8_t *in_tc, int nsamp, int nchans,
                           uint8_t *out_ct, size_t out_pitch)
{
    /* 64x64 blocks keep both the strided reads and the writes in cache. */
#pragma omp parallel for schedule(static) collapse(2)
    for (int cb = 0; cb < nchans; cb += 64)
        for (int tb = 0; tb < nsamp; tb += 64) {
            const int ce = cb + 64 < nchans ? cb + 64 : nchans;
            const int te = tb + 64 < nsamp ? tb + 64 : nsamp;
            for (int c = cb; c < ce; c++)
                for (int t = tb; t < te; t++)
                    out_ct[(size_t)c * out_pitch + t] = in_tc[(size_t)t * nchans + c];
        }
}

void oracle_corner_turn_u16(const uint16_t *in_tc, int nsamp, int nchans,
                            uint8_t *out_ct, size_t out_pitch)
{
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nchans; c++)
        for (int t = 0; t < nsamp; t++)
            out_ct[(size_t)c * out_pitch + t] = (uint8_t)in_tc[(size_t)t * nchans + c];
}

/* src/aa_device_binning_kernel.cu:43 computes (a+b)/2.0f in float and
 * src/aa_device_corner_turn_kernel.cu:96 truncates it back to an integer, so
 * one level is floor((a+b)/2). */
/* src/aa_device_zero_dm_kernel.cu:37-88.  The reference sums a time
 * sample's channels in a float accumulator; the values are integers and the
 * sum stays below 2^24 (nchans * 65535 < 2^24 for nchans <= 256 at 16 bits,
 * any nchans up to 65793 at 8 bits), so an int64 sum converted once is the
 * same float.  Then sum / nchans, then (x - mean) + bandpass[c], each in
 * float as written; results below 0 give 0, above the nbits maximum give
 * the maximum, else truncation (the kernel's conversion to unsigned short).
 * The reference clamps for nbits 4, 8 and 16 and leaves other widths
 * undefined; here every width clamps to 2^nbits - 1. */
static float zero_dm_value(float x, float mean, float bp, int nbits)
{
    const float r = (x - mean) + bp;
    const float mx = (float)((1u << (nbits > 16 ? 16 : nbits)) - 1u);
    if (r < 0.0f) return 0.0f;
    if (r > mx) return mx;
    return (float)(unsigned)r;
}

void oracle_zero_dm_u8(uint8_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nsamp; t++) {
        uint8_t *row = data_tc + (size_t)t * (size_t)nchans;
        int64_t sum = 0;
        for (int c = 0; c < nchans; c++) sum += row[c];
        const float mean = (float)sum / (float)nchans;
        for (int c = 0; c < nchans; c++)
            row[c] = (uint8_t)(unsigned)zero_dm_value((float)row[c], mean, bandpass[c], nbits);
    }
}

void oracle_zero_dm_u16(uint16_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nsamp; t++) {
        uint16_t *row = data_tc + (size_t)t * (size_t)nchans;
        int64_t sum = 0;
        for (int c = 0; c < nchans; c++) sum += row[c];
        const float mean = (float)sum / (float)nchans;
        for (int c = 0; c < nchans; c++)
            row[c] = (uint16_t)(unsigned)zero_dm_value((float)row[c], mean, bandpass[c], nbits);
    }
}

/* Zero-DM with outlier rejection (src/aa_device_zero_dm_outliers_kernel.cu:
 * 15-233, launched by src/aa_zero_dm_outliers.cpp:47-67 with 128 threads per
 * time sample, outlier_sigma 3).  Accumulator v (0..127) takes channels v,
 * v+128, ... in order: the first with Initiate (:15-19), the rest with
 * Add_one (:21-29).  The block reduction (Reduce_SM :42-70 for i = 64, 32,
 * then Reduce_WARP :72-91 for q = 16..1) merges accumulator v with v+i for
 * v < i, level by level, with Merge (:31-40) and the empty-accumulator
 * rules written there.  Types follow the declarations (float M, S; double
 * counts and factors); a product feeding an add is fused, as CUDA's default
 * contraction compiles those lines.  mean = M/j, stdev = sqrt(S/j) in double
 * (:118-121).  Passes: all channels (MSD_block), then up to 5 rejection
 * passes over channels with lo < x < hi, lo/hi = mean -+ 3 stdev fused
 * (:136-137), stopping when old_mean - mean < 0.001f (:203-207).  Output
 * (:209-231): (x - mean) + bandpass[c] in float, clamped, truncated; a NaN
 * (0/0 when no channel is inside the limits) converts to 0.  The reference
 * clamps for nbits 4, 8 and 16; here every width clamps to 2^nbits - 1. */
typedef struct { float M, S; int j; } zdo_acc;

static void zdo_push(zdo_acc *a, float x)
{
    if (a->j == 0) {
        a->M = x;
        a->S = 0.0f;
        a->j = 1;
        return;
    }
    a->j += 1;
    const double rj = (double)a->j;
    a->M = a->M + x;
    const float ft = (float)fma(rj, (double)x, -(double)a->M);
    const double c = 1.0 / (rj * (rj - 1.0));
    a->S = (float)fma(c * (double)ft, (double)ft, (double)a->S);
}

static void zdo_merge(zdo_acc *a, const zdo_acc *b)
{
    if (b->j <= 0) return;
    if (a->j == 0) {
        *a = *b;
        return;
    }
    const double rb = (double)b->j, ra = (double)a->j;
    const float ft = (float)fma(rb / ra, (double)a->M, -(double)b->M);
    const double coef = ra / (rb * (ra + rb));
    const float ss = a->S + b->S;
    a->S = (float)fma(coef * (double)ft, (double)ft, (double)ss);
    a->M = a->M + b->M;
    a->j += b->j;
}

/* One statistics pass; reject = 0: every channel. */
static void zdo_pass(const float *x, int nchans, int reject, float lo, float hi, float *mean, float *stdev)
{
    zdo_acc acc[128];
    for (int v = 0; v < 128; v++) acc[v].M = acc[v].S = 0.0f, acc[v].j = 0;
    for (int c = 0; c < nchans; c++)
        if (!reject || (x[c] > lo && x[c] < hi)) zdo_push(&acc[c % 128], x[c]);
    for (int i = 64; i > 0; i >>= 1)
        for (int v = 0; v < i; v++) zdo_merge(&acc[v], &acc[v + i]);
    *mean = (float)((double)acc[0].M / (double)acc[0].j);
    *stdev = (float)sqrt((double)acc[0].S / (double)acc[0].j);
}

static void zdo_row(float *x, int nchans, float *mean_out)
{
    float mean, stdev;
    zdo_pass(x, nchans, 0, 0.0f, 0.0f, &mean, &stdev);
    for (int f = 0; f < 5; f++) {
        const float old_mean = mean;
        const float lo = fmaf(-3.0f, stdev, mean), hi = fmaf(3.0f, stdev, mean);
        zdo_pass(x, nchans, 1, lo, hi, &mean, &stdev);
        if (old_mean - mean < 0.001f) break;
    }
    *mean_out = mean;
}

static float zdo_value(float x, float mean, float bp, int nbits)
{
    const float r = (x - mean) + bp;
    const float mx = (float)((1u << (nbits > 16 ? 16 : nbits)) - 1u);
    if (r < 0.0f) return 0.0f;
    if (r > mx) return mx;
    if (r != r) return 0.0f;
    return (float)(unsigned)r;
}

void oracle_zero_dm_outliers_u8(uint8_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits)
{
#pragma omp parallel
    {
        float *x = (float *)malloc((size_t)nchans * sizeof(float));
#pragma omp for schedule(static)
        for (int t = 0; t < nsamp; t++) {
            uint8_t *row = data_tc + (size_t)t * (size_t)nchans;
            for (int c = 0; c < nchans; c++) x[c] = (float)row[c];
            float mean;
            zdo_row(x, nchans, &mean);
            for (int c = 0; c < nchans; c++) row[c] = (uint8_t)(unsigned)zdo_value(x[c], mean, bandpass[c], nbits);
        }
        free(x);
    }
}

void oracle_zero_dm_outliers_u16(uint16_t *data_tc, int nsamp, int nchans, const float *bandpass, int nbits)
{
#pragma omp parallel
    {
        float *x = (float *)malloc((size_t)nchans * sizeof(float));
#pragma omp for schedule(static)
        for (int t = 0; t < nsamp; t++) {
            uint16_t *row = data_tc + (size_t)t * (size_t)nchans;
            for (int c = 0; c < nchans; c++) x[c] = (float)row[c];
            float mean;
            zdo_row(x, nchans, &mean);
            for (int c = 0; c < nchans; c++)
                row[c] = (uint16_t)(unsigned)zdo_value(x[c], mean, bandpass[c], nbits);
        }
        free(x);
    }
}

/* Column normalisation of a [rows][cols] array in place, the reference's
 * normalization_transposed_kernel (src/aa_device_zero_dm_kernel.cu:100-164):
 * for each column, 32 float partial sums, partial y adding rows y, y+32,
 * y+64, ... in row order from 0.0f; then p[y] += p[y+i] for i = 16, 8, 4,
 * 2, 1 (y < i); s = p[0] / (float)rows - factor; every value becomes
 * x - s in float.  Integer data (nbits 1..16) is clamped to [0, 2^nbits - 1]
 * and truncated; float data (nbits 32) keeps x - s.  Input DDTR
 * normalisation runs it on the [t][c] chunk with cols = nchans and factor
 * (2^nbits - 1)/2 (src/aa_zero_dm.cpp:49-60, called at
 * include/aa_permitted_pipelines_generic.hpp:731-735 with the size
 * arguments in that order); output DDTR normalisation on each [dm][t]
 * plane with cols = t_processed, factor 127.5 (src/aa_zero_dm.cpp:62-74,
 * include/aa_permitted_pipelines_generic.hpp:823-831). */
#define NORM_CB 64   /* columns per block: rows are walked in order, 64 columns at a time */

static float load_kind(const void *data, int kind, size_t i)
{
    return kind == 0 ? (float)((const uint8_t *)data)[i]
         : kind == 1 ? (float)((const uint16_t *)data)[i] : ((const float *)data)[i];
}

/* s[k] for columns c0 .. c0 + nc - 1.  Partial y of a column adds the rows
 * r = y (mod 32) in increasing r, starting from 0.0f. */
static void column_offsets(const void *data, int kind, size_t rows, size_t cols, size_t c0, size_t nc,
                           float factor, float *s)
{
    float p[32][NORM_CB];
    for (int y = 0; y < 32; y++)
        for (size_t k = 0; k < nc; k++) p[y][k] = 0.0f;
    for (size_t r = 0; r < rows; r++)
        for (size_t k = 0; k < nc; k++) p[r % 32][k] = p[r % 32][k] + load_kind(data, kind, r * cols + c0 + k);
    for (size_t k = 0; k < nc; k++) {
        for (int i = 16; i > 0; i >>= 1)
            for (int y = 0; y < i; y++) p[y][k] = p[y][k] + p[y + i][k];
        s[k] = p[0][k] / (float)rows - factor;
    }
}

static float norm_clamp(float r, int nbits)
{
    const float mx = (float)((1u << (nbits > 16 ? 16 : nbits)) - 1u);
    if (r < 0.0f) return 0.0f;
    if (r > mx) return mx;
    return (float)(unsigned)r;
}

static void normalize_columns(void *data, int kind, size_t rows, size_t cols, float factor, int nbits)
{
    if (rows == 0) return;
    const long nblk = (long)((cols + NORM_CB - 1) / NORM_CB);
#pragma omp parallel for schedule(dynamic)
    for (long b = 0; b < nblk; b++) {
        const size_t c0 = (size_t)b * NORM_CB;
        const size_t nc = cols - c0 < NORM_CB ? cols - c0 : NORM_CB;
        float s[NORM_CB];
        column_offsets(data, kind, rows, cols, c0, nc, factor, s);
        for (size_t r = 0; r < rows; r++)
            for (size_t k = 0; k < nc; k++) {
                const size_t i = r * cols + c0 + k;
                if (kind == 0)
                    ((uint8_t *)data)[i] = (uint8_t)(unsigned)norm_clamp((float)((uint8_t *)data)[i] - s[k], nbits);
                else if (kind == 1)
                    ((uint16_t *)data)[i] =
                        (uint16_t)(unsigned)norm_clamp((float)((uint16_t *)data)[i] - s[k], nbits);
                else
                    ((float *)data)[i] = ((float *)data)[i] - s[k];
            }
    }
}

void oracle_normalize_columns_u8(uint8_t *data, size_t rows, size_t cols, float factor, int nbits)
{
    normalize_columns(data, 0, rows, cols, factor, nbits);
}

void oracle_normalize_columns_u16(uint16_t *data, size_t rows, size_t cols, float factor, int nbits)
{
    normalize_columns(data, 1, rows, cols, factor, nbits);
}

void oracle_normalize_columns_f32(float *data, size_t rows, size_t cols, float factor)
{
    normalize_columns(data, 2, rows, cols, factor, 32);
}

/* The set_bandpass_average option (include/aa_permitted_pipelines_generic.hpp:
 * 533-551 -> src/aa_host_cpu_msd.cpp:7-39): bandpass[i] = (float)(sum / n)
 * with a double sum over input[i*nsamp .. i*nsamp + nsamp) for i < nchans,
 * n = nsamp.  The input buffer is [t][c], so "channel" i averages the i-th
 * run of nsamp consecutive samples of that buffer, not channel i; this is
 * the reference's arithmetic as written. */
void oracle_bandpass_average_u16(const uint16_t *in_tc, size_t nchans, size_t nsamp, float *bandpass)
{
    for (size_t i = 0; i < nchans; i++) {
        double sum = 0.0;
        for (size_t j = 0; j < nsamp; j++) sum += in_tc[i * nsamp + j];
        bandpass[i] = (float)(sum / (double)nsamp);
    }
}

void oracle_bin_u8(const uint8_t *in, size_t in_pitch, int nchans, int n_out,
                   uint8_t *out, size_t out_pitch)
{
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nchans; c++) {
        const uint8_t *row = in + (size_t)c * in_pitch;
        uint8_t *orow = out + (size_t)c * out_pitch;
        for (int t = 0; t < n_out; t++)
            orow[t] = (uint8_t)(((unsigned)row[2 * t] + (unsigned)row[2 * t + 1]) >> 1);
    }
}

/* src/aa_device_dedispersion_kernel.cu:422-450.  The channel sum of 8-bit
 * values is an exact integer (< 2^24 for nchans <= 65793), so accumulating
 * in int32 gives the same float as the reference's float accumulator; the
 * final scaling is (sum / nchans) / bin in float, left to right (:448). */
void oracle_dedisperse_range(const uint8_t *L, size_t pitch, int row_len,
                             int nchans, const float *dmshifts, int ndms,
                             int t_processed, float mstartdm, float mdmstep,
                             int inBin, float *out, size_t out_pitch)
{
    (void)row_len;
    int *idx = (int *)malloc(sizeof(int) * (size_t)(ndms > 0 ? ndms : 1));
    for (int d = 0; d < ndms; d++) idx[d] = d;
    oracle_dedisperse_rows(L, pitch, nchans, dmshifts, idx, ndms, t_processed, mstartdm,
                           mdmstep, inBin, out, out_pitch);
    free(idx);
}

/* Runner sequencing: include/aa_permitted_pipelines_generic.hpp:687-903.
 * Chunk k loads t_processed[0][k] + maxshift samples starting at sample inc
 * (:693); range r runs on the x2 cascade level whose bin equals inBin_r,
 * with tsamp_r = tsamp*inBin_r (:778) and the row length
 * t_processed[r][k] + maxshift/inBin_r (:782-786); inc advances by
 * t_processed[0][k] (:896-899).  Output column offset is inc/inBin_r
 * (src/aa_device_save_data.cu:45). */
int oracle_ddtr_run_u8(const oracle_strategy *s, const uint8_t *input_tc, float **out)
{
    if (!s->ok) return -1;
    const int nchans = s->nchans;
    const float tsamp_original = (float)s->tsamp;
    long inc = 0;
    for (int r = 0; r < s->nranges; r++) {
        int ib = s->dm[r].inBin;
        if (ib < 1 || (ib & (ib - 1)) != 0) return -2;            /* power of two only */
        if (r > 0 && ib < s->dm[r - 1].inBin) return -3;           /* non-decreasing */
    }
    for (int k = 0; k < s->num_tchunks; k++) {
        const int tp0 = oracle_strategy_t_processed(s, 0, k);
        const long len0 = (long)tp0 + s->maxshift;
        if (inc + len0 > s->nsamp) return -4;
        size_t pitch = (size_t)len0;
        uint8_t *level = (uint8_t *)malloc((size_t)nchans * pitch);
        oracle_corner_turn_u8(input_tc + (size_t)inc * nchans, (int)len0, nchans, level, pitch);
        long level_len = len0;
        int bin = 1;
        for (int r = 0; r < s->nranges; r++) {
            const int ib = s->dm[r].inBin;
            while (bin < ib) {
                long n_out = level_len / 2;
                uint8_t *next = (uint8_t *)malloc((size_t)nchans * (size_t)(n_out > 0 ? n_out : 1));
                oracle_bin_u8(level, pitch, nchans, (int)n_out, next, (size_t)n_out);
                free(level);
                level = next;
                pitch = (size_t)n_out;
                level_len = n_out;
                bin *= 2;
            }
            const int tp = oracle_strategy_t_processed(s, r, k);
            const int maxshift_r = s->maxshift / ib;
            if ((long)tp + maxshift_r > level_len) { free(level); return -5; }
            float mstartdm, mdmstep;
            oracle_range_scalars(tsamp_original, ib, s->dm[r].low, s->dm[r].step, &mstartdm, &mdmstep);
            long total_r = 0;
            for (int j = 0; j < s->num_tchunks; j++) total_r += oracle_strategy_t_processed(s, r, j);
            const long col = inc / ib;
            if (col + tp > total_r) { free(level); return -6; }
            oracle_dedisperse_range(level, pitch, tp + maxshift_r, nchans, s->dmshifts,
                                    s->ndms[r], tp, mstartdm, mdmstep, ib,
                                    out[r] + col, (size_t)total_r);
        }
        free(level);
        inc += tp0;
    }
    return 0;
}

/* SplitMix64 (Steele, Lea, Flood 2014) used as a counter-based generator. */
static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

/* Box-Muller on two counter draws per sample: sample index i uses counters
 * 2i and 2i+1 of the stream keyed by `seed`. */
void oracle_gen_noise_u8(uint64_t seed, int nsamp, int nchans, double mean,
                         double sigma, uint8_t *out_tc)
{
    const uint64_t key = splitmix64(seed);
    const long n = (long)nsamp * nchans;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double u1 = 1.0 - u01(splitmix64(key + 2 * (uint64_t)i));
        double u2 = u01(splitmix64(key + 2 * (uint64_t)i + 1));
        double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        double v = nearbyint(mean + sigma * z);
        if (v < 0.0) v = 0.0;
        if (v > 255.0) v = 255.0;
        out_tc[i] = (uint8_t)v;
    }
}

/* Shift convention of src/aa_host_fake_signal_generator.cpp:75-78:
 * temp = (float)(dm/tsamp) in double then narrowed, shift = floor(dmshift*temp). */
void oracle_inject_pulse_u8(uint8_t *data_tc, int nsamp, int nchans,
                            const float *dmshifts, double dm, double tsamp,
                            int t0, int width, int amplitude)
{
    const float temp = (float)(dm / tsamp);
    for (int c = 0; c < nchans; c++) {
        int shift = (int)floorf(dmshifts[c] * temp);
        for (int i = 0; i < width; i++) {
            long t = (long)t0 + shift + i;
            if (t < 0 || t >= nsamp) continue;
            uint8_t *p = data_tc + (size_t)t * nchans + c;
            int v = (int)*p + amplitude;
            *p = (uint8_t)(v > 255 ? 255 : v);
        }
    }
}

/* src/aa_host_fake_signal_generator.cpp:13-34 (inverse-Gaussian shape,
 * normalised to its maximum; ties keep the later index). */
int oracle_aa_mask(float sigma, int length, float *mask)
{
    const double PI = 3.14159265359;
    float maximum = 0.0f;
    const float g_mean = 1.0f;
    int maximum_pos = 0;
    float wid = (float)((2.0f * sigma) * pow(3.0 * log(30.0), 0.5));
    float step = wid * sigma / (float)length;
    for (int i = 0; i < length; i++) {
        float x = (float)(i + 1) * step;
        double xd = (double)x;
        mask[i] = (float)(sqrt((double)sigma / (2.0 * PI * pow(xd, 3.0))) *
                          exp((-(double)sigma * pow((double)(x - g_mean), 2.0)) /
                              (2.0 * pow((double)g_mean, 2.0) * xd)));
        if (maximum <= mask[i]) {
            maximum = mask[i];
            maximum_pos = i;
        }
    }
    for (int i = 0; i < length; i++) mask[i] = mask[i] / maximum;
    return maximum_pos;
}

/* src/aa_host_fake_signal_generator.cpp:36-99 (u16 [t][c] with a zero
 * background; the injected peak is exactly 255). */
int oracle_aa_fake_signal_u16(const float *dmshifts, int nchans, int nsamples,
                              double tsamp, int maxshift, double dm_pos,
                              int signal_start, int width, float sigma,
                              int period_ms, uint16_t *out)
{
    float *mask = (float *)malloc(sizeof(float) * (size_t)width);
    int maximum_pos = oracle_aa_mask(sigma, width, mask);
    if (signal_start + maxshift > nsamples) { free(mask); return -1; }
    memset(out, 0, sizeof(uint16_t) * (size_t)nsamples * nchans);
    int *shift_index = (int *)malloc(sizeof(int) * (size_t)nchans);
    if (signal_start == 0) signal_start = -maximum_pos - 2 * maxshift;
    float temp = (float)(dm_pos / tsamp);
    for (int i = 0; i < nchans; i++) shift_index[i] = (int)floorf(dmshifts[i] * temp);
    int repeats = 1, period = 0;
    if (period_ms > 0) {
        period = (int)(period_ms / 1000.0 / tsamp);
        repeats = (nsamples + 2 * maxshift) / period + 1;
    }
    for (int r = 0; r < repeats; r++)
        for (int i = 0; i < width; i++) {
            const int time_pos = (i - maximum_pos + signal_start + period * r) * nchans;
            for (int j = 0; j < nchans; j++) {
                const int g = j + nchans * shift_index[j] + time_pos;
                if (g < nsamples * nchans && g > 0)
                    out[g] = (uint16_t)(255.0f * mask[i]);
            }
        }
    free(shift_index);
    free(mask);
    return maximum_pos;
}

/* Selected DM rows of one range.  Work items are (row, block of 8192
 * samples) so that a handful of rows still spreads over every core and the
 * int32 accumulator of a block stays in L1. */
#define ORACLE_TBLOCK 8192
void oracle_dedisperse_rows(const uint8_t *L, size_t pitch, int nchans, const float *dmshifts,
                            const int *dm_idx, int nrows, int t_processed, float mstartdm,
                            float mdmstep, int inBin, float *out, size_t out_pitch)
{
    const int nblk = (t_processed + ORACLE_TBLOCK - 1) / ORACLE_TBLOCK;
    int *delays = (int *)malloc(sizeof(int) * (size_t)nchans * (size_t)(nrows > 0 ? nrows : 1));
    for (int k = 0; k < nrows; k++)
        for (int c = 0; c < nchans; c++)
            delays[(size_t)k * nchans + c] = oracle_delay(dmshifts[c], dm_idx[k], mstartdm, mdmstep);
    const float fn = (float)nchans, fb = (float)inBin;
#pragma omp parallel
    {
        int32_t acc[ORACLE_TBLOCK];
#pragma omp for schedule(dynamic, 1)
        for (long item = 0; item < (long)nrows * nblk; item++) {
            const int k = (int)(item / nblk);
            const int t0 = (int)(item % nblk) * ORACLE_TBLOCK;
            const int n = t_processed - t0 < ORACLE_TBLOCK ? t_processed - t0 : ORACLE_TBLOCK;
            const int *dk = delays + (size_t)k * nchans;
            memset(acc, 0, sizeof(int32_t) * (size_t)n);
            for (int c = 0; c < nchans; c++) {
                const uint8_t *row = L + (size_t)c * pitch + (size_t)(t0 + dk[c]);
                for (int t = 0; t < n; t++) acc[t] += row[t];
            }
            float *orow = out + (size_t)k * out_pitch + t0;
            for (int t = 0; t < n; t++) orow[t] = ((float)acc[t] / fn) / fb;
        }
    }
    free(delays);
}

/* ------------------------------------------------------------------------
 * Samples wider than 8 bits (16-bit files, and 32-bit files the reader
 * scales to u16; src/aa_sigproc_input.cpp:400-433).  The device rows stay
 * u16 and the cache kernel's float accumulator rounds once channel sums pass
 * 2^24, so the sum is kept exactly as that kernel forms it: one float
 * accumulator per (DM, sample), channels added in order c = 0 .. nchans-1
 * (src/aa_device_dedispersion_kernel.cu:422-450, `local_kernel +=
 * (float)d_input[shift]`), then / nchans / inBin.
 * ------------------------------------------------------------------------ */

/* [t][c] u16 -> [c][t] u16 (src/aa_corner_turn.cpp:118-121). */
void oracle_transpose_u16(const uint16_t *in_tc, int nsamp, int nchans, uint16_t *out_ct, size_t out_pitch)
{
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nchans; c++)
        for (int t = 0; t < nsamp; t++) out_ct[(size_t)c * out_pitch + t] = in_tc[(size_t)t * nchans + c];
}

/* One x2 level of u16 rows: the bin kernel's ((float)a + (float)b) / 2.0f
 * (exact for u16 operands) truncated back by swap
 * (src/aa_device_binning_kernel.cu:33-45, src/aa_device_corner_turn_kernel.cu:91-98). */
void oracle_bin_u16(const uint16_t *in, size_t in_pitch, int nchans, int n_out, uint16_t *out, size_t out_pitch)
{
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nchans; c++) {
        const uint16_t *row = in + (size_t)c * in_pitch;
        uint16_t *orow = out + (size_t)c * out_pitch;
        for (int t = 0; t < n_out; t++)
            orow[t] = (uint16_t)(((float)row[2 * t] + (float)row[2 * t + 1]) / 2.0f);
    }
}

void oracle_dedisperse_rows_u16(const uint16_t *L, size_t pitch, int nchans, const float *dmshifts,
                                const int *dm_idx, int nrows, int t_processed, float mstartdm, float mdmstep,
                                int inBin, float *out, size_t out_pitch)
{
    const int nblk = (t_processed + ORACLE_TBLOCK - 1) / ORACLE_TBLOCK;
    int *delays = (int *)malloc(sizeof(int) * (size_t)nchans * (size_t)(nrows > 0 ? nrows : 1));
    for (int k = 0; k < nrows; k++)
        for (int c = 0; c < nchans; c++)
            delays[(size_t)k * nchans + c] = oracle_delay(dmshifts[c], dm_idx[k], mstartdm, mdmstep);
    const float fn = (float)nchans, fb = (float)inBin;
#pragma omp parallel
    {
        float acc[ORACLE_TBLOCK];
#pragma omp for schedule(dynamic, 1)
        for (long item = 0; item < (long)nrows * nblk; item++) {
            const int k = (int)(item / nblk);
            const int t0 = (int)(item % nblk) * ORACLE_TBLOCK;
            const int n = t_processed - t0 < ORACLE_TBLOCK ? t_processed - t0 : ORACLE_TBLOCK;
            const int *dk = delays + (size_t)k * nchans;
            for (int t = 0; t < n; t++) acc[t] = 0.0f;
            for (int c = 0; c < nchans; c++) {   /* channel order: the float sum rounds */
                const uint16_t *row = L + (size_t)c * pitch + (size_t)(t0 + dk[c]);
                for (int t = 0; t < n; t++) acc[t] = acc[t] + (float)row[t];
            }
            float *orow = out + (size_t)k * out_pitch + t0;
            for (int t = 0; t < n; t++) orow[t] = (acc[t] / fn) / fb;
        }
    }
    free(delays);
}

/* oracle_ddtr_run_u8's sequencing on u16 levels (input_tc: u16 [t][c]). */
int oracle_ddtr_run_u16(const oracle_strategy *s, const uint16_t *input_tc, float **out)
{
    if (!s->ok) return -1;
    const int nchans = s->nchans;
    const float tsamp_original = (float)s->tsamp;
    long inc = 0;
    for (int r = 0; r < s->nranges; r++) {
        int ib = s->dm[r].inBin;
        if (ib < 1 || (ib & (ib - 1)) != 0) return -2;
        if (r > 0 && ib < s->dm[r - 1].inBin) return -3;
    }
    for (int k = 0; k < s->num_tchunks; k++) {
        const int tp0 = oracle_strategy_t_processed(s, 0, k);
        const long len0 = (long)tp0 + s->maxshift;
        if (inc + len0 > s->nsamp) return -4;
        size_t pitch = (size_t)len0;
        uint16_t *level = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)nchans * pitch);
        oracle_transpose_u16(input_tc + (size_t)inc * nchans, (int)len0, nchans, level, pitch);
        long level_len = len0;
        int bin = 1;
        for (int r = 0; r < s->nranges; r++) {
            const int ib = s->dm[r].inBin;
            while (bin < ib) {
                long n_out = level_len / 2;
                uint16_t *next = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)nchans * (size_t)(n_out > 0 ? n_out : 1));
                oracle_bin_u16(level, pitch, nchans, (int)n_out, next, (size_t)n_out);
                free(level);
                level = next;
                pitch = (size_t)n_out;
                level_len = n_out;
                bin *= 2;
            }
            const int tp = oracle_strategy_t_processed(s, r, k);
            const int maxshift_r = s->maxshift / ib;
            if ((long)tp + maxshift_r > level_len) { free(level); return -5; }
            float mstartdm, mdmstep;
            oracle_range_scalars(tsamp_original, ib, s->dm[r].low, s->dm[r].step, &mstartdm, &mdmstep);
            long total_r = 0;
            for (int j = 0; j < s->num_tchunks; j++) total_r += oracle_strategy_t_processed(s, r, j);
            const long col = inc / ib;
            if (col + tp > total_r) { free(level); return -6; }
            int *idx = (int *)malloc(sizeof(int) * (size_t)(s->ndms[r] > 0 ? s->ndms[r] : 1));
            for (int d = 0; d < s->ndms[r]; d++) idx[d] = d;
            oracle_dedisperse_rows_u16(level, pitch, nchans, s->dmshifts, idx, s->ndms[r], tp, mstartdm, mdmstep, ib,
                                       out[r] + col, (size_t)total_r);
            free(idx);
        }
        free(level);
        inc += tp0;
    }
    return 0;
}
