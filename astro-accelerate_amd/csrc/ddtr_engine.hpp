// Per-handle dedispersion engine: the device state that the reference keeps
// in __constant__ symbols and runner members
// (src/aa_device_dedispersion_kernel.cu:13-14, src/aa_device_load_data.cu:39-61,
// include/aa_permitted_pipelines_generic.hpp:687-840), held per handle so
// several plans may coexist in one process.
//
// Device layout (HBM): level k of the x2 cascade is nchans rows of u8
// samples, row pitch a multiple of 256 bytes with at least kRowPad bytes of
// zero-initialised tail, so that the widest LDS window of the last time tile
// of a row stays inside the allocation.  All levels stay resident for the
// chunk (cfg. 4096 ch x 1 M samples: 4.29 + 2.15 + 1.07 + 0.54 GB).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ddtr_kernels.h"
#include "ddtr_lds.h"

namespace aa_hip {

constexpr size_t kRowPad = 4096;    // bytes of slack after the longest row
constexpr size_t kPitchAlign = 256; // row pitch alignment in bytes

struct range_desc {
    float dm_low;   // strategy-redefined low DM of the range
    float dm_step;
    int ndms;
    int inBin;      // power of two
};

// Thread-local message of the most recent failure.
void set_last_error(const std::string &msg);
const char *last_error();

// LDS chunk shape for a range and the staged row width it needs; variant -1
// when no shape fits (the direct kernel then runs), -2 = not yet chosen.
struct lds_choice {
    int variant;
    int row_quads;
};
// First shape (or only `variant` when >= 0) whose rows hold the range's
// windows, tried in the order 4, 2 (pair pre-summation with 2048- and
// 1024-sample tiles, the fastest on the 8-range 4096-channel plan), 3, 0, 1.
lds_choice choose_lds(const float *h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, int variant);

// One range, resolved to a kernel: variant >= 0 = LDS kernel with that chunk
// shape, -1 = direct kernel.
struct range_launch {
    const uint8_t *L = nullptr;
    size_t pitch = 0;
    int nchans = 0;
    const float *d_dmshifts = nullptr;
    int ndms = 0, t_processed = 0, inBin = 1;
    float mstartdm = 0.0f, mdmstep = 0.0f;
    float *out = nullptr;
    size_t out_pitch = 0;
    int variant = -1;
    int row_quads = 0;
    const void *pair_recs = nullptr;  // chunk shapes 2-4: device plan tables
    int sample_bytes = 1;             // 2: u16 rows (samples wider than 8 bits; pitch still in bytes)
    long row_len = 0;                 // u16 rows: valid samples per row
    int wide_win = -1;                // u16 rows: dedisperse_u16's window, -1 = direct kernel
};

class ddtr_engine {
public:
    // Returns AA_HIP_* status; on success *out owns the engine.  sample_bits
    // 8: u8 levels (8-bit samples, or u16 input holding them); 16: u16
    // levels for samples wider than 8 bits (float sums in channel order).
    static int create(int device, int nchans, const float *dmshifts, int nranges,
                      const float *dm_low, const float *dm_step, const int *ndms,
                      const int *inBin, double tsamp, int maxshift, int max_t_processed0,
                      ddtr_engine **out, int sample_bits = 8);
    int sample_bits() const { return m_sample_bytes * 8; }
    ~ddtr_engine();

    int set_kernel(int kernel);
    int kernel() const { return m_kernel; }
    size_t device_bytes() const { return m_device_bytes; }
    int device() const { return m_device; }
    int nchans() const { return m_nchans; }
    int nranges() const { return (int)m_ranges.size(); }
    int maxshift() const { return m_maxshift; }
    const range_desc &range(int r) const { return m_ranges[(size_t)r]; }
    const float *d_dmshifts() const { return m_d_dmshifts; }
    const std::vector<float> &dmshifts() const { return m_dmshifts; }

    // mstartdm/mdmstep of range r exactly as the reference derives them
    // (include/aa_permitted_pipelines_generic.hpp:778, src/aa_dedisperse.cpp:150-153).
    void range_scalars(int r, float *mstartdm, float *mdmstep) const;

    int ingest(const void *d_input_tc, int input_bits, int t_processed0, hipStream_t s);
    // ingest() of an 8-bit chunk left to a batched launch: when the chunk
    // and this plan's levels take the wide corner turn, fills *job with it
    // (tile_end unset), marks its levels resident and sets *deferred; else
    // leaves everything as it was (the caller runs ingest()).  The job must
    // run before anything reads the levels.
    int ingest_job(const void *d_input_tc, int t_processed0, ct_job *job, bool *deferred);
    // Synchronises `s`, returns 1 if a u16 ingest met a sample above 255
    // since the last call (and clears the flag), 0 if not, <0 on error.
    int take_input_overflow(hipStream_t s);
    int run_range(int r, int t_processed, float *d_out, size_t out_pitch, hipStream_t s);
    // Everything run_range does before the launch (binning up to the range's
    // level on `s`, validation, kernel choice); the launch itself is left to
    // the caller (launch_range, or a batched launch).
    int prepare_range(int r, int t_processed, float *d_out, size_t out_pitch, hipStream_t s, range_launch *rl);
    // LDS chunk shape preferred for range r (-1 = first that fits, in the
    // order of choose_lds); used when it fits the range's delay spread.
    int set_range_variant(int r, int variant);
    // Times every LDS chunk shape that fits range r (reps launches each, after
    // one untimed launch; output written to d_out) and keeps the fastest as
    // the range's preference.  ms[v] = average launch time of shape v, or -1
    // when it does not fit.  *best = -1 when no shape fits.
    int autotune_range(int r, int t_processed, float *d_out, size_t out_pitch, int reps, hipStream_t s,
                       int *best, float *ms);
    int execute_chunk(const void *d_input_tc, int input_bits, const int *t_processed,
                      float *const *d_out, const size_t *out_pitch, hipStream_t s);

    // Level-k buffer of the current chunk (for tests and the batched path).
    const uint8_t *level_ptr(int k) const { return m_levels[(size_t)k].ptr; }
    size_t level_pitch(int k) const { return m_levels[(size_t)k].pitch; }
    long level_len(int k) const { return m_levels[(size_t)k].len; }

private:
    ddtr_engine() = default;
    int ensure_level(int k, hipStream_t s);

    struct level {
        uint8_t *ptr = nullptr;
        size_t pitch = 0;     // bytes between rows
        long capacity = 0;    // longest row this buffer holds
        long len = 0;         // valid samples per row for the current chunk
    };

    int m_device = 0;
    int m_sample_bytes = 1;  // bytes per level sample: 1 (u8) or 2 (u16)
    int m_nchans = 0;
    float m_tsamp = 0.0f;  // metadata tsamp narrowed to float
    int m_maxshift = 0;
    int m_kernel = 0;
    std::vector<range_desc> m_ranges;
    std::vector<float> m_dmshifts;
    float *m_d_dmshifts = nullptr;
    int *m_d_err = nullptr;
    std::vector<level> m_levels;
    int m_resident = -1;  // highest level valid for the current chunk
    std::vector<lds_choice> m_lds;     // per range: chosen LDS shape (-2 = not computed, -1 = direct)
    std::vector<int> m_variant_pref;   // per range: preferred LDS shape, -1 = any
    std::vector<int> m_wide_win;       // per range (u16 levels): dedisperse_u16 window, 0 = not computed
    struct plan_tables {
        int variant = -1;  // chunk shape the tables are for
        void *ptr = nullptr;
        size_t bytes = 0;
    };
    std::vector<plan_tables> m_tables;  // per range: chunk shape 2-4 plan tables (device), built on first use
    size_t m_device_bytes = 0;
};

// Validates a dedispersion launch against the row length actually valid in
// L: every (c, d) delay must lie in [0, row_len - t_processed].  Host only.
int check_delays(const float *dmshifts, int nchans, int ndms, float mstartdm, float mdmstep,
                 int t_processed, long row_len);

// Host copy of the device delay rule.
inline int host_delay(float dmshift, int d, float mstartdm, float mdmstep) {
    const float shift_temp = __builtin_fmaf((float)d, mdmstep, mstartdm);
    const float prod = dmshift * shift_temp;
    return (int)prod;
}

// Resolves the kernel for one range (no launch).  choice.variant == -2 picks
// the first LDS shape that fits.  Fails with UNSUPPORTED if kernel is
// AA_HIP_KERNEL_LDS and no shape fits.
int resolve_launch(int kernel, const uint8_t *L, size_t pitch, int nchans, const float *d_dmshifts,
                   const float *h_dmshifts, int ndms, int t_processed, float mstartdm, float mdmstep, int inBin,
                   float *out, size_t out_pitch, lds_choice choice, range_launch *rl);
lds_params lds_params_of(const range_launch &rl);
hipError_t launch_range(const range_launch &rl, hipStream_t s);

}  // namespace aa_hip
