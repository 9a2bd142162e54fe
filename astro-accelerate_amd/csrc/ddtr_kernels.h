// Host-side launchers for the DDTR HIP kernels (gfx950).  All launchers are
// asynchronous on the given stream, check their arguments on the host before
// launching, and return hipSuccess or the first error.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace aa_hip {

// Dedispersion kernel variants selectable by callers and the autotuner.
enum class ddtr_kernel : int {
    automatic = 0,  // pick the fastest variant valid for the geometry
    direct = 1,     // one thread per (DM, sample); global-memory gathers
    lds_swar = 2,   // LDS-staged channel windows, packed 16-bit integer sums
};

// Tile shape of the LDS kernel.  dms_per_block DMs x (64 * samples_per_lane)
// samples per work-group, channels staged chunk_chans at a time.
struct lds_tile {
    int dms_per_wave;      // DM trials handled by one wave (each lane: all of them)
    int waves;             // waves per work-group
    int chunk_chans;       // channels staged in LDS per pass
};

// [t][c] -> [c][t].  in_tc holds nsamp rows of nchans samples.
hipError_t launch_corner_turn_u8(const uint8_t* in_tc, int nsamp, int nchans,
                                 uint8_t* out_ct, size_t out_pitch, hipStream_t s);
// u16 input carrying 8-bit samples; a value above 255 sets *d_err to 1.
hipError_t launch_corner_turn_u16(const uint16_t* in_tc, int nsamp, int nchans,
                                  uint8_t* out_ct, size_t out_pitch, int* d_err,
                                  hipStream_t s);
// [t][c] (u8, or u16 carrying 8-bit samples: sample_bytes 1 or 2) -> level 0
// [c][t] at ptrs[0], plus the x2 floor-average levels 1..nlev (nlev <= 3)
// L_k[t] = (L_{k-1}[2t] + L_{k-1}[2t+1]) >> 1, t < nsamp >> k, in the same
// pass.  Rows of level k are pitches[k] bytes apart.  *fused = the number of
// binned levels written: nlev on the wide path (nchans a multiple of 16,
// 16-byte aligned input, rows and pitches), else 0 (level 0 only).
constexpr int kCtMaxLevels = 3;
hipError_t launch_corner_turn_levels(const void* in_tc, int sample_bytes, long nsamp, int nchans,
                                     uint8_t* const* ptrs, const size_t* pitches, int nlev, int* d_err,
                                     hipStream_t s, int* fused);
// Whether launch_corner_turn_levels writes all nlev binned levels (its wide
// path) for these buffers, and the work-groups that path launches.
bool corner_turn_levels_wide(const void* in_tc, int nchans, uint8_t* const* ptrs, const size_t* pitches, int nlev);
long corner_turn_tiles(long nsamp, int nchans);
// One u8 chunk of a batched corner turn: as launch_corner_turn_levels on its
// wide path (level 0 and nlev binned levels at ptr[k], rows pitch[k] bytes
// apart).  tile_end: work-groups of this job and the jobs before it.
struct ct_job {
    const uint8_t* in;
    long nsamp;
    int nchans, nlev;
    uint8_t* ptr[kCtMaxLevels + 1];
    size_t pitch[kCtMaxLevels + 1];
    long tile_end;
};
// Every job of d_jobs (device memory; all with this nlev, the wide path
// holding for each, tile_end filled in) in one launch of ntiles work-groups.
hipError_t launch_corner_turn_batch(const ct_job* d_jobs, int njobs, long ntiles, int nlev, hipStream_t s);
// u16 [t][c] -> u16 [c][t] (the reference's in-place corner turn, via temp).
hipError_t launch_transpose_u16(const uint16_t* in, int rows, int cols,
                                uint16_t* out, hipStream_t s);

// out[c][t] = (in[c][2t] + in[c][2t+1]) >> 1 for t < n_out.
// u16 rows (in_pitch elements apart) of 8-bit values -> u8 rows (out_pitch
// bytes apart), n samples each; a value above 255 sets *d_err (not cleared).
hipError_t launch_narrow_u16(const uint16_t* in, long in_pitch, int rows, long n, uint8_t* out, size_t out_pitch,
                             int* d_err, hipStream_t s);
// Zero-DM filter in place on [t][c] samples (bp: nchans device floats;
// values clamped to 2^nbits - 1).
hipError_t launch_zero_dm_u8(uint8_t* data_tc, int nsamp, int nchans, const float* bp, int nbits, hipStream_t s);
hipError_t launch_zero_dm_u16(uint16_t* data_tc, int nsamp, int nchans, const float* bp, int nbits, hipStream_t s);
// Zero-DM with outlier rejection in place on [t][c] samples (bp: nchans
// device floats; values clamped to 2^nbits - 1).
hipError_t launch_zero_dm_outliers_u8(uint8_t* data_tc, long nsamp, int nchans, const float* bp, int nbits,
                                      hipStream_t s);
hipError_t launch_zero_dm_outliers_u16(uint16_t* data_tc, long nsamp, int nchans, const float* bp, int nbits,
                                       hipStream_t s);
// Column normalisation of a [rows][cols] array in place (the reference's
// DDTR input/output normalisation kernel): per column s = (float sum of the
// column in the reference's order) / rows - factor, x' = x - s (integers
// clamped to [0, 2^nbits - 1], truncated).  `scratch` (may be null) holds
// normalize_columns_scratch_bytes(cols) device bytes; with it, tall integer
// arrays whose sums are exact take the row-segmented path.
size_t normalize_columns_scratch_bytes(long cols);
// Host: the set_bandpass_average bandpass of a [t][c] input of u8
// (bytes_per_sample 1) or u16 samples: out[i] = mean of samples
// i*nsamp .. i*nsamp + nsamp - 1, i < nchans (the reference's indexing).
void bandpass_average(const void* in_tc, int bytes_per_sample, size_t nchans, size_t nsamp, float* out);
hipError_t launch_normalize_columns_u8(uint8_t* data, long rows, long cols, float factor, int nbits, void* scratch,
                                       hipStream_t s);
hipError_t launch_normalize_columns_u16(uint16_t* data, long rows, long cols, float factor, int nbits,
                                        void* scratch, hipStream_t s);
hipError_t launch_normalize_columns_f32(float* data, long rows, long cols, float factor, hipStream_t s);

hipError_t launch_bin2_u8(const uint8_t* in, size_t in_pitch, int nchans, int n_out,
                          uint8_t* out, size_t out_pitch, hipStream_t s);
// Reference bin_gpu semantics on u16 rows of length nsamp packed into rows
// of nsamp/2 (src/aa_bin_gpu.cpp:8-44), in place through a temp buffer.
hipError_t launch_bin2_u16_packed(uint16_t* data, int nchans, int nsamp, float* temp,
                                  hipStream_t s);

// One DM range.  L: nchans rows of u8 (or u16) samples with row pitch
// `pitch` elements; out: ndms rows of t_processed floats, row pitch out_pitch.
hipError_t launch_dedisperse_direct_u8(const uint8_t* L, size_t pitch, int nchans,
                                       const float* d_dmshifts, int ndms, int t_processed,
                                       float mstartdm, float mdmstep, int inBin,
                                       float* out, size_t out_pitch, hipStream_t s);
hipError_t launch_dedisperse_direct_u16(const uint16_t* L, size_t pitch, int nchans,
                                        const float* d_dmshifts, int ndms, int t_processed,
                                        float mstartdm, float mdmstep, int inBin,
                                        float* out, size_t out_pitch, hipStream_t s);

// Samples wider than 8 bits (u16 rows, row pitches in samples): float sums
// in channel order, as the reference's cache kernel forms them.
// wide_window: the staged window (tile + the largest delay spread of a
// 16-DM tile, even) dedisperse_u16 needs for the range, or -1 when it
// exceeds the kernel's LDS rows (the direct kernel then runs).  Host only.
int wide_window(const float* h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep);
// row_len: valid samples per row (reads beyond it see zeros); win from
// wide_window.  Any even pitch and 4-byte aligned L take single dword loads.
hipError_t launch_dedisperse_wide_u16(const uint16_t* L, size_t pitch, long row_len, int nchans,
                                      const float* d_dmshifts, int ndms, int t_processed, float mstartdm,
                                      float mdmstep, int inBin, float* out, size_t out_pitch, int win, hipStream_t s);
// [t][c] u16 -> [c][t] u16, out rows out_pitch samples apart.
hipError_t launch_transpose_u16_pitch(const uint16_t* in_tc, long nsamp, int nchans, uint16_t* out_ct,
                                      size_t out_pitch, hipStream_t s);
// out[c][t] = (in[c][2t] + in[c][2t+1]) >> 1, t < n_out, u16 rows.
hipError_t launch_bin2_u16(const uint16_t* in, size_t in_pitch, int nchans, long n_out, uint16_t* out,
                           size_t out_pitch, hipStream_t s);

// First single-pulse-search pass on a DM-time plane (boxcar MSD per width
// 1..W, then best boxcar SNR and its width per start sample); `work` holds
// spd_workspace_bytes() device bytes.  snr/width may be null.
#define SPD_MAX_WIDTH 32
size_t spd_workspace_bytes();
hipError_t launch_spd_first_pass(const float* plane, size_t pitch, int ndms, int nsamp, int W, float* msd,
                                 float* snr, size_t snr_pitch, uint16_t* width, size_t width_pitch, void* work,
                                 hipStream_t s);
// The same pass with its workspace in a file-scope device array (one per
// device, spd_static_scratch_bytes() bytes): allocates nothing.  Calls that
// overlap on one device share that array and race.
size_t spd_static_scratch_bytes();
hipError_t launch_spd_first_pass_static(const float* plane, size_t pitch, int ndms, int nsamp, int W, float* msd,
                                        float* snr, size_t snr_pitch, uint16_t* width, size_t width_pitch,
                                        hipStream_t s);

// Synthetic 8-bit filterbank noise: N(mean, sigma^2) rounded and clipped to
// [0,255], SplitMix64 counter stream keyed by seed (value depends only on
// seed and the flat [t][c] index).
hipError_t launch_synth_noise_u8(uint64_t seed, int64_t n, double mean, double sigma,
                                 uint8_t* out_tc, hipStream_t s);
// Adds `amplitude` (saturating) to `width` samples at t0 + floor(dmshift[c]*
// (float)(dm/tsamp)) in each channel of a [t][c] u8 buffer.
hipError_t launch_inject_pulse_u8(uint8_t* data_tc, int nsamp, int nchans,
                                  const float* d_dmshifts, float dm_over_tsamp,
                                  int t0, int width, int amplitude, hipStream_t s);

}  // namespace aa_hip
