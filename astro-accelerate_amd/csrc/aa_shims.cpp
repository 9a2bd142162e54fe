// Drop-in device functions with the reference's signatures
// (include/aa_dedisperse.hpp:20, include/aa_device_load_data.hpp,
// include/aa_corner_turn.hpp, include/aa_bin_gpu.hpp), on the reference's
// u16 [c][t] device layout and the legacy default stream.  The reference
// carries i_nchans, i_nsamp, i_t_processed_s and dm_shifts between these
// calls in __constant__ memory; here they are one mutex-guarded host record
// plus a library-owned device copy of the dmshift table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "ddtr_kernels.h"

#include "aa_bin_gpu.hpp"
#include "aa_ddtr_hip.h"
#include "aa_corner_turn.hpp"
#include "aa_zero_dm.hpp"
#include "aa_zero_dm_outliers.hpp"
#include "aa_dedisperse.hpp"
#include "aa_device_load_data.hpp"
#include "aa_log.hpp"
#include "ddtr_engine.hpp"

namespace astroaccelerate {

namespace {
// Plan tables of one range's pair shapes, kept across dedisperse() calls
// while the range's geometry is unchanged.
struct shim_tables {
    int nchans = 0, ndms = 0, variant = -1;
    float mstartdm = 0.0f, mdmstep = 0.0f;
    size_t pitch = 0;
    void *d = nullptr;
    int row_quads = 0;
};
// The 16-bit kernel's window of one range, kept while its geometry holds.
struct shim_window {
    int nchans = 0, ndms = 0;
    float mstartdm = 0.0f, mdmstep = 0.0f;
    int win = 0;
};

struct shim_constants {
    int nchans = 0;
    long i_nsamp = 0;
    int t_processed = 0;
    std::vector<float> dmshifts;
    float *d_dmshifts = nullptr;
    int d_capacity = 0;
    int device = -1;
    uint8_t *d_u8 = nullptr;      // the u16 input narrowed to u8 rows for the LDS kernels
    size_t u8_capacity = 0;
    size_t u8_pitch = 0;
    int *d_err = nullptr;         // set when an 8-bit call meets a sample above 255
    std::map<int, shim_tables> tables;
    std::map<int, shim_window> windows;
};
std::mutex g_mutex;
shim_constants g_const;
}  // namespace

// src/aa_device_load_data.cu:39-56
void load_chunk_data(unsigned short *device_pointer, unsigned short const *const host_pointer, int t_processed,
                     int maxshift, int nchans, float *dmshifts, float *d_dm_shifts, int nbits) {
    (void)d_dm_shifts;
    (void)nbits;
    std::lock_guard<std::mutex> lock(g_mutex);
    const long length = (long)t_processed + maxshift;
    const size_t size = (size_t)nchans * (size_t)length * sizeof(unsigned short);
    hipError_t e = hipMemcpy(device_pointer, host_pointer, size, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        LOG(log_level::error, std::string("load_chunk_data: ") + hipGetErrorString(e));
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (g_const.d_dmshifts && (g_const.d_capacity < nchans || g_const.device != dev)) {
        (void)hipFree(g_const.d_dmshifts);
        g_const.d_dmshifts = nullptr;
    }
    if (!g_const.d_dmshifts) {
        if (hipMalloc(&g_const.d_dmshifts, sizeof(float) * (size_t)nchans) != hipSuccess) {
            LOG(log_level::error, "load_chunk_data: cannot allocate the dmshift table");
            g_const.d_dmshifts = nullptr;
            return;
        }
        g_const.d_capacity = nchans;
        g_const.device = dev;
    }
    (void)hipMemcpy(g_const.d_dmshifts, dmshifts, sizeof(float) * (size_t)nchans, hipMemcpyHostToDevice);
    // The cached plan tables hold delays computed from the previous table:
    // a new observation (other fch1/foff/nchans) invalidates all of them.
    if (g_const.dmshifts.size() != (size_t)nchans ||
        !std::equal(g_const.dmshifts.begin(), g_const.dmshifts.end(), dmshifts)) {
        for (auto &kv : g_const.tables)
            if (kv.second.d) (void)hipFree(kv.second.d);
        g_const.tables.clear();
        g_const.windows.clear();
    }
    g_const.dmshifts.assign(dmshifts, dmshifts + nchans);
    g_const.nchans = nchans;
    g_const.i_nsamp = length;
    g_const.t_processed = t_processed;
    // dedisperse() allocates nothing per call: the u8 rows of 8-bit data
    // (16-byte aligned, 4096 bytes of slack past each row) and the overflow
    // flag are sized here, for the chunk's longest rows.
    const size_t pitch = ((size_t)length + 4096 + 255) / 256 * 256;
    if (g_const.u8_capacity < pitch * (size_t)nchans) {
        if (g_const.d_u8) (void)hipFree(g_const.d_u8);
        g_const.d_u8 = nullptr;
        g_const.u8_capacity = 0;
        if (hipMalloc(&g_const.d_u8, pitch * (size_t)nchans) == hipSuccess)
            g_const.u8_capacity = pitch * (size_t)nchans;
        else
            LOG(log_level::error, "load_chunk_data: cannot allocate the 8-bit rows");
    }
    g_const.u8_pitch = pitch;
    if (!g_const.d_err) {
        if (hipMalloc(&g_const.d_err, sizeof(int)) != hipSuccess || hipMemset(g_const.d_err, 0, sizeof(int)) != hipSuccess) {
            LOG(log_level::error, "load_chunk_data: cannot allocate the input flag");
            g_const.d_err = nullptr;
        }
    } else {
        // The previous chunk's 8-bit calls flag samples above 255 without
        // waiting for the device; the flag is read here, where the copy
        // above has already synchronised.
        int flag = 0;
        if (hipMemcpy(&flag, g_const.d_err, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess && flag) {
            LOG(log_level::error, "dedisperse: samples above 255 met with nbits <= 8 in the previous chunk "
                                  "(their sums were taken modulo 256); pass the data's nbits (16)");
            (void)hipMemset(g_const.d_err, 0, sizeof(int));
        }
    }
}

// src/aa_device_load_data.cu:58-61
void set_dedispersion_constants(int t_processed, int maxshift) {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_const.i_nsamp = (long)t_processed + maxshift;
    g_const.t_processed = t_processed;
}

// src/aa_dedisperse.cpp:5-163.  The failsafe kernel's semantics for every
// geometry: its grid covers t < 8*(t_processed/8) and d < 64*(ndms/64), the
// output row stride is the current i_t_processed_s.
//  * nbits <= 8: the u16 rows hold 8-bit values; they are narrowed to the u8
//    rows sized by load_chunk_data (on every call: bin_gpu rewrites the rows
//    between ranges) and dedispersed by the LDS kernels, the first chunk
//    shape that fits, with the range's plan tables built on its first call
//    and kept while its geometry holds.  A sample above 255 is flagged on
//    the device and reported by the next load_chunk_data.
//  * nbits > 8: the u16 rows are dedispersed as they are, float sums in
//    channel order (the 16-bit LDS kernel, or the direct kernel when a tile's
//    delay spread exceeds its rows).
// Nothing waits for the device and nothing is allocated per call beyond a
// range's first.
int dedisperse(int i, int t_processed, int *inBin, float *dmshifts, unsigned short *d_input, float *d_output,
               float *d_dm_shifts, int nchans, float *tsamp, float *dm_low, float *dm_step, int const *const ndms,
               int nbits, int failsafe) {
    (void)dmshifts;
    (void)d_dm_shifts;
    (void)nchans;
    (void)failsafe;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (nbits < 1 || nbits > 32) {
        LOG(log_level::error, "dedisperse: nbits must be 1..32");
        return 1;
    }
    if (!g_const.d_dmshifts || g_const.nchans <= 0) {
        LOG(log_level::error, "dedisperse: load_chunk_data has not been called");
        return 1;
    }
    const float mstartdm = (float)(dm_low[i] / (*tsamp));
    const float mdmstep = (float)(dm_step[i] / (*tsamp));
    const int t_cov = (t_processed / 8) * 8;
    const int d_cov = (ndms[i] / 64) * 64;
    const int nch = g_const.nchans;
    const long n = g_const.i_nsamp;
    if (aa_hip::check_delays(g_const.dmshifts.data(), nch, d_cov, mstartdm, mdmstep, t_cov, n) != 0) {
        LOG(log_level::error, std::string("dedisperse: ") + aa_hip::last_error());
        return 1;
    }
    if (d_cov == 0 || t_cov == 0) return 0;
    hipError_t e = hipSuccess;
    if (nbits > 8) {
        shim_window &w = g_const.windows[i];
        if (w.win == 0 || w.nchans != nch || w.ndms != d_cov || w.mstartdm != mstartdm || w.mdmstep != mdmstep) {
            w = shim_window{nch, d_cov, mstartdm, mdmstep,
                            aa_hip::wide_window(g_const.dmshifts.data(), nch, d_cov, mstartdm, mdmstep)};
        }
        if (w.win > 0)
            e = aa_hip::launch_dedisperse_wide_u16(d_input, (size_t)n, n, nch, g_const.d_dmshifts, d_cov, t_cov,
                                                   mstartdm, mdmstep, inBin[i], d_output, (size_t)g_const.t_processed,
                                                   w.win, nullptr);
        else
            e = aa_hip::launch_dedisperse_direct_u16(d_input, (size_t)n, nch, g_const.d_dmshifts, d_cov, t_cov,
                                                     mstartdm, mdmstep, inBin[i], d_output,
                                                     (size_t)g_const.t_processed, nullptr);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) LOG(log_level::error, std::string("dedisperse: ") + hipGetErrorString(e));
        return e == hipSuccess ? 0 : 1;
    }
    const size_t pitch = g_const.u8_pitch;
    if (!g_const.d_u8 || !g_const.d_err || pitch * (size_t)nch > g_const.u8_capacity || (long)pitch < n + 4096) {
        LOG(log_level::error, "dedisperse: the 8-bit rows of load_chunk_data are missing or too short");
        return 1;
    }
    e = aa_hip::launch_narrow_u16(d_input, n, nch, n, g_const.d_u8, pitch, g_const.d_err, nullptr);
    if (e != hipSuccess) {
        LOG(log_level::error, std::string("dedisperse: ") + hipGetErrorString(e));
        return 1;
    }
    aa_hip::range_launch rl;
    if (aa_hip::resolve_launch(AA_HIP_KERNEL_AUTO, g_const.d_u8, pitch, nch, g_const.d_dmshifts,
                               g_const.dmshifts.data(), d_cov, t_cov, mstartdm, mdmstep, inBin[i], d_output,
                               (size_t)g_const.t_processed, aa_hip::lds_choice{-2, 0}, &rl) != AA_HIP_SUCCESS) {
        LOG(log_level::error, std::string("dedisperse: ") + aa_hip::last_error());
        return 1;
    }
    if (rl.variant >= 2) {
        shim_tables &tb = g_const.tables[i];
        if (!tb.d || tb.nchans != nch || tb.ndms != d_cov || tb.variant != rl.variant || tb.pitch != pitch ||
            tb.mstartdm != mstartdm || tb.mdmstep != mdmstep) {
            if (tb.d) (void)hipFree(tb.d);
            tb = shim_tables{};
            std::vector<uint8_t> bytes;
            const int q = aa_hip::lds_plan_tables(rl.variant, g_const.dmshifts.data(), nch, d_cov, mstartdm, mdmstep,
                                                  pitch, &bytes);
            if (q != rl.row_quads) {
                LOG(log_level::error, "dedisperse: plan tables disagree with the chosen row width");
                return 1;
            }
            e = hipMalloc(&tb.d, bytes.size());
            if (e == hipSuccess) e = hipMemcpy(tb.d, bytes.data(), bytes.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                if (tb.d) (void)hipFree(tb.d);
                tb = shim_tables{};
                LOG(log_level::error, std::string("dedisperse: plan tables: ") + hipGetErrorString(e));
                return 1;
            }
            tb.nchans = nch;
            tb.ndms = d_cov;
            tb.variant = rl.variant;
            tb.pitch = pitch;
            tb.mstartdm = mstartdm;
            tb.mdmstep = mdmstep;
            tb.row_quads = q;
        }
        rl.pair_recs = tb.d;
    }
    e = aa_hip::launch_range(rl, nullptr);
    if (e == hipSuccess) e = hipGetLastError();
    return e == hipSuccess ? 0 : 1;
}

// src/aa_corner_turn.cpp:118-121: [secondary][primary] -> [primary][secondary].
void corner_turn_SM_inplace(unsigned short *const d_input, unsigned short *const d_temporary,
                            const size_t primary_size, const size_t secondary_size) {
    hipError_t e = aa_hip::launch_transpose_u16(d_input, (int)secondary_size, (int)primary_size, d_temporary, nullptr);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_input, d_temporary, primary_size * secondary_size * sizeof(unsigned short),
                           hipMemcpyDeviceToDevice, nullptr);
    if (e != hipSuccess) LOG(log_level::error, std::string("corner_turn_SM_inplace: ") + hipGetErrorString(e));
}

// src/aa_bin_gpu.cpp:8-44
void bin_gpu(unsigned short *const d_input, float *const d_output, const int nchans, const int nsamp) {
    hipError_t e = aa_hip::launch_bin2_u16_packed(d_input, nchans, nsamp, d_output, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) LOG(log_level::error, std::string("bin_gpu: ") + hipGetErrorString(e));
}

void zero_dm(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits,
             float *normalization_factor)
{
    const hipError_t e = aa_hip::launch_zero_dm_u16(d_input, nsamp, nchans, normalization_factor, nbits, 0);
    if (e != hipSuccess) LOG(log_level::error, std::string("zero_dm: ") + hipGetErrorString(e));
    (void)hipDeviceSynchronize();
}

void zero_dm(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits)
{
    if (nchans <= 0 || nbits < 1 || nbits > 16) {
        LOG(log_level::error, "zero_dm: invalid nchans or nbits");
        return;
    }
    const std::vector<float> bp((size_t)nchans, (float)(((1u << nbits) - 1u) / 2.0));
    float *d_bp = nullptr;
    if (hipMalloc(&d_bp, bp.size() * sizeof(float)) != hipSuccess) {
        LOG(log_level::error, "zero_dm: could not allocate the normalisation factors");
        return;
    }
    if (hipMemcpy(d_bp, bp.data(), bp.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess)
        zero_dm(d_input, nchans, nsamp, nbits, d_bp);
    (void)hipFree(d_bp);
}

void zero_dm_outliers(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits,
                      float *normalization_factor)
{
    const hipError_t e =
        aa_hip::launch_zero_dm_outliers_u16(d_input, nsamp, nchans, normalization_factor, nbits, 0);
    if (e != hipSuccess) LOG(log_level::error, std::string("zero_dm_outliers: ") + hipGetErrorString(e));
    (void)hipDeviceSynchronize();
}

// The reference's 4-argument form passes its pointer's value where the
// allocation needs its address (src/aa_zero_dm_outliers.cpp:20); here the
// default factors are allocated properly.
void zero_dm_outliers(unsigned short *const d_input, const int nchans, const int nsamp, const int nbits)
{
    if (nchans <= 0 || nbits < 1 || nbits > 16) {
        LOG(log_level::error, "zero_dm_outliers: invalid nchans or nbits");
        return;
    }
    const std::vector<float> bp((size_t)nchans, (float)((std::pow(2.0, nbits) - 1.0) / 2.0));
    float *d_bp = nullptr;
    if (hipMalloc(&d_bp, bp.size() * sizeof(float)) != hipSuccess) {
        LOG(log_level::error, "zero_dm_outliers: could not allocate the normalisation factors");
        return;
    }
    if (hipMemcpy(d_bp, bp.data(), bp.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess)
        zero_dm_outliers(d_input, nchans, nsamp, nbits, d_bp);
    (void)hipFree(d_bp);
}

// zero_dm_normalization_dm (src/aa_zero_dm.cpp:49-60): normalises a u16
// array of `second` rows of `first` values per column, factor
// (2^nbits - 1)/2.  The reference's runner passes (nchans, nTimesamples) in
// that order (include/aa_permitted_pipelines_generic.hpp:734), so each
// channel of the [t][c] chunk is normalised over its time samples.
void zero_dm_normalization_dm(unsigned short *d_input, size_t first, size_t second, int nbits)
{
    if (nbits < 1 || nbits > 16 || first == 0) {
        LOG(log_level::error, "zero_dm_normalization_dm: invalid size or nbits");
        return;
    }
    void *scratch = nullptr;
    if (hipMalloc(&scratch, aa_hip::normalize_columns_scratch_bytes((long)first)) != hipSuccess) scratch = nullptr;
    const float factor = (float)((std::pow(2.0, nbits) - 1.0) / 2.0);
    const hipError_t e =
        aa_hip::launch_normalize_columns_u16(d_input, (long)second, (long)first, factor, nbits, scratch, 0);
    if (e != hipSuccess) LOG(log_level::error, std::string("zero_dm_normalization_dm: ") + hipGetErrorString(e));
    (void)hipDeviceSynchronize();
    if (scratch) (void)hipFree(scratch);
}

// post_DDTR_normalization (src/aa_zero_dm.cpp:62-74): each time sample of a
// [nDMs][nTimesamples] float plane moved by (its mean over DMs - 127.5).
void post_DDTR_normalization(float *d_input, size_t nTimesamples, size_t nDMs)
{
    if (nTimesamples == 0) return;
    const hipError_t e = aa_hip::launch_normalize_columns_f32(d_input, (long)nDMs, (long)nTimesamples, 127.5f, 0);
    if (e != hipSuccess) LOG(log_level::error, std::string("post_DDTR_normalization: ") + hipGetErrorString(e));
    (void)hipDeviceSynchronize();
}

}  // namespace astroaccelerate
