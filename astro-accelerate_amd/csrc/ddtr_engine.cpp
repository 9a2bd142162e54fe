// Dedispersion engine: corner turn -> x2 cascade -> per-range dedispersion,
// all on caller streams, with every launch validated on the host first.
#include "ddtr_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "aa_ddtr_hip.h"
#include "ddtr_lds.h"

namespace aa_hip {

namespace {
thread_local std::string g_last_error;

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int fail(int code, const std::string &msg) {
    set_last_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? AA_HIP_ERROR_OUT_OF_MEMORY : AA_HIP_ERROR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
int log2i(int v) {
    int k = 0;
    while ((1 << k) < v) ++k;
    return k;
}

struct device_guard {
    int prev = -1;
    explicit device_guard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~device_guard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};
}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

int check_delays(const float *dmshifts, int nchans, int ndms, float mstartdm, float mdmstep,
                 int t_processed, long row_len) {
    if (ndms <= 0 || t_processed <= 0) return AA_HIP_SUCCESS;
    // delay(c, d) is monotone in d for each c (fma of a non-negative step),
    // so the extremes over d are at d = 0 and d = ndms - 1.
    long lo = 0, hi = 0;
    for (int c = 0; c < nchans; ++c) {
        const long a = host_delay(dmshifts[c], 0, mstartdm, mdmstep);
        const long b = host_delay(dmshifts[c], ndms - 1, mstartdm, mdmstep);
        lo = std::min(lo, std::min(a, b));
        hi = std::max(hi, std::max(a, b));
    }
    if (lo < 0)
        return fail(AA_HIP_ERROR_UNSUPPORTED,
                    "negative dispersion delay (ascending channel frequencies are not supported)");
    if ((long)t_processed + hi > row_len)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT,
                    "rows too short: t_processed " + std::to_string(t_processed) + " + max delay " +
                        std::to_string(hi) + " > row length " + std::to_string(row_len));
    if (mdmstep < 0.0f) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "negative DM step");
    return AA_HIP_SUCCESS;
}

hipError_t dispatch_dedisperse(int kernel, const uint8_t *L, size_t pitch, int nchans,
                               const float *d_dmshifts, const float *h_dmshifts, int ndms,
                               int t_processed, float mstartdm, float mdmstep, int inBin,
                               float *out, size_t out_pitch, hipStream_t s, int row_quads) {
    if (kernel != AA_HIP_KERNEL_DIRECT && (pitch % 16) == 0 && ((uintptr_t)L % 16) == 0) {
        if (row_quads == 0) row_quads = lds_row_quads(h_dmshifts, nchans, ndms, mstartdm, mdmstep);
        if (row_quads > 0) {
            lds_params p{};
            p.L = L;
            p.pitch = pitch;
            p.nchans = nchans;
            p.dmshifts = d_dmshifts;
            p.ndms = ndms;
            p.t_processed = t_processed;
            p.mstartdm = mstartdm;
            p.mdmstep = mdmstep;
            p.fnchans = (float)nchans;
            p.fbin = (float)inBin;
            p.out = out;
            p.out_pitch = out_pitch;
            p.row_quads = row_quads;
            return launch_dedisperse_lds(p, s);
        }
        if (kernel == AA_HIP_KERNEL_LDS) return hipErrorInvalidValue;  // asked for LDS, geometry does not fit
    }
    if (kernel == AA_HIP_KERNEL_LDS) return hipErrorInvalidValue;
    return launch_dedisperse_direct_u8(L, pitch, nchans, d_dmshifts, ndms, t_processed, mstartdm,
                                       mdmstep, inBin, out, out_pitch, s);
}

int ddtr_engine::create(int device, int nchans, const float *dmshifts, int nranges,
                        const float *dm_low, const float *dm_step, const int *ndms,
                        const int *inBin, double tsamp, int maxshift, int max_t_processed0,
                        ddtr_engine **out) {
    if (!out) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (nchans <= 0 || nranges <= 0 || !dmshifts || !dm_low || !dm_step || !ndms || !inBin ||
        maxshift < 0 || max_t_processed0 <= 0 || !(tsamp > 0.0))
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "invalid plan geometry");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(AA_HIP_ERROR_NO_DEVICE, "no HIP device");
    }
    if (device < 0 || device >= count) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "bad device id");

    int max_bin = 1;
    for (int r = 0; r < nranges; ++r) {
        if (!is_pow2(inBin[r]))
            return fail(AA_HIP_ERROR_UNSUPPORTED, "inBin must be a power of two (range " + std::to_string(r) + ")");
        if (ndms[r] < 0) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "negative ndms");
        max_bin = std::max(max_bin, inBin[r]);
    }
    for (int c = 0; c < nchans; ++c)
        if (!(dmshifts[c] >= 0.0f) || !std::isfinite(dmshifts[c]))
            return fail(AA_HIP_ERROR_UNSUPPORTED,
                        "dmshift table must be finite and non-negative (descending channel frequencies)");

    device_guard guard(device);
    ddtr_engine *e = new ddtr_engine();
    e->m_device = device;
    e->m_nchans = nchans;
    e->m_tsamp = (float)tsamp;
    e->m_maxshift = maxshift;
    e->m_dmshifts.assign(dmshifts, dmshifts + nchans);
    for (int r = 0; r < nranges; ++r) e->m_ranges.push_back(range_desc{dm_low[r], dm_step[r], ndms[r], inBin[r]});

    hipError_t st = hipMalloc(&e->m_d_dmshifts, sizeof(float) * (size_t)nchans);
    if (st == hipSuccess) st = hipMemcpy(e->m_d_dmshifts, dmshifts, sizeof(float) * (size_t)nchans, hipMemcpyHostToDevice);
    if (st == hipSuccess) st = hipMalloc(&e->m_d_err, sizeof(int));
    if (st == hipSuccess) st = hipMemset(e->m_d_err, 0, sizeof(int));
    e->m_device_bytes = sizeof(float) * (size_t)nchans + sizeof(int);

    const int nlevels = log2i(max_bin) + 1;
    long cap = (long)max_t_processed0 + maxshift;
    for (int k = 0; k < nlevels && st == hipSuccess; ++k) {
        level lv;
        lv.capacity = cap;
        lv.pitch = round_up((size_t)cap + kRowPad, kPitchAlign);
        const size_t bytes = lv.pitch * (size_t)nchans;
        st = hipMalloc(&lv.ptr, bytes);
        if (st == hipSuccess) st = hipMemset(lv.ptr, 0, bytes);
        e->m_device_bytes += bytes;
        e->m_levels.push_back(lv);
        cap /= 2;
    }
    if (st == hipSuccess) st = hipDeviceSynchronize();
    if (st != hipSuccess) {
        const int code = hip_fail(st, "ddtr_engine allocation");
        delete e;
        return code;
    }
    *out = e;
    return AA_HIP_SUCCESS;
}

ddtr_engine::~ddtr_engine() {
    device_guard guard(m_device);
    for (auto &lv : m_levels)
        if (lv.ptr) (void)hipFree(lv.ptr);
    if (m_d_dmshifts) (void)hipFree(m_d_dmshifts);
    if (m_d_err) (void)hipFree(m_d_err);
}

int ddtr_engine::set_kernel(int kernel) {
    if (kernel < AA_HIP_KERNEL_AUTO || kernel > AA_HIP_KERNEL_LDS)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "unknown kernel variant");
    m_kernel = kernel;
    return AA_HIP_SUCCESS;
}

void ddtr_engine::range_scalars(int r, float *mstartdm, float *mdmstep) const {
    const range_desc &rd = m_ranges[(size_t)r];
    const float tsamp_r = m_tsamp * (float)rd.inBin;
    *mstartdm = rd.dm_low / tsamp_r;
    *mdmstep = rd.dm_step / tsamp_r;
}

int ddtr_engine::ingest(const void *d_input_tc, int input_bits, int t_processed0, hipStream_t s) {
    if (!d_input_tc) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "input is NULL");
    const long len0 = (long)t_processed0 + m_maxshift;
    if (t_processed0 <= 0 || len0 > m_levels[0].capacity)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "t_processed0 exceeds the handle's capacity");
    device_guard guard(m_device);
    hipError_t e;
    if (input_bits == 8) {
        e = launch_corner_turn_u8((const uint8_t *)d_input_tc, (int)len0, m_nchans, m_levels[0].ptr,
                                  m_levels[0].pitch, s);
    } else if (input_bits == 16) {
        e = launch_corner_turn_u16((const uint16_t *)d_input_tc, (int)len0, m_nchans, m_levels[0].ptr,
                                   m_levels[0].pitch, m_d_err, s);
    } else {
        return fail(AA_HIP_ERROR_UNSUPPORTED, "input_bits must be 8 or 16");
    }
    if (e != hipSuccess) return hip_fail(e, "corner turn");
    for (auto &lv : m_levels) lv.len = 0;
    m_levels[0].len = len0;
    m_resident = 0;
    return AA_HIP_SUCCESS;
}

int ddtr_engine::take_input_overflow(hipStream_t s) {
    device_guard guard(m_device);
    int flag = 0;
    hipError_t e = hipMemcpyAsync(&flag, m_d_err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && flag) e = hipMemsetAsync(m_d_err, 0, sizeof(int), s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "input flag");
    return flag ? 1 : 0;
}

int ddtr_engine::ensure_level(int k, hipStream_t s) {
    if (m_resident < 0) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "no chunk ingested");
    while (m_resident < k) {
        const level &src = m_levels[(size_t)m_resident];
        level &dst = m_levels[(size_t)m_resident + 1];
        const long n_out = src.len / 2;
        if (n_out > dst.capacity) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "binning level overflow");
        hipError_t e = launch_bin2_u8(src.ptr, src.pitch, m_nchans, (int)n_out, dst.ptr, dst.pitch, s);
        if (e != hipSuccess) return hip_fail(e, "binning");
        dst.len = n_out;
        ++m_resident;
    }
    return AA_HIP_SUCCESS;
}

int ddtr_engine::run_range(int r, int t_processed, float *d_out, size_t out_pitch, hipStream_t s) {
    if (r < 0 || r >= (int)m_ranges.size()) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "bad range index");
    if (!d_out || out_pitch < (size_t)std::max(t_processed, 0))
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "bad output buffer");
    const range_desc &rd = m_ranges[(size_t)r];
    const int k = log2i(rd.inBin);
    device_guard guard(m_device);
    int st = ensure_level(k, s);
    if (st != AA_HIP_SUCCESS) return st;
    const level &lv = m_levels[(size_t)k];
    // Row length the reference gives this range: t_processed + maxshift/inBin.
    const long i_nsamp = (long)t_processed + m_maxshift / rd.inBin;
    if (i_nsamp > lv.len)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "range " + std::to_string(r) + " needs " +
                                                       std::to_string(i_nsamp) + " samples per row, level holds " +
                                                       std::to_string(lv.len));
    float mstartdm, mdmstep;
    range_scalars(r, &mstartdm, &mdmstep);
    st = check_delays(m_dmshifts.data(), m_nchans, rd.ndms, mstartdm, mdmstep, t_processed, i_nsamp);
    if (st != AA_HIP_SUCCESS) return st;
    if (m_row_quads.size() != m_ranges.size()) m_row_quads.assign(m_ranges.size(), 0);
    if (m_row_quads[(size_t)r] == 0)
        m_row_quads[(size_t)r] = lds_row_quads(m_dmshifts.data(), m_nchans, rd.ndms, mstartdm, mdmstep);
    hipError_t e = dispatch_dedisperse(m_kernel, lv.ptr, lv.pitch, m_nchans, m_d_dmshifts, m_dmshifts.data(),
                                       rd.ndms, t_processed, mstartdm, mdmstep, rd.inBin, d_out, out_pitch, s,
                                       m_row_quads[(size_t)r]);
    if (e != hipSuccess) return hip_fail(e, "dedisperse");
    return AA_HIP_SUCCESS;
}

int ddtr_engine::execute_chunk(const void *d_input_tc, int input_bits, const int *t_processed,
                               float *const *d_out, const size_t *out_pitch, hipStream_t s) {
    if (!t_processed || !d_out || !out_pitch) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL array");
    int st = ingest(d_input_tc, input_bits, t_processed[0], s);
    for (int r = 0; r < (int)m_ranges.size() && st == AA_HIP_SUCCESS; ++r)
        st = run_range(r, t_processed[r], d_out[r], out_pitch[r], s);
    return st;
}

}  // namespace aa_hip
