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

lds_choice choose_lds(const float *h_dmshifts, int nchans, int ndms, float mstartdm, float mdmstep, int variant) {
    lds_choice c{-1, -1};
    static const int order[kLdsVariants] = {4, 5, 2, 3, 0, 1, 6};
    for (int v : order) {
        if (variant >= 0 && v != variant) continue;
        const int q = lds_row_quads(v, h_dmshifts, nchans, ndms, mstartdm, mdmstep);
        if (q > 0) {
            c.variant = v;
            c.row_quads = q;
            return c;
        }
    }
    return c;
}

int resolve_launch(int kernel, const uint8_t *L, size_t pitch, int nchans, const float *d_dmshifts,
                   const float *h_dmshifts, int ndms, int t_processed, float mstartdm, float mdmstep, int inBin,
                   float *out, size_t out_pitch, lds_choice choice, range_launch *rl) {
    *rl = range_launch{};
    rl->L = L;
    rl->pitch = pitch;
    rl->nchans = nchans;
    rl->d_dmshifts = d_dmshifts;
    rl->ndms = ndms;
    rl->t_processed = t_processed;
    rl->mstartdm = mstartdm;
    rl->mdmstep = mdmstep;
    rl->inBin = inBin;
    rl->out = out;
    rl->out_pitch = out_pitch;
    rl->variant = -1;
    const int forced = kernel == AA_HIP_KERNEL_LDS_C32    ? 0
                       : kernel == AA_HIP_KERNEL_LDS_C16  ? 1
                       : kernel == AA_HIP_KERNEL_LDS_PAIR ? 4   // the widest pair tile that fits: 4, 5, 2
                       : kernel == AA_HIP_KERNEL_LDS_QUAD ? 3
                                                          : -1;
    if (kernel != AA_HIP_KERNEL_DIRECT && (pitch % 16) == 0 && ((uintptr_t)L % 16) == 0) {
        if (forced >= 0) {
            choice = choose_lds(h_dmshifts, nchans, ndms, mstartdm, mdmstep, forced);
            if (choice.variant < 0 && forced == 4) choice = choose_lds(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 5);
            if (choice.variant < 0 && forced == 4) choice = choose_lds(h_dmshifts, nchans, ndms, mstartdm, mdmstep, 2);
        }
        else if (choice.variant == -2) choice = choose_lds(h_dmshifts, nchans, ndms, mstartdm, mdmstep, -1);
        if (choice.variant >= 0) {
            rl->variant = choice.variant;
            rl->row_quads = choice.row_quads;
            return AA_HIP_SUCCESS;
        }
    }
    if (kernel == AA_HIP_KERNEL_LDS || forced >= 0)
        return fail(AA_HIP_ERROR_UNSUPPORTED, "the LDS kernel does not fit this geometry (delay spread per tile "
                                              "or row alignment)");
    return AA_HIP_SUCCESS;
}

lds_params lds_params_of(const range_launch &rl) {
    lds_params p{};
    p.L = rl.L;
    p.pitch = rl.pitch;
    p.nchans = rl.nchans;
    p.dmshifts = rl.d_dmshifts;
    p.ndms = rl.ndms;
    p.t_processed = rl.t_processed;
    p.mstartdm = rl.mstartdm;
    p.mdmstep = rl.mdmstep;
    p.fnchans = (float)rl.nchans;
    p.fbin = (float)rl.inBin;
    const auto pow2 = [](long v) { return v > 0 && (v & (v - 1)) == 0; };
    p.fscale = pow2(rl.nchans) && pow2(rl.inBin) ? (float)(1.0 / ((double)rl.nchans * (double)rl.inBin)) : 0.0f;
    p.out = rl.out;
    p.out_pitch = rl.out_pitch;
    p.row_quads = rl.row_quads;
    p.pair_recs = rl.pair_recs;
    lds_job_tiles(p, rl.variant);
    return p;
}

hipError_t launch_range(const range_launch &rl, hipStream_t s) {
    if (rl.sample_bytes == 2) {
        const uint16_t *L = reinterpret_cast<const uint16_t *>(rl.L);
        if (rl.wide_win > 0)
            return launch_dedisperse_wide_u16(L, rl.pitch / 2, rl.row_len, rl.nchans, rl.d_dmshifts, rl.ndms,
                                              rl.t_processed, rl.mstartdm, rl.mdmstep, rl.inBin, rl.out, rl.out_pitch,
                                              rl.wide_win, s);
        return launch_dedisperse_direct_u16(L, rl.pitch / 2, rl.nchans, rl.d_dmshifts, rl.ndms, rl.t_processed,
                                            rl.mstartdm, rl.mdmstep, rl.inBin, rl.out, rl.out_pitch, s);
    }
    if (rl.variant >= 0) return launch_dedisperse_lds(lds_params_of(rl), rl.variant, s);
    return launch_dedisperse_direct_u8(rl.L, rl.pitch, rl.nchans, rl.d_dmshifts, rl.ndms, rl.t_processed,
                                       rl.mstartdm, rl.mdmstep, rl.inBin, rl.out, rl.out_pitch, s);
}

int ddtr_engine::create(int device, int nchans, const float *dmshifts, int nranges,
                        const float *dm_low, const float *dm_step, const int *ndms,
                        const int *inBin, double tsamp, int maxshift, int max_t_processed0,
                        ddtr_engine **out, int sample_bits) {
    if (!out) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (sample_bits != 8 && sample_bits != 16) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "sample_bits must be 8 or 16");
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
    e->m_sample_bytes = sample_bits / 8;
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
        lv.pitch = round_up(((size_t)cap + kRowPad) * (size_t)e->m_sample_bytes, kPitchAlign);
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
    for (const plan_tables &t : m_tables)
        if (t.ptr) (void)hipFree(t.ptr);
}

int ddtr_engine::set_kernel(int kernel) {
    if (kernel < AA_HIP_KERNEL_AUTO || kernel > AA_HIP_KERNEL_LDS_QUAD)
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
    if (input_bits != 8 && input_bits != 16) return fail(AA_HIP_ERROR_UNSUPPORTED, "input_bits must be 8 or 16");
    device_guard guard(m_device);
    if (m_sample_bytes == 2) {   // u16 levels: samples keep their width
        if (input_bits != 16) return fail(AA_HIP_ERROR_UNSUPPORTED, "a 16-bit plan takes 16-bit input");
        const hipError_t e = launch_transpose_u16_pitch((const uint16_t *)d_input_tc, len0, m_nchans,
                                                        (uint16_t *)m_levels[0].ptr, m_levels[0].pitch / 2, s);
        if (e != hipSuccess) return hip_fail(e, "corner turn");
        for (auto &lv : m_levels) lv.len = 0;
        m_levels[0].len = len0;
        m_resident = 0;
        return AA_HIP_SUCCESS;
    }
    // The corner turn writes the binned levels the plan needs in the same
    // pass when the shapes allow it; otherwise ensure_level bins on demand.
    const int nlev = std::min((int)m_levels.size() - 1, kCtMaxLevels);
    uint8_t *ptrs[kCtMaxLevels + 1];
    size_t pitches[kCtMaxLevels + 1];
    for (int k = 0; k <= nlev; ++k) {
        ptrs[k] = m_levels[(size_t)k].ptr;
        pitches[k] = m_levels[(size_t)k].pitch;
    }
    int fused = 0;
    const hipError_t e = launch_corner_turn_levels(d_input_tc, input_bits / 8, len0, m_nchans, ptrs, pitches, nlev,
                                                   m_d_err, s, &fused);
    if (e != hipSuccess) return hip_fail(e, "corner turn");
    for (auto &lv : m_levels) lv.len = 0;
    for (int k = 0; k <= fused; ++k) m_levels[(size_t)k].len = len0 >> k;
    m_resident = fused;
    return AA_HIP_SUCCESS;
}

int ddtr_engine::ingest_job(const void *d_input_tc, int t_processed0, ct_job *job, bool *deferred) {
    *deferred = false;
    if (!d_input_tc) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "input is NULL");
    const long len0 = (long)t_processed0 + m_maxshift;
    if (t_processed0 <= 0 || len0 > m_levels[0].capacity)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "t_processed0 exceeds the handle's capacity");
    if (m_sample_bytes != 1 || len0 > 0x7fffffffL) return AA_HIP_SUCCESS;   // ingest() handles these
    const int nlev = std::min((int)m_levels.size() - 1, kCtMaxLevels);
    ct_job j{};
    for (int k = 0; k <= nlev; ++k) {
        j.ptr[k] = m_levels[(size_t)k].ptr;
        j.pitch[k] = m_levels[(size_t)k].pitch;
        if ((size_t)(len0 >> k) > j.pitch[k]) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "level row too short");
    }
    if (!corner_turn_levels_wide(d_input_tc, m_nchans, j.ptr, j.pitch, nlev)) return AA_HIP_SUCCESS;
    j.in = static_cast<const uint8_t *>(d_input_tc);
    j.nsamp = len0;
    j.nchans = m_nchans;
    j.nlev = nlev;
    *job = j;
    for (auto &lv : m_levels) lv.len = 0;
    for (int k = 0; k <= nlev; ++k) m_levels[(size_t)k].len = len0 >> k;
    m_resident = nlev;
    *deferred = true;
    return AA_HIP_SUCCESS;
}

int ddtr_engine::take_input_overflow(hipStream_t s) {
    if (m_sample_bytes == 2) return 0;   // u16 levels hold any 16-bit value
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
        hipError_t e = m_sample_bytes == 2
                           ? launch_bin2_u16((const uint16_t *)src.ptr, src.pitch / 2, m_nchans, n_out,
                                             (uint16_t *)dst.ptr, dst.pitch / 2, s)
                           : launch_bin2_u8(src.ptr, src.pitch, m_nchans, (int)n_out, dst.ptr, dst.pitch, s);
        if (e != hipSuccess) return hip_fail(e, "binning");
        dst.len = n_out;
        ++m_resident;
    }
    return AA_HIP_SUCCESS;
}

int ddtr_engine::prepare_range(int r, int t_processed, float *d_out, size_t out_pitch, hipStream_t s,
                               range_launch *rl) {
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
    if (m_sample_bytes == 2) {   // u16 levels: dedisperse_u16, or the direct kernel when its window is too wide
        *rl = range_launch{};
        rl->L = lv.ptr;
        rl->pitch = lv.pitch;
        rl->nchans = m_nchans;
        rl->d_dmshifts = m_d_dmshifts;
        rl->ndms = rd.ndms;
        rl->t_processed = t_processed;
        rl->mstartdm = mstartdm;
        rl->mdmstep = mdmstep;
        rl->inBin = rd.inBin;
        rl->out = d_out;
        rl->out_pitch = out_pitch;
        rl->sample_bytes = 2;
        rl->row_len = i_nsamp;
        if (m_wide_win.size() != m_ranges.size()) m_wide_win.assign(m_ranges.size(), 0);
        int &w = m_wide_win[(size_t)r];
        if (w == 0) w = wide_window(m_dmshifts.data(), m_nchans, rd.ndms, mstartdm, mdmstep);
        rl->wide_win = m_kernel == AA_HIP_KERNEL_DIRECT ? -1 : w;
        if (rl->wide_win < 0 && m_kernel != AA_HIP_KERNEL_AUTO && m_kernel != AA_HIP_KERNEL_DIRECT)
            return fail(AA_HIP_ERROR_UNSUPPORTED, "the 16-bit LDS kernel does not fit this range's delay spread");
        return AA_HIP_SUCCESS;
    }
    if (m_lds.size() != m_ranges.size()) m_lds.assign(m_ranges.size(), lds_choice{-2, 0});
    lds_choice &c = m_lds[(size_t)r];
    if (c.variant == -2) {
        const int want = r < (int)m_variant_pref.size() ? m_variant_pref[(size_t)r] : -1;
        c = choose_lds(m_dmshifts.data(), m_nchans, rd.ndms, mstartdm, mdmstep, want);
        if (c.variant < 0 && want >= 0)  // preferred shape does not fit: any that does
            c = choose_lds(m_dmshifts.data(), m_nchans, rd.ndms, mstartdm, mdmstep, -1);
    }
    st = resolve_launch(m_kernel, lv.ptr, lv.pitch, m_nchans, m_d_dmshifts, m_dmshifts.data(), rd.ndms,
                        t_processed, mstartdm, mdmstep, rd.inBin, d_out, out_pitch, c, rl);
    if (st != AA_HIP_SUCCESS || rl->variant < 2) return st;
    // Chunk shapes 2 and 3 read plan tables; they depend on the range and
    // shape only, so they are built once (on the first call that needs them,
    // never while a later call is being captured into a graph).
    if (m_tables.size() != m_ranges.size()) m_tables.assign(m_ranges.size(), plan_tables{});
    plan_tables &pt = m_tables[(size_t)r];
    if (pt.variant != rl->variant) {
        if (pt.ptr) {
            (void)hipFree(pt.ptr);
            m_device_bytes -= pt.bytes;
            pt = plan_tables{};
        }
        std::vector<uint8_t> bytes;
        if (lds_plan_tables(rl->variant, m_dmshifts.data(), m_nchans, rd.ndms, mstartdm, mdmstep, lv.pitch, &bytes) !=
            rl->row_quads)
            return fail(AA_HIP_ERROR_UNSUPPORTED, "plan tables disagree with the chosen row width");
        void *t = nullptr;
        hipError_t e = hipMalloc(&t, bytes.size());
        if (e == hipSuccess) e = hipMemcpy(t, bytes.data(), bytes.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (t) (void)hipFree(t);
            return hip_fail(e, "plan tables");
        }
        pt.variant = rl->variant;
        pt.ptr = t;
        pt.bytes = bytes.size();
        m_device_bytes += bytes.size();
    }
    void *t = pt.ptr;
    rl->pair_recs = t;
    return AA_HIP_SUCCESS;
}

int ddtr_engine::set_range_variant(int r, int variant) {
    if (r < 0 || r >= (int)m_ranges.size()) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "bad range index");
    if (variant < -1 || variant >= kLdsVariants) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "bad LDS variant");
    if (m_variant_pref.size() != m_ranges.size()) m_variant_pref.assign(m_ranges.size(), -1);
    m_variant_pref[(size_t)r] = variant;
    if (m_lds.size() == m_ranges.size()) m_lds[(size_t)r] = lds_choice{-2, 0};
    return AA_HIP_SUCCESS;
}

int ddtr_engine::autotune_range(int r, int t_processed, float *d_out, size_t out_pitch, int reps, hipStream_t s,
                                int *best, float *ms) {
    if (!best || !ms || reps <= 0) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "bad autotune arguments");
    if (m_sample_bytes == 2) return fail(AA_HIP_ERROR_UNSUPPORTED, "16-bit plans have one kernel (nothing to tune)");
    if (r < 0 || r >= (int)m_ranges.size()) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "bad range index");
    device_guard guard(m_device);
    const int saved_kernel = m_kernel;
    m_kernel = AA_HIP_KERNEL_AUTO;
    *best = -1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int st = AA_HIP_SUCCESS;
    float best_ms = 0.0f;
    for (int v = 0; v < kLdsVariants && e == hipSuccess && st == AA_HIP_SUCCESS; ++v) {
        ms[v] = -1.0f;
        st = set_range_variant(r, v);
        range_launch rl;
        if (st == AA_HIP_SUCCESS) st = prepare_range(r, t_processed, d_out, out_pitch, s, &rl);
        if (st != AA_HIP_SUCCESS || rl.variant != v) continue;
        e = launch_range(rl, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch_range(rl, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) break;
        ms[v] = t / (float)reps;
        if (*best < 0 || ms[v] < best_ms) {
            *best = v;
            best_ms = ms[v];
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    m_kernel = saved_kernel;
    if (st == AA_HIP_SUCCESS && e != hipSuccess) st = hip_fail(e, "autotune");
    if (st == AA_HIP_SUCCESS) st = set_range_variant(r, *best);
    return st;
}

int ddtr_engine::run_range(int r, int t_processed, float *d_out, size_t out_pitch, hipStream_t s) {
    range_launch rl;
    const int st = prepare_range(r, t_processed, d_out, out_pitch, s, &rl);
    if (st != AA_HIP_SUCCESS) return st;
    device_guard guard(m_device);
    const hipError_t e = launch_range(rl, s);
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
