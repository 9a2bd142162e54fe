// extern "C" surface of the fractional-delay (differentiable) dedispersion:
// a plan holding the delay table and trial offsets on the device, forward
// and adjoint calls on caller-owned device buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "aa_ddtr_hip.h"
#include "ddtr_engine.hpp"
#include "frac_ddtr.h"
#include "frac_refine.h"

struct aa_hip_frac {
    int device = 0;
    int nchans = 0, ntrials = 0, nsamp = 0, max_batch = 0;
    int trials_per_wave = 0;     // forward / DM adjoint grouping
    int chans_per_wave = 0;      // input adjoint grouping
    float* d_ds = nullptr;
    float* d_trials = nullptr;
    double* d_part = nullptr;
    size_t part_entries = 0;
    size_t bytes = 0;
};

struct aa_hip_refine {
    int device = 0;
    int ncand = 0;
    int grid = 0;
    int loss_kind = 0;
    aa_hip::refine_cand* d_cands = nullptr;
    int* d_order = nullptr;
    int* d_counter = nullptr;
    double* d_power = nullptr;   // channel powers of every candidate (loss 1)
};

namespace {

int fail(int status, const std::string& msg) {
    aa_hip::set_last_error(msg);
    return status;
}

int from_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return AA_HIP_SUCCESS;
    aa_hip::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? AA_HIP_ERROR_OUT_OF_MEMORY : AA_HIP_ERROR_DEVICE;
}

void release(aa_hip_frac* p) {
    if (!p) return;
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != p->device) (void)hipSetDevice(p->device);
    if (p->d_ds) (void)hipFree(p->d_ds);
    if (p->d_trials) (void)hipFree(p->d_trials);
    if (p->d_part) (void)hipFree(p->d_part);
    if (cur != p->device) (void)hipSetDevice(cur);
    delete p;
}

// Channel groups (16 * q channels) for the input adjoint: the largest q whose
// delay spread over a group stays inside a staged row for |trial + dm| up to
// u_absmax.  Outside that, the kernel reads the window from global memory.
int chans_per_wave(const std::vector<float>& ds, float u_absmax) {
    const int C = (int)ds.size();
    for (int q = aa_hip::kFracMaxPerWave; q >= 1; q >>= 1) {
        bool ok = true;
        for (int c0 = 0; c0 < C && ok; c0 += 16 * q) {
            float lo = ds[(size_t)c0], hi = lo;
            for (int c = c0; c < std::min(C, c0 + 16 * q); ++c) {
                lo = std::min(lo, ds[(size_t)c]);
                hi = std::max(hi, ds[(size_t)c]);
            }
            ok = ((double)hi - lo) * u_absmax + 8.0 <= (double)aa_hip::kFracMaxOffset;
        }
        if (ok) return q;
    }
    return 1;
}

aa_hip::frac_args base_args(const aa_hip_frac* p, int nbatch, const float* d_dm) {
    aa_hip::frac_args a{};
    a.ds = p->d_ds;
    a.trials = p->d_trials;
    a.dm = d_dm;
    a.B = nbatch;
    a.C = p->nchans;
    a.T = p->nsamp;
    a.K = p->ntrials;
    a.n_ttiles = (p->nsamp + aa_hip::kFracTile - 1) / aa_hip::kFracTile;
    return a;
}

bool aligned16(const void* ptr) { return ((uintptr_t)ptr & 15u) == 0; }

}  // namespace

extern "C" {

int aa_hip_frac_create(int device, int nchans, const float* delay_per_dm, int ntrials, const float* trials,
                       int nsamp, int max_batch, float dm_absmax, aa_hip_frac** out) {
    if (!out) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL output handle");
    *out = nullptr;
    if (!delay_per_dm || !trials) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL table");
    if (nchans <= 0 || ntrials <= 0 || nsamp <= 0 || max_batch <= 0 || !(dm_absmax >= 0.0f))
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "invalid geometry");
    if (nsamp % 4 != 0) return fail(AA_HIP_ERROR_UNSUPPORTED, "nsamp must be a multiple of 4");
    std::vector<float> ds(delay_per_dm, delay_per_dm + nchans);
    std::vector<float> tr(trials, trials + ntrials);
    float ds_absmax = 0.0f, tr_absmax = 0.0f;
    for (float d : ds) {
        if (!std::isfinite(d)) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "non-finite delay table");
        ds_absmax = std::max(ds_absmax, std::fabs(d));
    }
    for (float t : tr) {
        if (!std::isfinite(t)) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "non-finite trial offset");
        tr_absmax = std::max(tr_absmax, std::fabs(t));
    }
    // every delay must stay far inside int range: |tau| < 2^24 samples
    if ((double)ds_absmax * ((double)tr_absmax + dm_absmax) >= 16777216.0)
        return fail(AA_HIP_ERROR_UNSUPPORTED, "delays beyond 2^24 samples");
    const int tq = aa_hip::frac_trials_per_wave(tr.data(), ntrials, ds_absmax);
    if (tq == 0)
        return fail(AA_HIP_ERROR_UNSUPPORTED,
                    "trial offsets too far apart: 16 consecutive trials must span at most " +
                        std::to_string(aa_hip::kFracMaxOffset - 8) + " samples of delay");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(AA_HIP_ERROR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "device out of range");

    auto* p = new (std::nothrow) aa_hip_frac();
    if (!p) return fail(AA_HIP_ERROR_OUT_OF_MEMORY, "host allocation");
    p->device = device;
    p->nchans = nchans;
    p->ntrials = ntrials;
    p->nsamp = nsamp;
    p->max_batch = max_batch;
    p->trials_per_wave = tq;
    p->chans_per_wave = chans_per_wave(ds, tr_absmax + dm_absmax);
    int cur = 0;
    (void)hipGetDevice(&cur);
    hipError_t e = hipSetDevice(device);
    const int ttiles = (nsamp + aa_hip::kFracTile - 1) / aa_hip::kFracTile;
    const int groups = (ntrials + 16 * tq - 1) / (16 * tq);
    p->part_entries = (size_t)max_batch * ttiles * groups;
    if (e == hipSuccess) e = hipMalloc(&p->d_ds, sizeof(float) * nchans);
    if (e == hipSuccess) e = hipMalloc(&p->d_trials, sizeof(float) * ntrials);
    if (e == hipSuccess) e = hipMalloc(&p->d_part, sizeof(double) * p->part_entries);
    if (e == hipSuccess) e = hipMemcpy(p->d_ds, ds.data(), sizeof(float) * nchans, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_trials, tr.data(), sizeof(float) * ntrials, hipMemcpyHostToDevice);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) {
        const int st = from_hip(e, "aa_hip_frac_create");
        release(p);
        return st;
    }
    p->bytes = sizeof(float) * (nchans + ntrials) + sizeof(double) * p->part_entries;
    *out = p;
    return AA_HIP_SUCCESS;
}

int aa_hip_frac_destroy(aa_hip_frac* h) {
    release(h);
    return AA_HIP_SUCCESS;
}

int aa_hip_frac_grouping(const aa_hip_frac* h, int* trials_per_wave, int* chans_per_wave) {
    if (!h) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL handle");
    if (trials_per_wave) *trials_per_wave = h->trials_per_wave;
    if (chans_per_wave) *chans_per_wave = h->chans_per_wave;
    return AA_HIP_SUCCESS;
}

int aa_hip_frac_forward(aa_hip_frac* h, const float* d_x, int nbatch, const float* d_dm, float* d_y,
                        void* stream) {
    if (!h || !d_x || !d_dm || !d_y) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (nbatch < 0 || nbatch > h->max_batch) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "batch exceeds the plan's");
    if (!aligned16(d_x) || !aligned16(d_y)) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "buffers must be 16-byte aligned");
    if (nbatch == 0) return AA_HIP_SUCCESS;
    aa_hip::frac_args a = base_args(h, nbatch, d_dm);
    a.x = d_x;
    a.y = d_y;
    a.group_q = h->trials_per_wave;
    a.n_groups = (h->ntrials + 16 * a.group_q - 1) / (16 * a.group_q);
    return from_hip(aa_hip::launch_frac_forward(a, (hipStream_t)stream), "frac forward");
}

int aa_hip_frac_adjoint(aa_hip_frac* h, const float* d_x, const float* d_gy, int nbatch, const float* d_dm,
                        float* d_gx, double* d_gdm, void* stream) {
    if (!h || !d_gy || !d_dm) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (d_gdm && !d_x) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "the DM adjoint needs the input");
    if (nbatch < 0 || nbatch > h->max_batch) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "batch exceeds the plan's");
    if ((d_x && !aligned16(d_x)) || !aligned16(d_gy) || (d_gx && !aligned16(d_gx)))
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "buffers must be 16-byte aligned");
    if (nbatch == 0) return AA_HIP_SUCCESS;
    const hipStream_t s = (hipStream_t)stream;
    if (d_gx) {
        aa_hip::frac_args a = base_args(h, nbatch, d_dm);
        a.gy = d_gy;
        a.gx = d_gx;
        a.group_q = h->chans_per_wave;
        a.n_groups = (h->nchans + 16 * a.group_q - 1) / (16 * a.group_q);
        const int st = from_hip(aa_hip::launch_frac_grad_x(a, s), "frac input adjoint");
        if (st != AA_HIP_SUCCESS) return st;
    }
    if (d_gdm) {
        aa_hip::frac_args a = base_args(h, nbatch, d_dm);
        a.x = d_x;
        a.gy = d_gy;
        a.gdm = d_gdm;
        a.gdm_part = h->d_part;
        a.group_q = h->trials_per_wave;
        a.n_groups = (h->ntrials + 16 * a.group_q - 1) / (16 * a.group_q);
        if (aa_hip::frac_grad_dm_parts(a) > h->part_entries)
            return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "workspace too small for the batch");
        const int st = from_hip(aa_hip::launch_frac_grad_dm(a, s), "frac DM adjoint");
        if (st != AA_HIP_SUCCESS) return st;
    }
    return AA_HIP_SUCCESS;
}

int aa_hip_refine_create(int device, int ncand, const int64_t* x_offsets, const int64_t* ds_offsets,
                         const int* nchans, const int* nsamp, aa_hip_refine** out) {
    if (!out) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL output handle");
    *out = nullptr;
    if (ncand <= 0 || !x_offsets || !ds_offsets || !nchans || !nsamp)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "invalid candidate table");
    std::vector<aa_hip::refine_cand> cands((size_t)ncand);
    int64_t power_entries = 0;
    for (int b = 0; b < ncand; ++b) {
        if (nchans[b] <= 0 || nchans[b] > aa_hip::kRefineMaxChannels || nsamp[b] <= 0 ||
            nsamp[b] > aa_hip::kRefineMaxSamples || x_offsets[b] < 0 || ds_offsets[b] < 0)
            return fail(AA_HIP_ERROR_UNSUPPORTED, "candidate " + std::to_string(b) +
                                                      ": windows up to " + std::to_string(aa_hip::kRefineMaxChannels) +
                                                      " channels x " + std::to_string(aa_hip::kRefineMaxSamples) +
                                                      " samples, non-negative offsets");
        cands[(size_t)b] = aa_hip::refine_cand{x_offsets[b], ds_offsets[b], power_entries, nchans[b], nsamp[b]};
        power_entries += nchans[b];
    }
    // largest windows first (longest work items start early)
    std::vector<int> order((size_t)ncand);
    for (int b = 0; b < ncand; ++b) order[(size_t)b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int i, int j) {
        return (long)cands[(size_t)i].C * cands[(size_t)i].T > (long)cands[(size_t)j].C * cands[(size_t)j].T;
    });
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(AA_HIP_ERROR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "device out of range");
    auto* p = new (std::nothrow) aa_hip_refine();
    if (!p) return fail(AA_HIP_ERROR_OUT_OF_MEMORY, "host allocation");
    p->device = device;
    p->ncand = ncand;
    int cur = 0;
    (void)hipGetDevice(&cur);
    hipError_t e = hipSetDevice(device);
    int ncu = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    p->grid = std::max(1, ncu) * aa_hip::refine_blocks_per_cu();
    if (e == hipSuccess) e = hipMalloc(&p->d_cands, sizeof(aa_hip::refine_cand) * (size_t)ncand);
    if (e == hipSuccess) e = hipMalloc(&p->d_order, sizeof(int) * (size_t)ncand);
    if (e == hipSuccess) e = hipMalloc(&p->d_counter, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_power, sizeof(double) * (size_t)power_entries);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_cands, cands.data(), sizeof(aa_hip::refine_cand) * (size_t)ncand, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_order, order.data(), sizeof(int) * (size_t)ncand, hipMemcpyHostToDevice);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) {
        const int st = from_hip(e, "aa_hip_refine_create");
        aa_hip_refine_destroy(p);
        return st;
    }
    *out = p;
    return AA_HIP_SUCCESS;
}

int aa_hip_refine_destroy(aa_hip_refine* h) {
    if (!h) return AA_HIP_SUCCESS;
    if (h->d_cands) (void)hipFree(h->d_cands);
    if (h->d_order) (void)hipFree(h->d_order);
    if (h->d_counter) (void)hipFree(h->d_counter);
    if (h->d_power) (void)hipFree(h->d_power);
    delete h;
    return AA_HIP_SUCCESS;
}

int aa_hip_refine_set_loss(aa_hip_refine* h, int loss_kind) {
    if (!h) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL handle");
    if (loss_kind != AA_HIP_REFINE_LOSS_ENERGY && loss_kind != AA_HIP_REFINE_LOSS_EXCESS_ENERGY)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "unknown refinement loss");
    h->loss_kind = loss_kind;
    return AA_HIP_SUCCESS;
}

int aa_hip_refine_run(aa_hip_refine* h, const float* d_x, const float* d_delay_per_dm, float* d_u, float* d_step,
                      double* d_prev_grad, double* d_loss, double* d_grad, int scan_n, float scan_lo,
                      float scan_step, int nsteps, float step_max, float step_min, void* stream) {
    if (!h || !d_x || !d_delay_per_dm || !d_u || !d_step || !d_prev_grad || !d_loss || !d_grad)
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (nsteps < 0 || !(step_max > 0.0f) || !(step_min > 0.0f) || step_min > step_max || scan_n < 0 ||
        !std::isfinite(scan_lo) || !std::isfinite(scan_step))
        return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "invalid step settings");
    aa_hip::refine_args a{};
    a.x = d_x;
    a.ds = d_delay_per_dm;
    a.cands = h->d_cands;
    a.order = h->d_order;
    a.ncand = h->ncand;
    a.u = d_u;
    a.step = d_step;
    a.prev_grad = d_prev_grad;
    a.loss = d_loss;
    a.grad = d_grad;
    a.nsteps = nsteps;
    a.step_max = step_max;
    a.step_min = step_min;
    a.scan_n = scan_n;
    a.scan_lo = scan_lo;
    a.scan_step = scan_step;
    a.counter = h->d_counter;
    a.loss_kind = h->loss_kind;
    a.power = h->d_power;
    return from_hip(aa_hip::launch_refine(a, h->grid, (hipStream_t)stream), "refine");
}

}  // extern "C"
