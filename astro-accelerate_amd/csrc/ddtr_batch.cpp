#include "ddtr_batch.hpp"

#include <algorithm>
#include <cmath>
#include <new>
#include <string>

#include "aa_ddtr_hip.h"

namespace aa_hip {

namespace {

int fail(int code, const std::string &msg) {
    set_last_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? AA_HIP_ERROR_OUT_OF_MEMORY : AA_HIP_ERROR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

struct device_guard {
    int prev = -1;
    explicit device_guard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
    }
    ~device_guard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

int ddtr_batch::create(int nbeams, ddtr_engine *const *beams, ddtr_batch **out) {
    if (!out) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (nbeams <= 0 || !beams) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "no beams");
    for (int b = 0; b < nbeams; ++b) {
        if (!beams[b]) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "beam " + std::to_string(b) + " is NULL");
        if (beams[b]->device() != beams[0]->device())
            return fail(AA_HIP_ERROR_UNSUPPORTED, "all beams of a batch must be on one device");
        if (beams[b]->sample_bits() != 8)
            return fail(AA_HIP_ERROR_UNSUPPORTED, "batches take 8-bit plans (beam " + std::to_string(b) + " is 16-bit)");
    }
    auto *p = new (std::nothrow) ddtr_batch();
    if (!p) return fail(AA_HIP_ERROR_OUT_OF_MEMORY, "host allocation");
    p->m_device = beams[0]->device();
    p->m_beams.assign(beams, beams + nbeams);
    for (auto *e : p->m_beams) p->m_nranges_total += e->nranges();
    device_guard guard(p->m_device);
    int ncu = 0;
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->m_device);
    // one 1024-thread work-group per CU (its LDS use fills the CU)
    p->m_grid = 8 * std::max(1, (ncu + 7) / 8);
    for (int f = 0; f < kFamilies; ++f) {
        if (e == hipSuccess) e = hipMalloc(&p->m_fam[f].d_bounds, 9 * sizeof(long));
        if (e == hipSuccess && f == 0) e = hipMalloc(&p->m_fam[f].d_counters, 8 * sizeof(int));
    }
    if (e != hipSuccess) {
        const int st = hip_fail(e, "batch create");
        delete p;
        return st;
    }
    *out = p;
    return AA_HIP_SUCCESS;
}

ddtr_batch::~ddtr_batch() {
    device_guard guard(m_device);
    release_graph();
    for (auto &f : m_fam) {
        if (f.d_jobs) (void)hipFree(f.d_jobs);
        if (f.d_bounds) (void)hipFree(f.d_bounds);
        if (f.d_counters) (void)hipFree(f.d_counters);
    }
    for (auto &f : m_ct)
        if (f.d_jobs) (void)hipFree(f.d_jobs);
    if (m_capture_stream) (void)hipStreamDestroy(m_capture_stream);
}

int ddtr_batch::family_of(int variant, int row_quads) {
    if (variant == 0 || variant == 1) return 0;
    if (variant == 6) {
        const int ni = std::max(2, (row_quads + 63) / 64);   // as launch_dedisperse_wide2 picks
        return row_quads <= 0 || ni > 5 ? -1 : 13 + (ni - 2);
    }
    const int J = variant == 4 ? 4 : variant == 5 ? 3 : variant == 2 ? 2 : 0;
    const int ni = std::max(J, (row_quads + 63) / 64);   // the instantiation launch_dedisperse_wide picks
    if (J == 0 || row_quads <= 0 || ni > 5) return -1;
    return 1 + 4 * (J - 2) + (ni - 2);
}

void ddtr_batch::release_graph() {
    if (m_exec) (void)hipGraphExecDestroy(m_exec);
    if (m_graph) (void)hipGraphDestroy(m_graph);
    m_exec = nullptr;
    m_graph = nullptr;
    m_key.clear();
}

// Launches the corner turns of every beam (batched where the wide turn
// applies), then resolves every range (binning deeper levels as needed) into
// the families' job lists (m_fam[*].jobs) and the separate launches
// (m_direct); the caller uploads the family job lists (outside any graph
// capture) before the persistent kernels run.
int ddtr_batch::enqueue(const void *const *inputs, const int *t_processed, float *const *d_out,
                        const size_t *out_pitch, hipStream_t s, bool upload) {
    std::vector<lds_job> jobs[kFamilies];
    std::vector<range_launch> direct;
    for (auto &f : m_ct) f.jobs.clear();
    m_ct_direct = 0;
    int flat = 0;
    for (size_t b = 0; b < m_beams.size(); ++b) {
        ddtr_engine *e = m_beams[b];
        ct_job j{};
        bool deferred = false;
        int st = e->ingest_job(inputs[b], t_processed[flat], &j, &deferred);
        if (st == AA_HIP_SUCCESS && !deferred) {
            st = e->ingest(inputs[b], 8, t_processed[flat], s);
            ++m_ct_direct;
        }
        if (st != AA_HIP_SUCCESS) return st;
        if (deferred) {
            ct_family &F = m_ct[j.nlev];
            j.tile_end = (F.jobs.empty() ? 0 : F.jobs.back().tile_end) + corner_turn_tiles(j.nsamp, j.nchans);
            F.jobs.push_back(j);
        }
        flat += e->nranges();
    }
    for (auto &F : m_ct) {
        if (F.jobs.empty()) continue;
        if (F.jobs.back().tile_end > 0x7fffffffL) return fail(AA_HIP_ERROR_UNSUPPORTED, "corner turns exceed 2^31 tiles");
        hipError_t e = hipSuccess;
        if (!upload && (!F.d_jobs || F.jobs.size() > F.cap)) return fail(AA_HIP_ERROR_DEVICE, "batch changed in capture");
        if (upload) {
            if (F.jobs.size() > F.cap) {
                if (F.d_jobs) (void)hipFree(F.d_jobs);
                F.d_jobs = nullptr;
                F.cap = 0;
                e = hipMalloc(&F.d_jobs, F.jobs.size() * sizeof(ct_job));
                if (e != hipSuccess) return hip_fail(e, "corner-turn job table");
                F.cap = F.jobs.size();
            }
            // from pageable memory: wait for it before F.jobs can change
            e = hipMemcpyAsync(F.d_jobs, F.jobs.data(), F.jobs.size() * sizeof(ct_job), hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        }
        if (e == hipSuccess)
            e = launch_corner_turn_batch(F.d_jobs, (int)F.jobs.size(), F.jobs.back().tile_end, F.jobs.back().nlev, s);
        if (e != hipSuccess) return hip_fail(e, "batched corner turn");
    }
    flat = 0;
    for (size_t b = 0; b < m_beams.size(); ++b) {
        ddtr_engine *e = m_beams[b];
        int st = AA_HIP_SUCCESS;
        for (int r = 0; r < e->nranges(); ++r, ++flat) {
            range_launch rl;
            st = e->prepare_range(r, t_processed[flat], d_out[flat], out_pitch[flat], s, &rl);
            if (st != AA_HIP_SUCCESS) return st;
            if (rl.ndms <= 0 || rl.t_processed <= 0) continue;
            const int f = rl.sample_bytes == 1 ? family_of(rl.variant, rl.row_quads) : -1;
            if (f >= 0) {
                lds_job j{};
                j.p = lds_params_of(rl);
                j.variant = rl.variant;
                lds_job_tiles(j.p, rl.variant);
                jobs[f].push_back(j);
            } else {
                direct.push_back(rl);
            }
        }
    }
    for (int f = 0; f < kFamilies; ++f) {
        // Largest tiles (most channels) first: the costliest work starts
        // early and the queue tails are made of the cheapest tiles.  Queue x
        // holds the x-th eighth of every job (lds_job).
        std::stable_sort(jobs[f].begin(), jobs[f].end(),
                         [](const lds_job &a, const lds_job &b) { return a.p.nchans > b.p.nchans; });
        long qlen[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        long total = 0;
        for (auto &j : jobs[f]) {
            const long n = (long)j.p.n_dm_tiles * j.p.n_t_tiles;
            for (int x = 0; x < 8; ++x) {
                qlen[x] += (x + 1) * n / 8 - x * n / 8;
                j.qend[x] = qlen[x];
            }
            total += n;
        }
        if (total >= 0x7fffffffL) return fail(AA_HIP_ERROR_UNSUPPORTED, "batch holds more than 2^31 tiles");
        m_fam[f].jobs.swap(jobs[f]);
        m_fam[f].ntiles = total;
        m_fam[f].bounds.assign(9, 0);
        for (int x = 0; x < 8; ++x) m_fam[f].bounds[(size_t)x + 1] = m_fam[f].bounds[(size_t)x] + qlen[x];
    }
    m_direct.swap(direct);
    return AA_HIP_SUCCESS;
}

// The persistent byte-row launch and the pair families' launches (those
// with jobs), then the separate launches.  Reads only device-side job tables
// (the host copies of the queue bounds size the grids): capturable.
hipError_t ddtr_batch::launch_families(hipStream_t s) {
    hipError_t e = hipSuccess;
    if (!m_fam[0].jobs.empty())
        e = launch_dedisperse_lds_batch(m_fam[0].d_jobs, (int)m_fam[0].jobs.size(), m_fam[0].d_bounds,
                                        m_fam[0].d_counters, m_grid, s);
    for (int f = 1; e == hipSuccess && f < kFamilies; ++f) {
        if (m_fam[f].jobs.empty()) continue;
        long longest = 0;
        for (int x = 0; x < 8; ++x) longest = std::max(longest, m_fam[f].bounds[(size_t)x + 1] - m_fam[f].bounds[(size_t)x]);
        const int njobs = (int)m_fam[f].jobs.size();
        if (f >= 13) {   // shape 6
            e = launch_dedisperse_wide2_batch(m_fam[f].d_jobs, njobs, m_fam[f].d_bounds, longest, 2 + (f - 13), s);
            continue;
        }
        const int J = 2 + (f - 1) / 4, ni = 2 + (f - 1) % 4;
        e = launch_dedisperse_wide_batch(m_fam[f].d_jobs, njobs, m_fam[f].d_bounds, longest, J, ni, s);
    }
    for (size_t i = 0; e == hipSuccess && i < m_direct.size(); ++i) e = launch_range(m_direct[i], s);
    return e;
}

int ddtr_batch::execute(const void *const *inputs, const int *t_processed, float *const *d_out,
                        const size_t *out_pitch, bool use_graph, hipStream_t s) {
    if (!inputs || !t_processed || !d_out || !out_pitch) return fail(AA_HIP_ERROR_INVALID_ARGUMENT, "NULL array");
    device_guard guard(m_device);
    std::vector<uintptr_t> key;
    key.reserve(m_beams.size() + 3 * (size_t)m_nranges_total + 1);
    key.push_back((uintptr_t)s);
    for (size_t b = 0; b < m_beams.size(); ++b) key.push_back((uintptr_t)inputs[b]);
    for (int i = 0; i < m_nranges_total; ++i) {
        key.push_back((uintptr_t)d_out[i]);
        key.push_back((uintptr_t)out_pitch[i]);
        key.push_back((uintptr_t)(unsigned)t_processed[i]);
    }
    if (use_graph && m_exec && key == m_key) {
        const hipError_t e = hipGraphLaunch(m_exec, s);
        return e == hipSuccess ? AA_HIP_SUCCESS : hip_fail(e, "batch graph launch");
    }
    release_graph();  // the eager pass rewrites the job tables a captured graph reads

    // Eager pass: builds and uploads the job lists, and computes this call.
    int st = enqueue(inputs, t_processed, d_out, out_pitch, s, true);
    if (st != AA_HIP_SUCCESS) return st;
    hipError_t e = hipSuccess;
    for (int f = 0; f < kFamilies && e == hipSuccess; ++f) {
        family &F = m_fam[f];
        if (F.jobs.empty()) continue;
        if (F.jobs.size() > F.jobs_cap) {
            if (F.d_jobs) (void)hipFree(F.d_jobs);
            F.d_jobs = nullptr;
            F.jobs_cap = 0;
            e = hipMalloc(&F.d_jobs, F.jobs.size() * sizeof(lds_job));
            if (e != hipSuccess) return hip_fail(e, "batch job table");
            F.jobs_cap = F.jobs.size();
        }
        e = hipMemcpyAsync(F.d_jobs, F.jobs.data(), F.jobs.size() * sizeof(lds_job), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(F.d_bounds, F.bounds.data(), 9 * sizeof(long), hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // host vectors may change before the copies run
    if (e == hipSuccess) e = launch_families(s);
    if (e != hipSuccess) return hip_fail(e, "batch launch");
    if (!use_graph) return AA_HIP_SUCCESS;

    // Capture the same sequence (the job tables are already on the device
    // and are the same for these buffers); later calls replay it.
    // The legacy default stream cannot be captured: capture on a stream of
    // the batch's own (the graph is launched on the caller's stream).
    hipStream_t cs = s;
    if (!cs) {
        if (!m_capture_stream) {
            e = hipStreamCreateWithFlags(&m_capture_stream, hipStreamNonBlocking);
            if (e != hipSuccess) return hip_fail(e, "batch capture stream");
        }
        e = hipStreamSynchronize(s);  // the eager pass above ran on s
        if (e != hipSuccess) return hip_fail(e, "batch capture");
        cs = m_capture_stream;
    }
    size_t nj[kFamilies], nct[kCtMaxLevels + 1];
    for (int f = 0; f < kFamilies; ++f) nj[f] = m_fam[f].jobs.size();
    for (int k = 0; k <= kCtMaxLevels; ++k) nct[k] = m_ct[k].jobs.size();
    const size_t nd = m_direct.size();
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return hip_fail(e, "batch capture");
    st = enqueue(inputs, t_processed, d_out, out_pitch, cs, false);
    bool same = m_direct.size() == nd;
    for (int f = 0; f < kFamilies; ++f) same = same && m_fam[f].jobs.size() == nj[f];
    for (int k = 0; k <= kCtMaxLevels; ++k) same = same && m_ct[k].jobs.size() == nct[k];
    if (st == AA_HIP_SUCCESS && !same) st = fail(AA_HIP_ERROR_DEVICE, "batch changed in capture");
    if (st == AA_HIP_SUCCESS) e = launch_families(cs);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    if (st != AA_HIP_SUCCESS) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e == hipSuccess) e = ec;
    if (e == hipSuccess) e = hipGraphInstantiate(&m_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        m_exec = nullptr;
        return hip_fail(e, "batch graph");
    }
    m_graph = g;
    m_key.swap(key);
    return AA_HIP_SUCCESS;
}

}  // namespace aa_hip
