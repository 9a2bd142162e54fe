// SIGPROC filterbank reader (see include/aa_sigproc_input.hpp for the format
// rules it follows) and its C ABI (include/aa_ddtr_hip.h, aa_hip_sigproc_*).
#include "aa_sigproc_input.hpp"

#include <hip/hip_runtime.h>

#include <unistd.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>

#include "aa_ddtr_hip.h"
#include "ddtr_engine.hpp"

namespace astroaccelerate {

aa_sigproc_input::aa_sigproc_input(const std::string &path) : m_path(path) {}

aa_sigproc_input::~aa_sigproc_input()
{
    close();
    free_u8();
}

bool aa_sigproc_input::open()
{
    if (m_fp) return true;
    m_fp = std::fopen(m_path.c_str(), "rb");
    return m_fp != nullptr;
}

bool aa_sigproc_input::close()
{
    if (!m_fp) return true;
    const bool ok = std::fclose(m_fp) == 0;
    m_fp = nullptr;
    return ok;
}

void aa_sigproc_input::free_u8()
{
    if (m_u8) {
        if (m_u8_pinned) (void)hipHostFree(m_u8);
        else std::free(m_u8);
    }
    m_u8 = nullptr;
    m_u8_size = 0;
    m_u8_pinned = false;
}

aa_filterbank_metadata aa_sigproc_input::read_metadata()
{
    if (!open()) return aa_filterbank_metadata();
    aa_filterbank_metadata m;
    m_header_ok = read_header(m);
    m_header_is_read = true;
    return m_header_ok ? m : aa_filterbank_metadata();
}

// Keyword records up to HEADER_END (src/aa_sigproc_input.cpp:85-339).
bool aa_sigproc_input::read_header(aa_filterbank_metadata &metadata)
{
    static const char *const kInts[] = {"telescope_id", "machine_id", "data_type", "barycentric",
                                        "pulsarcentric", "nbits", "nsamples", "nchans", "nifs"};
    static const char *const kDoubles[] = {"az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp",
                                           "fch1", "foff", "fchannel", "refdm", "period"};
    int iv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dv[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (std::fseek(m_fp, 0, SEEK_SET) != 0) return false;
    for (;;) {
        int32_t n = 0;
        if (std::fread(&n, sizeof(n), 1, m_fp) != 1) return false;
        if (n <= 1 || n >= 80) continue;
        char key[80];
        if (std::fread(key, (size_t)n, 1, m_fp) != 1) return false;
        key[n] = '\0';
        if (std::strcmp(key, "HEADER_END") == 0) break;
        for (int i = 0; i < 9; ++i)
            if (std::strcmp(key, kInts[i]) == 0 && std::fread(&iv[i], sizeof(int), 1, m_fp) != 1) return false;
        for (int i = 0; i < 11; ++i)
            if (std::strcmp(key, kDoubles[i]) == 0 && std::fread(&dv[i], sizeof(double), 1, m_fp) != 1) return false;
    }
    const int nbits = iv[5], nchans = iv[7], nifs = iv[8];
    if (nifs != 1) {
        aa_hip::set_last_error("SIGPROC input: only one IF channel (nifs = 1) is supported");
        return false;
    }
    if (nchans <= 0 || (nbits != 1 && nbits != 2 && nbits != 4 && nbits != 8 && nbits != 16 && nbits != 32) ||
        ((long)nchans * nbits) % 8 != 0) {
        aa_hip::set_last_error("SIGPROC input: unsupported nbits / nchans");
        return false;
    }
    m_data_start = std::ftell(m_fp);
    if (m_data_start < 0 || std::fseek(m_fp, 0, SEEK_END) != 0) return false;
    const long end = std::ftell(m_fp);
    if (end < m_data_start || std::fseek(m_fp, m_data_start, SEEK_SET) != 0) return false;
    const long record = (long)nchans * nbits / 8;
    const long nsamp = (end - m_data_start) / record;
    m_meta = aa_filterbank_metadata(iv[0], iv[1], iv[2], "unknown", "unknown", iv[3], iv[4], dv[0], dv[1], dv[2],
                                    dv[3], dv[4], dv[5], nbits, (int)nsamp, dv[6], dv[7], 0, dv[8], 0, nchans, nifs,
                                    dv[9], dv[10]);
    metadata = m_meta;
    return true;
}

// `bytes` bytes from the data start into dst, in up to 8 slices read by
// their own threads (pread on the file's descriptor): one thread moves page
// cache to memory at a few GB/s.
bool aa_sigproc_input::read_parallel(uint8_t *dst, size_t bytes)
{
    const int fd = fileno(m_fp);
    const size_t kMinSlice = (size_t)64 << 20;
    const int nthreads = (int)std::max<size_t>(1, std::min<size_t>(8, bytes / kMinSlice));
    std::vector<std::thread> pool;
    std::vector<char> ok((size_t)nthreads, 0);
    for (int i = 0; i < nthreads; ++i) {
        const size_t a = bytes * (size_t)i / (size_t)nthreads, b = bytes * (size_t)(i + 1) / (size_t)nthreads;
        pool.emplace_back([=, &ok]() {
            size_t done = a;
            while (done < b) {
                const ssize_t n = pread(fd, dst + done, b - done, (off_t)(m_data_start + (long)done));
                if (n <= 0) return;
                done += (size_t)n;
            }
            ok[(size_t)i] = 1;
        });
    }
    for (auto &t : pool) t.join();
    return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

// Whole records from the data start into u8 (nbits <= 8) or u16 (any nbits),
// a block of records at a time (src/aa_sigproc_input.cpp:341-541).
bool aa_sigproc_input::read_records(uint8_t *u8, unsigned short *u16)
{
    const int nbits = m_meta.nbits();
    const size_t nchans = (size_t)m_meta.nchans();
    const size_t nsamp = (size_t)m_meta.nsamples();
    const size_t record = nchans * (size_t)nbits / 8;
    if (std::fseek(m_fp, m_data_start, SEEK_SET) != 0) return false;
    if (nbits == 8 && u8) return read_parallel(u8, record * nsamp);
    float fmin = 0.0f, bin = 1.0f;
    if (nbits == 32) {   // value range first (src/aa_sigproc_input.cpp:362-378)
        float mx = -10000000000.0f, mn = 10000000000.0f;
        std::vector<float> rec(nchans);
        for (size_t t = 0; t < nsamp; ++t) {
            if (std::fread(rec.data(), sizeof(float), nchans, m_fp) != nchans) return false;
            for (float x : rec) {
                mx = x > mx ? x : mx;
                mn = x < mn ? x : mn;
            }
        }
        fmin = mn;
        bin = (mx - mn) / 65535.0f;
        if (std::fseek(m_fp, m_data_start, SEEK_SET) != 0) return false;
    }
    const size_t block = std::max<size_t>(1, ((size_t)64 << 20) / record);
    std::vector<uint8_t> buf(block * record);
    for (size_t t0 = 0; t0 < nsamp; t0 += block) {
        const size_t nt = std::min(block, nsamp - t0);
        if (std::fread(buf.data(), record, nt, m_fp) != nt) return false;
        const size_t nv = nt * nchans;
        const size_t o = t0 * nchans;
        auto put = [&](size_t i, unsigned v) {
            if (u8) u8[o + i] = (uint8_t)v;
            else u16[o + i] = (unsigned short)v;
        };
        const uint8_t *b = buf.data();
        switch (nbits) {
        case 1:   // least significant bit first
            for (size_t i = 0; i < nv; ++i) put(i, (b[i >> 3] >> (i & 7)) & 1u);
            break;
        case 2:   // most significant pair first
            for (size_t i = 0; i < nv; ++i) put(i, (b[i >> 2] >> (6 - 2 * (i & 3))) & 3u);
            break;
        case 4:   // high nibble first
            for (size_t i = 0; i < nv; ++i) put(i, (b[i >> 1] >> ((i & 1) ? 0 : 4)) & 15u);
            break;
        case 8:
            for (size_t i = 0; i < nv; ++i) put(i, b[i]);
            break;
        case 16: {
            const unsigned short *w = reinterpret_cast<const unsigned short *>(b);
            for (size_t i = 0; i < nv; ++i) u16[o + i] = w[i];
            break;
        }
        default: {   // 32-bit float, scaled
            const float *f = reinterpret_cast<const float *>(b);
            for (size_t i = 0; i < nv; ++i) u16[o + i] = (unsigned short)((f[i] - fmin) / bin);
            break;
        }
        }
    }
    return true;
}

bool aa_sigproc_input::read_signal()
{
    if (!m_fp || !m_header_is_read || !m_header_ok || m_data_is_read) return false;
    m_input_buffer.assign((size_t)m_meta.nsamples() * (size_t)m_meta.nchans(), 0);
    if (!read_records(nullptr, m_input_buffer.data())) return false;
    m_data_is_read = true;
    return true;
}

bool aa_sigproc_input::read_signal_u8(bool pinned)
{
    if (!m_fp || !m_header_is_read || !m_header_ok || m_data_is_read) return false;
    if (m_meta.nbits() > 8) {
        aa_hip::set_last_error("SIGPROC input: the u8 path holds nbits 1, 2, 4 or 8");
        return false;
    }
    free_u8();
    const size_t n = (size_t)m_meta.nsamples() * (size_t)m_meta.nchans();
    void *p = nullptr;
    if (pinned) {
        if (hipHostMalloc(&p, std::max<size_t>(n, 1), hipHostMallocDefault) != hipSuccess) {
            aa_hip::set_last_error("SIGPROC input: pinned host allocation failed");
            return false;
        }
    } else {
        p = std::malloc(std::max<size_t>(n, 1));
        if (!p) {
            aa_hip::set_last_error("SIGPROC input: host allocation failed");
            return false;
        }
    }
    m_u8 = static_cast<uint8_t *>(p);
    m_u8_size = n;
    m_u8_pinned = pinned;
    if (!read_records(m_u8, nullptr)) {
        free_u8();
        aa_hip::set_last_error("SIGPROC input: short read");
        return false;
    }
    m_data_is_read = true;
    return true;
}

}  // namespace astroaccelerate

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct aa_hip_sigproc {
    astroaccelerate::aa_sigproc_input in;
    explicit aa_hip_sigproc(const char *path) : in(path) {}
};

extern "C" {

int aa_hip_sigproc_open(const char *path, aa_hip_sigproc **out, aa_hip_filterbank_header *hdr)
{
    if (!path || !out) {
        aa_hip::set_last_error("aa_hip_sigproc_open: null argument");
        return AA_HIP_ERROR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    aa_hip_sigproc *h = new aa_hip_sigproc(path);
    if (!h->in.open()) {
        aa_hip::set_last_error(std::string("aa_hip_sigproc_open: cannot open ") + path);
        delete h;
        return AA_HIP_ERROR_INVALID_ARGUMENT;
    }
    const astroaccelerate::aa_filterbank_metadata m = h->in.read_metadata();
    if (!h->in.header_ok()) {
        delete h;
        return AA_HIP_ERROR_UNSUPPORTED;
    }
    if (hdr) {
        hdr->telescope_id = m.telescope_id();
        hdr->machine_id = m.machine_id();
        hdr->data_type = m.data_type();
        hdr->barycentric = m.barycentric();
        hdr->pulsarcentric = m.pulsarcentric();
        hdr->nbits = m.nbits();
        hdr->nsamples = m.nsamples();
        hdr->nchans = m.nchans();
        hdr->nifs = m.nifs();
        hdr->az_start = m.az_start();
        hdr->za_start = m.za_start();
        hdr->src_raj = m.src_raj();
        hdr->src_dej = m.src_dej();
        hdr->tstart = m.tstart();
        hdr->tsamp = m.tsamp();
        hdr->fch1 = m.fch1();
        hdr->foff = m.foff();
        hdr->fchannel = m.fchannel();
        hdr->refdm = m.refdm();
        hdr->period = m.period();
    }
    *out = h;
    return AA_HIP_SUCCESS;
}

int aa_hip_sigproc_read_u8(aa_hip_sigproc *h, int pinned, const uint8_t **data, size_t *bytes)
{
    if (!h || !data || !bytes) {
        aa_hip::set_last_error("aa_hip_sigproc_read_u8: null argument");
        return AA_HIP_ERROR_INVALID_ARGUMENT;
    }
    if (!h->in.did_read_signal() && !h->in.read_signal_u8(pinned != 0)) return AA_HIP_ERROR_UNSUPPORTED;
    if (!h->in.input_buffer_u8()) {
        aa_hip::set_last_error("aa_hip_sigproc_read_u8: the data were read through the u16 path");
        return AA_HIP_ERROR_INVALID_ARGUMENT;
    }
    *data = h->in.input_buffer_u8();
    *bytes = h->in.input_buffer_u8_size();
    return AA_HIP_SUCCESS;
}

int aa_hip_sigproc_read_u16(aa_hip_sigproc *h, const uint16_t **data, size_t *count)
{
    if (!h || !data || !count) {
        aa_hip::set_last_error("aa_hip_sigproc_read_u16: null argument");
        return AA_HIP_ERROR_INVALID_ARGUMENT;
    }
    if (!h->in.did_read_signal() && !h->in.read_signal()) return AA_HIP_ERROR_UNSUPPORTED;
    if (h->in.input_buffer().empty() && h->in.input_buffer_u8()) {
        aa_hip::set_last_error("aa_hip_sigproc_read_u16: the data were read through the u8 path");
        return AA_HIP_ERROR_INVALID_ARGUMENT;
    }
    *data = reinterpret_cast<const uint16_t *>(h->in.input_buffer().data());
    *count = h->in.input_buffer().size();
    return AA_HIP_SUCCESS;
}

void aa_hip_sigproc_close(aa_hip_sigproc *h) { delete h; }

// The reference's Python FFI of the reader (src/aa_py_sigproc_input.cpp:6-33):
// the object is an aa_sigproc_input; the buffer is the u16-widened data.
void *aa_py_sigproc_input(char const *path)
{
    if (!path) return nullptr;
    return new (std::nothrow) astroaccelerate::aa_sigproc_input(path);
}

void aa_py_sigproc_input_del(void const *obj)
{
    delete static_cast<astroaccelerate::aa_sigproc_input const *>(obj);
}

aa_py_filterbank_metadata_struct aa_py_sigproc_input_read_metadata(void *obj)
{
    aa_py_filterbank_metadata_struct out{};
    if (!obj) return out;
    const astroaccelerate::aa_filterbank_metadata m =
        static_cast<astroaccelerate::aa_sigproc_input *>(obj)->read_metadata();
    out.m_tstart = m.tstart();
    out.m_tsamp = m.tsamp();
    out.m_fch1 = m.fch1();
    out.m_foff = m.foff();
    out.m_nbits = m.nbits();
    out.m_nsamples = m.nsamples();
    out.m_nchans = m.nchans();
    out.m_src_raj = m.src_raj();
    return out;
}

bool aa_py_sigproc_input_read_signal(void *obj)
{
    return obj && static_cast<astroaccelerate::aa_sigproc_input *>(obj)->read_signal();
}

unsigned short const *aa_py_sigproc_input_input_buffer(void const *obj)
{
    if (!obj) return nullptr;
    const auto &b = static_cast<astroaccelerate::aa_sigproc_input const *>(obj)->input_buffer();
    return b.empty() ? nullptr : b.data();
}

}  // extern "C"
