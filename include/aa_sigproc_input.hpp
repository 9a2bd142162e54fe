// SIGPROC filterbank reader: the reference's aa_sigproc_input interface
// (include/aa_sigproc_input.hpp:17-79; src/aa_sigproc_input.cpp), plus an
// 8-bit path that keeps the samples as u8 instead of widening them to u16.
//
// Header: the keyword/value records of a SIGPROC file up to HEADER_END, read
// as the reference reads them (src/aa_sigproc_input.cpp:85-339): int
// keywords telescope_id, machine_id, data_type, barycentric,
// pulsarcentric, nbits, nsamples, nchans, nifs; double keywords az_start,
// za_start, src_raj, src_dej, tstart, tsamp, fch1, foff, fchannel, refdm,
// period; any other string (and a length outside 2..79) is skipped with no
// value; nifs must be 1.  nsamples is taken from the data size (whole
// records of nchans samples at nbits each), not from the header.
//
// Data ([t][c], one record of nchans samples per time step; a partial last
// record is dropped):
//   - 8 and 16 bits: as stored;
//   - 4 bits: two channels per byte, high nibble first;
//   - 2 bits: four channels per byte, most significant pair first;
//   - 1 bit: eight channels per byte, least significant bit first;
//   - 32-bit float: scaled to 0..65535 by (x - min) / ((max - min) / 65535),
//     truncated (u16 path only).
// read_signal() fills the u16 buffer, as the reference does; read_signal_u8()
// fills a u8 buffer for nbits <= 8 (the dedispersion path's native input),
// optionally in pinned host memory so that its host-to-device copies run
// asynchronously at full link speed.
#ifndef ASTRO_ACCELERATE_AA_SIGPROC_INPUT_HPP
#define ASTRO_ACCELERATE_AA_SIGPROC_INPUT_HPP

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "aa_filterbank_metadata.hpp"

namespace astroaccelerate {

class aa_sigproc_input {
public:
    explicit aa_sigproc_input(const std::string &path);
    virtual ~aa_sigproc_input();
    aa_sigproc_input(const aa_sigproc_input &) = delete;
    aa_sigproc_input &operator=(const aa_sigproc_input &) = delete;

    bool open();
    bool close();

    // Metadata of the file (opened if needed); a default-constructed object
    // when the file cannot be opened or its header is invalid.
    aa_filterbank_metadata read_metadata();

    // Reads the data once, after read_metadata(), widened to u16 as the
    // reference does; false if called again or before the header was read.
    bool read_signal();
    const std::vector<unsigned short> &input_buffer() const { return m_input_buffer; }
    std::vector<unsigned short> &input_buffer_modifiable() { return m_input_buffer; }
    bool did_read_signal() const { return m_data_is_read; }

    // Reads the data once as u8 (nbits 1, 2, 4 or 8 only), into pinned host
    // memory when `pinned` (needs a GPU), else into ordinary memory.
    bool read_signal_u8(bool pinned);
    const uint8_t *input_buffer_u8() const { return m_u8; }
    size_t input_buffer_u8_size() const { return m_u8_size; }

    bool header_ok() const { return m_header_ok; }

private:
    bool read_header(aa_filterbank_metadata &metadata);
    bool read_records(uint8_t *u8, unsigned short *u16);
    bool read_parallel(uint8_t *dst, size_t bytes);
    void free_u8();

    std::string m_path;
    FILE *m_fp = nullptr;
    bool m_header_is_read = false;
    bool m_header_ok = false;
    bool m_data_is_read = false;
    long m_data_start = 0;
    aa_filterbank_metadata m_meta;
    std::vector<unsigned short> m_input_buffer;
    uint8_t *m_u8 = nullptr;
    size_t m_u8_size = 0;
    bool m_u8_pinned = false;
};

}  // namespace astroaccelerate

#endif  // ASTRO_ACCELERATE_AA_SIGPROC_INPUT_HPP
