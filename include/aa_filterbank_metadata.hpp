// Filterbank metadata: same constructors and accessors as the reference
// (include/aa_filterbank_metadata.hpp:33-279).  Geometry is held in double;
// the strategy narrows fch1, foff and tsamp to float where the reference does.
#ifndef ASTRO_ACCELERATE_AA_FILTERBANK_METADATA_HPP
#define ASTRO_ACCELERATE_AA_FILTERBANK_METADATA_HPP

#include <string>

namespace astroaccelerate {

class aa_filterbank_metadata {
public:
    aa_filterbank_metadata() = default;

    // Full SIGPROC header (include/aa_filterbank_metadata.hpp:33-93 order).
    aa_filterbank_metadata(int telescope_id, int machine_id, int data_type, const std::string &rawdatafile,
                           const std::string &source_name, int barycentric, int pulsarcentric,
                           double az_start, double za_start, double src_raj, double src_dej,
                           double tstart, double tsamp, int nbits, int nsamples, double fch1,
                           double foff, char FREQUENCY_START, double fchannel, char FREQUENCY_END,
                           int nchans, int nifs, double refdm, double period)
        : m_az_start(az_start), m_za_start(za_start), m_src_raj(src_raj), m_src_dej(src_dej),
          m_tstart(tstart), m_tsamp(tsamp), m_refdm(refdm), m_period(period), m_fch1(fch1),
          m_foff(foff), m_fchannel(fchannel), m_telescope_id(telescope_id),
          m_machine_id(machine_id), m_data_type(data_type), m_barycentric(barycentric),
          m_pulsarcentric(pulsarcentric), m_nbits(nbits), m_nsamples(nsamples),
          m_nchans(nchans), m_nifs(nifs), m_FREQUENCY_START(FREQUENCY_START),
          m_FREQUENCY_END(FREQUENCY_END), m_rawdatafile(rawdatafile), m_source_name(source_name) {}

    // Reduced constructor (include/aa_filterbank_metadata.hpp:146-178).
    aa_filterbank_metadata(const double &tstart, const double &tsamp, const int &nbits,
                           const int &nsamples, const double &fch1, const double &foff,
                           const int &nchans)
        : m_tstart(tstart), m_tsamp(tsamp), m_fch1(fch1), m_foff(foff), m_nbits(nbits),
          m_nsamples(nsamples), m_nchans(nchans) {}

    double az_start() const { return m_az_start; }
    double za_start() const { return m_za_start; }
    double src_raj() const { return m_src_raj; }
    double src_dej() const { return m_src_dej; }
    double tstart() const { return m_tstart; }
    double tsamp() const { return m_tsamp; }
    double refdm() const { return m_refdm; }
    double period() const { return m_period; }
    double fch1() const { return m_fch1; }
    double foff() const { return m_foff; }
    double fchannel() const { return m_fchannel; }
    int telescope_id() const { return m_telescope_id; }
    int machine_id() const { return m_machine_id; }
    int data_type() const { return m_data_type; }
    int barycentric() const { return m_barycentric; }
    int pulsarcentric() const { return m_pulsarcentric; }
    int nbits() const { return m_nbits; }
    int nsamples() const { return m_nsamples; }
    int nchans() const { return m_nchans; }
    int nifs() const { return m_nifs; }
    char FREQUENCY_START() const { return m_FREQUENCY_START; }
    char FREQUENCY_END() const { return m_FREQUENCY_END; }
    std::string rawdatafile() const { return m_rawdatafile; }
    std::string source_name() const { return m_source_name; }
    double sky_frequency() const { return m_fch1; }

private:
    double m_az_start = 0, m_za_start = 0, m_src_raj = 0, m_src_dej = 0;
    double m_tstart = 0, m_tsamp = 0, m_refdm = 0, m_period = 0;
    double m_fch1 = 0, m_foff = 0, m_fchannel = 0;
    int m_telescope_id = 0, m_machine_id = 0, m_data_type = 0, m_barycentric = 0;
    int m_pulsarcentric = 0, m_nbits = 0, m_nsamples = 0, m_nchans = 0, m_nifs = 0;
    char m_FREQUENCY_START = 0, m_FREQUENCY_END = 0;
    std::string m_rawdatafile, m_source_name;
};

}  // namespace astroaccelerate
#endif
