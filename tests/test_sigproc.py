"""SIGPROC filterbank reader (aa_hip_sigproc_*, aa_sigproc_input) against
the restated reference reader (oracle/sigproc_ref.py).

The restatement is pinned first by known answers the reference's rules fix
by hand (bit order of sub-byte samples, float scaling with an exact bin
width, keyword skipping, nsamples from the data size); then the product's
reader must return the same header and samples, for every bit width, on the
u16 path (the reference's widening) and the u8 path (no widening).  All of
it runs on the CPU (unpinned buffers).
"""
import struct

import numpy as np
import pytest

import sigproc_ref


def write_fil(path, nbits, nchans, payload: bytes, nifs=1, extra_strings=True, **fields):
    """A SIGPROC file through the package's writer (header records, then the payload)."""
    import py_astro_accelerate as aa
    strings = {"source_name": "FAKE_PSR", "rawdatafile": "x.raw"} if extra_strings else None
    base = dict(nsamples=7, src_raj=123456.7, src_dej=-654321.0, az_start=12.5, za_start=33.25)
    base.update(fields)
    aa.write_sigproc(path, nbits, nchans, payload, nifs=nifs, strings=strings, **base)


# ------------------------------------------------------------ oracle pinning
def test_oracle_sub_byte_known_answers():
    b = bytes([0b10110100])
    assert sigproc_ref.unpack(b, 1, 8).tolist() == [[0, 0, 1, 0, 1, 1, 0, 1]]   # least significant bit first
    assert sigproc_ref.unpack(b, 2, 4).tolist() == [[2, 3, 1, 0]]               # most significant pair first
    assert sigproc_ref.unpack(b, 4, 2).tolist() == [[11, 4]]                    # high nibble first
    assert sigproc_ref.unpack(bytes([7, 200]), 8, 2).tolist() == [[7, 200]]
    assert sigproc_ref.unpack(struct.pack("<2H", 513, 65535), 16, 2).tolist() == [[513, 65535]]


def test_oracle_float_scaling_exact_bin():
    # min 0, max 65535: the bin width is exactly 1, so samples truncate to themselves
    vals = [0.0, 1.5, 65535.0, 12345.9]
    got = sigproc_ref.unpack(struct.pack("<4f", *vals), 32, 4)
    assert got.tolist() == [[0, 1, 65535, 12345]]


def test_oracle_header_rules(tmp_path):
    p = str(tmp_path / "a.fil")
    write_fil(p, 8, 4, bytes(range(4 * 3)) + b"\x01\x02", tsamp=1e-3)   # 3 whole records + a partial one
    hdr, data = sigproc_ref.read(p)
    assert hdr["nchans"] == 4 and hdr["nbits"] == 8 and hdr["tsamp"] == 1e-3 and hdr["telescope_id"] == 4
    assert hdr["nsamples"] == 3 and data.shape == (3, 4)                   # from the size, not the header's 7
    assert data[2].tolist() == [8, 9, 10, 11]
    q = str(tmp_path / "b.fil")
    write_fil(q, 8, 4, bytes(8), nifs=2)
    with pytest.raises(ValueError):
        sigproc_ref.read(q)


# ------------------------------------------------------- product vs oracle
@pytest.mark.parametrize("nbits", [1, 2, 4, 8, 16, 32])
def test_reader_matches_oracle(aa, tmp_path, nbits):
    rng = np.random.default_rng(nbits)
    nchans, nsamp = 64, 301
    if nbits == 32:
        payload = rng.normal(3.0, 2.0, size=nsamp * nchans).astype("<f4").tobytes()
    elif nbits == 16:
        payload = rng.integers(0, 65536, size=nsamp * nchans, dtype=np.uint16).astype("<u2").tobytes()
    else:
        payload = rng.integers(0, 256, size=nsamp * nchans * nbits // 8, dtype=np.uint8).tobytes()
    p = str(tmp_path / f"r{nbits}.fil")
    write_fil(p, nbits, nchans, payload + b"\x05")   # a trailing partial record is dropped
    want_hdr, want = sigproc_ref.read(p)

    r = aa.aa_sigproc_input(p)
    hdr = r.read_metadata()
    for k in sigproc_ref.INT_KEYS + sigproc_ref.DOUBLE_KEYS:
        assert hdr[k] == want_hdr[k], k
    got = r.read_signal()
    assert got.dtype == np.uint16 and got.shape == (nsamp, nchans)
    np.testing.assert_array_equal(got, want)
    with pytest.raises(aa.AAHipError):   # read once, through one path
        r.read_signal_u8()
    r.close()

    if nbits <= 8:
        r8 = aa.aa_sigproc_input(p)
        r8.read_metadata()
        got8 = r8.read_signal_u8(pinned=False)
        assert got8.dtype == np.uint8
        np.testing.assert_array_equal(got8, want.astype(np.uint8))
        r8.close()
    else:
        r8 = aa.aa_sigproc_input(p)
        r8.read_metadata()
        with pytest.raises(aa.AAHipError):   # the u8 path holds nbits <= 8 only
            r8.read_signal_u8()


def test_reader_rejects_bad_files(aa, tmp_path):
    p = str(tmp_path / "nifs2.fil")
    write_fil(p, 8, 4, bytes(16), nifs=2)
    with pytest.raises(aa.AAHipError):
        aa.aa_sigproc_input(p).read_metadata()
    q = str(tmp_path / "odd.fil")
    write_fil(q, 4, 3, bytes(16))            # 3 channels of 4 bits: not whole bytes per record
    with pytest.raises(aa.AAHipError):
        aa.aa_sigproc_input(q).read_metadata()
    with pytest.raises(aa.AAHipError):
        aa.aa_sigproc_input(str(tmp_path / "missing.fil")).read_metadata()


def test_reader_keeps_geometry_for_the_pipeline(aa, tmp_path):
    p = str(tmp_path / "g.fil")
    write_fil(p, 8, 16, bytes(16 * 10), fch1=1400.25, foff=-0.25, tsamp=128e-6, tstart=58000.0)
    m = aa.aa_sigproc_input(p).metadata()
    assert (m.m_nchans, m.m_nsamples, m.m_nbits) == (16, 10, 8)
    assert (m.m_fch1, m.m_foff, m.m_tsamp, m.m_tstart) == (1400.25, -0.25, 128e-6, 58000.0)
