"""SIGPROC filterbank reading, restated from the reference's reader.

TEST INFRASTRUCTURE ONLY: the checker for the product's reader
(aa_hip_sigproc_*), imported by tests/ only.

Follows src/aa_sigproc_input.cpp:
  - header records (:85-339): a 4-byte length, then that many characters;
    lengths outside 2..79 are skipped with nothing else read; HEADER_END ends
    the header; the int keywords (telescope_id, machine_id, data_type,
    barycentric, pulsarcentric, nbits, nsamples, nchans, nifs) and the double
    keywords (az_start, za_start, src_raj, src_dej, tstart, tsamp, fch1,
    foff, fchannel, refdm, period) are followed by their value; any other
    string is skipped; nifs must be 1 (:340-344);
  - nsamples from the data size (:355-388): bytes * 8 / (nchans * nbits);
  - data (:394-541), one record of nchans samples per time step: 32-bit
    floats scaled by (x - min) / ((max - min) / 65535) in float and
    truncated; 16 and 8 bits as stored; 4 bits high nibble first
    (:469-475); 2 bits most significant pair first (:494-499); 1 bit least
    significant bit first (:517-523).
"""
from __future__ import annotations

import struct
from typing import Dict, Tuple

import numpy as np

INT_KEYS = ("telescope_id", "machine_id", "data_type", "barycentric", "pulsarcentric", "nbits", "nsamples",
            "nchans", "nifs")
DOUBLE_KEYS = ("az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp", "fch1", "foff", "fchannel",
               "refdm", "period")


def read_header(raw: bytes) -> Tuple[Dict[str, float], int]:
    """(header fields, offset of the first data byte); raises ValueError on a bad header."""
    hdr = {k: 0 for k in INT_KEYS}
    hdr.update({k: 0.0 for k in DOUBLE_KEYS})
    pos = 0
    while True:
        if pos + 4 > len(raw):
            raise ValueError("no HEADER_END")
        (n,) = struct.unpack_from("<i", raw, pos)
        pos += 4
        if n <= 1 or n >= 80:
            continue
        key = raw[pos:pos + n].decode("latin-1")
        pos += n
        if key == "HEADER_END":
            break
        if key in INT_KEYS:
            (hdr[key],) = struct.unpack_from("<i", raw, pos)
            pos += 4
        elif key in DOUBLE_KEYS:
            (hdr[key],) = struct.unpack_from("<d", raw, pos)
            pos += 8
    if hdr["nifs"] != 1:
        raise ValueError("nifs must be 1")
    return hdr, pos


def unpack(payload: bytes, nbits: int, nchans: int) -> np.ndarray:
    """[nsamples, nchans] samples of whole records (u16; floats scaled)."""
    rec = nchans * nbits // 8
    nsamp = len(payload) // rec
    b = np.frombuffer(payload[:nsamp * rec], dtype=np.uint8)
    if nbits == 8:
        v = b.astype(np.uint16)
    elif nbits == 16:
        v = np.frombuffer(payload[:nsamp * rec], dtype="<u2").astype(np.uint16)
    elif nbits == 4:
        v = np.stack([(b >> 4) & 15, b & 15], axis=1).reshape(-1).astype(np.uint16)
    elif nbits == 2:
        v = np.stack([(b >> 6) & 3, (b >> 4) & 3, (b >> 2) & 3, b & 3], axis=1).reshape(-1).astype(np.uint16)
    elif nbits == 1:
        v = np.stack([(b >> i) & 1 for i in range(8)], axis=1).reshape(-1).astype(np.uint16)
    elif nbits == 32:
        f = np.frombuffer(payload[:nsamp * rec], dtype="<f4")
        mx = np.float32(max(np.float32(-10000000000.0), f.max())) if f.size else np.float32(-1e10)
        mn = np.float32(min(np.float32(10000000000.0), f.min())) if f.size else np.float32(1e10)
        binw = np.float32((mx - mn) / np.float32(65535.0))
        v = ((f - mn) / binw).astype(np.float32)
        v = np.trunc(v).astype(np.int64).astype(np.uint16)
    else:
        raise ValueError("nbits")
    return v.reshape(nsamp, nchans)


def read(path: str) -> Tuple[Dict[str, float], np.ndarray]:
    with open(path, "rb") as fh:
        raw = fh.read()
    hdr, pos = read_header(raw)
    data = unpack(raw[pos:], hdr["nbits"], hdr["nchans"])
    hdr["nsamples"] = data.shape[0]
    return hdr, data
