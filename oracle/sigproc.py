"""CPU restatement of SigprocStokesWriter's output (src/lib/src/SigprocStokesWriter.cpp),
in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports this module.
"""
import struct

import numpy as np

F32 = np.float32


def pack(stokes, n_pols_write=1, n_bits=32, crop_min=0.0, crop_max=255.0):
    """The data bytes of one Stokes blob [t][s][pol][c] (:184-240): for t,
    for p < n_pols_write, subbands last to first, channels last to first;
    32 bits: the floats; 8 bits: _float2int (:285-290) in float arithmetic,
    the int's low byte written."""
    x = np.asarray(stokes, F32)[:, ::-1, :n_pols_write, ::-1].transpose(0, 2, 1, 3)
    if n_bits == 32:
        return np.ascontiguousarray(x).tobytes()
    assert n_bits == 8
    cmin, cmax = F32(crop_min), F32(crop_max)
    n_range = F32(255.0)
    delta = F32(cmax - cmin)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        f = np.where(x > cmax, cmax, x).astype(F32)
        y = np.rint(((f - cmin) * n_range).astype(F32) / delta).astype(F32)
        ok = np.isfinite(y) & (y >= -2.0 ** 31) & (y < 2.0 ** 31)
        q = np.where(ok, np.where(ok, y, 0).astype(np.int64), -(2 ** 31))
    q = np.where(f < cmin, 0, q)
    return np.ascontiguousarray((q & 0xFF).astype(np.uint8)).tobytes()


def _string(s):
    b = s.encode()
    return struct.pack("<i", len(b)) + b


def header(machine_id, telescope_id, source_name, raj, dej, fch1, foff, nchans, tsamp, nbits, tstart, nifs):
    """writeHeader (:123-158): length-prefixed keywords, int32 and double
    values, HEADER_START ... HEADER_END."""
    out = _string("HEADER_START")
    out += _string("machine_id") + struct.pack("<i", machine_id)
    out += _string("telescope_id") + struct.pack("<i", telescope_id)
    out += _string("data_type") + struct.pack("<i", 1)
    out += _string("source_name") + _string(source_name)
    out += _string("src_raj") + struct.pack("<d", float(F32(raj)))
    out += _string("src_dej") + struct.pack("<d", float(F32(dej)))
    out += _string("fch1") + struct.pack("<d", float(F32(fch1)))
    out += _string("foff") + struct.pack("<d", float(F32(foff)))
    out += _string("nchans") + struct.pack("<i", nchans)
    out += _string("tsamp") + struct.pack("<d", float(F32(tsamp)))
    out += _string("nbits") + struct.pack("<i", nbits)
    out += _string("tstart") + struct.pack("<d", tstart)
    out += _string("nifs") + struct.pack("<i", nifs)
    out += _string("HEADER_END")
    return out
