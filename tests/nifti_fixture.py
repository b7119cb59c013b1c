"""Test helper: a minimal NIfTI-1 single-file writer, field by field at the
standard's byte offsets (independent of the library's own writer)."""
import gzip
import struct

import numpy as np

DT = {2: ("u1", 8), 4: ("i2", 16), 8: ("i4", 32), 16: ("f4", 32), 64: ("f8", 64)}


def make_nifti(path, vol, datatype, big_endian=False, vox_offset=352):
    """vol is [nz][ny][nx] (x fastest); datatype a NIfTI code from DT."""
    e = ">" if big_endian else "<"
    nz, ny, nx = vol.shape
    h = bytearray(vox_offset)
    struct.pack_into(e + "i", h, 0, 348)
    struct.pack_into(e + "8h", h, 40, 3, nx, ny, nz, 1, 1, 1, 1)
    code, bits = DT[datatype]
    struct.pack_into(e + "hh", h, 70, datatype, bits)
    struct.pack_into(e + "8f", h, 76, 1, 1, 1, 1, 0, 0, 0, 0)
    struct.pack_into(e + "f", h, 108, float(vox_offset))
    h[344:348] = b"n+1\0"
    raw = bytes(h) + np.ascontiguousarray(vol, dtype=np.dtype(e + code)).tobytes()
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "wb") as f:
        f.write(raw)
