"""File formats of the reference program, through the C ABI (no GPU needed):
NIfTI-1 reading and writing (read_data.cpp:60-245, nifti1_read_write.cpp:
198-322) and the design file (read_data.cpp:262-316).

The reader is checked against images written here by an independent
struct-based writer (all supported datatypes, both byte orders, plain and
gzip), and against the reference's own bundled mask when that tree is
present; the writer is checked field by field against the header that
write_nifti_file fills in.
"""
import gzip
import os
import struct

import numpy as np
import pytest

import bsglmm_amd as B
from bsglmm_amd import BsgError
from nifti_fixture import make_nifti

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MASK = "/root/reference/bsglmm/images/mask.nii.gz"

@pytest.mark.parametrize("datatype", [2, 4, 8, 16, 64])
@pytest.mark.parametrize("big_endian", [False, True])
@pytest.mark.parametrize("ext", [".nii", ".nii.gz"])
def test_reader_types_byte_orders_and_compression(tmp_path, datatype, big_endian, ext):
    rng = np.random.default_rng(datatype)
    vol = rng.integers(0, 200, size=(4, 5, 6)).astype(np.float64)
    if datatype in (16, 64):
        vol += rng.random(vol.shape)
    p = str(tmp_path / ("img" + ext))
    make_nifti(p, vol, datatype, big_endian)
    assert B.nifti_info(p) == ((6, 5, 4), datatype)
    got = B.nifti_read(p)
    want = vol.astype(np.float32).astype(np.float64) if datatype == 16 else vol
    assert np.array_equal(got, want)


def test_reader_honours_vox_offset_and_rejects_other_files(tmp_path):
    vol = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    p = str(tmp_path / "ext.nii")
    make_nifti(p, vol, 64, vox_offset=368)  # an extension block before the data
    assert np.array_equal(B.nifti_read(p), vol)
    bad = tmp_path / "junk.nii"
    bad.write_bytes(b"\0" * 400)
    with pytest.raises(BsgError, match="not a NIfTI-1"):
        B.nifti_info(str(bad))
    with pytest.raises(BsgError, match="single file NIFTI"):
        B.nifti_info(str(tmp_path / "x.hdr"))
    with pytest.raises(BsgError, match="Cannot open"):
        B.nifti_info(str(tmp_path / "missing.nii"))


def test_name_fallback_between_nii_and_nii_gz(tmp_path):
    """nifti_expand (read_data.cpp:97-110): a missing .nii.gz falls back to .nii
    and a missing .nii to .nii.gz."""
    vol = np.ones((2, 2, 2))
    make_nifti(str(tmp_path / "a.nii"), vol, 64)
    make_nifti(str(tmp_path / "b.nii.gz"), 2 * vol, 64)
    assert B.nifti_resolve(str(tmp_path / "a.nii.gz")) == str(tmp_path / "a.nii")
    assert B.nifti_resolve(str(tmp_path / "b.nii")) == str(tmp_path / "b.nii.gz")
    assert np.array_equal(B.nifti_read(str(tmp_path / "b.nii")), 2 * vol)


@pytest.mark.parametrize("ext", [".nii", ".nii.gz"])
def test_writer_header_matches_write_nifti_file(tmp_path, ext):
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((3, 4, 5))
    p = str(tmp_path / ("out" + ext))
    B.nifti_write(p, vol)
    raw = gzip.open(p).read() if ext.endswith(".gz") else open(p, "rb").read()
    assert len(raw) == 352 + 8 * vol.size
    assert struct.unpack_from("<i", raw, 0)[0] == 348
    assert struct.unpack_from("<8h", raw, 40) == (3, 5, 4, 3, 1, 0, 0, 0)
    assert struct.unpack_from("<hh", raw, 70) == (64, 64)          # FLOAT64, bitpix 64
    assert struct.unpack_from("<8f", raw, 76) == (0, 2, 2, 2, 0, 0, 0, 0)  # 2 mm voxels
    assert struct.unpack_from("<fff", raw, 108) == (352.0, 0.0, 0.0)     # vox_offset, slope, inter
    assert raw[123] == 2                                           # NIFTI_UNITS_MM
    assert raw[344:348] == b"n+1\0" and raw[348:352] == b"\0\0\0\0"
    assert np.array_equal(np.frombuffer(raw[352:], "<f8").reshape(vol.shape), vol)
    assert np.array_equal(B.nifti_read(p), vol)  # and our reader round-trips it


@pytest.mark.skipif(not os.path.exists(REF_MASK), reason="reference tree not present")
def test_reads_the_reference_bundled_mask():
    """The reference's own images/mask.nii.gz (FLOAT64, as read_nifti1_mask
    requires) read here equals the padded mask held bit-packed in
    tests/golden/bundled_mask.npz, whose layout counts test_layout pins."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "bundled_mask.npz"))
    gx, gy, gz = (int(v) for v in g["grid"])
    (nx, ny, nz), dt = B.nifti_info(REF_MASK)
    assert dt == 64 and (nx, ny, nz) == (gx, gy, gz)
    vol = B.nifti_read(REF_MASK)
    m = np.zeros((nz + 2, ny + 2, nx + 2), np.uint8)
    m[1:-1, 1:-1, 1:-1] = vol.astype(np.uint8)
    want = np.unpackbits(g["bits"])[: int(g["nsize"])]
    assert np.array_equal(m.reshape(-1), want)


def test_design_file(tmp_path):
    p = tmp_path / "design.txt"
    p.write_text("ID grpA grpB age Image\n"
                 "s1 1 0 0.5 a.nii.gz\n"
                 "s2\t0\t1\t-1.25\tb.nii\n"
                 "s3 1 0 2e-1 c.nii.gz\n")
    covar, names, imgs = B.design_read(str(p), 3)
    assert names == ["grpA", "grpB", "age"]
    assert imgs == ["./images/a.nii.gz", "./images/b.nii", "./images/c.nii.gz"]
    assert np.array_equal(covar, np.array([[1, 0, 1], [0, 1, 0], [0.5, -1.25, 0.2]], np.float32))
    (tmp_path / "ragged.txt").write_text("ID a Image\ns1 1 x.nii\ns2 0\n")
    with pytest.raises(BsgError, match="whole rows"):
        B.design_read(str(tmp_path / "ragged.txt"), 1)
