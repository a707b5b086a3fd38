"""The -F output formats of the host library (mmc_host_save_field) against
the layouts the reference writes: mesh_saveweight / mcx_savedata
(src/mmc_mesh.c:1726-1800, src/mmc_utils.c:919-993), the NIfTI-1 header of
mcx_savenii (src/mmc_utils.c:515-607), and the JNIfTI text/binary documents of
mcx_savejnii / mcx_savebnii with the JData array annotation of
mcx_jdataencode (src/mmc_utils.c:622-905, 3071-3185).  Every file is parsed
back here with an independent reader and must hold the field exactly."""
import base64
import ctypes as C
import json
import os
import struct
import zlib

import numpy as np
import pytest

from mmc_amd import _lib

DATALEN, MAXGATE, TSTEP = 37, 3, 5e-10


@pytest.fixture(scope="module")
def field():
    rng = np.random.default_rng(3)
    return rng.random(DATALEN * MAXGATE) * 10.0 ** rng.integers(-12, 3, DATALEN * MAXGATE)


def save(tmp_path, fmt, f, compress=1, outputtype=1):
    L = _lib.lib()
    sess = str(tmp_path / "sess")
    fid = L.mmc_host_format_id(fmt.encode())
    assert fid >= 0
    rc = L.mmc_host_save_field(sess.encode(), fid, _lib.ptr(f, C.c_double), DATALEN, MAXGATE, 1, TSTEP,
                               outputtype, compress)
    assert rc == 0
    return sess


def test_format_ids_follow_the_reference_enum():
    L = _lib.lib()
    names = ["ascii", "bin", "nii", "hdr", "mc2", "tx3", "jnii", "bnii"]   # enum TOutputFormat
    assert [L.mmc_host_format_id(n.encode()) for n in names] == list(range(8))
    assert L.mmc_host_format_id(b"nifti") == -1


def test_ascii(tmp_path, field):
    sess = save(tmp_path, "ascii", field)
    rows = [l.split("\t") for l in open(sess + ".dat").read().splitlines()]
    assert len(rows) == DATALEN * MAXGATE
    assert [int(r[0]) for r in rows] == list(range(1, DATALEN + 1)) * MAXGATE
    assert np.allclose([float(r[1]) for r in rows], field, rtol=1e-6, atol=0)


@pytest.mark.parametrize("fmt", ["bin", "mc2"])
def test_raw_doubles(tmp_path, field, fmt):
    sess = save(tmp_path, fmt, field)
    assert np.array_equal(np.fromfile(sess + "." + fmt, np.float64), field)


def nifti_header(raw):
    h = {}
    h["sizeof_hdr"] = struct.unpack_from("<i", raw, 0)[0]
    h["dim"] = struct.unpack_from("<8h", raw, 40)
    h["datatype"], h["bitpix"] = struct.unpack_from("<2h", raw, 70)
    h["pixdim"] = struct.unpack_from("<8f", raw, 76)
    h["vox_offset"], h["scl_slope"] = struct.unpack_from("<2f", raw, 108)
    h["xyzt_units"] = raw[123]
    h["magic"] = raw[344:348]
    return h


@pytest.mark.parametrize("fmt", ["nii", "hdr"])
def test_nifti_and_analyze(tmp_path, field, fmt):
    sess = save(tmp_path, fmt, field)
    raw = open(sess + "." + fmt, "rb").read()
    h = nifti_header(raw)
    assert h["sizeof_hdr"] == 348
    assert h["dim"][:5] == (4, 1, MAXGATE, DATALEN, 1)
    assert (h["datatype"], h["bitpix"]) == (64, 64)                 # NIFTI_TYPE_FLOAT64
    assert h["pixdim"][1:4] == (1.0, 1.0, 1.0)
    assert h["pixdim"][4] == pytest.approx(TSTEP * 1e6)
    assert h["xyzt_units"] == 2 | 24                                 # mm | usec
    if fmt == "nii":
        assert h["magic"] == b"n+1\0" and h["vox_offset"] == 352.0
        assert raw[348:352] == b"\0\0\0\0"
        data = np.frombuffer(raw[352:], np.float64)
    else:
        assert h["magic"] == b"ni1\0" and h["vox_offset"] == 0.0 and len(raw) == 348
        data = np.fromfile(sess + ".img", np.float64)
    assert np.array_equal(data, field)


def test_tx3(tmp_path, field):
    sess = save(tmp_path, "tx3", field)
    raw = open(sess + ".tx3", "rb").read()
    assert struct.unpack_from("<I3i", raw, 0) == (0x8814, 1, MAXGATE, DATALEN)
    assert np.array_equal(np.frombuffer(raw[16:], np.float64), field)


def check_jnifti(doc, field, compressed, dims=(DATALEN, MAXGATE, 1), vox=(1.0, TSTEP, 1.0)):
    # mesh data: Dim [datalen, maxgate, srcnum] (lastdim 3 of mcx_savedata,
    # src/mmc_utils.c:941-956); grid data: [nx, ny, nz, maxgate, srcnum]
    hdr, dat = doc["NIFTIHeader"], doc["NIFTIData"]
    assert list(hdr["Dim"]) == list(dims)
    assert hdr["DataType"] == "double" and hdr["BitDepth"] == 64
    assert hdr["LastSliceID"] == MAXGATE
    assert np.allclose(hdr["VoxelSize"], vox, rtol=1e-6)
    assert hdr["Description"] == "MMC volumetric output: Fluence (J/mm^2)"
    assert dat["_ArrayType_"] == "double" and list(dat["_ArraySize_"]) == list(dims)
    if compressed:
        assert dat["_ArrayZipType_"] == "zlib" and dat["_ArrayZipSize_"] == int(np.prod(dims))
        payload = zlib.decompress(dat["_ArrayZipData_"])
    else:
        payload = dat["_ArrayData_"]
    assert np.array_equal(np.frombuffer(payload, np.float64), field)


@pytest.mark.parametrize("compress", [1, 0])
def test_jnii(tmp_path, field, compress):
    sess = save(tmp_path, "jnii", field, compress=compress)
    doc = json.load(open(sess + ".jnii"))
    assert doc["_DataInfo_"]["JNIFTIVersion"] == "0.5"
    dat = doc["NIFTIData"]
    key = "_ArrayZipData_" if compress else "_ArrayData_"
    dat[key] = base64.b64decode(dat[key])
    check_jnifti(doc, field, compress)


def bjdata_load(raw):
    """Minimal BJData draft-2 reader (little-endian; objects, arrays,
    optimized typed arrays, strings and scalars) for the test."""
    fmt = {"U": "<B", "i": "<b", "u": "<H", "I": "<h", "m": "<I", "l": "<i", "M": "<Q", "L": "<q",
           "d": "<f", "D": "<d"}
    npt = {"U": np.uint8, "m": np.uint32, "l": np.int32, "d": np.float32, "D": np.float64}
    pos = 0

    def scalar(m):
        nonlocal pos
        v = struct.unpack_from(fmt[m], raw, pos)[0]
        pos += struct.calcsize(fmt[m])
        return v

    def length():
        nonlocal pos
        m = chr(raw[pos]); pos += 1
        return scalar(m)

    def string():
        nonlocal pos
        n = length()
        s = raw[pos:pos + n].decode(); pos += n
        return s

    def value():
        nonlocal pos
        m = chr(raw[pos]); pos += 1
        if m == "{":
            obj = {}
            while chr(raw[pos]) != "}":
                k = string()
                obj[k] = value()
            pos += 1
            return obj
        if m == "[":
            if chr(raw[pos]) == "$":
                t = chr(raw[pos + 1]); assert chr(raw[pos + 2]) == "#"
                pos += 3
                n = length()
                a = np.frombuffer(raw, npt[t], n, pos)
                pos += n * a.itemsize
                return a.tobytes() if t == "U" else a
            out = []
            while chr(raw[pos]) != "]":
                out.append(value())
            pos += 1
            return out
        if m == "S":
            return string()
        if m == "C":
            c = chr(raw[pos]); pos += 1
            return c
        return scalar(m)

    doc = value()
    assert pos == len(raw)
    return doc


@pytest.mark.parametrize("compress", [1, 0])
def test_bnii(tmp_path, field, compress):
    sess = save(tmp_path, "bnii", field, compress=compress)
    doc = bjdata_load(open(sess + ".bnii", "rb").read())
    assert doc["_DataInfo_"]["SerialFormat"] == "https://neurojson.org/bjdata/draft2"
    if not compress:
        doc["NIFTIData"]["_ArrayData_"] = doc["NIFTIData"]["_ArrayData_"].tobytes()
    check_jnifti(doc, field, compress)


# ---- dual-grid data: cfg->dim voxels of edge cfg->steps (mcx_savedata with
# the grid's dims, src/mmc_utils.c:515-545, 937-963)
GRID, STEP = (4, 3, 5), 0.5


def save_grid(tmp_path, fmt, f, srcnum=1, compress=1):
    L = _lib.lib()
    sess = str(tmp_path / "grid")
    dim = np.array(GRID, np.int32)
    rc = L.mmc_host_save_grid(sess.encode(), L.mmc_host_format_id(fmt.encode()), _lib.ptr(f, C.c_double),
                              _lib.ptr(dim, C.c_int32), C.c_float(STEP), MAXGATE, srcnum, C.c_float(TSTEP), 1,
                              compress)
    assert rc == 0
    return sess


@pytest.fixture(scope="module")
def gfield():
    return np.random.default_rng(4).random(int(np.prod(GRID)) * MAXGATE)


def test_grid_nifti_is_a_volume_series(tmp_path, gfield):
    sess = save_grid(tmp_path, "nii", gfield)
    raw = open(sess + ".nii", "rb").read()
    h = nifti_header(raw)
    assert h["dim"][:5] == (4,) + GRID + (MAXGATE,)
    assert h["pixdim"][1:4] == (STEP, STEP, STEP)
    assert np.array_equal(np.frombuffer(raw[352:], np.float64), gfield)


def test_grid_tx3_and_jnifti_dims(tmp_path, gfield):
    sess = save_grid(tmp_path, "tx3", gfield)
    raw = open(sess + ".tx3", "rb").read()
    assert struct.unpack_from("<I3i", raw, 0) == (0x8814,) + GRID
    sess = save_grid(tmp_path, "jnii", gfield)
    doc = json.load(open(sess + ".jnii"))
    doc["NIFTIData"]["_ArrayZipData_"] = base64.b64decode(doc["NIFTIData"]["_ArrayZipData_"])
    check_jnifti(doc, gfield, 1, dims=GRID + (MAXGATE, 1), vox=(STEP, STEP, STEP, TSTEP, 1.0))
    sess = save_grid(tmp_path, "bnii", gfield, compress=0)
    doc = bjdata_load(open(sess + ".bnii", "rb").read())
    doc["NIFTIData"]["_ArrayData_"] = doc["NIFTIData"]["_ArrayData_"].tobytes()
    check_jnifti(doc, gfield, 0, dims=GRID + (MAXGATE, 1), vox=(STEP, STEP, STEP, TSTEP, 1.0))


def test_text_output_of_shared_patterns_names_the_pattern(tmp_path):
    # mesh_saveweight writes "value<TAB>pattern<TAB>%e" rows when srcnum > 1
    # (src/mmc_mesh.c:1782-1792), the field laid out [gate][value][pattern]
    L = _lib.lib()
    f = np.arange(DATALEN * MAXGATE * 2, dtype=np.float64) + 0.5
    sess = str(tmp_path / "pat")
    assert L.mmc_host_save_field(sess.encode(), 0, _lib.ptr(f, C.c_double), DATALEN, MAXGATE, 2, TSTEP, 1, 1) == 0
    rows = [l.split("\t") for l in open(sess + ".dat").read().splitlines()]
    assert len(rows) == DATALEN * MAXGATE * 2
    j = np.tile(np.repeat(np.arange(1, DATALEN + 1), 2), MAXGATE)
    k = np.tile([1, 2], DATALEN * MAXGATE)
    assert np.array_equal([int(r[0]) for r in rows], j) and np.array_equal([int(r[1]) for r in rows], k)
    assert np.allclose([float(r[2]) for r in rows], f, rtol=1e-6)
