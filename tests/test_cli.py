"""The command-line front end (mmc_amd/bin/mmc): the reference's JSON input
and flags, and its .dat / .bin / .mch outputs."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

MMC = os.path.join(ROOT, "mmc_amd", "bin", "mmc")


def run(args, cwd):
    return subprocess.run([MMC] + args, cwd=cwd, capture_output=True, text=True, timeout=600)


def test_help_and_errors(tmp_path):
    assert run(["-h"], tmp_path).returncode == 0
    bad = tmp_path / "bad.json"
    bad.write_text('{"Session": {"Photons": 10,, }}')
    r = run(["-f", str(bad)], tmp_path)
    assert r.returncode != 0 and "JSON error" in r.stderr
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "--zip", "lz4"], tmp_path)
    assert r.returncode != 0 and "unsupported compression" in r.stderr
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "-F", "nifti"], tmp_path)
    assert r.returncode != 0 and "unsupported output format" in r.stderr
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "--implicit", "1"], tmp_path)
    assert r.returncode != 0 and "unsupported option" in r.stderr
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "-E", str(tmp_path / "none.mch")], tmp_path)
    assert r.returncode != 0
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "-G", "-1"], tmp_path)
    assert r.returncode != 0 and "CPU path" in r.stderr
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "-M", "H"], tmp_path)
    assert r.returncode != 0


def read_mch(path):
    raw = open(path, "rb").read()
    head = struct.unpack("4s7If I f I i I 2i", raw[:64])
    assert head[0] == b"MCXH"
    colcount, saved = head[4], head[7]
    data = np.frombuffer(raw[64:64 + 4 * colcount * saved], np.float32).reshape(saved, colcount)
    return head, data


def test_legacy_inp_input_and_json_dump(tmp_path):
    # mcx_loadconfig's line format (src/mmc_utils.c:2425-2670), checked
    # against the values written in tests/golden/onecube/onecube.inp, then
    # written back by --dumpjson (mcx_savejdata) and re-read unchanged
    r = run(["-f", os.path.join(GOLDEN, "onecube", "onecube.inp"), "--dumpjson", "dump.json"], tmp_path)
    assert r.returncode == 0, r.stderr
    doc = json.load(open(tmp_path / "dump.json"))
    assert doc["Session"]["Photons"] == 100 and doc["Session"]["RNGSeed"] == 17182818
    src = doc["Optode"]["Source"]
    assert src["Type"] == "pencil" and src["Pos"] == [2, 8, 0] and src["Dir"][:3] == [0, 0, 1]
    f = doc["Forward"]
    assert f["T0"] == 0 and f["T1"] == pytest.approx(5e-9) and f["Dt"] == pytest.approx(5e-10)
    assert doc["Mesh"]["MeshID"] == "onecube" and doc["Mesh"]["InitElem"] == 3
    assert [d["Pos"] + [d["R"]] for d in doc["Optode"]["Detector"]] == [[2, 6, 0, 1], [2, 4, 0, 1], [2, 2, 0, 1]]
    node = np.loadtxt(os.path.join(GOLDEN, "onecube", "node_onecube.dat"), skiprows=1)[:, 1:]
    elem = np.loadtxt(os.path.join(GOLDEN, "onecube", "elem_onecube.dat"), skiprows=1, dtype=int)[:, 1:]
    assert np.array_equal(np.array(doc["Mesh"]["MeshNode"]), node)
    assert np.array_equal(np.array(doc["Mesh"]["MeshElem"]), elem)
    med = doc["Domain"]["Media"][1]
    assert [med["mua"], med["mus"], med["g"], med["n"]] == pytest.approx([0.005, 1.010101, 0.01, 1.0], rel=1e-5)
    # the dump is a complete input: reading it back gives the same dump
    r = run(["-f", str(tmp_path / "dump.json"), "--dumpjson", "again.json"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert open(tmp_path / "again.json").read() == open(tmp_path / "dump.json").read()


@pytest.mark.gpu
def test_inp_and_its_json_dump_give_the_same_run(tmp_path):
    inp = os.path.join(GOLDEN, "onecube", "onecube.inp")
    assert run(["-f", inp, "--dumpjson", "d.json"], tmp_path).returncode == 0
    for name, src in (("a", inp), ("b", str(tmp_path / "d.json"))):
        r = run(["-f", src, "-n", "5000", "-M", "S", "-C", "0", "-F", "bin", "-s", name], tmp_path)
        assert r.returncode == 0, r.stderr
    a, b = np.fromfile(tmp_path / "a.bin"), np.fromfile(tmp_path / "b.bin")
    assert a.sum() > 0 and np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(a)


@pytest.mark.gpu
def test_cube2_json_element_output_and_detectors(tmp_path):
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "-n", "20000", "-M", "S", "-C", "0", "-d", "1",
             "-s", "c2"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "photon/ms" in r.stdout
    dat = np.loadtxt(tmp_path / "c2.dat")
    assert dat.shape == (6 * 50, 2)                   # 6 elements x 50 gates
    assert np.array_equal(dat[:6, 0], np.arange(1, 7))
    assert np.all(dat[:, 1] >= 0) and dat[:, 1].sum() > 0
    head, rec = read_mch(tmp_path / "c2.mch")
    assert head[2] == 1 and head[3] == 4 and head[4] == 2 * 1 + 2   # maxmedia, detnum, colcount
    assert rec.shape[0] == head[7] > 0
    assert set(np.unique(rec[:, 0]).astype(int)) <= {1, 2, 3, 4}
    assert np.all(rec[:, -1] == 1.0)                  # launch weight


@pytest.mark.gpu
def test_cube2_default_is_the_dual_grid(tmp_path):
    # the GPU builds default to -M G (src/mmc_utils.c:241-243): 60 mm cube,
    # 2 mm voxels -> int(60.000002 / 2) + 1 = 31 per axis, 50 gates
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "-n", "20000", "--gridsize", "2", "-F", "nii",
             "-s", "g2"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "dual grid: 31 x 31 x 31" in r.stdout
    raw = open(tmp_path / "g2.nii", "rb").read()
    assert struct.unpack_from("<8h", raw, 40)[:5] == (4, 31, 31, 31, 50)
    assert struct.unpack_from("<3f", raw, 80) == (2.0, 2.0, 2.0)
    f = np.frombuffer(raw[352:], np.float64).reshape(50, 31, 31, 31)
    assert np.all(f >= 0) and f[0].sum() > 0
    # the source at (30, 30, 0) points +z: the first gate's fluence peaks in
    # the voxel column under it, near the surface
    z, y, x = np.unravel_index(np.argmax(f[0]), f[0].shape)
    assert (x, y) in {(14, 14), (14, 15), (15, 14), (15, 15)} and z <= 2


@pytest.mark.gpu
def test_onecube_mesh_files_nodal_binary(tmp_path):
    r = run(["-f", os.path.join(GOLDEN, "onecube", "onecube.json"), "-n", "5000", "-M", "P", "-C", "1",
             "-F", "bin", "-s", "oc"], tmp_path)
    assert r.returncode == 0, r.stderr
    f = np.fromfile(tmp_path / "oc.bin", np.float64)
    assert f.shape == (8 * 10,)                       # 8 nodes x 10 gates
    assert np.all(f >= 0) and f.sum() > 0


@pytest.mark.gpu
def test_every_output_format_holds_the_same_field(tmp_path):
    # one run per -F format; the fields agree to the FP32 atomic summation order
    import json
    import zlib
    import base64
    fields = {}
    for fmt in ("bin", "nii", "hdr", "mc2", "tx3", "jnii"):
        r = run(["-f", os.path.join(GOLDEN, "onecube", "onecube.json"), "-n", "20000", "-M", "P", "-C", "1",
                 "-F", fmt, "-s", "oc_" + fmt], tmp_path)
        assert r.returncode == 0, r.stderr
        p = tmp_path / ("oc_" + fmt + "." + fmt)
        if fmt in ("bin", "mc2"):
            fields[fmt] = np.fromfile(p, np.float64)
        elif fmt == "nii":
            fields[fmt] = np.frombuffer(p.read_bytes()[352:], np.float64)
        elif fmt == "hdr":
            fields[fmt] = np.fromfile(tmp_path / "oc_hdr.img", np.float64)
        elif fmt == "tx3":
            fields[fmt] = np.frombuffer(p.read_bytes()[16:], np.float64)
        else:
            d = json.load(open(p))["NIFTIData"]
            assert d["_ArraySize_"] == [8, 10, 1]
            fields[fmt] = np.frombuffer(zlib.decompress(base64.b64decode(d["_ArrayZipData_"])), np.float64)
    ref = fields["bin"]
    assert ref.shape == (80,) and ref.sum() > 0
    for fmt, f in fields.items():
        assert np.allclose(f, ref, rtol=1e-5, atol=1e-7 * ref.max()), fmt


@pytest.mark.gpu
def test_saved_seeds_replay_the_detected_photons(tmp_path):
    # -q 1 appends each saved photon's seed; -E file.mch relaunches them
    # (src/mmc_mesh.c:815-898) and -O L deposits the weighted path length
    cube2 = os.path.join(GOLDEN, "cube2", "cube2.json")
    r = run(["-f", cube2, "-n", "20000", "-M", "S", "-C", "0", "-d", "1", "-q", "1", "-s", "orig"], tmp_path)
    assert r.returncode == 0, r.stderr
    head, rec = read_mch(tmp_path / "orig.mch")
    n, col = head[7], head[4]
    raw = (tmp_path / "orig.mch").read_bytes()
    assert head[9] == 16 and len(raw) == 64 + 4 * n * col + 16 * n     # seedbyte, seed block
    r = run(["-f", cube2, "-M", "S", "-C", "0", "-d", "1", "-E", str(tmp_path / "orig.mch"), "-s", "rep"], tmp_path)
    assert r.returncode == 0, r.stderr
    head2, rec2 = read_mch(tmp_path / "rep.mch")
    assert head2[7] == n                                   # every replayed photon is detected again
    key = lambda a: a[np.lexsort(a.T[::-1])]
    assert np.array_equal(key(rec2), key(rec))
    r = run(["-f", cube2, "-M", "S", "-C", "0", "-E", str(tmp_path / "orig.mch"), "-O", "L", "-U", "0",
             "-F", "bin", "-s", "wl"], tmp_path)
    assert r.returncode == 0, r.stderr
    f = np.fromfile(tmp_path / "wl.bin", np.float64)
    # one medium: replay weight w0*exp(-mua*ppath), deposit = weight * path
    w = rec[:, -1].astype(np.float64) * np.exp(-0.005 * rec[:, 2].astype(np.float64))
    assert abs(f.sum() / (w * rec[:, 2]).sum() - 1) < 1e-4


@pytest.mark.gpu
def test_adjoint_jacobians_from_the_command_line(tmp_path):
    # -O w on the dual grid: the source, then each detector (given a Dir) run
    # as a source; <session>_jmua and _jd hold [s][d][voxel] float values
    cfg = json.load(open(os.path.join(GOLDEN, "cube2", "cube2.json")))
    cfg["Optode"]["Detector"] = [{"Pos": [30.0, 20.0, 0.0], "R": 0.0, "Dir": [0, 0, 1]},
                                 {"Pos": [30.0, 40.0, 0.0], "R": 0.0, "Dir": [0, 0, 1]}]
    js = tmp_path / "adj.json"
    js.write_text(json.dumps(cfg))
    r = run(["-f", str(js), "-n", "60000", "--gridsize", "2", "-O", "w", "-F", "bin", "-s", "adj"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("adjoint slot") == 3
    jm = np.fromfile(tmp_path / "adj_jmua.bin", np.float64).reshape(2, 31, 31, 31)
    jd = np.fromfile(tmp_path / "adj_jd.bin", np.float64).reshape(2, 31, 31, 31)
    assert np.all(jm <= 0) and jm.min() < 0 and np.abs(jd).max() > 0
    # each pair is most sensitive between its source and detector, near the surface
    for d, y in ((0, 20.0), (1, 40.0)):
        z, yy, x = np.unravel_index(np.argmin(jm[d]), jm[d].shape)
        assert z <= 3 and abs(x * 2 - 30) <= 6 and min(yy * 2, 60) <= max(30, y) + 4
    r = run(["-f", str(js), "-n", "30000", "--gridsize", "2", "-O", "a", "-F", "bin", "-s", "only"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "only_jmua.bin").exists() and not (tmp_path / "only_jd.bin").exists()
    del cfg["Optode"]["Detector"][1]["Dir"]
    js.write_text(json.dumps(cfg))
    r = run(["-f", str(js), "-n", "1000", "-O", "a", "-s", "bad"], tmp_path)
    assert r.returncode != 0 and "Dir" in r.stderr


@pytest.mark.gpu
def test_momentum_transfer_columns_in_the_mch(tmp_path):
    # -m 1: a third block of maxmedia columns after the partial paths, and the
    # M bit in savedetflag
    r = run(["-f", os.path.join(GOLDEN, "cube2", "cube2.json"), "-n", "20000", "-M", "S", "-C", "0", "-d", "1",
             "-m", "1", "-s", "mom"], tmp_path)
    assert r.returncode == 0, r.stderr
    head, data = read_mch(tmp_path / "mom.mch")
    prop = head[2]
    assert head[4] == 1 + 3 * prop + 1 and head[-3] & 0x08
    nscat, mom = data[:, 1:1 + prop], data[:, 1 + 2 * prop:1 + 3 * prop]
    assert len(data) > 0 and np.all(mom >= 0) and np.all(mom <= 2 * nscat + 1e-3) and mom.sum() > 0


@pytest.mark.gpu
def test_area_detector_image(tmp_path):
    # -x 2 with Param1/Param2 on the first detector: <session>.img holds the
    # detected weights binned by exit pixel and gate (float [gate][y][x])
    cfg = json.load(open(os.path.join(GOLDEN, "cube2", "cube2.json")))
    cfg["Optode"]["Detector"][0].update({"Param1": [20.0, 0.0, 0.0, 10], "Param2": [0.0, 20.0, 0.0, 5]})
    cfg["Optode"]["Detector"][0]["Pos"] = [20.0, 20.0, 0.0]
    cfg["Optode"]["Detector"][0]["R"] = 15.0
    js = tmp_path / "img.json"
    js.write_text(json.dumps(cfg))
    r = run(["-f", str(js), "-n", "50000", "-M", "S", "-C", "0", "-d", "1", "-x", "2", "-s", "img"], tmp_path)
    assert r.returncode == 0, r.stderr
    head, data = read_mch(tmp_path / "img.mch")
    img = np.fromfile(tmp_path / "img.img", np.float32)
    maxgate = 50
    assert img.size == maxgate * 5 * 10 and img.sum() > 0
    prop = head[2]
    x, y = data[:, -7], data[:, -6]
    # pixel index truncated toward zero, as the reference's float-to-int conversion
    xi = np.trunc((x - np.float32(20)) / np.float32(20) * np.float32(10))
    yi = np.trunc((y - np.float32(20)) / np.float32(20) * np.float32(5))
    inside = (xi >= 0) & (xi <= 9) & (yi >= 0) & (yi <= 4)
    med = np.array([m["mua"] for m in cfg["Domain"]["Media"]], np.float32)
    w = data[:, -1] * np.exp(-(data[:, 1 + prop:1 + 2 * prop] * med[1:prop + 1]).sum(1))
    assert img.sum() == pytest.approx(w[inside].sum(), rel=1e-4)


def test_inp_wide_field_detector_plane(tmp_path):
    # one detector of radius 0 is a wide-field plane: two axis lines (pixel
    # counts in w) and the pattern file follow (src/mmc_utils.c:2612-2660)
    import shutil
    for f in ("node_onecube.dat", "elem_onecube.dat", "face_onecube.dat", "prop_onecube.dat"):
        src = os.path.join(GOLDEN, "onecube", f)
        if os.path.exists(src):
            shutil.copy(src, tmp_path / f)
    inp = tmp_path / "wf.inp"
    inp.write_text("100\n17182818\n2 8 0.0\n0. 0. 1.\n0.e+00 5.e-09 5e-10\nonecube\n3\n1 0\n0.0 0.0 0.0\n"
                   "pencil\n0 0 0 0\n0 0 0 0\n4 0 0 8\n0 6 0 3\ndetpat.bin\n")
    r = run(["-f", str(inp), "--dumpjson", "wf.json"], tmp_path)
    assert r.returncode == 0, r.stderr
    d = json.load(open(tmp_path / "wf.json"))["Optode"]["Detector"][0]
    assert d["R"] == 0 and d["Param1"] == [4, 0, 0, 8] and d["Param2"] == [0, 6, 0, 3]
