"""Area-detector post-processing of detected records (host library):
mmc_host_detimage (-x 2, mesh_getdetimage, src/mmc_mesh.c:2029-2078) and
mmc_host_detweight (mesh_getdetweight, src/mmc_mesh.c:2121-2138).

Pinned by hand-built records: pixel = floor((exit - origin) / span * pixels),
gate = floor(int(sum n_j L_j) * R_C0 / tstep) clamped to the last gate, value
w0 * prod exp(-mua_j L_j unitinmm); records off the plane are skipped (image)
or refused (pattern weight)."""
import ctypes as C

import numpy as np

from mmc_amd import _lib

R_C0 = np.float32(3.335640951981520e-12)


def rec(detid, ppath, exitxy, w0, prop=2):
    r = np.zeros(2 * prop + 8, np.float32)
    r[0] = detid
    r[1 + prop:1 + 2 * prop] = ppath
    r[-7], r[-6] = exitxy
    r[-1] = w0
    return r


def test_detector_image_and_pattern_weight():
    L = _lib.lib()
    prop = 2
    med = np.array([[0, 0, 1, 1], [0.01, 1, 0.9, 1.4], [0.02, 1, 0.9, 1.3]], np.float32)
    rows = np.stack([rec(1, [10.0, 5.0], (1.5, 0.5), 1.0), rec(1, [30.0, 10.0], (3.9, 3.2), 0.5),
                     rec(1, [1.0, 1.0], (5.5, 1.0), 1.0),                      # outside: x beyond the plane
                     rec(1, [2000.0, 2000.0], (0.1, 0.1), 1.0)]).astype(np.float32)   # late: last gate
    p1 = np.array([4.0, 0, 0, 4], np.float32)      # 4 mm, 4 pixels along x
    p2 = np.array([0, 4.0, 0, 2], np.float32)      # 4 mm, 2 pixels along y
    org = np.array([0.0, 0.0], np.float32)
    tstep, maxgate, unit = np.float32(1e-10), 5, np.float32(1.0)
    img = np.zeros((maxgate, 2, 4), np.float32)
    rc = L.mmc_host_detimage(_lib.ptr(img, C.c_float), _lib.ptr(rows, C.c_float), len(rows), rows.shape[1], prop,
                             _lib.ptr(med, C.c_float), C.c_float(unit), C.c_float(tstep), maxgate,
                             _lib.ptr(org, C.c_float), _lib.ptr(p1, C.c_float), _lib.ptr(p2, C.c_float))
    assert rc == 0
    want = np.zeros_like(img)
    for r in rows:
        pp = r[1 + prop:1 + 2 * prop].astype(np.float64)
        x, y = int((r[-7] - 0) / 4 * 4), int((r[-6] - 0) / 4 * 2)
        if not (0 <= x < 4 and 0 <= y < 2):
            continue
        path = int(pp[0] * 1.4 + pp[1] * 1.3)
        g = min(int(path * 3.335640951981520e-12 / 1e-10), maxgate - 1)
        want[g, y, x] += r[-1] * np.exp(-(pp * med[1:, 0]).sum())
    assert np.allclose(img, want, rtol=1e-5, atol=0)
    assert img[0, 0, 1] > 0 and img[1, 1, 3] > 0 and img[maxgate - 1, 0, 0] > 0 and np.count_nonzero(img) == 3
    # pattern weight at the exit pixel; a record off the plane is refused
    pat = np.arange(8, dtype=np.float32) + 1
    w = np.zeros(len(rows), np.float32)
    inside = rows[[0, 1, 3]]
    rc = L.mmc_host_detweight(_lib.ptr(w, C.c_float), _lib.ptr(inside, C.c_float), len(inside), rows.shape[1],
                              _lib.ptr(org, C.c_float), _lib.ptr(p1, C.c_float), _lib.ptr(p2, C.c_float),
                              _lib.ptr(pat, C.c_float))
    assert rc == 0 and list(w[:3]) == [pat[0 * 4 + 1], pat[1 * 4 + 3], pat[0]]
    rc = L.mmc_host_detweight(_lib.ptr(w, C.c_float), _lib.ptr(rows, C.c_float), len(rows), rows.shape[1],
                              _lib.ptr(org, C.c_float), _lib.ptr(p1, C.c_float), _lib.ptr(p2, C.c_float),
                              _lib.ptr(pat, C.c_float))
    assert rc == -2
