"""Wide-field detector elements (mesh label -2; mesh_srcdetelem and
mesh_loadmedia, src/mmc_mesh.c:390-427, 511-527; detection at
src/mmc_raytrace.c:2013-2014): the elements take the background's optical
properties, and a photon that leaves the mesh through one of them is detected
with that element's id as its detector id.

Pin: a layer of detector elements on top of a tissue box changes nothing
inside the tissue (the tissue/detector face refracts and reflects exactly as
the tissue/outside face with the same outside index), so the photons the
oracle detects through the layer are exactly those that leave the bare box
through its top face, with the same scattering counts and partial paths.
The path a photon then travels through detector elements lands in record
column 2 prop + 1, as in the reference (src/mmc_raytrace.c:1869 with type
prop+1): without momentum or exit columns that is the launch-weight column,
which becomes 1 + the straight flight through the layer.
"""
import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import meshgen
from mmc_amd.replay import seeds_of

MED = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.8, 1.37]], np.float32)


def box(nz, det_layer):
    node, elem = meshgen.t6_lattice(10, 10, nz, 1.0)
    etype = np.ones(len(elem), np.int32)
    if det_layer:
        cz = meshgen.centroids(node, elem)[:, 2]
        etype[cz > nz - 1] = -2
    return node, elem, etype, MED


def sim(n=4000):
    return M.SimConfig(nphoton=n, seed=8, method="S", srcpos=(5.25, 4.75, 0.0), srcdir=(0.0, 0.0, 1.0),
                       tstart=0.0, tend=5e-9, tstep=5e-9, outputtype="x", issavedet=1,
                       detpos=np.array([[5.0, 5.0, 0.0, 1.0]], np.float32), maxdetphoton=100000)


def test_detector_layer_detects_the_photons_leaving_through_the_top():
    cfg = sim()
    sd = seeds_of(cfg.seed, np.arange(cfg.nphoton))
    with_det = O.OracleMesh(*box(6, True), nodal=False)
    bare = O.OracleMesh(*box(5, False), nodal=False)
    assert with_det.isextdet == 1 and with_det.etype.max() == 2
    od = O.run(with_det, cfg, photonseed=sd, records=(0, cfg.nphoton), maxdet=cfg.nphoton, math="portable")
    cfg_bare = sim()
    cfg_bare.detpos = None                       # the bare box: no detectors at all
    cfg_bare.issavedet = 0
    ob = O.run(bare, cfg_bare, photonseed=sd, records=(0, cfg.nphoton), math="portable")
    p = ob.photons["p"]
    inner = lambda v: (v > 1e-4) & (v < 10 - 1e-4)
    top = (ob.photons["reason"] == 1) & (np.abs(p[:, 2] - 5.0) < 1e-4) & inner(p[:, 0]) & inner(p[:, 1])
    det = od.photons["detid"] > 0
    assert top.sum() > 100
    assert np.array_equal(det, top)
    # detector ids are the detector elements the photons left through
    assert np.all(with_det.etype[od.photons["detid"][det] - 1] == 2)
    assert od.detected == det.sum() and len(od.detdata) == det.sum()
    # records: detector id, scattering count, partial path (tissue 1) and
    # launch weight plus the path through the detector layer, which the
    # photon crosses in a straight line from where it left the bare box
    rec = od.detdata[np.argsort(od.detphoton)]
    ids = np.sort(od.detphoton)
    assert np.array_equal(rec[:, 0].astype(np.int64), od.photons["detid"][ids])
    assert np.array_equal(rec[:, 1], ob.photons["nscat"][ids])
    flight = np.linalg.norm(od.photons["p"][ids].astype(np.float64) - ob.photons["p"][ids], axis=1)
    assert np.median(flight) > 1.0                       # most leave through the top, some the sides
    assert np.allclose(rec[:, 3], 1.0 + flight, rtol=0, atol=1e-4)
    # the photons' paths through the tissue are unchanged by the layer
    assert np.array_equal(od.photons["nscat"], ob.photons["nscat"])
    assert np.array_equal(od.photons["w"][det], ob.photons["w"][top])


@pytest.mark.gpu
@pytest.mark.parametrize("cols", ["plain", "momentum", "exit"])
def test_gpu_detector_elements_match_the_oracle(cols):
    # the detector-layer path lands in the weight, first momentum or (then
    # rewritten) first exit column
    arrays = box(6, True)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    om = O.OracleMesh(*arrays, nodal=False)
    assert tm.isextdet == 1
    cfg = sim(20000)
    cfg.ismomentum = int(cols == "momentum")
    cfg.issaveexit = int(cols == "exit")
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              maxdet=cfg.nphoton, math="portable")
    assert np.array_equal(g.photons["detid"], o.photons["detid"])
    assert np.array_equal(g.photons["eid"], o.photons["eid"])
    assert g.detected == o.detected > 0
    gi, oi = np.argsort(g.detphoton), np.argsort(o.detphoton)
    assert np.array_equal(g.detphoton[gi], o.detphoton[oi])
    assert np.allclose(g.detdata[gi], o.detdata[oi], rtol=1e-5, atol=1e-5)
    # and against the reference's arithmetic (the host C library's libm)
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              maxdet=cfg.nphoton, math="glibc")
    assert (g.photons["eid"] == o.photons["eid"]).mean() >= 0.999
    assert (g.photons["detid"] == o.photons["detid"]).mean() >= 0.999
