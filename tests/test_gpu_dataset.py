"""Dataset files from the GPU engines: the batch engine, the server and the
surrogate's data generator write records in the reference's layouts, and
reading them back gives the simulated values unchanged."""
import numpy as np
import pytest

from pyxopto_amd import dataset, serve, surrogate
from pyxopto_amd.mcml import mc
from pyxopto_amd.mcml.batch import Batch, Instance
from pyxopto_amd.mcvox import mc as mcv

from helpers import XINIT

pytestmark = pytest.mark.gpu


def _instance(mua, musr, k):
    pf = mc.mcpf.Hg(0.8)
    L, D = mc.mclayer, mc.mcdetector
    layers = [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=pf),
              L.Layer(d=np.inf, n=1.33, mua=mua, mus=musr / 0.2, pf=pf),
              L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=pf)]
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 5e-3, 50)),
                      bottom=D.Total(), specular=D.Total())
    return Instance(layers, mc.mcsource.Line(), det, nphotons=50000,
                    seed=100 + k)


def test_batch_results_written_and_read_back(tmp_path):
    grid = [(10.0, 500.0), (100.0, 1000.0), (1000.0, 2000.0)]
    insts = [_instance(m, s, k) for k, (m, s) in enumerate(grid)]
    b = Batch(insts, rnginit=XINIT)
    res, _ = b.run()
    for (mua, musr), inst, det in zip(grid, insts, res):
        path = dataset.save(str(tmp_path / dataset.mcml_filename(mua, musr)),
                            dataset.mcml_record(inst, det))
        back = dataset.load(path)
        np.testing.assert_array_equal(back['reflectance'],
                                      det.top.reflectance)
        assert back['total_reflectance'] == pytest.approx(
            det.top.raw.sum() / det.top.nphotons, rel=0, abs=0)
        objs = dataset.objects(back)
        assert objs['layers'][1].mua == mua
        # the configuration read back reproduces the run
        again = Batch([Instance(objs['layers'], objs['source'],
                                objs['detectors'], nphotons=50000,
                                seed=inst.seed)], rnginit=XINIT)
        res2, _ = again.run()
        np.testing.assert_array_equal(res2[0].top.raw, det.top.raw)


def test_served_requests_to_records(tmp_path):
    inst = _instance(50.0, 800.0, 7)
    G = mcv.mcgeometry
    M = mcv.mcmaterial
    mats = M.Materials([M.Material(1.0, 0.0, 0.0, mcv.mcpf.Hg(0.9)),
                        M.Material(1.4, 100.0, 1e4, mcv.mcpf.Hg(0.9))])
    vox = G.Voxels(G.Axis(-1e-3, 1e-3, 10), G.Axis(-1e-3, 1e-3, 10),
                   G.Axis(0.0, 2e-3, 10))
    A = mcv.mcfluence.Axis
    flu = mcv.mcfluence.Fluence(A(-1e-3, 1e-3, 10), A(-1e-3, 1e-3, 10),
                                A(0.0, 2e-3, 10))
    D = mcv.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 1e-3, 10)),
                      bottom=D.Total())
    vsim = mcv.Mc(vox, mats, mcv.mcsource.Line(), det, fluence=flu,
                  rnginit=XINIT)
    vsim.voxels.material[:] = 1
    with serve.Server() as srv:
        r1 = srv.submit_layered(inst)
        r2 = srv.submit_voxel(vsim, 40000, seed=3)
        lres = r1.result(timeout=120)
        _, fres, dres = r2.result(timeout=120)
    p1 = dataset.save(str(tmp_path / 'layered.npz'),
                      dataset.mcml_record(inst, lres))
    p2 = dataset.save(str(tmp_path / dataset.mcvox_fluence_filename(
        200e-6, 500e-6)), dataset.mcvox_fluence_record(vsim, fres, dres,
                                                       40000))
    b1, b2 = dataset.load(p1), dataset.load(p2)
    np.testing.assert_array_equal(b1['reflectance'], lres.top.reflectance)
    np.testing.assert_array_equal(b2['fluence_data'], fres.data)
    assert b2['fluence_data'].sum() > 0


def test_surrogate_data_generator_writes_dataset_files(tmp_path):
    rows = surrogate.sample_box(4, seed=5)
    names = surrogate.write_dataset(str(tmp_path), rows, 20000, seed=9)
    R, _ = surrogate.simulate(rows, 20000, seed=9)
    assert len(set(names)) == 4
    for k, name in enumerate(names):
        back = dataset.load(name)
        assert back['method'] == 'aw'
        # the same keyed packets as the training-data path
        np.testing.assert_array_equal(back['reflectance'], R[k])
        layers = dataset.objects(back)['layers']
        assert layers[1].mua == pytest.approx(np.exp(rows[k][0]))
