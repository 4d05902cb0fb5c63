"""GPU parity of the stepping methods and fluence accumulators of the
layered (xo_mcml_run) and voxelized (xo_mcvox_run) kernels: bit-exact with
the oracle under the static schedule for albedo weight, albedo rejection and
microscopic Beer-Lambert with Cartesian and r-z fluence, and statistical
parity of the production (dynamic) schedule."""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcml import mc
from oracle import pyoracle as O

from helpers import XINIT, batch_mean_se, oracle_batches, z_scores
from test_gpu_mcvox import skin_vessel_sim

pytestmark = pytest.mark.gpu

METHODS = [N.XO_METHOD_AW, N.XO_METHOD_AR, N.XO_METHOD_MBL]


def _sim(method, geometry):
    L, D = mc.mclayer, mc.mcdetector
    pf = mc.mcpf.Hg(0.85)
    layers = L.Layers([
        L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=pf),
        L.Layer(d=5e-4, n=1.37, mua=500.0, mus=1.5e4, pf=pf),
        L.Layer(d=1e-3, n=1.40, mua=50.0, mus=2e4, pf=pf),
        L.Layer(d=np.inf, n=1.45, mua=0.0, mus=0.0, pf=pf)])
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 3e-3, 60)),
                      bottom=D.Total(), specular=D.Total())
    A = mc.mcfluence.Axis
    if geometry == 'xyz':
        flu = mc.mcfluence.Fluence(A(-2e-3, 2e-3, 40), A(-2e-3, 2e-3, 40),
                                   A(0.0, 1.5e-3, 30))
    else:
        flu = mc.mcfluence.FluenceRz(A(0.0, 2e-3, 40), A(0.0, 1.5e-3, 30))
    return mc.Mc(layers, mc.mcsource.Line(), det, fluence=flu,
                 options=[mc.mcoptions.McMethod(method)], rnginit=XINIT)


@pytest.mark.parametrize('method', METHODS)
@pytest.mark.parametrize('geometry', ['xyz', 'rz'])
def test_methods_static_schedule_bit_exact(method, geometry):
    print('case', method, geometry, flush=True)
    sim = _sim(method, geometry)
    cfg, lut, naccum = sim.pack()
    lanes, npk = 8192, 40000
    ctx = N.Context(0, 0, XINIT)
    gpu, rep = ctx.run_mcml(cfg, lut, npk, naccum, lanes,
                            N.XO_FLAG_STATIC_SCHEDULE)
    gx, _ = ctx.get_rng(lanes)
    x, a = N.init_rng(lanes, XINIT)
    ref, counts, rx = O.mcml_run(cfg, lut, x, a, npk, naccum, O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    assert int(rep.deposits) == int(counts[3])
    np.testing.assert_array_equal(gpu, ref)
    np.testing.assert_array_equal(gx, rx)
    assert gpu[cfg.fluence.offset:].sum() > 0  # fluence bins were reached
    ctx.close()


@pytest.mark.parametrize('method', [N.XO_METHOD_AR, N.XO_METHOD_MBL])
def test_methods_dynamic_schedule_statistics(method):
    sim = _sim(method, 'rz')
    cfg, lut, naccum = sim.pack()
    k, npk = 40, 20000
    ctx = N.Context(0, 0, XINIT)
    g = np.array([ctx.run_mcml(cfg, lut, npk, naccum)[0] for _ in range(k)],
                 dtype=np.float64) / N.XO_ACCU_K / npk

    def one(x, a):
        acc, _, _ = O.mcml_run(cfg, lut, x, a, npk, naccum, O.MATH_LIBM)
        return acc.astype(np.float64) / N.XO_ACCU_K / npk
    o = oracle_batches(one, k, 256)
    mg, sg = batch_mean_se(g)
    mo, so = batch_mean_se(o)
    ok = (g > 0).all(axis=0) & (o > 0).all(axis=0)
    z = z_scores(mg[ok], sg[ok], mo[ok], so[ok])
    assert np.abs(z).max() < 4.5, np.abs(z).max()
    for i in (60, 61):  # transmittance, specular (after 60 radial bins)
        assert abs(mg[i] - mo[i]) <= 3 * np.hypot(sg[i], so[i]) + 1e-12
    ctx.close()


@pytest.mark.parametrize('method', [N.XO_METHOD_AR, N.XO_METHOD_MBL])
@pytest.mark.parametrize('geometry', ['xyz', 'rz'])
def test_voxel_methods_static_schedule_bit_exact(method, geometry):
    print('case voxel', method, geometry, flush=True)
    sim = skin_vessel_sim(method=method, rz=geometry == 'rz')
    cfg, lut, naccum = sim.pack()
    assert cfg.method == method
    lanes, npk = 8192, 40000
    ctx = N.Context(0, 0, XINIT)
    gpu, rep = ctx.run_mcvox(cfg, sim.voxels.material, lut, npk, naccum,
                             lanes, N.XO_FLAG_STATIC_SCHEDULE)
    gx, _ = ctx.get_rng(lanes)
    x, a = N.init_rng(lanes, XINIT)
    ref, counts, rx = O.mcvox_run(cfg, sim.voxels.material, lut, x, a, npk,
                                  naccum, O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    assert int(rep.voxel_crossings) == int(counts[2])
    assert int(rep.deposits) == int(counts[3])
    np.testing.assert_array_equal(gpu, ref)
    np.testing.assert_array_equal(gx, rx)
    assert gpu[cfg.fluence.offset:].sum() > 0
    ctx.close()


def test_voxel_mbl_dynamic_schedule_statistics():
    sim = skin_vessel_sim(n=20, method=N.XO_METHOD_MBL, rz=True)
    cfg, lut, naccum = sim.pack()
    vox = sim.voxels.material
    k, npk = 40, 10000
    ctx = N.Context(0, 0, XINIT)
    g = np.array([ctx.run_mcvox(cfg, vox if i == 0 else None, lut, npk,
                                naccum)[0] for i in range(k)],
                 dtype=np.float64) / N.XO_ACCU_K / npk

    def one(x, a):
        acc, _, _ = O.mcvox_run(cfg, vox, lut, x, a, npk, naccum, O.MATH_LIBM)
        return acc.astype(np.float64) / N.XO_ACCU_K / npk
    o = oracle_batches(one, k, 256)
    mg, sg = batch_mean_se(g)
    mo, so = batch_mean_se(o)
    ok = (g > 0).all(axis=0) & (o > 0).all(axis=0)
    z = z_scores(mg[ok], sg[ok], mo[ok], so[ok])
    assert np.abs(z).max() < 4.5, np.abs(z).max()
    ctx.close()


@pytest.mark.parametrize('kernel,method,geometry', [
    ('mcml', N.XO_METHOD_AW, 'xyz'), ('mcml', N.XO_METHOD_MBL, 'rz'),
    ('mcml', N.XO_METHOD_AR, 'rz'), ('mcvox', N.XO_METHOD_AW, 'xyz'),
    ('mcvox', N.XO_METHOD_MBL, 'xyz'), ('mcvox', N.XO_METHOD_AR, 'rz')])
def test_fluence_rate_static_schedule_bit_exact(kernel, method, geometry):
    """Fluence-rate mode (weight / mua of the current layer or material,
    fluence.py:139-141) in both kernels, the voxel one also through its LDS
    box of Cartesian bins: bit-exact with the oracle."""
    print('case fluence rate', kernel, method, geometry, flush=True)
    lanes, npk = 8192, 40000
    ctx = N.Context(0, 0, XINIT)
    x, a = N.init_rng(lanes, XINIT)
    if kernel == 'mcml':
        sim = _sim(method, geometry)
        cfg, lut, naccum = sim.pack()
        cfg.fluence.rate = 1
        gpu, rep = ctx.run_mcml(cfg, lut, npk, naccum, lanes,
                                N.XO_FLAG_STATIC_SCHEDULE)
        ref, counts, _ = O.mcml_run(cfg, lut, x, a, npk, naccum,
                                    O.MATH_KERNEL)
    else:
        sim = skin_vessel_sim(method=method, rz=geometry == 'rz')
        cfg, lut, naccum = sim.pack()
        cfg.fluence.rate = 1
        vox = sim.voxels.material
        gpu, rep = ctx.run_mcvox(cfg, vox, lut, npk, naccum, lanes,
                                 N.XO_FLAG_STATIC_SCHEDULE)
        ref, counts, _ = O.mcvox_run(cfg, vox, lut, x, a, npk, naccum,
                                     O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(gpu, ref)
    assert gpu[cfg.fluence.offset:].sum() > 0
    ctx.close()
