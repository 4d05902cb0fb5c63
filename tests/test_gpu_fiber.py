"""GPU parity of the UniformFiberNI source (xopto/mcml/mcsource/fiberni.py):
bit-exact against the oracle in the layered kernel and in the batch
engine, and the reference's own acceptance test at its full size — 400 runs
of 1e7 packets against the CUDAMCML fixtures of
xopto/mcml/test/validate.py:348-511 (single layer) and :692-850 (double
layer), passing when the mean relative error of the fiber-collected
reflectance is below 0.5 %."""
import os

import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd import workloads as W
from pyxopto_amd.mcml import mc
from pyxopto_amd.mcml.batch import Batch, Instance
from pyxopto_amd.util.convolve import fiber_reflectance
from oracle import pyoracle as O

from helpers import XINIT

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def _fiber_sim_with_specular(mua=300.0, musr=2000.0):
    D = mc.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 1.7e-3, 340),
                                   cosmin=0.98845),
                      bottom=D.Total(), specular=D.Total())
    sim = mc.Mc(W.fiber_layers(mua, musr), W.fiber_source(), det,
                rnginit=XINIT)
    sim.rmax = 5e-3
    return sim


def test_fiber_static_schedule_bit_exact():
    sim = _fiber_sim_with_specular()
    lanes, npk = 8192, 60000
    cfg, lut, naccum = sim.pack()
    ctx = N.Context(0, 0, XINIT)
    gpu, rep = ctx.run_mcml(cfg, lut, npk, naccum, lanes,
                            N.XO_FLAG_STATIC_SCHEDULE)
    gx, _ = ctx.get_rng(lanes)
    x, a = N.init_rng(lanes, XINIT)
    ref, counts, ox = O.mcml_run(cfg, lut, x, a, npk, naccum, O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(gpu, ref)
    np.testing.assert_array_equal(gx, ox)
    assert gpu[-1] > 0  # the specular detector saw the core reflection


def test_fiber_batch_bit_exact():
    insts = [Instance(W.fiber_layers(mua, musr), W.fiber_source(),
                      W.fiber_detectors(), nphotons=20000 + 997 * k,
                      seed=77 + k, rmax=5e-3)
             for k, (mua, musr) in enumerate([(10.0, 800.0), (500.0, 3000.0),
                                              (2000.0, 6000.0)])]
    insts.append(Instance(W.fiber_layers(100.0, 1000.0), mc.mcsource.Line(),
                          W.fiber_detectors(), nphotons=15000, seed=5,
                          rmax=5e-3))
    b = Batch(insts, rnginit=XINIT)
    arr, lut, naccum = b.pack()
    acc, rep = b.context.run_batch(arr, lut, b.opts(chunk=512), naccum)
    ref, counts = O.batch_run(arr, lut, naccum, math_mode=O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(acc, ref)


def _fixture_run(name, layers_for):
    f = np.load(os.path.join(GOLDEN, name + '.npz'))
    mua, musr = np.meshgrid(f['lut_mua'], f['lut_musr'], indexing='ij')
    insts = [Instance(layers_for(f, m, s), W.fiber_source(),
                      W.fiber_detectors(), nphotons=W.FIBER_PACKETS,
                      seed=1000 + k, rmax=float(f['rmax']))
             for k, (m, s) in enumerate(zip(mua.ravel(), musr.ravel()))]
    b = Batch(insts, rnginit=XINIT)
    res, rep = b.run(b.opts(chunk=8192, block=256))
    r = res[0].top.r
    R = np.array([d.top.reflectance for d in res])
    sim = fiber_reflectance(r, R, f['sds'], float(f['fiber_dcore']))
    err = (sim - f['reflectance'].reshape(sim.shape)) / \
        f['reflectance'].reshape(sim.shape)
    print('{}: mean relative error {:.3f} %, max |error| {:.2f} %, '
          '{:.2f} s for {} packets'.format(name, 100 * err.mean(),
                                          100 * np.abs(err).max(),
                                          rep.kernel_ms / 1e3, rep.packets))
    return err


def test_single_layer_fiber_fixture_full_size():
    g = 0.85

    def layers(f, m, s):
        return W.fiber_layers(m, s, g)
    err = _fixture_run('cuda_singlelayer_uniformfiber', layers)
    assert abs(err.mean()) < 0.005


def test_double_layer_fiber_fixture_full_size():
    def layers(f, m, s):
        L = mc.mclayer
        g = f['layer_g']
        n = f['layer_n']
        pf = [mc.mcpf.Hg(float(v)) for v in g]
        return L.Layers([
            L.Layer(d=np.inf, n=float(n[0]), mua=0.0, mus=0.0, pf=pf[0]),
            L.Layer(d=float(f['layer_d'][1]), n=float(n[1]),
                    mua=float(f['layer1_mua']),
                    mus=float(f['layer1_musr']) / (1.0 - g[1]), pf=pf[1]),
            L.Layer(d=float(f['layer_d'][2]), n=float(n[2]), mua=float(m),
                    mus=float(s) / (1.0 - g[2]), pf=pf[2]),
            L.Layer(d=np.inf, n=float(n[3]), mua=0.0, mus=0.0, pf=pf[3])])
    err = _fixture_run('cuda_doublelayer_uniformfiber', layers)
    assert abs(err.mean()) < 0.005


@pytest.mark.parametrize('direction', [(0.0, 0.0, 1.0), (0.3, -0.2, 0.9)])
def test_uniform_fiber_static_schedule_bit_exact(direction):
    fib = mc.fiberutil.MultimodeFiber(400e-6, 420e-6, 1.45, 0.22)
    src = mc.mcsource.UniformFiber(fib, position=(1e-4, 0.0, 0.0),
                                   direction=direction)
    D = mc.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 3e-3, 60)),
                      bottom=D.Total(), specular=D.Total())
    sim = mc.Mc(W.fiber_layers(300.0, 2000.0), src, det, rnginit=XINIT)
    sim.rmax = 5e-3
    lanes, npk = 8192, 60000
    cfg, lut, naccum = sim.pack()
    ctx = N.Context(0, 0, XINIT)
    gpu, rep = ctx.run_mcml(cfg, lut, npk, naccum, lanes,
                            N.XO_FLAG_STATIC_SCHEDULE)
    x, a = N.init_rng(lanes, XINIT)
    ref, counts, _ = O.mcml_run(cfg, lut, x, a, npk, naccum, O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(gpu, ref)
    assert gpu[-1] > 0
    ctx.close()


def test_uniform_fiber_batch_bit_exact():
    """Angled UniformFiber instances, mixed with a UniformFiberNI one, in
    the batch engine: bit-exact with the oracle's batch run."""
    fib = mc.fiberutil.MultimodeFiber(400e-6, 420e-6, 1.45, 0.22)
    D = mc.mcdetector
    dirs = [(0.0, 0.0, 1.0), (0.3, -0.2, 0.9), (-0.5, 0.1, 0.8)]
    insts = []
    for k, d in enumerate(dirs):
        src = mc.mcsource.UniformFiber(fib, position=(1e-4 * k, 0.0, 0.0),
                                       direction=d)
        det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 3e-3, 60)),
                          bottom=D.Total(), specular=D.Total())
        insts.append(Instance(W.fiber_layers(300.0 + 200 * k, 2000.0), src,
                              det, nphotons=15000 + 313 * k, seed=91 + k,
                              rmax=5e-3))
    insts.append(Instance(W.fiber_layers(100.0, 1000.0), W.fiber_source(),
                          W.fiber_detectors(), nphotons=12000, seed=5,
                          rmax=5e-3))
    b = Batch(insts, rnginit=XINIT)
    arr, lut, naccum = b.pack()
    acc, rep = b.context.run_batch(arr, lut, b.opts(chunk=512), naccum)
    ref, counts = O.batch_run(arr, lut, naccum, math_mode=O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(acc, ref)
    # the angled fibers' specular totals are non-zero
    for k in range(len(dirs)):
        base, n = int(arr[k].accum_base), int(arr[k].accum_count)
        assert acc[base + n - 1] > 0
