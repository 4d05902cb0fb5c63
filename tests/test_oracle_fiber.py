"""The oracle's UniformFiberNI source against the reference's CUDAMCML
fiber fixture (xopto/mcml/test/validate.py:348-511): reflectance collected
by 200 um fibers at five separations from a layer under a 200 um, NA 0.22
source fiber, for points of the 20 x 20 grid of mua and mus'. The
reference passes its run when the mean relative error over the grid is
below 0.5 % at 1e7 packets per point; at 1e6 packets on eight points the
mean must be below 2.5 % (about four standard errors)."""
import numpy as np

from pyxopto_amd import _native as N
from pyxopto_amd import workloads as W
from pyxopto_amd.util.convolve import fiber_reflectance
from oracle import pyoracle as O

import os
FIXTURE = os.path.join(os.path.dirname(__file__), 'golden',
                       'cuda_singlelayer_uniformfiber.npz')


def test_benchmark_constants_match_the_fixture():
    f = np.load(FIXTURE)
    np.testing.assert_allclose(W.FIBER_GRID_MUA, f['lut_mua'])
    np.testing.assert_allclose(W.FIBER_GRID_MUSR, f['lut_musr'])
    np.testing.assert_allclose(W.FIBER_SDS, f['sds'])
    sim = W.fiber_sim()
    cfg, _, naccum = sim.pack()
    assert naccum == int(f['axis_n'])
    assert cfg.source_kind == N.XO_SRC_FIBER_NI
    assert abs(cfg.fiber.radius - f['source_dcore'] / 2) < 1e-9
    assert abs(cfg.fiber.n - f['source_ncore']) < 1e-6
    assert abs(cfg.top.cos_min - f['cosmin']) < 1e-7
    assert [sim.layers[i].n for i in range(3)] == list(f['layer_n'])
    assert sim.rmax == f['rmax']


def test_fiber_source_matches_cudamcml_fixture():
    f = np.load(FIXTURE)
    ref = f['reflectance'].reshape(400, 5)
    edges = np.linspace(0.0, 1.7e-3, 341)
    area = np.pi * (edges[1:] ** 2 - edges[:-1] ** 2)
    r = 0.5 * (edges[1:] + edges[:-1])
    sim = W.fiber_sim()
    mua, musr = W.fiber_grid()
    x, a = N.init_rng(2048, 0x0123456789ABCDEF)
    n = 1_000_000
    errs = []
    for k in (0, 21, 57, 133, 210, 266, 342, 399):
        sim.layers[1].mua = mua[k]
        sim.layers[1].mus = musr[k] / (1.0 - 0.85)
        cfg, lut, naccum = sim.pack()
        acc, _, x = O.mcml_run(cfg, lut, x, a, n, naccum, O.MATH_LIBM)
        R = acc.astype(np.float64) / N.XO_ACCU_K / (n * area)
        fr = fiber_reflectance(r, R, W.FIBER_SDS, 200e-6)[0]
        errs.append((fr - ref[k]) / ref[k])
    errs = np.array(errs)
    assert abs(errs.mean()) < 0.025, errs.mean()
    assert np.abs(errs).max() < 0.2
