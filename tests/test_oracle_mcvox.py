"""Pinning of the voxelized oracle: the reference's own layered-vs-voxel
strategy (xopto/mcvox/test/validate.py:47-49, 263-266) — the CUDAMCML fixture
run as a 1x1x1 voxel grid — plus the closed-form deposition of a
non-scattering absorbing column and a layered/voxel cross-check of a
two-layer slab."""
import os

import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcml import mc as mcml
from pyxopto_amd.mcvox import mc
from oracle import pyoracle as O

from helpers import (XINIT, batch_mean_se, bins_to_test, oracle_batches,
                     z_scores)

GOLD = os.path.join(os.path.dirname(__file__), 'golden',
                    'cuda_singlelayer_linesource_radial.npz')


def fixture_voxel_sim():
    f = np.load(GOLD, allow_pickle=False)
    rmax = float(f['rmax'])
    M = mc.mcmaterial
    materials = M.Materials([
        M.Material(n=f['layer_n'][0], mua=f['layer_mua'][0],
                   mus=f['layer_mus'][0], pf=mc.mcpf.Hg(f['layer_g'][0])),
        M.Material(n=f['layer_n'][1], mua=f['layer_mua'][1],
                   mus=f['layer_mus'][1], pf=mc.mcpf.Hg(f['layer_g'][1]))])
    G = mc.mcgeometry
    voxels = G.Voxels(G.Axis(-2 * rmax, 2 * rmax, 1),
                      G.Axis(-2 * rmax, 2 * rmax, 1),
                      G.Axis(0.0, float(f['layer_d'][1]), 1))
    D = mc.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(float(f['axis_start']),
                                                float(f['axis_stop']),
                                                int(f['axis_n'])),
                                   cosmin=float(f['cosmin'])))
    sim = mc.Mc(voxels, materials, mc.mcsource.Line(), det, rnginit=XINIT)
    sim.rmax = rmax
    sim.voxels.material[:] = 1
    return sim, f


def test_voxel_oracle_reproduces_cudamcml_fixture():
    sim, f = fixture_voxel_sim()
    cfg, lut, naccum = sim.pack()
    npk = 500000
    x, a = N.init_rng(16384, XINIT)
    acc, counts, _ = O.mcvox_run(cfg, sim.voxels.material, lut, x, a, npk,
                                 naccum, O.MATH_LIBM)
    raw = acc / N.XO_ACCU_K
    e = sim.detectors.top.edges
    refl = raw / (npk * np.pi * (e[1:] ** 2 - e[:-1] ** 2))
    rel = (refl - f['reflectance']) / f['reflectance'] * 100.0
    assert abs(rel.mean()) < 0.5, rel.mean()


def column_sim(mua=300.0, nz=40, depth=4e-3, n=1.33, sigma=None):
    M = mc.mcmaterial
    materials = M.Materials([
        M.Material(n=n, mua=0.0, mus=0.0, pf=mc.mcpf.Hg(0.5)),
        M.Material(n=n, mua=mua, mus=0.0, pf=mc.mcpf.Hg(0.5))])
    G = mc.mcgeometry
    voxels = G.Voxels(G.Axis(-1e-3, 1e-3, 4), G.Axis(-1e-3, 1e-3, 4),
                      G.Axis(0.0, depth, nz))
    fl = mc.mcfluence.Fluence(G.Axis(-1e-3, 1e-3, 4), G.Axis(-1e-3, 1e-3, 4),
                              G.Axis(0.0, depth, nz))
    D = mc.mcdetector
    det = D.Detectors(top=D.Total(), bottom=D.Total(), specular=D.Total())
    src = mc.mcsource.GaussianBeam(sigma) if sigma else mc.mcsource.Line()
    sim = mc.Mc(voxels, materials, src, det, fluence=fl, rnginit=XINIT)
    sim.voxels.material[:] = 1
    return sim


@pytest.mark.parametrize('source', ['line', 'gaussian'])
def test_nonscattering_column_deposition_closed_form(source):
    mua, nz, depth = 300.0, 40, 4e-3
    sim = column_sim(mua, nz, depth, sigma=1e-4 if source == 'gaussian'
                     else None)
    cfg, lut, naccum = sim.pack()
    npk = 300000
    x, a = N.init_rng(4096, XINIT)
    acc, counts, _ = O.mcvox_run(cfg, sim.voxels.material, lut, x, a, npk,
                                 naccum, O.MATH_LIBM)
    off, n = sim.allocation(sim.fluence)
    dep = acc[off:off + n].reshape(nz, 4, 4).sum(axis=(1, 2)) / N.XO_ACCU_K
    z = np.linspace(0.0, depth, nz + 1)
    p = np.exp(-mua * z[:-1]) - np.exp(-mua * z[1:])
    # every packet deposits its full weight (1, matched indices) once
    expect = npk * p
    sigma = np.sqrt(npk * p * (1 - p))
    assert np.all(np.abs(dep - expect) < 4 * sigma + 1.0), (dep - expect) / sigma
    # transmitted weight is exp(-mua*depth)
    t = acc[1] / N.XO_ACCU_K / npk
    pt = np.exp(-mua * depth)
    assert abs(t - pt) < 4 * np.sqrt(pt * (1 - pt) / npk)
    assert acc[0] == 0 and acc[2] == 0       # no reflection, no specular


def two_layer_pair():
    """The same 2-layer slab as a layered stack and as a 1x1x2 voxel grid."""
    mats = [(1.0, 0.0, 0.0, 0.0), (1.37, 200.0, 1.5e4, 0.85),
            (1.44, 20.0, 1e4, 0.9)]
    d1, d2 = 3e-4, 7e-4
    L = mcml.mclayer
    layers = L.Layers([
        L.Layer(d=np.inf, n=1.0, mua=0, mus=0, pf=mcml.mcpf.Hg(0)),
        L.Layer(d=d1, n=1.37, mua=200.0, mus=1.5e4, pf=mcml.mcpf.Hg(0.85)),
        L.Layer(d=d2, n=1.44, mua=20.0, mus=1e4, pf=mcml.mcpf.Hg(0.9)),
        L.Layer(d=np.inf, n=1.0, mua=0, mus=0, pf=mcml.mcpf.Hg(0))])
    D = mcml.mcdetector
    dl = D.Detectors(top=D.Radial(D.RadialAxis(0, 3e-3, 30)), bottom=D.Total())
    lay = mcml.Mc(layers, mcml.mcsource.Line(), dl, rnginit=XINIT)
    M = mc.mcmaterial
    materials = M.Materials([M.Material(n=m[0], mua=m[1], mus=m[2],
                                        pf=mc.mcpf.Hg(m[3])) for m in mats])
    G = mc.mcgeometry
    big = 0.2
    voxels = G.Voxels(G.Axis(-big, big, 1), G.Axis(-big, big, 1),
                      G.Axis(0.0, d1 + d2, 2))
    # unequal layers: a 2-voxel z axis has equal steps, so use 10 voxels of
    # 0.1 mm: 3 of material 1 and 7 of material 2
    voxels = G.Voxels(G.Axis(-big, big, 1), G.Axis(-big, big, 1),
                      G.Axis(0.0, d1 + d2, 10))
    dv = mc.mcdetector.Detectors(
        top=mc.mcdetector.Radial(mc.mcdetector.RadialAxis(0, 3e-3, 30)),
        bottom=mc.mcdetector.Total())
    vox = mc.Mc(voxels, materials, mc.mcsource.Line(), dv, rnginit=XINIT)
    vox.voxels.material[:3] = 1
    vox.voxels.material[3:] = 2
    return lay, vox


def test_voxel_oracle_matches_layered_oracle():
    lay, vox = two_layer_pair()
    cl, ll, nl = lay.pack()
    cv, lv, nv = vox.pack()
    # 50 batches per side: the batch-mean standard errors then carry ~98
    # degrees of freedom, so |z| < 4 over 31 bins rarely fails by chance
    npk, k, lanes = 20000, 50, 1024
    g = oracle_batches(lambda x, a: O.mcml_run(cl, ll, x, a, npk, nl)[0]
                       / N.XO_ACCU_K / npk, k, lanes, xinit=777777777)
    v = oracle_batches(lambda x, a: O.mcvox_run(cv, vox.voxels.material, lv,
                                                x, a, npk, nv)[0]
                       / N.XO_ACCU_K / npk, k, lanes, xinit=888888888)
    m1, s1 = batch_mean_se(g)
    m2, s2 = batch_mean_se(v)
    z = z_scores(m1, s1, m2, s2)
    ok = bins_to_test(g, v)
    assert ok.sum() >= 20
    assert np.all(np.abs(z[ok]) < 4.0), z[ok]
