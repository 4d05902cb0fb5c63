"""GPU parity of the layered-media kernel (xo_mcml_run through the C ABI).

1. Bit-exact: with the static packet schedule, the GPU must reproduce the
   oracle's integer accumulators, step counts and final RNG states exactly
   when the oracle evaluates log/sin/cos/cbrt with the kernels' functions.
2. Statistical: the production (dynamic schedule) kernel against the oracle
   with the C library's math: |z| < 4 per bin over batch means, total diffuse
   reflectance and transmittance within 3 sigma.
"""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcml import mc
from oracle import pyoracle as O

from helpers import (XINIT, batch_mean_se, config1_sim, oracle_batches,
                     slab_layers, standard_detectors, bins_to_test, z_scores)

pytestmark = pytest.mark.gpu


def _sims():
    from pyxopto_amd.mcbase import pf as hostpf
    lut_pf = mc.mcpf.LutEx(hostpf.MHg, (0.85, 0.6), 2000)
    L = mc.mclayer
    multi = L.Layers([
        L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=mc.mcpf.Hg(0.0)),
        L.Layer(d=1e-4, n=1.34, mua=1600.0, mus=4e4, pf=mc.mcpf.Hg(0.8)),
        L.Layer(d=2e-4, n=1.40, mua=50.0, mus=2.5e4, pf=mc.mcpf.Hg(0.9)),
        L.Layer(d=5e-4, n=1.36, mua=300.0, mus=1e4, pf=mc.mcpf.Hg(0.0)),
        L.Layer(d=np.inf, n=1.45, mua=0.0, mus=0.0, pf=mc.mcpf.Hg(0.0))])
    return {
        'hg_slab': config1_sim(),
        'mhg_slab': mc.Mc(slab_layers(lambda: mc.mcpf.MHg(0.8, 0.7)),
                          mc.mcsource.Line(), standard_detectors(),
                          rnginit=XINIT),
        'lut_slab': mc.Mc(slab_layers(lambda: lut_pf), mc.mcsource.Line(),
                          standard_detectors(), rnginit=XINIT),
        'multilayer': mc.Mc(multi, mc.mcsource.Line(),
                            standard_detectors(cosmin=0.5, logscale=True),
                            rnginit=XINIT),
        'oblique_rmax': mc.Mc(slab_layers(lambda: mc.mcpf.Hg(0.9), d=2e-3),
                              mc.mcsource.Line(direction=(0.3, 0.0, 0.954)),
                              standard_detectors(), rnginit=XINIT),
    }


SIMS = None


def sims():
    global SIMS
    if SIMS is None:
        SIMS = _sims()
        SIMS['oblique_rmax'].rmax = 2e-3
    return SIMS


@pytest.mark.parametrize('name', ['hg_slab', 'mhg_slab', 'lut_slab',
                                  'multilayer', 'oblique_rmax'])
def test_static_schedule_bit_exact(name):
    sim = sims()[name]
    print('case', name, flush=True)
    lanes, npk = 8192, 60000
    cfg, lut, naccum = sim.pack()
    ctx = N.Context(0, 0, XINIT)
    gpu, rep = ctx.run_mcml(cfg, lut, npk, naccum, lanes,
                            N.XO_FLAG_STATIC_SCHEDULE)
    gx, _ = ctx.get_rng(lanes)
    x, a = N.init_rng(lanes, XINIT)
    ref, counts, rx = O.mcml_run(cfg, lut, x, a, npk, naccum, O.MATH_KERNEL)
    assert rep.packets == npk
    assert int(rep.steps) == int(counts[0])
    assert int(rep.boundary_steps) == int(counts[1])
    np.testing.assert_array_equal(gpu, ref)
    np.testing.assert_array_equal(gx, rx)
    ctx.close()


def _extreme_anisotropy_sim(pf_of):
    L = mc.mclayer
    layers = L.Layers([
        L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=pf_of(0.0)),
        L.Layer(d=2e-4, n=1.37, mua=100.0, mus=2e4, pf=pf_of(1.0)),
        L.Layer(d=2e-4, n=1.40, mua=100.0, mus=2e4, pf=pf_of(-1.0)),
        L.Layer(d=np.inf, n=1.45, mua=0.0, mus=0.0, pf=pf_of(0.0))])
    return mc.Mc(layers, mc.mcsource.Line(), standard_detectors(),
                 rnginit=XINIT)


@pytest.mark.parametrize('kind', ['hg', 'mhg'])
def test_static_schedule_bit_exact_at_g_plus_minus_one(kind):
    """|g| = 1 makes the HG numerator 1 - g^2 zero and, for one extreme
    deviate, the denominator too: the kernels' range-restricted division
    must still give the oracle's bits."""
    pf_of = (lambda g: mc.mcpf.Hg(g)) if kind == 'hg' else \
        (lambda g: mc.mcpf.MHg(g, 0.9))
    sim = _extreme_anisotropy_sim(pf_of)
    print('case', kind, flush=True)
    lanes, npk = 4096, 20000
    cfg, lut, naccum = sim.pack()
    ctx = N.Context(0, 0, XINIT)
    gpu, rep = ctx.run_mcml(cfg, lut, npk, naccum, lanes,
                            N.XO_FLAG_STATIC_SCHEDULE)
    gx, _ = ctx.get_rng(lanes)
    x, a = N.init_rng(lanes, XINIT)
    ref, counts, rx = O.mcml_run(cfg, lut, x, a, npk, naccum, O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(gpu, ref)
    np.testing.assert_array_equal(gx, rx)
    ctx.close()


def test_anisotropy_outside_unit_interval_is_rejected():
    sim = _extreme_anisotropy_sim(lambda g: mc.mcpf.Hg(g))
    cfg, lut, naccum = sim.pack()
    cfg.layers[1].pf.g = 1.5
    ctx = N.Context(0, 0, XINIT)
    with pytest.raises(RuntimeError, match='anisotropy'):
        ctx.run_mcml(cfg, lut, 1000, naccum, 1024, N.XO_FLAG_STATIC_SCHEDULE)
    ctx.close()


def test_static_schedule_global_bins_match_lds_bins():
    sim = sims()['hg_slab']
    cfg, lut, naccum = sim.pack()
    ctx = N.Context(0, 0, XINIT)
    a1, _ = ctx.run_mcml(cfg, lut, 30000, naccum, 4096,
                         N.XO_FLAG_STATIC_SCHEDULE)
    ctx.set_rng(*N.init_rng(4096, XINIT))
    a2, _ = ctx.run_mcml(cfg, lut, 30000, naccum, 4096,
                         N.XO_FLAG_STATIC_SCHEDULE | N.XO_FLAG_NO_LDS_BINS)
    np.testing.assert_array_equal(a1, a2)
    ctx.close()


def _normalized(sim, accum, npk):
    return accum.astype(np.float64) / N.XO_ACCU_K / npk


@pytest.mark.parametrize('name', ['hg_slab', 'lut_slab', 'multilayer'])
def test_dynamic_schedule_statistical_parity(name):
    sim = sims()[name]
    print('case', name, flush=True)
    k, npk = 40, 100000
    cfg, lut, naccum = sim.pack()
    ctx = N.Context(0, 0, XINIT)
    g = []
    for _ in range(k):
        acc, rep = ctx.run_mcml(cfg, lut, npk, naccum, 0, 0)
        assert rep.packets == npk
        g.append(_normalized(sim, acc, npk))
    g = np.asarray(g)
    lanes = 1024
    o = oracle_batches(
        lambda x, a: _normalized(sim, O.mcml_run(cfg, lut, x, a, npk, naccum,
                                                 O.MATH_LIBM)[0], npk),
        k, lanes)
    mg, sg = batch_mean_se(g)
    mo, so = batch_mean_se(o)
    z = z_scores(mg, sg, mo, so)
    ok = bins_to_test(g, o)
    assert ok.sum() >= 10
    assert np.all(np.abs(z[ok]) < 4.0), (np.flatnonzero(np.abs(z) >= 4),
                                         z[np.abs(z) >= 4])
    # total diffuse reflectance (all radial bins) and transmittance
    nb = sim.detectors.top.n
    for sl in (slice(0, nb), slice(nb, nb + 1)):
        tg, to = g[:, sl].sum(axis=1), o[:, sl].sum(axis=1)
        m1, s1 = batch_mean_se(tg)
        m2, s2 = batch_mean_se(to)
        assert abs(m1 - m2) < 3.0 * np.sqrt(s1 ** 2 + s2 ** 2)
    ctx.close()


def test_mc_run_api_matches_oracle_statistics():
    """Mc.run end to end: reflectance within 3 sigma of the oracle."""
    sim = config1_sim()
    res = []
    out = None
    for _ in range(8):
        _, _, det = sim.run(250000)
        res.append(det.top.raw.sum() / det.top.nphotons)
        out = sim.run(250000, out=(None, None, det))[2]
    assert out.top.nphotons == 500000
    assert sim.run_report['threads'] > 0
    cfg, lut, naccum = sim.pack()
    o = oracle_batches(lambda x, a: O.mcml_run(
        cfg, lut, x, a, 250000, naccum)[0][:100].sum() / N.XO_ACCU_K / 250000,
        8, 4096)
    m1, s1 = batch_mean_se(np.asarray(res))
    m2, s2 = batch_mean_se(o)
    assert abs(m1 - m2) < 3.0 * np.sqrt(s1 ** 2 + s2 ** 2)


def test_run_copyseeds_downloads_the_generator_states():
    """Mc.run(copyseeds=True) copies the device's mutable generator states
    into rng_seeds_x (xopto/mcml/mc.py:962-966); without it rng_seeds_x
    keeps the initial states. Under the static schedule the states of the
    used lanes are the oracle's final states."""
    sim = config1_sim()
    x0 = sim.rng_seeds_x.copy()
    lanes, npk = 4096, 10000
    sim.run(npk, maxthreads=lanes, static_schedule=True)
    np.testing.assert_array_equal(sim.rng_seeds_x, x0)
    sim.run(npk, maxthreads=lanes, static_schedule=True, copyseeds=True)
    cfg, lut, naccum = sim.pack()
    x, a = N.init_rng(lanes, XINIT)
    _, _, x1 = O.mcml_run(cfg, lut, x, a, npk, naccum, O.MATH_KERNEL)
    _, _, x2 = O.mcml_run(cfg, lut, x1, a, npk, naccum, O.MATH_KERNEL)
    np.testing.assert_array_equal(sim.rng_seeds_x[:lanes], x2)
    np.testing.assert_array_equal(sim.rng_seeds_x[lanes:], x0[lanes:])
    np.testing.assert_array_equal(sim.rng_seeds_a[:lanes], a)
    sim.context.close()
