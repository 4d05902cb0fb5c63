"""GPU parity of the hardware-math build (XO_FLAG_FAST_MATH, what
Mc(cl_build_options=['-cl-fast-relaxed-math']) selects, as the reference's
own build option does: xopto/cl/cloptions.py:137-152).

The fast build is not bit-exact with the oracle, so it is held to the
statistical bar of the production kernels:
1. the device arithmetic (v_log_f32, v_exp_f32, v_sin/cos_f32 in turns,
   v_sqrt_f32, v_rsq_f32, v_rcp_f32) against float64 on its arguments'
   ranges, within stated error bounds;
2. the layered, voxel and batch kernels against the oracle with the C
   library's math: |z| < 4 per bin over batch means, totals within 3 sigma;
3. the reference's CUDAMCML fiber fixtures at full size (400 runs of 1e7
   packets, xopto/mcml/test/validate.py:348-511, :692-850): mean relative
   error of the fiber reflectance below 0.5 %.
"""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcml import mc
from pyxopto_amd.mcml.batch import Batch, Instance
from oracle import pyoracle as O

from helpers import (XINIT, batch_mean_se, bins_to_test, oracle_batches,
                     z_scores)

pytestmark = pytest.mark.gpu

FAST = N.XO_FLAG_FAST_MATH


def _ctx():
    return N.Context(0, 0, XINIT)


# ---------------------------------------------------------------- 1
def _grid(lo, hi, n, rng, log=False):
    if log:
        v = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    else:
        v = rng.uniform(lo, hi, n)
    return v.astype(np.float32)


def test_fast_math_accuracy_against_float64():
    rng = np.random.default_rng(11)
    n = 1 << 20
    u = (rng.integers(1, 2 ** 32, n, dtype=np.uint64).astype(np.float64) *
         2.0 ** -32).astype(np.float32)          # uniform deviates
    t = rng.uniform(0.0, 1.0, n).astype(np.float32)
    cases = {
        # which: (arguments, float64 reference, bound kind, bound)
        N.XO_MATH_LOG: (u, lambda v: np.log(v), 'abs_or_rel', 4e-7),
        N.XO_MATH_SIN_TURNS: (t, lambda v: np.sin(2 * np.pi * v), 'abs', 2e-6),
        N.XO_MATH_COS_TURNS: (t, lambda v: np.cos(2 * np.pi * v), 'abs', 2e-6),
        N.XO_MATH_SQRT: (u, np.sqrt, 'rel', 4e-7),
        N.XO_MATH_RCP: (_grid(1e-3, 1e3, n, rng, log=True),
                        lambda v: 1.0 / v, 'rel', 4e-7),
        N.XO_MATH_RSQRT: (_grid(0.9, 1.1, n, rng),
                          lambda v: 1.0 / np.sqrt(v), 'rel', 4e-7),
        N.XO_MATH_CBRT: (_grid(-1.0, 1.0, n, rng), np.cbrt, 'rel', 2e-6),
        N.XO_MATH_EXP: (_grid(-40.0, 0.0, n, rng), np.exp, 'rel', 2e-6),
    }
    ctx = _ctx()
    report = {}
    for which, (x, ref, kind, bound) in cases.items():
        for fast in (False, True):
            y = ctx.probe_math(which, x, fast).astype(np.float64)
            r = ref(x.astype(np.float64))
            err = np.abs(y - r)
            if kind == 'abs':
                e = err
            elif kind == 'rel':
                e = err / np.abs(r)
            else:
                e = err / np.maximum(np.abs(r), 1.0)
            report[(which, fast)] = float(e.max())
            assert np.all(np.isfinite(y))
            assert e.max() < bound, (which, fast, e.max())
    print('max errors (which, fast): {}'.format(report), flush=True)
    ctx.close()


def test_fast_math_special_values():
    ctx = _ctx()
    # ln 0 = -inf (a zero deviate gives an infinite step, clamped by the
    # kernels), cbrt keeps the sign and zero
    assert ctx.probe_math(N.XO_MATH_LOG, [0.0], True)[0] == -np.inf
    c = ctx.probe_math(N.XO_MATH_CBRT, [0.0, -0.0, -0.125, 0.125], True)
    assert c[0] == 0.0 and np.signbit(c[1])
    np.testing.assert_allclose(c[2:], [-0.5, 0.5], rtol=2e-6)
    ctx.close()


# ---------------------------------------------------------------- 2
def _normalized(accum, npk):
    return accum.astype(np.float64) / N.XO_ACCU_K / npk


def _assert_parity(g, o, nbins_radial, min_bins=10):
    mg, sg = batch_mean_se(g)
    mo, so = batch_mean_se(o)
    z = z_scores(mg, sg, mo, so)
    ok = bins_to_test(g, o)
    assert ok.sum() >= min_bins
    assert np.all(np.abs(z[ok]) < 4.0), (np.flatnonzero(np.abs(z) >= 4),
                                         z[np.abs(z) >= 4])
    for sl in (slice(0, nbins_radial), slice(nbins_radial, nbins_radial + 1)):
        tg, to = g[:, sl].sum(axis=1), o[:, sl].sum(axis=1)
        m1, s1 = batch_mean_se(tg)
        m2, s2 = batch_mean_se(to)
        assert abs(m1 - m2) < 3.0 * np.sqrt(s1 ** 2 + s2 ** 2)


@pytest.mark.parametrize('name', ['hg_slab', 'mhg_slab', 'lut_slab',
                                  'multilayer', 'oblique_rmax'])
def test_fast_layered_statistical_parity(name):
    from test_gpu_mcml import sims
    sim = sims()[name]
    print('case', name, flush=True)
    k, npk = 40, 100000
    cfg, lut, naccum = sim.pack()
    ctx = _ctx()
    g = []
    for _ in range(k):
        acc, rep = ctx.run_mcml(cfg, lut, npk, naccum, 0, FAST)
        assert rep.packets == npk
        g.append(_normalized(acc, npk))
    o = oracle_batches(
        lambda x, a: _normalized(O.mcml_run(cfg, lut, x, a, npk, naccum,
                                            O.MATH_LIBM)[0], npk), k, 1024)
    _assert_parity(np.asarray(g), o, sim.detectors.top.n)
    ctx.close()


@pytest.mark.parametrize('method', [0, 2])
def test_fast_voxel_statistical_parity(method):
    from test_gpu_mcvox import skin_vessel_sim
    sim = skin_vessel_sim(method=method)
    print('case skin vessel method', method, flush=True)
    cfg, lut, naccum = sim.pack()
    k, npk = 40, 20000
    vox = sim.voxels.material
    ctx = _ctx()
    g = []
    for i in range(k):
        acc, rep = ctx.run_mcvox(cfg, vox if i == 0 else None, lut, npk,
                                 naccum, 0, FAST)
        assert rep.packets == npk
        g.append(_normalized(acc, npk))
    g = np.asarray(g)
    o = oracle_batches(lambda x, a: _normalized(O.mcvox_run(
        cfg, vox, lut, x, a, npk, naccum)[0], npk), k, 512)
    off, n = sim.allocation(sim.fluence)

    def reduce(b):
        det = b[:, :off]
        prof = b[:, off:off + n].reshape(b.shape[0], 40, 40, 40).sum(
            axis=(2, 3))
        return np.concatenate([det, prof], axis=1)
    gr, orr = reduce(g), reduce(o)
    _assert_parity(gr, orr, 20, min_bins=40)
    ctx.close()


def _batch_instances(copies):
    """Four layered instances (HG, MHG-LUT, 1 and 3 turbid layers, thin and
    thick), each repeated `copies` times with its own seed."""
    from pyxopto_amd.mcbase import pf as hostpf
    L = mc.mclayer
    D = mc.mcdetector
    lut_pf = mc.mcpf.LutEx(hostpf.MHg, (0.9, 0.4), 2000)

    def det():
        return D.Detectors(top=D.Radial(D.RadialAxis(0.0, 5e-3, 50)),
                           bottom=D.Total(), specular=D.Total())
    stacks = [
        [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=mc.mcpf.Hg(0.8)),
         L.Layer(d=1e-3, n=1.37, mua=100.0, mus=1e4, pf=mc.mcpf.Hg(0.8)),
         L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=mc.mcpf.Hg(0.8))],
        [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=lut_pf),
         L.Layer(d=2e-3, n=1.45, mua=50.0, mus=2e4, pf=lut_pf),
         L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=lut_pf)],
        [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=mc.mcpf.Hg(0.9)),
         L.Layer(d=1e-4, n=1.34, mua=1600.0, mus=4e4, pf=mc.mcpf.Hg(0.8)),
         L.Layer(d=2e-4, n=1.40, mua=50.0, mus=2.5e4, pf=mc.mcpf.Hg(0.9)),
         L.Layer(d=2e-3, n=1.36, mua=300.0, mus=1e4, pf=mc.mcpf.Hg(0.7)),
         L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=mc.mcpf.Hg(0.9))],
        [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=lut_pf),
         L.Layer(d=5e-5, n=1.5, mua=5000.0, mus=1e3, pf=lut_pf),
         L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=lut_pf)],
    ]
    # one kind of phase function per batch run: HG and LUT stacks apart
    out = {'hg': [], 'lut': []}
    for s, layers in enumerate(stacks):
        kind = 'lut' if s in (1, 3) else 'hg'
        for c in range(copies):
            out[kind].append((s, Instance(L.Layers(layers), mc.mcsource.Line(),
                                          det(), nphotons=4000,
                                          seed=1000 * s + c)))
    return out


@pytest.mark.parametrize('kind', ['hg', 'lut'])
def test_fast_batch_statistical_parity(kind):
    copies = 40
    insts = _batch_instances(copies)[kind]
    print('case batch', kind, len(insts), 'instances', flush=True)
    b = Batch([i for _, i in insts], rnginit=XINIT,
              cl_build_options=['-cl-fast-relaxed-math', '-cl-mad-enable'])
    assert b.fast_math
    arr, lut, naccum = b.pack()
    opts = b.opts(chunk=512)
    assert opts.flags & FAST
    acc, rep = b.context.run_batch(arr, lut, opts, naccum)
    ref, _ = O.batch_run(arr, lut, naccum, math_mode=O.MATH_LIBM)
    stacks = sorted({s for s, _ in insts})
    for s in stacks:
        rows = [j for j, (si, _) in enumerate(insts) if si == s]
        base = [int(arr[j].accum_base) for j in rows]
        cnt = int(arr[rows[0]].accum_count)
        npk = int(arr[rows[0]].npackets)
        g = np.asarray([_normalized(acc[o:o + cnt], npk) for o in base])
        o = np.asarray([_normalized(ref[o:o + cnt], npk) for o in base])
        _assert_parity(g, o, 50, min_bins=3)
    b.context.close()


def test_mc_cl_build_options_select_fast_math():
    from helpers import config1_sim
    from pyxopto_amd.mcml import mc as mcml
    exact = config1_sim()
    fast = mcml.Mc(exact.layers, exact.source, exact.detectors,
                   rnginit=XINIT, cl_build_options=['-cl-fast-relaxed-math'])
    assert fast.fast_math and not exact.fast_math
    r = []
    for sim in (exact, fast):
        vals = []
        for _ in range(8):
            _, _, det = sim.run(250000)
            vals.append(det.top.raw.sum() / det.top.nphotons)
        r.append(batch_mean_se(np.asarray(vals)))
    (m1, s1), (m2, s2) = r
    assert abs(m1 - m2) < 3.0 * np.sqrt(s1 ** 2 + s2 ** 2)


# ---------------------------------------------------------------- 3
def _fast_fixture(name, layers_for):
    import test_gpu_fiber as F
    orig = F.Batch

    class FastBatch(orig):
        def __init__(self, instances, **kw):
            kw.setdefault('cl_build_options', ['-cl-fast-relaxed-math'])
            super().__init__(instances, **kw)
    F.Batch = FastBatch
    try:
        return F._fixture_run(name, layers_for)
    finally:
        F.Batch = orig


def test_fast_single_layer_fiber_fixture_full_size():
    from pyxopto_amd import workloads as W
    err = _fast_fixture('cuda_singlelayer_uniformfiber',
                        lambda f, m, s: W.fiber_layers(m, s, 0.85))
    assert abs(err.mean()) < 0.005


def test_fast_double_layer_fiber_fixture_full_size():
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
    err = _fast_fixture('cuda_doublelayer_uniformfiber', layers)
    assert abs(err.mean()) < 0.005
