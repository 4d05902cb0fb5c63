"""GPU parity on the workload the benchmark times: the 4096-instance mixed
layered batch of workloads.layered_batch (1-5 turbid layers, thickness
0.05-5 mm, mua 1-5000 /m, mus 1e3-5e4 /m, HG or 2000-point MHG lookup
tables, budgets 1e4-1e6).

1. Its corners (workloads.layered_batch_extremes: the optically thickest
   stacks, the lowest absorption, the highest scattering, the thickest
   layers, five-layer lookup-table stacks, plus evenly spaced instances;
   48 in all) with budgets capped at 2e4 packets: the exact build equals
   the oracle bit for bit under the production launch (integers and step
   counts), and both builds pass |z| < 4 per bin against the oracle with
   the C library's math on four of them.
2. At full size (8.8e8 packets): the fast build's integers do not depend on
   the launch shape (two shapes, equal arrays), every instance conserves
   weight (specular + reflected + transmitted <= 1), and the batch's total
   diffuse reflectance agrees with the exact build's within 3 standard
   errors of the instances' spread.
"""
import os

import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd import workloads as W
from pyxopto_amd.mcml.batch import Batch
from oracle import pyoracle as O

from helpers import XINIT, batch_mean_se, bins_to_test, z_scores

pytestmark = pytest.mark.gpu

FAST = ['-cl-fast-relaxed-math']
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _extremes(cap=20000):
    idx = W.layered_batch_extremes()
    insts = W.layered_batch(indices=idx)
    for inst in insts:
        inst.nphotons = min(inst.nphotons, cap)
    return idx, insts


def test_bench_workload_corners_bit_exact():
    idx, insts = _extremes()
    print('case', len(insts), 'corner instances', idx, flush=True)
    b = Batch(insts, rnginit=XINIT)
    arr, lut, naccum = b.pack()
    ref, counts = O.batch_run(arr, lut, naccum, math_mode=O.MATH_KERNEL)
    for chunk in (2048, 8192):
        acc, rep = b.context.run_batch(arr, lut, b.opts(chunk=chunk), naccum)
        assert int(rep.steps) == int(counts[0])
        assert int(rep.boundary_steps) == int(counts[1])
        np.testing.assert_array_equal(acc, ref)
    b.context.close()


def _corner_picks(insts):
    """The thickest stack (sum of mus d), the lowest absorption, a
    five-layer lookup-table stack and the highest scattering."""
    def turbid(inst):
        return [inst.layers[k] for k in range(1, len(inst.layers) - 1)]
    tau = [sum(l.mus * l.d for l in turbid(i)) for i in insts]
    mua = [min(l.mua for l in turbid(i)) for i in insts]
    mus = [max(l.mus for l in turbid(i)) for i in insts]
    lut5 = [k for k, i in enumerate(insts)
            if len(i.layers) == 7 and i.layers.pf_kind == N.XO_PF_LUT]
    picks = [int(np.argmax(tau)), int(np.argmin(mua)), lut5[0],
             int(np.argmax(mus))]
    return sorted(set(picks))


@pytest.mark.parametrize('fast', [False, True])
def test_bench_workload_corners_statistical(fast):
    from pyxopto_amd.mcml.batch import Instance
    _, insts = _extremes()
    picks = _corner_picks(insts)
    copies, npk = 40, 2000
    rows = [(p, Instance(insts[p].layers, insts[p].source, insts[p].detectors,
                         nphotons=npk, seed=7919 * (p + 1) + c))
            for p in picks for c in range(copies)]
    print('case corners', picks, 'fast', fast, flush=True)
    b = Batch([i for _, i in rows], rnginit=XINIT,
              cl_build_options=FAST if fast else None)
    arr, lut, naccum = b.pack()
    acc, _ = b.context.run_batch(arr, lut, b.opts(chunk=512), naccum)
    ref, _ = O.batch_run(arr, lut, naccum, math_mode=O.MATH_LIBM)
    for p in picks:
        js = [j for j, (q, _) in enumerate(rows) if q == p]
        cnt = int(arr[js[0]].accum_count)

        def bins(a):
            return np.asarray([a[int(arr[j].accum_base):][:cnt] for j in js],
                              dtype=np.float64) / N.XO_ACCU_K / npk
        g, o = bins(acc), bins(ref)
        m1, s1 = batch_mean_se(g)
        m2, s2 = batch_mean_se(o)
        z = z_scores(m1, s1, m2, s2)
        ok = bins_to_test(g, o)
        assert np.all(np.abs(z[ok]) < 4.0), (p, z[ok])
        for sl in (slice(0, 100), slice(100, 101)):
            a1, b1 = batch_mean_se(g[:, sl].sum(axis=1))
            a2, b2 = batch_mean_se(o[:, sl].sum(axis=1))
            assert abs(a1 - a2) < 3.0 * np.sqrt(b1 ** 2 + b2 ** 2)
    b.context.close()


def test_bench_workload_full_size_properties():
    insts = W.layered_batch(cache_dir=os.path.join(REPO, 'pyxopto_amd',
                                                   'data'))
    npk = sum(i.nphotons for i in insts)
    print('case full batch', len(insts), 'instances', npk, 'packets',
          flush=True)
    fast = Batch(insts, rnginit=XINIT, cl_build_options=FAST)
    arr, lut, naccum = fast.pack()
    a1, r1 = fast.context.run_batch(arr, lut, fast.opts(chunk=8192), naccum)
    a2, r2 = fast.context.run_batch(arr, lut,
                                    fast.opts(chunk=2048, block=64), naccum)
    assert r1.packets == npk and r2.packets == npk
    np.testing.assert_array_equal(a1, a2)
    exact = Batch(insts, rnginit=XINIT)
    e1, _ = exact.context.run_batch(arr, lut, exact.opts(chunk=8192), naccum)
    rd_fast, rd_exact = [], []
    for i in range(len(insts)):
        base, cnt = int(arr[i].accum_base), int(arr[i].accum_count)
        n = float(arr[i].npackets)
        v = a1[base:base + cnt].astype(np.float64) / N.XO_ACCU_K / n
        # radial bins, transmittance, specular: at most the launched weight
        assert v.sum() <= 1.0 + 1e-3, (i, v.sum())
        rd_fast.append(v[:100].sum())
        rd_exact.append(e1[base:base + 100].astype(np.float64).sum() /
                        N.XO_ACCU_K / n)
    d = np.asarray(rd_fast) - np.asarray(rd_exact)
    m, s = batch_mean_se(d)
    assert abs(m) < 3.0 * s + 1e-6, (m, s)
    fast.context.close()
    exact.context.close()
