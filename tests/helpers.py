"""Shared builders and statistics for the tests.

Statistical parity: the reference keeps no per-bin variance, so the standard
error of each bin comes from K independent batches (batch means); a bin
passes when |z| = |m_gpu - m_ref| / sqrt(se_gpu^2 + se_ref^2) < 4, and totals
within 3 combined standard errors. The standard errors are themselves
estimates (t-distributed with about 2K-2 degrees of freedom), so the tests
use K >= 40 batches per side: with K = 10 a 100-bin comparison would fail
by chance in several percent of runs.
"""
import numpy as np

from pyxopto_amd import _native as N
from pyxopto_amd.mcml import mc

XINIT = 0x5EED000000000001
XINIT_ORACLE = 0x0123456789ABCDEF


def slab_layers(pf_factory, n_above=1.0, n_slab=1.37, n_below=1.45,
                d=1e-3, mua=100.0, mus=1e4):
    L = mc.mclayer
    return L.Layers([
        L.Layer(d=np.inf, n=n_above, mua=0.0, mus=0.0, pf=pf_factory()),
        L.Layer(d=d, n=n_slab, mua=mua, mus=mus, pf=pf_factory()),
        L.Layer(d=np.inf, n=n_below, mua=0.0, mus=0.0, pf=pf_factory())])


def standard_detectors(n=100, rmax=5e-3, cosmin=0.0, logscale=False):
    D = mc.mcdetector
    start = 1e-5 if logscale else 0.0
    return D.Detectors(top=D.Radial(D.RadialAxis(start, rmax, n, logscale),
                                    cosmin=cosmin),
                       bottom=D.Total(), specular=D.Total())


def config1_sim(rnginit=XINIT):
    """1 mm slab, n 1.37 between 1.0 and 1.45, mua 1/cm, mus 100/cm,
    HG g=0.8, radial 0-5 mm in 100 bins, total R/T."""
    layers = slab_layers(lambda: mc.mcpf.Hg(0.8))
    return mc.Mc(layers, mc.mcsource.Line(), standard_detectors(),
                 rnginit=rnginit)


def batch_mean_se(batches):
    b = np.asarray(batches, dtype=np.float64)
    k = b.shape[0]
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(k)


def z_scores(m1, se1, m2, se2):
    se = np.sqrt(se1 ** 2 + se2 ** 2)
    z = np.zeros_like(m1)
    ok = se > 0
    z[ok] = (m1[ok] - m2[ok]) / se[ok]
    both_zero = ~ok & (m1 == m2)
    z[~ok & ~both_zero] = np.inf
    return z


def bins_to_test(b1, b2, min_nonzero=1.0):
    """Bins where every batch of both sides collected weight."""
    return (np.asarray(b1) > 0).all(axis=0) & (np.asarray(b2) > 0).all(axis=0)


def oracle_batches(run_one, k, lanes, xinit=XINIT_ORACLE):
    """K oracle batches on disjoint seed rows; run_one(x, a) -> vector."""
    x, a = N.init_rng(lanes * k, xinit)
    out = []
    for i in range(k):
        sl = slice(i * lanes, (i + 1) * lanes)
        out.append(run_one(x[sl], a[sl]))
    return np.asarray(out)
