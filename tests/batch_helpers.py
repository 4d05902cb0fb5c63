"""Small mixed batches for the batched-engine tests."""
import numpy as np

from pyxopto_amd.mcbase import pf as hostpf
from pyxopto_amd.mcml import batch as mbatch
from pyxopto_amd.mcml import mc


def small_batch(n=40, seed=7, budget=(2000, 20000), with_mhg=True):
    rng = np.random.default_rng(seed)
    L, D = mc.mclayer, mc.mcdetector
    lut_cache = {}
    out = []
    for i in range(n):
        nl = int(rng.integers(1, 6))
        kind = ['hg', 'lut', 'mhg'][i % (3 if with_mhg else 2)]
        pfs = []
        for k in range(nl):
            g = float(np.round(rng.uniform(0.6, 0.95), 2))
            b = float(np.round(rng.uniform(0.0, 1.0), 1))
            if kind == 'hg':
                pfs.append(mc.mcpf.Hg(g))
            elif kind == 'mhg':
                pfs.append(mc.mcpf.MHg(g, b))
            else:
                if (g, b) not in lut_cache:
                    lut_cache[(g, b)] = hostpf.MHg(g, b).mclut(500)
                pfs.append(mc.mcpf.Lut(*lut_cache[(g, b)]))
        amb = pfs[0] if kind != 'hg' else mc.mcpf.Hg(0.0)
        layers = [L.Layer(d=np.inf, n=1.0, mua=0, mus=0, pf=amb)]
        for k in range(nl):
            layers.append(L.Layer(
                d=float(np.exp(rng.uniform(np.log(5e-5), np.log(2e-3)))),
                n=float(rng.uniform(1.33, 1.5)),
                mua=float(np.exp(rng.uniform(np.log(50), np.log(5000)))),
                mus=float(np.exp(rng.uniform(np.log(1e3), np.log(1e4)))),
                pf=pfs[k]))
        layers.append(L.Layer(d=np.inf, n=1.0, mua=0, mus=0, pf=amb))
        det = D.Detectors(top=D.Radial(D.RadialAxis(0, 5e-3, 100)),
                          bottom=D.Total(), specular=D.Total())
        npk = int(rng.integers(budget[0], budget[1]))
        out.append(mbatch.Instance(layers, mc.mcsource.Line(), det,
                                   nphotons=npk,
                                   seed=int(rng.integers(1, 2 ** 62)),
                                   rmax=float('inf') if i % 5 else 3e-3))
    return out
