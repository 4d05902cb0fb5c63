"""Synthetic workloads used by the benchmark, the smoke test and the tests.

- ``slab_1mm``: one turbid slab 1 mm thick (n 1.37, mua 1/cm, mus 100/cm,
  HG g 0.8) between air (n 1.0) and glass (n 1.45), pencil beam, radial
  reflectance 0-5 mm in 100 bins, total transmittance and specular
  reflectance; 1e6 packets.
- ``skin_vessel_200``: 200^3 voxels of 10 um; epidermis (voxel centres at
  z <= 100 um: mua 16/cm, mus 400/cm, g 0.8, n 1.34) over dermis (mua
  0.5/cm, mus 250/cm, g 0.9, n 1.40) with a horizontal vessel along y,
  radius 100 um, axis at 400 um depth (mua 15/cm, mus 300/cm, g 0.98,
  n 1.40 as the dermis); Gaussian beam sigma 50 um at normal incidence;
  energy deposition in every voxel; 1e8 packets. The surrounding medium is
  n 1.0 with no absorption or scattering.
- ``layered_batch``: 4096 layered instances with 1-5 turbid layers and
  uneven packet budgets for the batched engine (see its docstring).
- ``fiber_*``: the reference's published benchmark (400 single-layer runs
  under a multimode fiber, 1e7 packets each), see FIBER_* below.
- ``skin_two_layer``: 64^3 voxels of 31.25 um (a 2 mm cube), epidermis
  100 um over dermis (the skin_vessel_200 materials without the vessel),
  Gaussian beam sigma 50 um, deposition in every voxel (64^3 u64 bins,
  2 MiB); the voxel requests of the serving workload, 1e6 packets each.
"""
import os

import numpy as np

from .mcml import mc as mcml
from .mcvox import mc as mcvox

SLAB_PACKETS = 1_000_000
SKIN_VESSEL_PACKETS = 100_000_000


def slab_1mm(rnginit: int = 0x5EED000000000001, cl_build_options=None):
    L = mcml.mclayer
    layers = L.Layers([
        L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=mcml.mcpf.Hg(0.8)),
        L.Layer(d=1e-3, n=1.37, mua=100.0, mus=1e4, pf=mcml.mcpf.Hg(0.8)),
        L.Layer(d=np.inf, n=1.45, mua=0.0, mus=0.0, pf=mcml.mcpf.Hg(0.8))])
    D = mcml.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 5e-3, 100)),
                      bottom=D.Total(), specular=D.Total())
    return mcml.Mc(layers, mcml.mcsource.Line(), det, rnginit=rnginit,
                   cl_build_options=cl_build_options)


def skin_vessel(n: int = 200, voxel: float = 10e-6,
                rnginit: int = 0x5EED000000000001, epidermis: float = 100e-6,
                vessel_depth: float = 400e-6, vessel_radius: float = 100e-6,
                cl_build_options=None):
    half = 0.5 * n * voxel
    M = mcvox.mcmaterial
    materials = M.Materials([
        M.Material(n=1.0, mua=0.0, mus=0.0, pf=mcvox.mcpf.Hg(0.8)),
        M.Material(n=1.34, mua=1600.0, mus=4e4, pf=mcvox.mcpf.Hg(0.8)),
        M.Material(n=1.40, mua=50.0, mus=2.5e4, pf=mcvox.mcpf.Hg(0.9)),
        M.Material(n=1.40, mua=1500.0, mus=3e4, pf=mcvox.mcpf.Hg(0.98))])
    G = mcvox.mcgeometry
    ax = (G.Axis(-half, half, n), G.Axis(-half, half, n),
          G.Axis(0.0, 2 * half, n))
    voxels = G.Voxels(*ax)
    fluence = mcvox.mcfluence.Fluence(G.Axis(-half, half, n),
                                      G.Axis(-half, half, n),
                                      G.Axis(0.0, 2 * half, n))
    D = mcvox.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, half, 100)),
                      bottom=D.Total(), specular=D.Total())
    sim = mcvox.Mc(voxels, materials, mcvox.mcsource.GaussianBeam(50e-6), det,
                   fluence=fluence, rnginit=rnginit,
                   cl_build_options=cl_build_options)
    z, _, x = sim.voxels.meshgrid()
    mat = sim.voxels.material
    mat[...] = 2
    mat[z <= epidermis] = 1
    mat[x ** 2 + (z - vessel_depth) ** 2 <= vessel_radius ** 2] = 3
    sim.voxels.update()
    return sim


def skin_two_layer(n: int = 64, size: float = 2e-3,
                   rnginit: int = 0x5EED000000000001,
                   epidermis: float = 100e-6, cl_build_options=None):
    half = 0.5 * size
    M = mcvox.mcmaterial
    materials = M.Materials([
        M.Material(n=1.0, mua=0.0, mus=0.0, pf=mcvox.mcpf.Hg(0.8)),
        M.Material(n=1.34, mua=1600.0, mus=4e4, pf=mcvox.mcpf.Hg(0.8)),
        M.Material(n=1.40, mua=50.0, mus=2.5e4, pf=mcvox.mcpf.Hg(0.9))])
    G = mcvox.mcgeometry
    voxels = G.Voxels(G.Axis(-half, half, n), G.Axis(-half, half, n),
                      G.Axis(0.0, size, n))
    fluence = mcvox.mcfluence.Fluence(G.Axis(-half, half, n),
                                      G.Axis(-half, half, n),
                                      G.Axis(0.0, size, n))
    D = mcvox.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, half, 100)),
                      bottom=D.Total(), specular=D.Total())
    sim = mcvox.Mc(voxels, materials, mcvox.mcsource.GaussianBeam(50e-6), det,
                   fluence=fluence, rnginit=rnginit,
                   cl_build_options=cl_build_options)
    z, _, _ = sim.voxels.meshgrid()
    mat = sim.voxels.material
    mat[...] = 2
    mat[z <= epidermis] = 1
    sim.voxels.update()
    return sim


# The reference's published benchmark (test/performance.py:36, 115-130 with
# xopto/mcml/test/validate.py:348-511): a 20 x 20 grid of mua 1-2500 /m and
# mus' 500-6000 /m of one 8 mm layer (n 1.33, HG g 0.85) under a 200 um
# core, NA 0.22 fiber (n 1.452, also the medium above), radial detector
# 0-1.7 mm in 340 bins collecting within the fiber acceptance, rmax 5 mm,
# 1e7 packets per grid point.
FIBER_GRID_MUA = np.linspace(1.0, 2500.0, 20)
FIBER_GRID_MUSR = np.linspace(500.0, 6000.0, 20)
FIBER_PACKETS = 10_000_000
FIBER_SDS = np.array([220e-6, 440e-6, 660e-6, 880e-6, 1100e-6])


def fiber_layers(mua: float = 0.0, musr: float = 0.0, g: float = 0.85):
    L = mcml.mclayer
    pf = mcml.mcpf.Hg(g)
    return L.Layers([
        L.Layer(d=np.inf, n=1.452, mua=0.0, mus=0.0, pf=pf),
        L.Layer(d=8e-3, n=1.33, mua=mua, mus=musr / (1.0 - g), pf=pf),
        L.Layer(d=np.inf, n=1.33, mua=0.0, mus=0.0, pf=pf)])


def fiber_source():
    fib = mcml.fiberutil.MultimodeFiber(dcore=200e-6, dcladding=200e-6,
                                        ncore=1.452, na=0.22)
    return mcml.mcsource.UniformFiberNI(fib)


def fiber_detectors():
    D = mcml.mcdetector
    cosmin = (1.0 - (0.22 / 1.452) ** 2) ** 0.5
    return D.Detectors(top=D.Radial(D.RadialAxis(0.0, 1.7e-3, 340),
                                    cosmin=cosmin))


def fiber_grid():
    """(mua, mus') of the 400 runs, in the reference's order (mua major)."""
    mua, musr = np.meshgrid(FIBER_GRID_MUA, FIBER_GRID_MUSR, indexing='ij')
    return mua.ravel(), musr.ravel()


def fiber_sim(rnginit: int = 0x5EED000000000001, cl_build_options=None):
    """The benchmark's simulator (mua, mus' set per run, as the reference
    test does)."""
    sim = mcml.Mc(fiber_layers(), fiber_source(), fiber_detectors(),
                  rnginit=rnginit, cl_build_options=cl_build_options)
    sim.rmax = 5e-3
    return sim


def fiber_batch(packets: int = FIBER_PACKETS, seed: int = 20261204):
    """The 400 runs of the benchmark as instances of one batch."""
    from .mcml import batch as mbatch
    out = []
    for k, (mua, musr) in enumerate(zip(*fiber_grid())):
        out.append(mbatch.Instance(fiber_layers(mua, musr), fiber_source(),
                                   fiber_detectors(), nphotons=packets,
                                   seed=seed * 1000 + k, rmax=5e-3))
    return out


VOXEL_REQUEST_PACKETS = 1_000_000
BATCH_INSTANCES = 4096


def _lut_params(args):
    from .mcbase import pf as hostpf
    g, b = args
    return hostpf.MHg(g, b).mclut(2000)


def _luts_cached(lut_jobs, cache_file, processes):
    """LUT parameters and tables of the (g, beta) pairs, from cache_file
    when it holds exactly these pairs (an .npz of plain arrays)."""
    keys = np.array(lut_jobs, dtype=np.float64).reshape(-1, 2)
    if cache_file and os.path.isfile(cache_file):
        with np.load(cache_file, allow_pickle=False) as z:
            if z['keys'].shape == keys.shape and np.array_equal(z['keys'],
                                                                 keys):
                return [(list(p), t) for p, t in zip(z['params'],
                                                     z['tables'])]
    if processes == 1 or not lut_jobs:
        results = [_lut_params(j) for j in lut_jobs]
    else:
        import multiprocessing as mp
        procs = processes or min(16, len(os.sched_getaffinity(0)))
        with mp.get_context('spawn').Pool(procs) as pool:
            results = pool.map(_lut_params, lut_jobs, chunksize=16)
    if cache_file and results:
        os.makedirs(os.path.dirname(cache_file), exist_ok=True)
        tmp = cache_file + '.tmp.npz'
        np.savez(tmp, keys=keys,
                 params=np.array([r[0] for r in results], dtype=np.float64),
                 tables=np.array([r[1] for r in results]))
        os.replace(tmp, cache_file)
    return results


def layered_batch_extremes(n: int = BATCH_INSTANCES, seed: int = 20261202,
                           per_kind: int = 8):
    """Indices of instances of layered_batch(n, seed) at the corners of
    its parameter box: the largest total optical thickness (sum of mus d),
    the lowest absorption, the highest scattering, the thickest layer,
    five-layer lookup-table stacks, plus evenly spaced ones; per_kind of
    each, without repeats, in increasing order."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        nl = int(rng.integers(1, 6))
        d = np.exp(rng.uniform(np.log(0.05e-3), np.log(5e-3), nl))
        mua = np.exp(rng.uniform(np.log(1.0), np.log(5000.0), nl))
        mus = np.exp(rng.uniform(np.log(1e3), np.log(5e4), nl))
        rng.uniform(1.33, 1.50, nl)
        use_lut = bool(rng.integers(0, 2))
        rng.uniform(0.6, 0.95, nl)
        rng.uniform(0.0, 1.0, nl)
        rng.uniform(np.log(1e4), np.log(1e6))
        rng.integers(1, 2 ** 63)
        rows.append((float((mus * d).sum()), float(mua.min()),
                     float(mus.max()), float(d.max()), nl == 5 and use_lut))
    t = np.array([r[:4] for r in rows])
    picks = []
    for order in (np.argsort(-t[:, 0]), np.argsort(t[:, 1]),
                  np.argsort(-t[:, 2]), np.argsort(-t[:, 3]),
                  np.array([i for i, r in enumerate(rows) if r[4]]),
                  np.linspace(0, n - 1, per_kind).astype(int)):
        k = 0
        for i in order:
            if k == per_kind:
                break
            if int(i) not in picks:
                picks.append(int(i))
                k += 1
    return sorted(picks)


def layered_batch(n: int = BATCH_INSTANCES, seed: int = 20261202,
                  budget_scale: float = 1.0, processes: int = None,
                  cache_dir: str = None, indices=None):
    """``n`` layered instances drawn from a seeded generator: 1-5 turbid
    layers each (thickness log-uniform 0.05-5 mm, mua log-uniform
    0.01-50 /cm, mus log-uniform 10-500 /cm, n uniform 1.33-1.50) between
    air above and below; phase function per instance either HG (g uniform
    0.6-0.95 per layer) or a 2000-point lookup table of modified HG (g
    uniform 0.6-0.95, beta uniform 0-1 per layer); packet budget
    log-uniform 1e4-1e6 (times budget_scale); 100-bin radial reflectance
    0-5 mm plus total transmittance and specular reflectance each.
    The lookup tables are built by a pool of `processes` spawned workers
    (1: in this process) or read back from an .npz in cache_dir written by
    an earlier call. Returns a list of batch.Instance; with `indices`, only
    those instances of the same draw (and only their tables are built)."""
    from .mcml import batch as mbatch
    rng = np.random.default_rng(seed)
    specs = []
    lut_jobs = []
    for i in range(n):
        nl = int(rng.integers(1, 6))
        d = np.exp(rng.uniform(np.log(0.05e-3), np.log(5e-3), nl))
        mua = np.exp(rng.uniform(np.log(1.0), np.log(5000.0), nl))
        mus = np.exp(rng.uniform(np.log(1e3), np.log(5e4), nl))
        nn = rng.uniform(1.33, 1.50, nl)
        use_lut = bool(rng.integers(0, 2))
        g = rng.uniform(0.6, 0.95, nl)
        beta = rng.uniform(0.0, 1.0, nl)
        budget = int(np.exp(rng.uniform(np.log(1e4), np.log(1e6))) *
                     budget_scale)
        inst_seed = int(rng.integers(1, 2 ** 63))
        specs.append((nl, d, mua, mus, nn, use_lut, g, beta, max(budget, 1),
                      inst_seed))
    if indices is not None:
        specs = [specs[int(i)] for i in indices]
    for nl, _, _, _, _, use_lut, g, beta, _, _ in specs:
        if use_lut:
            lut_jobs.extend((float(g[k]), float(beta[k])) for k in range(nl))
    luts = {}
    if lut_jobs:
        cache_file = None if cache_dir is None or indices is not None else \
            os.path.join(cache_dir,
                         'layered_batch_luts_{}_{}.npz'.format(n, seed))
        luts = dict(zip(lut_jobs, _luts_cached(lut_jobs, cache_file,
                                               processes)))
    L = mcml.mclayer
    D = mcml.mcdetector
    out = []
    for nl, d, mua, mus, nn, use_lut, g, beta, budget, inst_seed in specs:
        if use_lut:
            pfs = [mcml.mcpf.Lut(*luts[(float(g[k]), float(beta[k]))])
                   for k in range(nl)]
            amb = pfs[0]
        else:
            pfs = [mcml.mcpf.Hg(float(g[k])) for k in range(nl)]
            amb = mcml.mcpf.Hg(0.0)
        layers = [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=amb)]
        layers += [L.Layer(d=float(d[k]), n=float(nn[k]), mua=float(mua[k]),
                           mus=float(mus[k]), pf=pfs[k]) for k in range(nl)]
        layers += [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=amb)]
        det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 5e-3, 100)),
                          bottom=D.Total(), specular=D.Total())
        out.append(mbatch.Instance(layers, mcml.mcsource.Line(), det,
                                   nphotons=budget, seed=inst_seed))
    return out
