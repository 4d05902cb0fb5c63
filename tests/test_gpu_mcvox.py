"""GPU parity of the voxelized kernel (xo_mcvox_run through the C ABI):
bit-exact against the oracle with the static schedule, statistical with the
production schedule, and the layered-vs-voxel cross-check on the GPU."""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcvox import mc
from oracle import pyoracle as O

from helpers import (XINIT, batch_mean_se, bins_to_test, oracle_batches,
                     z_scores)
from test_oracle_mcvox import column_sim, fixture_voxel_sim, two_layer_pair

pytestmark = pytest.mark.gpu


def skin_vessel_sim(n=40, size=0.8e-3, lut=False, method=0, rz=False):
    """Two-layer skin with a vessel (the mcvox 200^3 workload, reduced);
    `method` is the McMethod value, `rz` swaps the Cartesian fluence for
    an r-z one."""
    M = mc.mcmaterial
    if lut:
        from pyxopto_amd.mcbase import pf as hostpf
        pfs = [mc.mcpf.LutEx(hostpf.MHg, (g, 0.8), 1000)
               for g in (0.8, 0.8, 0.9, 0.95)]
    else:
        pfs = [mc.mcpf.Hg(g) for g in (0.8, 0.8, 0.9, 0.98)]
    materials = M.Materials([
        M.Material(n=1.0, mua=0.0, mus=0.0, pf=pfs[0]),
        M.Material(n=1.34, mua=1600.0, mus=4e4, pf=pfs[1]),
        M.Material(n=1.40, mua=50.0, mus=2.5e4, pf=pfs[2]),
        M.Material(n=1.40, mua=1500.0, mus=3e4, pf=pfs[3])])
    G = mc.mcgeometry
    voxels = G.Voxels(G.Axis(-size / 2, size / 2, n),
                      G.Axis(-size / 2, size / 2, n), G.Axis(0.0, size, n))
    if rz:
        fl = mc.mcfluence.FluenceRz(G.Axis(0.0, size / 2, n // 2),
                                    G.Axis(0.0, size, n))
    else:
        fl = mc.mcfluence.Fluence(G.Axis(-size / 2, size / 2, n),
                                  G.Axis(-size / 2, size / 2, n),
                                  G.Axis(0.0, size, n))
    D = mc.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0, size / 2, 20)),
                      bottom=D.Total(), specular=D.Total())
    sim = mc.Mc(voxels, materials, mc.mcsource.GaussianBeam(50e-6), det,
                fluence=fl, options=[mc.mcoptions.McMethod(method)],
                rnginit=XINIT)
    z, y, x = sim.voxels.meshgrid()
    mat = sim.voxels.material
    mat[...] = 2
    mat[z <= 100e-6] = 1
    mat[x ** 2 + (z - 0.4e-3) ** 2 <= 100e-6 ** 2] = 3
    sim.voxels.update()
    return sim


CASES = {
    'skin_vessel_hg': lambda: skin_vessel_sim(),
    'skin_vessel_lut': lambda: skin_vessel_sim(lut=True),
    'fixture_1x1x1': lambda: fixture_voxel_sim()[0],
    'column_gauss': lambda: column_sim(sigma=1e-4),
}


@pytest.mark.parametrize('name', list(CASES))
def test_static_schedule_bit_exact(name):
    print('case', name, flush=True)
    sim = CASES[name]()
    lanes, npk = 8192, 40000
    cfg, lut, naccum = sim.pack()
    ctx = N.Context(0, 0, XINIT)
    gpu, rep = ctx.run_mcvox(cfg, sim.voxels.material, lut, npk, naccum,
                             lanes, N.XO_FLAG_STATIC_SCHEDULE)
    gx, _ = ctx.get_rng(lanes)
    x, a = N.init_rng(lanes, XINIT)
    ref, counts, rx = O.mcvox_run(cfg, sim.voxels.material, lut, x, a, npk,
                                  naccum, O.MATH_KERNEL)
    assert rep.packets == npk
    assert int(rep.steps) == int(counts[0])
    assert int(rep.boundary_steps) == int(counts[1])
    assert int(rep.voxel_crossings) == int(counts[2])
    assert int(rep.deposits) == int(counts[3])
    np.testing.assert_array_equal(gpu, ref)
    np.testing.assert_array_equal(gx, rx)
    # a second run without re-uploading the grid continues the streams
    gpu2, _ = ctx.run_mcvox(cfg, None, lut, npk, naccum, lanes,
                            N.XO_FLAG_STATIC_SCHEDULE)
    ref2, _, _ = O.mcvox_run(cfg, sim.voxels.material, lut, rx, a, npk,
                             naccum, O.MATH_KERNEL)
    np.testing.assert_array_equal(gpu2, ref2)
    ctx.close()


def test_dynamic_schedule_statistical_parity_skin_vessel():
    sim = skin_vessel_sim()
    cfg, lut, naccum = sim.pack()
    k, npk = 40, 20000
    ctx = N.Context(0, 0, XINIT)
    g = []
    vox = sim.voxels.material
    for i in range(k):
        acc, rep = ctx.run_mcvox(cfg, vox if i == 0 else None, lut, npk,
                                 naccum, 0, 0)
        assert rep.packets == npk
        g.append(acc.astype(np.float64) / N.XO_ACCU_K / npk)
    g = np.asarray(g)
    o = oracle_batches(lambda x, a: O.mcvox_run(
        cfg, vox, lut, x, a, npk, naccum)[0] / N.XO_ACCU_K / npk, k, 512)
    # detectors exactly; fluence through its z profile (per-voxel counts of
    # the 40^3 grid are too sparse for a z-test at this size)
    off, n = sim.allocation(sim.fluence)
    def reduce(b):
        det = b[:, :off]
        prof = b[:, off:off + n].reshape(b.shape[0], 40, 40, 40).sum(axis=(2, 3))
        return np.concatenate([det, prof], axis=1)
    gr, orr = reduce(g), reduce(o)
    m1, s1 = batch_mean_se(gr)
    m2, s2 = batch_mean_se(orr)
    z = z_scores(m1, s1, m2, s2)
    ok = bins_to_test(gr, orr)
    assert ok.sum() >= 40
    assert np.all(np.abs(z[ok]) < 4.0), z[ok]
    tot1, tot2 = gr[:, :20].sum(axis=1), orr[:, :20].sum(axis=1)
    a1, b1 = batch_mean_se(tot1)
    a2, b2 = batch_mean_se(tot2)
    assert abs(a1 - a2) < 3 * np.sqrt(b1 ** 2 + b2 ** 2)
    ctx.close()


def test_gpu_voxel_matches_gpu_layered():
    lay, vox = two_layer_pair()
    k, npk = 40, 100000
    g, v = [], []
    for i in range(k):
        g.append(lay.run(npk)[2].top.raw / npk)
        v.append(vox.run(npk)[2].top.raw / npk)
    m1, s1 = batch_mean_se(np.asarray(g))
    m2, s2 = batch_mean_se(np.asarray(v))
    z = z_scores(m1, s1, m2, s2)
    assert np.all(np.abs(z) < 4.0), z


def test_mcvox_api_fluence_normalization():
    sim = column_sim(mua=300.0, nz=40, depth=4e-3)
    _, fl, det = sim.run(200000)
    # deposition density integrates to the absorbed fraction
    absorbed = fl.data.sum() * fl.dx * fl.dy * fl.dz
    assert abs(absorbed - (1 - np.exp(-300.0 * 4e-3))) < 5e-3
    assert fl.nphotons == 200000


def test_full_size_skin_vessel_keyed_runs_independent_of_launch_shape():
    """The 200^3 skin-vessel deposition at its full size (1e8 packets) with
    packet-keyed generators: two launches with different lane counts give
    the same integer accumulators bit for bit (every packet draws from its
    own generator, and u64 adds commute)."""
    from pyxopto_amd import workloads as W
    sim = W.skin_vessel()
    cfg, lut, naccum = sim.pack()
    vox = sim.voxels.material
    cfg.packet_seed, cfg.first_packet = 0x5EED1234, 0
    npk = 100_000_000
    print('skin_vessel 200^3, 1e8 packets, two launch shapes', flush=True)
    ctx = N.Context(0, 0, XINIT)
    a1, r1 = ctx.run_mcvox(cfg, vox, lut, npk, naccum, 0, N.XO_FLAG_PACKET_KEYED)
    a2, r2 = ctx.run_mcvox(cfg, vox, lut, npk, naccum, 65536,
                           N.XO_FLAG_PACKET_KEYED)
    ctx.close()
    assert r1.packets == npk and r2.packets == npk
    assert int(r1.steps) == int(r2.steps)
    np.testing.assert_array_equal(a1, a2)
    off = int(cfg.fluence.offset)
    n = 200 ** 3
    absorbed = a1[off:off + n].astype(np.float64).sum() / N.XO_ACCU_K / npk
    print('absorbed fraction', absorbed, flush=True)
    assert 0.1 < absorbed < 1.0, absorbed
