"""Host logic of the drop-in API and the C ABI surface (no GPU needed)."""
import ctypes
import re
import os

import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcbase import mcoptions
from pyxopto_amd.mcml import mc

from helpers import XINIT, config1_sim, slab_layers

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), 'include',
                      'xomc.h')


def test_library_exports_every_declared_symbol():
    text = open(HEADER).read()
    declared = set(re.findall(r'^[a-z0-9_ ]+\**\s*\**(xo_[a-z0-9_]+)\s*\(', text,
                              re.M))
    assert len(declared) >= 15
    lib = ctypes.CDLL(N.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(N.PROTOTYPES) == declared


def test_struct_sizes_match_library():
    sizes = (ctypes.c_uint32 * 64)()
    count = N.lib().xo_struct_sizes(sizes, 64)
    assert count == len(N.STRUCTS_IN_ABI_ORDER)
    for i, s in enumerate(N.STRUCTS_IN_ABI_ORDER):
        assert ctypes.sizeof(s) == sizes[i], s.__name__
    assert N.lib().xo_abi_version() == 1


def test_layer_stack_packing():
    sim = config1_sim()
    cfg, lut, naccum = sim.pack()
    assert cfg.num_layers == 3 and lut is None and naccum == 102
    l0, l1, l2 = cfg.layers[0], cfg.layers[1], cfg.layers[2]
    assert l0.top == -np.inf and l0.bottom == 0.0 and l2.bottom == np.inf
    assert l1.top == 0.0 and l1.bottom == np.float32(1e-3)
    assert l2.top == l1.bottom
    # inv_mut = 1/(mua+mus), inf if 0; mua_inv_mut = 1 if mus == 0
    assert l1.inv_mut == np.float32(1.0 / (100.0 + 1e4))
    assert l1.mua_inv_mut == np.float32(100.0 / (100.0 + 1e4))
    assert l0.inv_mut == np.inf and l0.mua_inv_mut == 1.0
    # critical cosines: 1.37 -> 1.0 on top, 1.37 -> 1.45 none at the bottom
    assert l1.cos_critical_top == np.float32((1 - (1 / 1.37) ** 2) ** 0.5)
    assert l1.cos_critical_bottom == 0.0
    assert l2.cos_critical_top == np.float32((1 - (1.37 / 1.45) ** 2) ** 0.5)
    # specular reflectance of the line source ((n0 - n1)/(n0 + n1))^2
    assert abs(cfg.source.reflectance - (0.37 / 2.37) ** 2) < 1e-7
    assert list(cfg.source.direction_sample) == [0.0, 0.0, 1.0]
    # detectors: radial bins first, then bottom and specular totals
    assert (cfg.top.offset, cfg.bottom.offset, cfg.specular.offset) == \
        (0, 100, 101)
    assert cfg.top.kind == N.XO_DET_RADIAL and cfg.top.n == 100
    assert cfg.top.inv_dr == np.float32(100 / 5e-3)


def test_cumulative_bounds_round_through_binary32():
    L = mc.mclayer
    layers = L.Layers([L.Layer(d=np.inf, n=1, mua=0, mus=0, pf=mc.mcpf.Hg(0))]
                      + [L.Layer(d=0.1e-3 * (i + 1), n=1.4, mua=1, mus=10,
                                 pf=mc.mcpf.Hg(0)) for i in range(5)]
                      + [L.Layer(d=np.inf, n=1, mua=0, mus=0,
                                 pf=mc.mcpf.Hg(0))])
    sim = mc.Mc(layers, mc.mcsource.Line(), None, rnginit=XINIT)
    cfg, _, _ = sim.pack()
    bottom = np.float32(0.0)
    for i in range(1, 6):
        bottom = np.float32(float(bottom) + layers[i].d)
        assert cfg.layers[i].bottom == bottom
        assert cfg.layers[i + 1].top == bottom


def test_oblique_line_source_refraction():
    layers = slab_layers(lambda: mc.mcpf.Hg(0.8))
    d = np.array([0.3, 0.0, 0.954])
    d = d / np.linalg.norm(d)
    sim = mc.Mc(layers, mc.mcsource.Line(position=(1e-3, 0, -1e-3),
                                         direction=tuple(d)), None,
                rnginit=XINIT)
    cfg, _, _ = sim.pack()
    s = cfg.source
    # projected onto z = 0 along the direction
    assert abs(s.position[0] - (1e-3 + d[0] * (1e-3 / d[2]))) < 1e-9
    assert s.position[2] == 0.0
    # Snell: sin2 = n0/n1 sin1, direction unit length
    ds = np.array(list(s.direction_sample), dtype=np.float64)
    assert abs(np.linalg.norm(ds) - 1) < 1e-6
    assert abs(ds[0] - d[0] * 1.0 / 1.37) < 1e-7
    assert list(s.direction_reflected) == pytest.approx([d[0], 0, -d[2]])


def test_lut_phase_function_packing_and_dedup():
    from pyxopto_amd.mcbase import pf as hostpf
    lp = mc.mcpf.LutEx(hostpf.MHg, (0.8, 0.5), 500)
    layers = slab_layers(lambda: lp)
    sim = mc.Mc(layers, mc.mcsource.Line(), None, rnginit=XINIT)
    cfg, lut, _ = sim.pack()
    assert cfg.pf_kind == N.XO_PF_LUT
    assert lut.size == 500                     # one table shared by 3 layers
    assert cfg.layers[1].pf.offset == 0 and cfg.layers[1].pf.size == 500
    assert lut[0] == pytest.approx(-1.0, abs=1e-3)
    assert lut[-1] == pytest.approx(1.0, abs=1e-3)
    assert np.all(np.diff(lut) >= 0)


def test_errors_mirror_reference():
    layers = slab_layers(lambda: mc.mcpf.Hg(0.8))
    with pytest.raises(ValueError, match='not valid'):
        mc.Mc(layers, mc.mcsource.Line(), None, rnginit=0)
    with pytest.raises(ValueError, match='not valid'):
        mc.Mc(layers, mc.mcsource.Line(), None, rnginit=(4294967117 << 32))
    sim = mc.Mc(layers, mc.mcsource.Line(), None, rnginit=XINIT)
    with pytest.raises(ValueError):
        sim.rmax = 0.0
    with pytest.raises(ValueError):
        sim.run(-5)
    L = mc.mclayer
    with pytest.raises(TypeError):
        L.Layers([L.Layer(d=1, n=1, mua=0, mus=0, pf=mc.mcpf.Hg(0)),
                  L.Layer(d=1, n=1, mua=0, mus=0, pf=mc.mcpf.MHg(0, 1)),
                  L.Layer(d=1, n=1, mua=0, mus=0, pf=mc.mcpf.Hg(0))])
    with pytest.raises(ValueError):
        L.Layers([L.Layer(d=1, n=1, mua=0, mus=0, pf=mc.mcpf.Hg(0))])
    with pytest.raises(ValueError):
        sim.layer(1).pf = mc.mcpf.MHg(0.1, 0.2)
    with pytest.raises(ValueError):
        mcoptions.McMethod('mbl')
    with pytest.raises(ValueError):
        mc.mcsource.Line(direction=(0, 0, 0))


def test_options_reach_the_packed_config():
    layers = slab_layers(lambda: mc.mcpf.Hg(0.8))
    sim = mc.Mc(layers, mc.mcsource.Line(), None, rnginit=XINIT,
                options=[mcoptions.McMinimumPacketWeight(1e-3),
                         mcoptions.McPacketLotteryChance(0.05),
                         mcoptions.McUseNativeMath.on])
    cfg, _, _ = sim.pack()
    assert cfg.w_min == np.float32(1e-3)
    assert cfg.lottery_chance == np.float32(0.05)
    assert cfg.use_lottery == 1 and cfg.method == 0


def test_normalization_closed_forms():
    D = mc.mcdetector
    r = D.Radial(D.RadialAxis(0.0, 1e-3, 4))
    r.update_data(None, [np.array([N.XO_ACCU_K] * 4, dtype=np.uint64)], 10)
    e = np.linspace(0, 1e-3, 5)
    np.testing.assert_allclose(r.reflectance,
                               1.0 / (10 * np.pi * (e[1:] ** 2 - e[:-1] ** 2)))
    t = D.Total()
    t.update_data(None, [np.array([2 * N.XO_ACCU_K], dtype=np.uint64)], 4)
    assert t.reflectance[0] == pytest.approx(0.5)
    ax = D.RadialAxis(0.0, 2.0, 2)
    np.testing.assert_allclose(ax.centers, [2 / 3 * 1 / 1, 2 / 3 * 7 / 3])


def test_no_gpu_raises_loudly_without_device():
    if N.device_count() > 0:
        pytest.skip('a device is present')
    with pytest.raises(N.NativeError, match='no HIP device'):
        N.Context(0, 0, XINIT)


def test_rng_seeds_before_any_run_are_the_initial_states():
    """rng_seeds_x / rng_seeds_a hold init_RNG's states for every seed row
    (xopto/mcbase/mcworker.py:1219-1237) without touching a device."""
    sim = config1_sim()
    x, a = N.init_rng(64, sim.rnginit)
    assert sim.rng_seeds_x.shape == (N.XO_MAX_SEEDS,)
    np.testing.assert_array_equal(sim.rng_seeds_x[:64], x)
    np.testing.assert_array_equal(sim.rng_seeds_a[:64], a)
