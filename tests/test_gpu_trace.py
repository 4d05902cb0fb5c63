"""GPU parity of the traced layered kernel (xo_mcml_run_trace) and the
sampling-volume kernel (xo_sampling_volume_run): bit-exact with the oracle
under the static schedule (trace entries, event counts, accumulators, RNG
states) and for packet-keyed runs under the production schedule; the
sampling volume of a trace bit-exact with the oracle's; the Mc-level API
(Trace with a Filter, Mc.sampling_volume) on top."""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcml import mc
from oracle import pyoracle as O

from helpers import XINIT, slab_layers

pytestmark = pytest.mark.gpu


def _sim(method=0, pf=lambda: mc.mcpf.Hg(0.8), mua=300.0):
    layers = slab_layers(pf, n_above=1.0, n_slab=1.37, n_below=1.45, d=1e-3,
                         mua=mua, mus=2e4)
    D = mc.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 2e-3, 20)),
                      bottom=D.Total(), specular=D.Total())
    return mc.Mc(layers, mc.mcsource.Line(), det,
                 options=[mc.mcoptions.McMethod(method)], rnginit=XINIT)


def _xt(maxlen=1500, options=N.XO_TRACE_ALL, mask=None, plon=True):
    t = N.XoTrace()
    t.max_events, t.options, t.plon = maxlen, options, int(plon)
    if mask is not None:
        t.use_events, t.event_mask = 1, mask
    return t


CASES = {
    'aw_all': (0, _xt()),
    'ar_all': (1, _xt()),
    'mbl_all': (2, _xt()),
    'aw_overflow': (0, _xt(6)),
    'aw_end': (0, _xt(1, N.XO_TRACE_END)),
    'aw_start_end_no_pl': (0, _xt(2, N.XO_TRACE_START | N.XO_TRACE_END,
                                  plon=False)),
    'mbl_scattering_mask': (2, _xt(500, mask=N.XO_EVENT_SCATTERING)),
    'aw_boundary_mask': (0, _xt(500, mask=N.XO_EVENT_BOUNDARY_HIT |
                                N.XO_EVENT_TERMINATED)),
}


@pytest.mark.parametrize('name', list(CASES))
def test_trace_static_schedule_bit_exact(name):
    print('case', name, flush=True)
    method, trace = CASES[name]
    sim = _sim(method)
    cfg, lut, naccum = sim.pack()
    lanes, npk = 4096, 20000
    ctx = N.Context(0, 0, XINIT)
    acc, rep, data, n = ctx.run_mcml_trace(cfg, lut, npk, naccum, trace,
                                           lanes, N.XO_FLAG_STATIC_SCHEDULE)
    gx, _ = ctx.get_rng(lanes)
    x, a = N.init_rng(lanes, XINIT)
    racc, counts, rx, rdata, rn = O.mcml_run_trace(
        cfg, lut, x, a, npk, naccum, trace, O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(n, rn)
    np.testing.assert_array_equal(data, rdata)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(gx, rx)
    # the untraced kernel gives the same accumulators
    ctx2 = N.Context(0, 0, XINIT)
    acc0, rep0 = ctx2.run_mcml(cfg, lut, npk, naccum, lanes,
                               N.XO_FLAG_STATIC_SCHEDULE)
    np.testing.assert_array_equal(acc0, acc)
    ctx.close()
    ctx2.close()


def test_keyed_trace_rows_independent_of_schedule():
    sim = _sim(0, pf=lambda: mc.mcpf.MHg(0.8, 0.6))
    cfg, lut, naccum = sim.pack()
    cfg.packet_seed, cfg.first_packet = 0xABCDEF12345, 0
    npk, trace = 30000, _xt(400)
    ctx = N.Context(0, 0, XINIT)
    acc, rep, data, n = ctx.run_mcml_trace(cfg, lut, npk, naccum, trace, 0,
                                           N.XO_FLAG_PACKET_KEYED)
    x, a = N.init_rng(1024, XINIT)
    racc, counts, _, rdata, rn = O.mcml_run_trace(
        cfg, lut, x, a, npk, naccum, trace, O.MATH_KERNEL, keyed=True)
    np.testing.assert_array_equal(n, rn)
    np.testing.assert_array_equal(data, rdata)
    np.testing.assert_array_equal(acc, racc)
    ctx.close()


@pytest.mark.parametrize('end_field', [6, 7])
def test_sampling_volume_bit_exact(end_field):
    sim = _sim(0)
    cfg, lut, naccum = sim.pack()
    x, a = N.init_rng(512, XINIT)
    _, _, _, data, n = O.mcml_run_trace(cfg, lut, x, a, 4000, naccum, _xt(),
                                        O.MATH_KERNEL)
    n[7] = 0   # a packet without events is skipped
    sv = N.XoSamplingVolume()
    for i, (tl, sz, sh) in enumerate(zip((-1e-3, -1e-3, 0.0),
                                         (2e-5, 2e-5, 2e-5), (100, 100, 50))):
        sv.top_left[i], sv.voxel_size[i], sv.shape[i] = tl, sz, sh
    sv.multiplier, sv.k, sv.end_field, sv.offset = 5e4, N.XO_ACCU_K, \
        end_field, 0
    nbins = 100 * 100 * 50
    ctx = N.Context(0, 0, XINIT)
    acc, total, rep = ctx.sampling_volume(sv, data, n, nbins)
    racc, rtotal, rcounts = O.sampling_volume(sv, data, n, nbins)
    assert total == rtotal
    assert int(rep.steps) == int(rcounts[0]) and rep.truncated == 0
    np.testing.assert_array_equal(acc, racc)
    assert acc.sum() > 0
    ctx.close()


def test_mc_trace_filter_and_sampling_volume():
    T, F = mc.mctrace.Trace, mc.mctrace.Filter
    # packets that leave through the top within 0.5 mm of the source
    f = F(z=(-1e-9, 1e-9), pz=(-1.0, 0.0), r=(0.0, 5e-4, (0.0, 0.0)))
    layers = slab_layers(lambda: mc.mcpf.Hg(0.8), n_above=1.0, n_slab=1.37,
                         n_below=1.45, d=1e-3, mua=300.0, mus=2e4)
    sim = mc.Mc(layers, mc.mcsource.Line(), None,
                trace=T(maxlen=1000, filter=f), rnginit=XINIT)
    trace, _, _ = sim.run(20000, static_schedule=True, maxthreads=4096)
    assert 0 < len(trace) < 20000
    t = trace.terminal
    assert np.all(t['z'] == 0.0) and np.all(t['pz'] < 0)
    assert np.all(t['x'] ** 2 + t['y'] ** 2 <= (5e-4) ** 2 * (1 + 1e-6))
    # the same run without a filter, filtered on the host, gives the same
    # packets
    sim2 = mc.Mc(layers, mc.mcsource.Line(), None, trace=T(maxlen=1000),
                 rnginit=XINIT)
    full, _, _ = sim2.run(20000, static_schedule=True, maxthreads=4096)
    sel, dropped = f(full, update=False)
    assert dropped == trace.dropped
    np.testing.assert_array_equal(sel.n, trace.n)
    np.testing.assert_array_equal(sel.data, trace.data)
    A = mc.mcsv.Axis
    sv = mc.mcsv.SamplingVolume(A(-1e-3, 1e-3, 40), A(-1e-3, 1e-3, 40),
                                A(0.0, 1e-3, 20), weight_field='w')
    sv = sim.sampling_volume(trace, sv)
    assert sv.data.shape == (20, 40, 40)
    assert sv.weight == pytest.approx(
        float(np.sum(trace.terminal['w'], dtype=np.float64)), rel=1e-5)
    # the sampling volume of reflected packets sits under the source
    zprof = sv.data.sum(axis=(1, 2))
    assert zprof[0] > 0 and zprof[0] > zprof[-1]


VOX_CASES = {
    'vox_aw_all': (0, _xt(1200)),
    'vox_mbl_all': (2, _xt(1200)),
    'vox_ar_end': (1, _xt(1, N.XO_TRACE_END)),
    'vox_aw_refraction_mask': (0, _xt(300, mask=N.XO_EVENT_REFRACTION)),
}


@pytest.mark.parametrize('name', list(VOX_CASES))
def test_voxel_trace_static_schedule_bit_exact(name):
    from test_gpu_mcvox import skin_vessel_sim
    print('case', name, flush=True)
    method, trace = VOX_CASES[name]
    sim = skin_vessel_sim(n=20, method=method)
    cfg, lut, naccum = sim.pack()
    vox = sim.voxels.material
    lanes, npk = 4096, 12000
    ctx = N.Context(0, 0, XINIT)
    acc, rep, data, n = ctx.run_mcvox(cfg, vox, lut, npk, naccum, lanes,
                                      N.XO_FLAG_STATIC_SCHEDULE, trace=trace)
    x, a = N.init_rng(lanes, XINIT)
    racc, counts, rx, rdata, rn = O.mcvox_run_trace(
        cfg, vox, lut, x, a, npk, naccum, trace, O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    np.testing.assert_array_equal(n, rn)
    np.testing.assert_array_equal(data, rdata)
    np.testing.assert_array_equal(acc, racc)
    ctx.close()


def test_voxel_mc_trace_and_sampling_volume():
    from test_gpu_mcvox import skin_vessel_sim
    from pyxopto_amd.mcvox import mc as mcv
    base = skin_vessel_sim(n=20)
    sim = mcv.Mc(base.voxels, base.materials, base.source, base.detectors,
                 trace=mcv.mctrace.Trace(500), rnginit=XINIT)
    sim.voxels.material[...] = base.voxels.material
    trace, _, det = sim.run(5000)
    assert len(trace) == 5000
    A = mcv.mcsv.Axis
    sv = mcv.mcsv.SamplingVolume(A(-4e-4, 4e-4, 20), A(-4e-4, 4e-4, 20),
                                 A(0.0, 8e-4, 20), weight_field='w')
    sv = sim.sampling_volume(trace, sv)
    assert sv.data.sum() > 0
    assert sv.weight == pytest.approx(
        float(np.sum(trace.terminal['w'], dtype=np.float64)), rel=1e-5)
