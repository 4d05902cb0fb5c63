"""Serving engine on the GPU: every request's integers are the same alone
and in the mix (bit-exact with the batch engine for layered requests and
with a keyed plain run for voxel requests), requests that do not fit yet
wait for memory instead of failing, cancellation releases memory, and the
captured tick graph gives the result of eager launches on every replay.
Requests larger than the server are rejected at once."""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd import serve as S
from pyxopto_amd import workloads as W
from pyxopto_amd.mcml.batch import Batch

from batch_helpers import small_batch

pytestmark = pytest.mark.gpu

XINIT = 0x5EED000000000001


def _alone_batch(inst):
    """Integers of one instance run alone by the batch engine."""
    b = Batch([inst], rnginit=XINIT)
    arr, lut, naccum = b.pack()
    acc, _ = b.context.run_batch(arr, lut, b.opts(chunk=512), naccum)
    return acc


def _packed(inst):
    b = Batch([inst], rnginit=XINIT)
    arr, lut, _ = b.pack()
    return arr[0], lut


@pytest.fixture(scope='module')
def voxel_sim():
    return W.skin_two_layer(n=24, size=1.2e-3)


def test_layered_and_voxel_requests_same_alone_and_in_mix(voxel_sim):
    insts = small_batch(24, seed=11, budget=(2000, 60000))
    seed_vox = 0xC0FFEE
    nvox = 150000
    cfg, lut, naccum = voxel_sim.pack()
    vox_alone, _ = voxel_sim.context.run_mcvox(
        cfg_with_seed(cfg, seed_vox), voxel_sim.voxels.material, lut, nvox,
        naccum, 0, N.XO_FLAG_PACKET_KEYED)
    # small ticks so requests span several ticks and overlap each other
    with S.Server(tick_packets=200000, vox_tick_packets=40000,
                  chunk=512) as srv:
        reqs = []
        for k, inst in enumerate(insts):
            arr, l = _packed(inst)
            reqs.append(srv.submit_packed_layered(arr, l))
            if k == 5:
                vreq = srv.submit_packed_voxel(
                    cfg, voxel_sim.voxels.material, lut, naccum, nvox,
                    seed_vox)
        assert srv.drain(120)
        st = srv.stats()
        assert st['completed'] == len(insts) + 1
        assert st['ticks'] >= 2
        assert st['accum_used'] == 0 and st['voxel_used'] == 0
    np.testing.assert_array_equal(vreq.result(), vox_alone)
    for inst, r in zip(insts, reqs):
        np.testing.assert_array_equal(r.result(), _alone_batch(inst))


def cfg_with_seed(cfg, seed):
    out = N.XoMcvoxCfg.from_buffer_copy(cfg)
    out.packet_seed = seed
    out.first_packet = 0
    return out


def test_keyed_plain_layered_run_equals_batch_engine():
    inst = small_batch(3, seed=5, budget=(30000, 40000))[1]
    from pyxopto_amd.mcml import mc
    sim = mc.Mc(inst.layers, inst.source, inst.detectors, rnginit=XINIT)
    sim.rmax = inst.rmax
    _, _, det = sim.run(inst.nphotons, packet_seed=inst.seed)
    acc = _alone_batch(inst)
    got = np.concatenate([det.top.raw, det.bottom.raw, det.specular.raw])
    np.testing.assert_array_equal(
        np.rint(got * N.XO_ACCU_K).astype(np.uint64), acc)


def test_admission_waits_for_accumulator_memory():
    insts = small_batch(6, seed=3, budget=(5000, 8000))
    # room for one request's 102 bins at a time
    with S.Server(accum_capacity=150, tick_packets=4000, chunk=256) as srv:
        reqs = [srv.submit_packed_layered(*_packed(i)) for i in insts]
        assert srv.drain(120)
        assert srv.stats()['admitted'] == len(insts)
    for inst, r in zip(insts, reqs):
        np.testing.assert_array_equal(r.result(), _alone_batch(inst))
    lat = S.latency_summary(reqs)
    assert lat['n'] == len(insts)


def test_request_larger_than_the_server_is_rejected():
    insts = small_batch(2, seed=3)
    with S.Server(accum_capacity=50) as srv:
        # 102 bins can never fit in 50: an error at once, not a wait that
        # would hold back every later request
        r = srv.submit_packed_layered(*_packed(insts[0]))
        assert r.wait(30)
        with pytest.raises(ValueError, match='server holds'):
            r.result()
    with S.Server(accum_capacity=150) as srv:
        r = srv.submit_packed_layered(*_packed(insts[1]))
        np.testing.assert_array_equal(r.result(60), _alone_batch(insts[1]))


def test_cancel_releases_memory_and_others_complete():
    insts = small_batch(4, seed=9, budget=(3000, 5000))
    # far more packets than the test runs: still in flight when cancelled
    big = small_batch(1, seed=2, budget=(2_000_000_000, 2_000_000_001))[0]
    with S.Server(tick_packets=100000, chunk=512) as srv:
        rb = srv.submit_packed_layered(*_packed(big))
        rs = [srv.submit_packed_layered(*_packed(i)) for i in insts]
        for r in rs:
            assert r.wait(60)
        assert not rb.done()
        rb.cancel()
        assert rb.wait(60)
        with pytest.raises(RuntimeError, match='cancelled'):
            rb.result()
        assert srv.drain(60)
        assert srv.stats()['cancelled'] == 1
        st = srv.stats()
        assert st['accum_used'] == 0 and st['lut_used'] == 0
    for inst, r in zip(insts, rs):
        np.testing.assert_array_equal(r.result(), _alone_batch(inst))


def test_zero_packet_request_completes():
    inst = small_batch(1, seed=4)[0]
    inst.nphotons = 0
    with S.Server() as srv:
        r = srv.submit_packed_layered(*_packed(inst))
        assert r.wait(30)
        assert not r.result().any()


def test_graph_replays_equal_eager_ticks():
    insts = small_batch(10, seed=21, budget=(20000, 40000))
    out = {}
    for use_graph in (1, 0):
        with S.Server(use_graph=use_graph, tick_packets=30000,
                      chunk=256) as srv:
            reqs = [srv.submit_packed_layered(*_packed(i)) for i in insts]
            assert srv.drain(120)
            assert srv.stats()['ticks'] >= 8  # many replays of each graph
            out[use_graph] = [r.result() for r in reqs]
    for a, b in zip(out[1], out[0]):
        np.testing.assert_array_equal(a, b)


def test_high_level_submit_returns_detectors(voxel_sim):
    inst = small_batch(2, seed=13)[1]
    with S.Server() as srv:
        r1 = srv.submit_layered(inst)
        r2 = srv.submit_voxel(voxel_sim, 20000, seed=99)
        det = r1.result(60)
        _, flu, vdet = r2.result(60)
    assert det.top.nphotons == inst.nphotons
    assert flu is not None and vdet.top.nphotons == 20000
    _, flu_alone, _ = voxel_sim.run(20000, packet_seed=99)
    np.testing.assert_array_equal(flu.raw, flu_alone.raw)
