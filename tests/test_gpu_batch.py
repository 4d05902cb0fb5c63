"""Batched engine on the GPU: bit-exact with the oracle whatever the
schedule, the same result alone and in the mix, and at every launch
shape."""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcml import mc
from pyxopto_amd.mcml.batch import Batch
from oracle import pyoracle as O

from batch_helpers import small_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mixed():
    insts = small_batch(40)
    b = Batch(insts, rnginit=0x5EED000000000001)
    return b, b.pack()


def test_batch_bit_exact_with_oracle(mixed):
    b, (arr, lut, naccum) = mixed
    acc, rep = b.context.run_batch(arr, lut, b.opts(chunk=1024), naccum)
    ref, counts = O.batch_run(arr, lut, naccum, math_mode=O.MATH_KERNEL)
    assert rep.packets == sum(i.nphotons for i in b.instances)
    assert int(rep.steps) == int(counts[0])
    assert int(rep.boundary_steps) == int(counts[1])
    np.testing.assert_array_equal(acc, ref)


@pytest.mark.parametrize('method', [1, 2])
def test_batch_methods_bit_exact_with_oracle(method):
    """Albedo rejection (1) and microscopic Beer-Lambert (2) in the batch
    kernel against the oracle's layered packet loop with the same method."""
    from pyxopto_amd.mcml import mc
    print('method', method, flush=True)
    insts = small_batch(24)
    b = Batch(insts, options=[mc.mcoptions.McMethod(method)],
              rnginit=0x5EED000000000001)
    arr, lut, naccum = b.pack()
    opts = b.opts(chunk=512)
    assert opts.method == method
    acc, rep = b.context.run_batch(arr, lut, opts, naccum)
    ref, counts = O.batch_run(arr, lut, naccum, math_mode=O.MATH_KERNEL,
                              method=method)
    assert int(rep.steps) == int(counts[0])
    assert int(rep.boundary_steps) == int(counts[1])
    np.testing.assert_array_equal(acc, ref)
    # the method changes the result
    acc0, _ = b.context.run_batch(arr, lut, b.opts(chunk=512).__class__(
        **{f: getattr(opts, f) for f, _ in opts._fields_ if f != 'method'}),
        naccum)
    assert not np.array_equal(acc0, acc)


@pytest.mark.parametrize('chunk,block,bpc', [(64, 64, 1), (333, 128, 0),
                                             (8192, 256, 2)])
def test_launch_shape_does_not_change_results(mixed, chunk, block, bpc):
    b, (arr, lut, naccum) = mixed
    a1, _ = b.context.run_batch(arr, lut, b.opts(1024, 256, 0), naccum)
    a2, _ = b.context.run_batch(arr, lut, b.opts(chunk, block, bpc), naccum)
    np.testing.assert_array_equal(a1, a2)


def test_instance_alone_equals_instance_in_mix(mixed):
    b, (arr, lut, naccum) = mixed
    mix, _ = b.context.run_batch(arr, lut, b.opts(), naccum)
    for i in (0, 7, 21, 39):
        alone = Batch([b.instances[i]], rnginit=0x5EED000000000001)
        res, _ = alone.run(alone.opts(chunk=256))
        base = arr[i].accum_base
        got = np.concatenate([res[0].top.raw, res[0].bottom.raw,
                              res[0].specular.raw]) * N.XO_ACCU_K
        np.testing.assert_array_equal(np.rint(got).astype(np.uint64),
                                      mix[base:base + 102])


def test_batch_api_autotuned_run(mixed):
    b, _ = mixed
    res, rep = b.run()
    assert len(res) == len(b.instances)
    assert res[0].top.nphotons == b.instances[0].nphotons
    assert rep.truncated == 0


def test_batch_rejects_unknown_method_and_backward_fiber(mixed):
    b, (arr, lut, naccum) = mixed
    opts = b.opts(chunk=1024)
    opts.method = 7
    with pytest.raises(RuntimeError, match='stepping method'):
        b.context.run_batch(arr, lut, opts, naccum)
    one = (N.XoBatchInstance * 1)()
    one[0] = arr[0]
    one[0].source_kind = N.XO_SRC_FIBER
    one[0].fiber_axis.direction[2] = -1.0
    with pytest.raises(RuntimeError, match='fiber direction'):
        b.context.run_batch(one, lut, b.opts(chunk=1024), naccum)


@pytest.mark.parametrize('fast', [False, True])
def test_batch_plan_replays_equal_batch_run(fast):
    """A captured batch launch (xo_batch_plan_*): every replay returns the
    integers of xo_batch_run, with eager batch runs on the same stream
    between the replays."""
    b = Batch(small_batch(40, seed=11), rnginit=0x5EED000000000001,
              cl_build_options=['-cl-fast-relaxed-math'] if fast else None)
    arr, lut, naccum = b.pack()
    opts = b.opts(chunk=1024)
    ctx = b.context
    ref, rep_ref = ctx.run_batch(arr, lut, opts, naccum)
    plan = ctx.batch_plan(arr, lut, opts, naccum)
    part = (type(arr[0]) * 5)(*arr[:5])
    for k in range(3):
        print('case plan replay', k, 'fast', fast, flush=True)
        acc, rep = plan.run()
        np.testing.assert_array_equal(acc, ref)
        assert rep.steps == rep_ref.steps and rep.packets == rep_ref.packets
        other, _ = ctx.run_batch(part, lut, b.opts(chunk=512), naccum)
        assert other.sum() > 0
    res, rep = b.plan(opts).run()
    assert rep.steps == rep_ref.steps
    ctx.close()        # closes the plans first


def _with_fluence(insts):
    """Give every third instance a Cartesian grid, every third an r-z grid
    (in fluence-rate mode), leave the rest without one."""
    A = mc.mcfluence.Axis
    for k, inst in enumerate(insts):
        if k % 3 == 0:
            inst.fluence = mc.mcfluence.Fluence(
                A(-1e-3, 1e-3, 8), A(-1e-3, 1e-3, 8), A(0.0, 2e-3, 10))
        elif k % 3 == 1:
            inst.fluence = mc.mcfluence.FluenceRz(A(0.0, 2e-3, 12),
                                                  A(0.0, 2e-3, 10),
                                                  mode='fluence')
    return insts


@pytest.mark.parametrize('method', [0, 1, 2])
def test_batch_fluence_bit_exact_with_oracle(method):
    """Per-instance fluence grids in the batch engine (the layered kernel's
    deposits, fluence.py:102-147 / fluencerz.py:104-160, for each stepping
    method): bit-exact with the oracle under the production schedule, and
    each instance's integers are those it gets alone."""
    print('case batch fluence method', method, flush=True)
    insts = _with_fluence(small_batch(24, seed=21))
    b = Batch(insts, options=[mc.mcoptions.McMethod(method)],
              rnginit=0x5EED000000000001)
    arr, lut, naccum = b.pack()
    assert sum(int(arr[i].fluence.enabled != 0) for i in range(len(arr))) == 16
    acc, rep = b.context.run_batch(arr, lut, b.opts(chunk=1024), naccum)
    ref, counts = O.batch_run(arr, lut, naccum, method=method,
                              math_mode=O.MATH_KERNEL)
    assert int(rep.steps) == int(counts[0])
    assert int(rep.deposits) == int(counts[3]) > 0
    np.testing.assert_array_equal(acc, ref)
    res = b.fluence_results(acc)
    assert res[2] is None and res[0].raw.sum() > 0 and res[1].raw.sum() > 0
    # instance 4 alone
    one = Batch([insts[4]], options=[mc.mcoptions.McMethod(method)],
                rnginit=0x5EED000000000001)
    a1, l1, n1 = one.pack()
    alone, _ = one.context.run_batch(a1, l1, one.opts(chunk=256), n1)
    base, cnt = int(arr[4].accum_base), int(arr[4].accum_count)
    np.testing.assert_array_equal(alone[:cnt], acc[base:base + cnt])
    one.context.close()
    b.context.close()
