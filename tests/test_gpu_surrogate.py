"""Surrogate network on the GPU: the bf16 MFMA product against a float64
product of the same bf16 operands (every transpose case, ragged tiles), the
gradient against the CPU restatement, captured training steps against
eager ones, and what the trained network gains over Monte Carlo alone."""
import time

import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd import surrogate as S
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    return N.Context(0, 1, S.XINIT)


def _gemm(ctx, ta, tb, A, B, M_, N_, K):
    import ctypes
    C = np.zeros((M_, N_), dtype=np.float32)
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    lda = A.shape[1]
    ldb = B.shape[1]
    f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    N.check(N.lib().xo_gemm_bf16(ctx._h, int(ta), int(tb), M_, N_, K, f(A),
                                 lda, f(B), ldb, f(C), N_), 'xo_gemm_bf16')
    return C


@pytest.mark.parametrize('ta,tb', [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize('M_,N_,K', [(16, 16, 32), (64, 64, 64), (70, 50, 45),
                                     (5, 130, 7), (257, 100, 256)])
def test_bf16_product_layout(ctx, ta, tb, M_, N_, K):
    rng = np.random.default_rng(M_ * 131 + N_ * 7 + K + 11 * ta + 3 * tb)
    A = rng.normal(0, 1, (K, M_) if ta else (M_, K)).astype(np.float32)
    B = rng.normal(0, 1, (N_, K) if tb else (K, N_)).astype(np.float32)
    C = _gemm(ctx, ta, tb, A, B, M_, N_, K)
    a = M.bf16_round(A).astype(np.float64)
    b = M.bf16_round(B).astype(np.float64)
    a = a.T if ta else a
    b = b.T if tb else b
    ref = a @ b
    bound = np.abs(a) @ np.abs(b)
    # fp32 accumulation of exact bf16 products
    assert np.all(np.abs(C - ref) <= 1e-6 * K * bound + 1e-30)
    # a transposed operand would be far off
    assert np.abs(C - ref).max() < 1e-3 * np.abs(ref).max()


@pytest.mark.parametrize('dims,n', [(S.DIMS, 384), (S.DIMS, 1000),
                                    ((5, 256, 100), 384),
                                    ((5, 256, 256, 256, 256, 256, 128), 500),
                                    ((5, 64, 64, 10), 384)])
def test_gradient_matches_cpu_restatement(ctx, dims, n):
    """S.DIMS, (5, 256, 100) and the six-layer network run the fused
    training kernels (the fewest and the most layers they take; 1000 and
    500 rows: a ragged last row block), (5, 64, 64, 10) the layer-by-layer
    products."""
    net = S.Network(dims, ctx, max_batch=1024)
    rng = np.random.default_rng(5)
    p = M.init_params(dims, seed=6)
    p[-dims[-1]:] = rng.normal(0, 0.1, dims[-1])
    net.set_params(p)
    x = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    t = rng.normal(0, 1, (n, dims[-1])).astype(np.float32)
    loss, g = net.gradient(x, t)
    lref, gref = M.loss_and_gradient(p, x, t, dims, emulate_bf16=True)
    assert abs(loss - lref) / lref < 1e-4
    layers = M.unpack(np.arange(p.size), dims)
    for w_idx, b_idx in layers:
        for idx in (w_idx.ravel(), b_idx.ravel()):
            gg, gr = g[idx], gref[idx]
            rel = np.linalg.norm(gg - gr) / max(np.linalg.norm(gr), 1e-30)
            assert rel < 1e-2, rel
            assert np.dot(gg, gr) / (np.linalg.norm(gg) * np.linalg.norm(gr)) \
                > 0.9999
    y, _ = net.forward(x)
    yref, _ = M.forward(p, x, dims, emulate_bf16=True)
    # fp32 against float64 accumulation can flip the bf16 rounding of a few
    # activations; each flip moves an output by about one bf16 step
    assert np.linalg.norm(y - yref) / np.linalg.norm(yref) < 1e-3
    assert np.abs(y - yref).max() < 5e-3 * max(1.0, np.abs(yref).max())
    net.close()


def test_ragged_batch_after_a_full_one(ctx):
    """A batch of 1000 rows after one of 1024 on the same network: the
    rows past the batch in the last 32-row k step must carry zero
    gradients, not what the larger batch left there."""
    dims = S.DIMS
    net = S.Network(dims, ctx, max_batch=1024)
    rng = np.random.default_rng(31)
    p = M.init_params(dims, seed=32)
    net.set_params(p)
    x = rng.uniform(-1, 1, (1024, 5)).astype(np.float32)
    t = rng.normal(0, 1, (1024, 100)).astype(np.float32)
    net.gradient(x, t)
    loss, g = net.gradient(x[:1000], t[:1000])
    lref, gref = M.loss_and_gradient(p, x[:1000], t[:1000], dims, emulate_bf16=True)
    assert abs(loss - lref) / lref < 1e-4
    for w_idx, b_idx in M.unpack(np.arange(p.size), dims):
        for idx in (w_idx.ravel(), b_idx.ravel()):
            rel = np.linalg.norm(g[idx] - gref[idx]) / np.linalg.norm(gref[idx])
            assert rel < 1e-2, rel
    net.close()


@pytest.mark.parametrize('rows', [1, 63, 64, 1000, 4096])
def test_fused_inference_equals_layer_by_layer(ctx, rows):
    """forward() runs the fused inference kernel, gradient() the fused
    training kernels (row blocks of 32 instead of 64, every activation kept
    for the backward pass). Same products in the same order: the network
    outputs agree to the bit, which the gradient's loss against the
    inference outputs shows as exactly zero."""
    dims = S.DIMS
    net = S.Network(dims, ctx, max_batch=4096)
    rng = np.random.default_rng(rows)
    p = M.init_params(dims, seed=rows + 1)
    p[-100:] = rng.normal(0, 0.1, 100)
    net.set_params(p)
    x = rng.uniform(-1, 1, (rows, 5)).astype(np.float32)
    y, _ = net.forward(x)
    loss, _ = net.gradient(x, y)
    assert loss == 0.0
    yref, _ = M.forward(p, x, dims, emulate_bf16=True)
    assert np.linalg.norm(y - yref) / np.linalg.norm(yref) < 1e-3
    net.close()


@pytest.mark.parametrize('dims', [(5, 64, 64, 10), S.DIMS, (5, 256, 100),
                                  (5, 256, 256, 256, 256, 256, 128)])
def test_captured_steps_equal_eager_steps(ctx, dims):
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (1000, 5)).astype(np.float32)
    t = np.sin(x @ rng.normal(0, 1, (5, dims[-1]))).astype(np.float32)
    p0 = M.init_params(dims, seed=8)
    out = []
    for use_graph in (False, True, True):
        net = S.Network(dims, ctx, max_batch=256)
        net.set_params(p0)
        net.set_data(x, t)
        # 70 steps: one 64-step graph and a 6-step graph; then 5 steps (a
        # 5-step graph captured in place of the 6-step one) and 70 again
        l1, _ = net.train(70, 128, 1e-3, seed=3, use_graph=use_graph)
        l2, _ = net.train(5, 128, 1e-3, seed=3, use_graph=use_graph)
        l3, _ = net.train(70, 128, 1e-3, seed=3, use_graph=use_graph)
        out.append((net.get_params(), np.concatenate([l1, l2, l3])))
        net.close()
    for params, losses in out[1:]:
        np.testing.assert_array_equal(params, out[0][0])
        np.testing.assert_allclose(losses, out[0][1], rtol=1e-5)


def _sample_rows(seed, step, batch, n):
    """The minibatch rows of a training step: row i of step s is
    mix64(seed + golden * ((s << 32) + i + 1)) mod n (xo_mlp_train)."""
    m = (1 << 64) - 1
    rows = []
    for i in range(batch):
        z = (seed + 0x9E3779B97F4A7C15 * ((step << 32) + i + 1)) & m
        z ^= z >> 30
        z = (z * 0xBF58476D1CE4E5B9) & m
        z ^= z >> 27
        z = (z * 0x94D049BB133111EB) & m
        z ^= z >> 31
        rows.append(z % n)
    return np.array(rows)


def test_fused_adam_step_and_weight_copies(ctx):
    """One fused step from fresh moments moves each parameter by
    -lr g / (|g| + eps) with g the gradient of the step's minibatch (the
    CPU restatement's, on the rows the step draws); after more steps the
    outputs (fragment copies of W_l) and the gradient (copies of W_l and
    W_l^T), both rewritten by the step's Adam epilogue, equal the CPU
    restatement's on the fp32 weights read back."""
    dims = S.DIMS
    rng = np.random.default_rng(21)
    x = rng.uniform(-1, 1, (3000, 5)).astype(np.float32)
    t = np.tanh(x @ rng.normal(0, 1, (5, 100))).astype(np.float32)
    p0 = M.init_params(dims, seed=22)
    net = S.Network(dims, ctx, max_batch=1024)
    net.set_params(p0)
    net.set_data(x, t)
    lr, seed, B = 1e-3, 5, 1000
    net.train(1, B, lr, seed=seed)
    p1 = net.get_params()
    rows = _sample_rows(seed, 0, B, len(x))
    _, g = M.loss_and_gradient(p0, x[rows], t[rows], dims, emulate_bf16=True)
    g = g.astype(np.float32)
    step_ref = -lr * g / (np.abs(g) + 1e-8)
    d = (p1.astype(np.float64) - p0) - step_ref
    # the bf16 products can flip the sign of a gradient near zero
    assert np.mean(np.abs(d) > 0.05 * lr) < 0.03
    assert np.all(np.abs(p1 - p0) <= lr * (1 + 1e-3))
    net.train(4, B, lr, seed=seed + 1)
    p = net.get_params()
    xq = rng.uniform(-1, 1, (700, 5)).astype(np.float32)
    tq = rng.normal(0, 1, (700, 100)).astype(np.float32)
    y, _ = net.forward(xq)
    yref, _ = M.forward(p, xq, dims, emulate_bf16=True)
    assert np.linalg.norm(y - yref) / np.linalg.norm(yref) < 1e-3
    loss, gq = net.gradient(xq, tq)
    lref, gqref = M.loss_and_gradient(p, xq, tq, dims, emulate_bf16=True)
    assert abs(loss - lref) / lref < 1e-4
    for w_idx, b_idx in M.unpack(np.arange(p.size), dims):
        for idx in (w_idx.ravel(), b_idx.ravel()):
            rel = np.linalg.norm(gq[idx] - gqref[idx]) / np.linalg.norm(gqref[idx])
            assert rel < 1e-2, rel
    net.close()


def test_step_counter_across_calls_gradients_and_resets(ctx):
    """The device step counter (it picks each step's minibatch rows and
    Adam's bias correction) advances once per step across train calls, is
    left alone by a gradient evaluation between them, and starts again at
    zero after set_params: 3 + 4 steps with a gradient in between, and 7
    steps after a reset, give the losses and weights of 7 steps in one
    call, bit for bit."""
    dims = S.DIMS
    rng = np.random.default_rng(31)
    x = rng.uniform(-1, 1, (2000, 5)).astype(np.float32)
    t = np.tanh(x @ rng.normal(0, 1, (5, 100))).astype(np.float32)
    xq = rng.uniform(-1, 1, (300, 5)).astype(np.float32)
    tq = rng.normal(0, 1, (300, 100)).astype(np.float32)
    p0 = M.init_params(dims, seed=32)
    lr, seed, B = 1e-3, 9, 512
    ref = S.Network(dims, ctx, max_batch=512)
    ref.set_params(p0)
    ref.set_data(x, t)
    l_ref, _ = ref.train(7, B, lr, seed=seed)
    p_ref = ref.get_params()
    ref.close()
    net = S.Network(dims, ctx, max_batch=512)
    net.set_params(p0)
    net.set_data(x, t)
    l1, _ = net.train(3, B, lr, seed=seed)
    net.gradient(xq, tq)
    l2, _ = net.train(4, B, lr, seed=seed)
    np.testing.assert_array_equal(np.concatenate([l1, l2]), l_ref)
    np.testing.assert_array_equal(net.get_params(), p_ref)
    net.set_params(p0)
    l3, _ = net.train(7, B, lr, seed=seed)
    np.testing.assert_array_equal(l3, l_ref)
    np.testing.assert_array_equal(net.get_params(), p_ref)
    net.close()


def test_training_fits_a_smooth_function(ctx):
    dims = S.DIMS
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (4096, 5)).astype(np.float32)
    t = np.tanh(x @ rng.normal(0, 1, (5, 100))).astype(np.float32)
    net = S.Network(dims, ctx, max_batch=1024)
    net.set_params(M.init_params(dims, seed=10))
    net.set_data(x, t)
    losses, ms = net.train(600, 512, 2e-3, seed=4)
    assert losses[-20:].mean() < 0.05 * losses[:5].mean()
    assert ms > 0
    net.close()


def test_surrogate_gains_over_monte_carlo(ctx):
    """Trained on cheap noisy simulations, the network answers a query in a
    small fraction of the Monte Carlo time and is closer to a converged
    Monte Carlo profile than a Monte Carlo run of the training budget.
    Outside the training box the service answers with Monte Carlo."""
    box = {'mua': (50.0, 200.0), 'musr': (800.0, 1500.0), 'g': (0.8, 0.9),
           'beta': (0.5, 1.0), 'n': (1.35, 1.45)}
    train_rows = S.sample_box(768, box, seed=1)
    R_train, t_gen = S.simulate(train_rows, 20000, seed=11, context=ctx)
    sur = S.ReflectanceSurrogate(box, context=ctx, seed=3, max_batch=1024)
    losses, _ = sur.fit(train_rows, R_train, steps=1500, batch=256, lr=1e-3)
    assert losses[-50:].mean() < 0.2 * losses[:10].mean()

    test_rows = S.sample_box(16, box, seed=2)
    R_ref, _ = S.simulate(test_rows, 400000, seed=12, context=ctx)
    R_cheap, t_cheap = S.simulate(test_rows, 20000, seed=13, context=ctx)
    t0 = time.perf_counter()
    R_sur, _ = sur.predict(test_rows)
    t_sur = time.perf_counter() - t0
    err_sur = np.median(S.profile_error(R_sur, R_ref))
    err_cheap = np.median(S.profile_error(R_cheap, R_ref))
    print('surrogate error {:.4f} (MC at the training budget {:.4f}); '
          'time {:.2e} s vs {:.2e} s'.format(err_sur, err_cheap, t_sur,
                                             t_cheap))
    assert err_sur < err_cheap
    assert t_sur < t_cheap

    svc = S.Service(sur, mc_packets=20000, tolerance=1.0, context=ctx)
    outside = test_rows[:2].copy()
    outside[:, 0] = np.log(5000.0)           # mua beyond the box
    res = svc.query(np.vstack([test_rows[2:4], outside]))
    assert [r[1] for r in res] == ['surrogate', 'surrogate', 'mc_outside',
                                   'mc_outside']
    strict = S.Service(sur, mc_packets=20000, tolerance=0.0, context=ctx)
    res = strict.query(test_rows[:2], spot_check=True)
    assert all(r[1] == 'mc_fallback' for r in res)


def test_context_close_releases_its_networks():
    """A network holds its context's device: closing the context closes the
    network first, and later runs on a new context are unaffected."""
    c = N.Context(0, 1, S.XINIT)
    net = S.Network((5, 16, 4), c, max_batch=64)
    c.close()
    assert net._h is None
    c2 = N.Context(0, 1, S.XINIT)
    net2 = S.Network((5, 16, 4), c2, max_batch=64)
    net2.set_params(M.init_params((5, 16, 4), seed=1))
    y, _ = net2.forward(np.zeros((3, 5), np.float32))
    assert np.all(np.isfinite(y))
    c2.close()
