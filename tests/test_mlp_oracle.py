"""The CPU restatement of the surrogate network, pinned before it is used
as the checker: bf16 rounding against torch's independent conversion and
hand-picked ties, backpropagation against central finite differences."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M


def test_bf16_round_matches_torch_and_ties():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 1000),
        rng.normal(0, 1e30, 1000), [0.0, -0.0, 1.0, -2.5, np.inf, -np.inf]
    ]).astype(np.float32)
    ours = M.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(ours.view(np.uint32), ref.view(np.uint32))
    # ties go to the even mantissa: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    t = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20],
                 dtype=np.float32)
    np.testing.assert_array_equal(M.bf16_round(t),
                                  np.array([1.0, 1 + 2 ** -6, 1 + 2 ** -7],
                                           dtype=np.float32))
    assert np.isnan(M.bf16_round(np.array([np.nan], np.float32)))[0]


def test_gradient_matches_finite_differences():
    dims = (3, 7, 6, 4)
    rng = np.random.default_rng(1)
    p = M.init_params(dims, seed=2).astype(np.float64)
    p += rng.normal(0, 0.1, p.size)   # nonzero biases
    x = rng.normal(0, 1, (9, 3))
    t = rng.normal(0, 1, (9, 4))
    loss, g = M.loss_and_gradient(p, x, t, dims, emulate_bf16=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd[i] = (M.loss_and_gradient(pp, x, t, dims, False)[0] -
                 M.loss_and_gradient(pm, x, t, dims, False)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)
    # a dropped term would fail: the bias gradients are not all zero
    assert np.abs(g).max() > 1e-3


def test_unpack_rejects_wrong_count():
    with pytest.raises(ValueError):
        M.unpack(np.zeros(10), (3, 4))


def test_bf16_emulation_stays_close_to_float64():
    dims = (5, 32, 32, 10)
    rng = np.random.default_rng(3)
    p = M.init_params(dims, seed=4)
    x = rng.uniform(-1, 1, (64, 5))
    t = rng.normal(0, 1, (64, 10))
    l64, g64 = M.loss_and_gradient(p, x, t, dims, False)
    l16, g16 = M.loss_and_gradient(p, x, t, dims, True)
    assert abs(l16 - l64) / l64 < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2
