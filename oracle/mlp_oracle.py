"""CPU restatement of the reflectance surrogate's network (test
infrastructure only: tests/ use it as the checker of the bf16 MFMA kernels
in pyxopto_amd/csrc/xo_mlp.hip).

The reference has no such network on its path (its regression helpers in
xopto/util/regression/ann.py call TensorFlow, which is absent here); the
arithmetic restated is the textbook fully connected ReLU network and its
backpropagation. ``emulate_bf16=True`` rounds the operands of every product
to bf16 (round to nearest even) exactly where the device does and
accumulates in float64, so the device and this restatement differ only by
fp32 accumulation order; ``emulate_bf16=False`` is plain float64 and is
pinned against central finite differences (tests/test_mlp_oracle.py).
"""
import numpy as np


def bf16_round(a):
    """float32 values rounded to the nearest bf16 (ties to even), as
    float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)
    out = np.where(nan, np.float32(np.nan), out)
    return out


def unpack(params, dims):
    """[(W_l [in][out], b_l [out]), ...] from the packed parameter vector."""
    layers, off = [], 0
    for i, o in zip(dims[:-1], dims[1:]):
        w = np.asarray(params[off:off + i * o]).reshape(i, o)
        off += i * o
        b = np.asarray(params[off:off + o])
        off += o
        layers.append((w, b))
    if off != len(params):
        raise ValueError('parameter count does not match the widths')
    return layers


def forward(params, x, dims, emulate_bf16=True):
    """Network output (float64) and the hidden activations."""
    rnd = bf16_round if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    layers = unpack(params, dims)
    h = rnd(x).astype(np.float64)
    acts = [h]
    for k, (w, b) in enumerate(layers):
        z = h @ rnd(w).astype(np.float64) + np.asarray(b, np.float64)
        if k + 1 < len(layers):
            h = np.maximum(z, 0.0)
            if emulate_bf16:
                h = bf16_round(h.astype(np.float32)).astype(np.float64)
            acts.append(h)
        else:
            return z, acts
    raise ValueError('no layers')


def loss_and_gradient(params, x, t, dims, emulate_bf16=True):
    """Mean squared error over every output and its gradient, packed like
    the parameters (dL/dW_l = H_{l-1}^T dZ_l, dL/db_l = column sums of
    dZ_l, dZ_{l-1} = (dZ_l W_l^T) * (H_{l-1} > 0))."""
    y, acts = forward(params, x, dims, emulate_bf16)
    layers = unpack(params, dims)
    t = np.asarray(t, np.float64)
    count = y.size
    d = y - t
    loss = float(np.sum(d * d) / count)
    dz = 2.0 * d / count
    if emulate_bf16:
        dz = bf16_round(dz.astype(np.float32)).astype(np.float64)
    grads = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        w, _ = layers[k]
        h_in = acts[k]
        gw = h_in.T @ dz
        gb = dz.sum(axis=0)
        grads[k] = (gw, gb)
        if k == 0:
            break
        wk = bf16_round(w).astype(np.float64) if emulate_bf16 else \
            np.asarray(w, np.float64)
        dh = (dz @ wk.T) * (acts[k] > 0)
        if emulate_bf16:
            dh = bf16_round(dh.astype(np.float32)).astype(np.float64)
        dz = dh
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat


def init_params(dims, seed=0):
    """He-normal weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in zip(dims[:-1], dims[1:]):
        parts.append(rng.normal(0.0, np.sqrt(2.0 / i), size=i * o))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)
