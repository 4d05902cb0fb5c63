"""Learned surrogate of the radial reflectance of a semi-infinite turbid
medium, with Monte Carlo as its fallback and as the measure of its error.

Inputs: (ln mua, ln mus', g, beta, n) of a medium with a modified
Henyey-Greenstein phase function (mus' = mus (1 - beta g), the first moment
of MHg being beta g, xopto/pf/mhg.py:103-122) under a normal pencil beam
from air. Output: ln R(r) over the 100 bins of 0-5 mm of a Radial detector
(the last bin also collects r > 5 mm, radial.py:117-150), R in 1/m^2.

The network is 5-256-256-256-100 with ReLU, trained and evaluated by the
library's bf16 MFMA kernels (xo_mlp_*), fp32 accumulation, Adam. Training
data come from the batched transport engine (one instance per sample, every
packet keyed on the sample's seed). A query outside the training box, or
whose Monte Carlo spot check disagrees by more than the tolerance, is
answered by Monte Carlo.

The reference keeps no such model: its inverse-problem helpers fit
regressors to simulated datasets (xopto/util/regression/ann.py, TensorFlow).
"""
import ctypes
import time
from ctypes import POINTER, c_double, c_float, c_uint32, c_uint64
from typing import Dict, Optional, Tuple

import numpy as np

from . import _native as N

DIMS = (5, 256, 256, 256, 100)
NBINS = 100
RMAX = 5e-3
DEFAULT_BOX = {'mua': (10.0, 1000.0),      # 1/m
               'musr': (300.0, 3000.0),    # 1/m
               'g': (0.6, 0.95),
               'beta': (0.0, 1.0),
               'n': (1.33, 1.50)}
XINIT = 0x5EED000000000001


def _fptr(a):
    return a.ctypes.data_as(POINTER(c_float))


class Network:
    """A fully connected ReLU network on the device (xo_mlp)."""

    def __init__(self, dims=DIMS, context: N.Context = None,
                 max_batch: int = 4096):
        self.dims = tuple(int(d) for d in dims)
        self._ctx = context if context is not None else N.Context(0, 1, XINIT)
        d = (c_uint32 * len(self.dims))(*self.dims)
        h = N.lib().xo_mlp_create(self._ctx._h, d, len(self.dims) - 1,
                                  c_uint32(max_batch))
        if not h:
            raise N.NativeError('xo_mlp_create: ' + N.last_error())
        self._h = h
        self.max_batch = int(max_batch)
        self.nparams = int(N.lib().xo_mlp_param_count(h))

    def close(self):
        if getattr(self, '_h', None):
            N.lib().xo_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        N.check(N.lib().xo_mlp_set_params(self._h, _fptr(p), c_uint64(p.size)),
                'xo_mlp_set_params')

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        N.check(N.lib().xo_mlp_get_params(self._h, _fptr(p), c_uint64(p.size)),
                'xo_mlp_get_params')
        return p

    def forward(self, x):
        """(outputs, kernel ms)."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        y = np.empty((x.shape[0], self.dims[-1]), dtype=np.float32)
        ms = c_double()
        N.check(N.lib().xo_mlp_forward(self._h, _fptr(x), c_uint64(x.shape[0]),
                                       _fptr(y), ctypes.byref(ms)),
                'xo_mlp_forward')
        return y, ms.value

    def gradient(self, x, t):
        """(mean squared error, gradient) of one batch."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.dims[-1])
        g = np.empty(self.nparams, dtype=np.float32)
        loss = c_float()
        N.check(N.lib().xo_mlp_gradient(self._h, _fptr(x), _fptr(t),
                                        c_uint32(x.shape[0]), _fptr(g),
                                        ctypes.byref(loss)),
                'xo_mlp_gradient')
        return float(loss.value), g

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.dims[-1])
        N.check(N.lib().xo_mlp_set_data(self._h, _fptr(x), _fptr(t),
                                        c_uint32(x.shape[0])),
                'xo_mlp_set_data')

    def train(self, steps: int, batch: int, lr: float, seed: int = 1,
              use_graph: bool = True):
        """(per-step minibatch losses, kernel ms)."""
        losses = np.empty(max(int(steps), 1), dtype=np.float32)
        ms = c_double()
        N.check(N.lib().xo_mlp_train(self._h, c_uint32(steps), c_uint32(batch),
                                     c_float(lr), c_uint64(seed),
                                     int(bool(use_graph)), _fptr(losses),
                                     ctypes.byref(ms)),
                'xo_mlp_train')
        return losses[:steps], ms.value


def init_params(dims=DIMS, seed: int = 0):
    """He-normal weights, zero biases, packed W_1, b_1, ..."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in zip(dims[:-1], dims[1:]):
        parts.append(rng.normal(0.0, np.sqrt(2.0 / i), size=i * o))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


def ring_areas():
    edges = np.linspace(0.0, RMAX, NBINS + 1)
    return np.pi * (edges[1:] ** 2 - edges[:-1] ** 2)


def profile_error(R, R_ref) -> np.ndarray:
    """Area-weighted mean absolute relative error of radial profiles:
    sum |R - R_ref| A / sum R_ref A (per row)."""
    A = ring_areas()
    R = np.atleast_2d(R)
    R_ref = np.atleast_2d(R_ref)
    return np.abs(R - R_ref) @ A / np.maximum(R_ref @ A, 1e-300)


def sample_box(n: int, box=DEFAULT_BOX, seed: int = 0) -> np.ndarray:
    """n raw feature rows (ln mua, ln mus', g, beta, n), uniform in the box
    (log-uniform in mua and mus')."""
    rng = np.random.default_rng(seed)
    lo = np.array([np.log(box['mua'][0]), np.log(box['musr'][0]),
                   box['g'][0], box['beta'][0], box['n'][0]])
    hi = np.array([np.log(box['mua'][1]), np.log(box['musr'][1]),
                   box['g'][1], box['beta'][1], box['n'][1]])
    return lo + (hi - lo) * rng.random((n, 5))


def _instance(row, packets: int, seed: int):
    from .mcml import mc, batch as mbatch
    lmua, lmusr, g, beta, n = (float(v) for v in row)
    mua, musr = np.exp(lmua), np.exp(lmusr)
    mus = musr / (1.0 - beta * g)
    pf = mc.mcpf.MHg(g, beta)
    L, D = mc.mclayer, mc.mcdetector
    layers = [L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=pf),
              L.Layer(d=np.inf, n=n, mua=mua, mus=mus, pf=pf),
              L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=pf)]
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, RMAX, NBINS)),
                      bottom=D.Total(), specular=D.Total())
    return mbatch.Instance(layers, mc.mcsource.Line(), det, nphotons=packets,
                           seed=seed)


def simulate(rows, packets: int, seed: int = 1,
             context: N.Context = None) -> Tuple[np.ndarray, float]:
    """Monte Carlo radial reflectance [len(rows)][100] of each feature row
    (one batch launch, packet generators keyed on (seed, row)); returns
    (R, seconds)."""
    from .mcml.batch import Batch
    rows = np.atleast_2d(rows)
    insts = [_instance(r, packets, (seed * 0x9E3779B97F4A7C15 + k + 1) &
                       0xFFFFFFFFFFFFFFFF) for k, r in enumerate(rows)]
    b = Batch(insts, rnginit=XINIT)
    if context is not None:
        b._ctx = context
    packed = b.pack()
    t0 = time.perf_counter()
    res, _ = b.run(b.opts(chunk=2048, block=256), packed=packed)
    dt = time.perf_counter() - t0
    return np.array([r.top.reflectance for r in res]), dt


def write_dataset(folder: str, rows, packets: int, seed: int = 1,
                  context: N.Context = None) -> list:
    """Simulate the feature rows (one batch launch) and write one file per
    row in the layered dataset layout (analytical_pf.template.py:226-246;
    ``dataset.mcml_record``), named by mua and mus' like the reference's
    files. Returns the file names."""
    import os
    from . import dataset
    from .mcml.batch import Batch
    rows = np.atleast_2d(rows)
    insts = [_instance(r, packets, (seed * 0x9E3779B97F4A7C15 + k + 1) &
                       0xFFFFFFFFFFFFFFFF) for k, r in enumerate(rows)]
    b = Batch(insts, rnginit=XINIT)
    if context is not None:
        b._ctx = context
    res, _ = b.run(b.opts(chunk=2048, block=256))
    names = []
    for row, inst, det in zip(rows, insts, res):
        mua, musr = np.exp(row[0]), np.exp(row[1])
        name = os.path.join(folder, dataset.mcml_filename(mua, musr))
        names.append(dataset.save(name, dataset.mcml_record(inst, det, 'aw')))
    return names


class ReflectanceSurrogate:
    """Network plus the feature and target scalings of one training box."""

    def __init__(self, box: Dict = None, dims=DIMS, seed: int = 0,
                 context: N.Context = None, max_batch: int = 4096):
        self.box = dict(box or DEFAULT_BOX)
        self.net = Network(dims, context, max_batch)
        self.net.set_params(init_params(dims, seed))
        self._lo = np.array([np.log(self.box['mua'][0]),
                             np.log(self.box['musr'][0]), self.box['g'][0],
                             self.box['beta'][0], self.box['n'][0]])
        self._hi = np.array([np.log(self.box['mua'][1]),
                             np.log(self.box['musr'][1]), self.box['g'][1],
                             self.box['beta'][1], self.box['n'][1]])
        self.y_mean = np.zeros(dims[-1])
        self.y_std = np.ones(dims[-1])
        self.floor = 1e-3

    def inside(self, rows) -> np.ndarray:
        rows = np.atleast_2d(rows)
        return np.all((rows >= self._lo - 1e-12) & (rows <= self._hi + 1e-12),
                      axis=1)

    def _x(self, rows):
        return (2.0 * (np.atleast_2d(rows) - self._lo) /
                (self._hi - self._lo) - 1.0).astype(np.float32)

    def _y(self, R):
        return (np.log(np.maximum(R, self.floor)) - self.y_mean) / self.y_std

    def fit(self, rows, R, steps: int = 3000, batch: int = 1024,
            lr: float = 1e-3, seed: int = 1):
        """Train on (rows, R); returns (losses, kernel ms)."""
        logR = np.log(np.maximum(R, self.floor))
        self.y_mean = logR.mean(axis=0)
        self.y_std = np.maximum(logR.std(axis=0), 1e-6)
        self.net.set_data(self._x(rows), self._y(R).astype(np.float32))
        return self.net.train(steps, min(batch, self.net.max_batch), lr, seed)

    def predict(self, rows):
        """(R [rows][100], kernel ms)."""
        y, ms = self.net.forward(self._x(rows))
        return np.exp(y * self.y_std + self.y_mean), ms

    def save(self, path):
        np.savez(path, params=self.net.get_params(), y_mean=self.y_mean,
                 y_std=self.y_std, lo=self._lo, hi=self._hi,
                 floor=np.float64(self.floor))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.net.set_params(z['params'])
            self.y_mean, self.y_std = z['y_mean'], z['y_std']
            self._lo, self._hi = z['lo'], z['hi']
            self.floor = float(z['floor'])


class Service:
    """Answers reflectance queries with the surrogate where it is trusted,
    with Monte Carlo otherwise: outside the training box always, and inside
    it whenever a spot check (a Monte Carlo run of mc_packets on the same
    query) disagrees by more than `tolerance` in profile_error."""

    def __init__(self, surrogate: ReflectanceSurrogate,
                 mc_packets: int = 1_000_000, tolerance: float = 0.02,
                 context: N.Context = None):
        self.surrogate = surrogate
        self.mc_packets = int(mc_packets)
        self.tolerance = float(tolerance)
        self._ctx = context
        self.counts = {'surrogate': 0, 'mc_outside': 0, 'mc_fallback': 0}
        self._seed = 1

    def query(self, rows, spot_check: bool = False):
        """[(R, source, spot-check error or None)] for each feature row."""
        rows = np.atleast_2d(rows)
        inside = self.surrogate.inside(rows)
        out = [None] * len(rows)
        idx_in = np.nonzero(inside)[0]
        idx_out = np.nonzero(~inside)[0]
        if len(idx_in):
            R_s, _ = self.surrogate.predict(rows[idx_in])
            if spot_check:
                R_mc, _ = simulate(rows[idx_in], self.mc_packets, self._seed,
                                   self._ctx)
                self._seed += 1
                err = profile_error(R_s, R_mc)
                for k, i in enumerate(idx_in):
                    if err[k] > self.tolerance:
                        out[i] = (R_mc[k], 'mc_fallback', float(err[k]))
                        self.counts['mc_fallback'] += 1
                    else:
                        out[i] = (R_s[k], 'surrogate', float(err[k]))
                        self.counts['surrogate'] += 1
            else:
                for k, i in enumerate(idx_in):
                    out[i] = (R_s[k], 'surrogate', None)
                    self.counts['surrogate'] += 1
        if len(idx_out):
            R_mc, _ = simulate(rows[idx_out], self.mc_packets, self._seed,
                               self._ctx)
            self._seed += 1
            for k, i in enumerate(idx_out):
                out[i] = (R_mc[k], 'mc_outside', None)
                self.counts['mc_outside'] += 1
        return out
