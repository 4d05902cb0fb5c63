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
import json
import os
import sys
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
# The box of the shipped network: mua 0.01-50 /cm, g 0.6-0.95, beta 0-1 and
# n 1.33-1.50 as the 4096-instance layered batch draws its layers; mus'
# 3-60 /cm, which holds the reference's validation grid (mus' 5-60 /cm,
# xopto/mcml/test/validate.py:348-366).
TISSUE_BOX = {'mua': (1.0, 5000.0),
              'musr': (300.0, 6000.0),
              'g': (0.6, 0.95),
              'beta': (0.0, 1.0),
              'n': (1.33, 1.50)}
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                           'weights')
TISSUE_WEIGHTS = os.path.join(WEIGHTS_DIR, 'reflectance_semiinfinite_mhg.npz')
TISSUE_RECORD = os.path.join(WEIGHTS_DIR, 'reflectance_semiinfinite_mhg.json')
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
        self._ctx._dependents.add(self)
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


def _row_seed(seed: int, k: int) -> int:
    return (seed * 0x9E3779B97F4A7C15 + k + 1) & 0xFFFFFFFFFFFFFFFF


def simulate(rows, packets: int, seed: int = 1,
             context: N.Context = None, fast_math: bool = True,
             first_row: int = 0) -> Tuple[np.ndarray, float]:
    """Monte Carlo radial reflectance [len(rows)][100] of each feature row
    (one batch launch, packet generators keyed on (seed, first_row + k) for
    row k); returns (R, seconds of the run call). fast_math: the
    production arithmetic (the reference's -cl-fast-relaxed-math build)."""
    R, stats = _simulate(rows, packets, seed, context, fast_math, first_row)
    return R, stats['run_s']


def _simulate(rows, packets, seed, context, fast_math, first_row=0):
    from .mcml.batch import Batch
    rows = np.atleast_2d(rows)
    insts = [_instance(r, packets, _row_seed(seed, first_row + k))
             for k, r in enumerate(rows)]
    b = Batch(insts, rnginit=XINIT,
              cl_build_options=['-cl-fast-relaxed-math'] if fast_math
              else None)
    if context is not None:
        b._ctx = context
    packed = b.pack()
    t0 = time.perf_counter()
    res, rep = b.run(b.opts(chunk=8192 if packets >= 65536 else 2048,
                            block=256), packed=packed)
    run_s = time.perf_counter() - t0
    R = np.array([r.top.reflectance for r in res])
    return R, {'run_s': run_s, 'kernel_s': rep.kernel_ms * 1e-3,
               'steps': int(rep.steps), 'packets': int(rep.packets)}


def write_dataset(folder: str, rows, packets: int, seed: int = 1,
                  context: N.Context = None, fast_math: bool = True) -> list:
    """Simulate the feature rows (one batch launch) and write one file per
    row in the layered dataset layout (analytical_pf.template.py:226-246;
    ``dataset.mcml_record``), named by mua and mus' like the reference's
    files. Returns the file names."""
    import os
    from . import dataset
    from .mcml.batch import Batch
    rows = np.atleast_2d(rows)
    insts = [_instance(r, packets, _row_seed(seed, k))
             for k, r in enumerate(rows)]
    b = Batch(insts, rnginit=XINIT,
              cl_build_options=['-cl-fast-relaxed-math'] if fast_math
              else None)
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

    def fit_schedule(self, rows, R, schedule, batch: int = 4096,
                     seed: int = 1, log=None, power_weighted: bool = True):
        """Train on (rows, R) through phases of (steps, learning rate) with
        Adam state carried across them; returns (losses, kernel ms).

        power_weighted: each bin's squared log error counts in proportion
        to the mean share of the reflected power that bin holds (R A / sum
        R A over the training rows, floored at 2 % of the largest share),
        the weight with which a bin's relative error enters profile_error:
        the targets of bin j are scaled by sqrt(w_j) through y_std, so the
        network and its mean squared error are unchanged."""
        logR = np.log(np.maximum(R, self.floor))
        self.y_mean = logR.mean(axis=0)
        self.y_std = np.maximum(logR.std(axis=0), 1e-6)
        if power_weighted:
            A = ring_areas()
            share = (R * A) / np.maximum((R * A).sum(axis=1, keepdims=True),
                                         1e-300)
            wgt = share.mean(axis=0)
            wgt = np.maximum(wgt, 0.02 * wgt.max())
            wgt = wgt / wgt.mean()
            self.y_std = self.y_std / np.sqrt(wgt)
        self.net.set_data(self._x(rows), self._y(R).astype(np.float32))
        losses, ms = [], 0.0
        for k, (steps, lr) in enumerate(schedule):
            part, t = self.net.train(int(steps), min(batch, self.net.max_batch),
                                     float(lr), seed + k)
            losses.append(part)
            ms += t
            if log:
                log('train: phase {} ({} steps, lr {:g}): loss {:.3e} -> '
                    '{:.3e}, {:.1f} s'.format(k, steps, lr, part[:20].mean(),
                                              part[-100:].mean(), t * 1e-3))
        return np.concatenate(losses), ms

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
        self.box = {'mua': (float(np.exp(self._lo[0])), float(np.exp(self._hi[0]))),
                    'musr': (float(np.exp(self._lo[1])), float(np.exp(self._hi[1]))),
                    'g': (float(self._lo[2]), float(self._hi[2])),
                    'beta': (float(self._lo[3]), float(self._hi[3])),
                    'n': (float(self._lo[4]), float(self._hi[4]))}


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


def generate_training_set(n: int, packets: int, box=TISSUE_BOX,
                          seed: int = 20261203, shard: int = 2048,
                          context: N.Context = None, log=None):
    """n feature rows sampled in the box and their Monte Carlo radial
    reflectance at `packets` packets each, simulated by the batch engine
    `shard` rows per launch (row k's packets are keyed on (seed, k), so a
    row's result does not depend on the shard it runs in). Returns
    (rows, R float32 [n][100], totals)."""
    rows = sample_box(n, box, seed=seed)
    R = np.empty((n, NBINS), dtype=np.float32)
    tot = {'packets': 0, 'steps': 0, 'kernel_s': 0.0}
    t0 = time.perf_counter()
    for first in range(0, n, shard):
        part = rows[first:first + shard]
        Rs, st = _simulate(part, packets, seed, context, True, first)
        R[first:first + len(part)] = Rs
        for k in tot:
            tot[k] += st[k]
        if log:
            log('generate: rows {}-{} of {}: {:.3g} packets/s, {:.3g} steps/s, '
                '{:.0f} s so far'.format(first, first + len(part), n,
                                         st['packets'] / st['kernel_s'],
                                         st['steps'] / st['kernel_s'],
                                         time.perf_counter() - t0))
    tot['wall_s'] = time.perf_counter() - t0
    return rows, R, tot


DEFAULT_SCHEDULE = ((30000, 1e-3), (30000, 3e-4), (20000, 1e-4),
                    (10000, 3e-5))


def build_tissue_surrogate(samples: int = 200000, packets: int = 1000000,
                           out_dir: str = WEIGHTS_DIR, seed: int = 20261203,
                           schedule=DEFAULT_SCHEDULE, batch: int = 4096,
                           validation: int = 2048, log=print):
    """Generate the training set over TISSUE_BOX, train the network, and
    write its weights (TISSUE_WEIGHTS layout) and a JSON record of the
    generation, the training and the validation errors into out_dir. The
    last `validation` rows are held out; their Monte Carlo noise floor is
    measured with a second, independent simulation of the same rows."""
    ctx = N.Context(0, 0, XINIT)
    rows, R, gen = generate_training_set(samples, packets, TISSUE_BOX, seed,
                                         context=ctx, log=log)
    ntr = samples - validation
    sur = ReflectanceSurrogate(TISSUE_BOX, context=ctx, seed=seed & 0xFFFF,
                               max_batch=batch)
    losses, train_ms = sur.fit_schedule(rows[:ntr], R[:ntr], schedule,
                                        batch=batch, seed=seed, log=log)
    vrows, vR = rows[ntr:], R[ntr:]
    R_sur, _ = sur.predict(vrows)
    R_mc2, _ = simulate(vrows, packets, seed + 1, ctx)
    err_sur = profile_error(R_sur, vR)
    err_mc = profile_error(R_mc2, vR)

    def q(e):
        return {'median': float(np.median(e)), 'mean': float(e.mean()),
                'p90': float(np.quantile(e, 0.9)),
                'p99': float(np.quantile(e, 0.99)), 'max': float(e.max()),
                'frac_above_0.02': float(np.mean(e > 0.02))}

    os.makedirs(out_dir, exist_ok=True)
    wpath = os.path.join(out_dir, os.path.basename(TISSUE_WEIGHTS))
    sur.save(wpath)
    record = {
        'network': '5-256-256-256-100 relu, bf16 MFMA products, fp32 '
                   'accumulation, Adam on fp32 master weights',
        'inputs': 'ln mua [1/m], ln musr [1/m], g, beta, n (scaled to '
                  '[-1, 1] over the box)',
        'outputs': 'ln R over 100 radial bins 0-5 mm [1/m^2], standardized '
                   'per bin',
        'medium': 'semi-infinite, modified Henyey-Greenstein phase '
                  'function, air above, normal pencil beam (Line source)',
        'box': TISSUE_BOX,
        'samples': samples, 'packets_per_sample': packets, 'seed': seed,
        'engine': 'batch kernel, hardware arithmetic '
                  '(-cl-fast-relaxed-math), packets keyed on (seed, row)',
        'generation': gen,
        'training': {'rows': ntr, 'batch': batch,
                     'schedule': [list(p) for p in schedule],
                     'steps': int(len(losses)), 'kernel_s': train_ms * 1e-3,
                     'loss_first': float(losses[:100].mean()),
                     'loss_last': float(losses[-500:].mean())},
        'validation': {'rows': int(validation),
                       'error': 'area-weighted mean absolute relative error '
                                'of the profile against the Monte Carlo run '
                                'of the same row',
                       'surrogate_vs_mc': q(err_sur),
                       'mc_vs_mc_same_budget': q(err_mc)},
    }
    with open(os.path.join(out_dir, os.path.basename(TISSUE_RECORD)), 'w') as f:
        json.dump(record, f, indent=1)
    log('validation: surrogate vs MC median {:.4f} p90 {:.4f}; MC vs MC '
        'median {:.4f} p90 {:.4f}'.format(
            record['validation']['surrogate_vs_mc']['median'],
            record['validation']['surrogate_vs_mc']['p90'],
            record['validation']['mc_vs_mc_same_budget']['median'],
            record['validation']['mc_vs_mc_same_budget']['p90']))
    ctx.close()
    return record


def load_tissue_surrogate(context: N.Context = None,
                          path: str = TISSUE_WEIGHTS) -> ReflectanceSurrogate:
    """The shipped network of TISSUE_BOX (weights written by
    build_tissue_surrogate)."""
    sur = ReflectanceSurrogate(TISSUE_BOX, context=context)
    sur.load(path)
    return sur


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(
        description='Generate the Monte Carlo training set of the '
                    'reflectance surrogate over its box, train the network '
                    'and write its weights and record.')
    ap.add_argument('--samples', type=int, default=200000)
    ap.add_argument('--packets', type=int, default=1000000)
    ap.add_argument('--out', default=WEIGHTS_DIR)
    ap.add_argument('--seed', type=int, default=20261203)
    ap.add_argument('--steps-scale', type=float, default=1.0,
                    help='scale of every phase of the training schedule')
    ap.add_argument('--validation', type=int, default=2048)
    a = ap.parse_args(argv)
    sched = tuple((max(1, int(s * a.steps_scale)), lr)
                  for s, lr in DEFAULT_SCHEDULE)

    def log(msg):
        print(msg, flush=True)
    build_tissue_surrogate(a.samples, a.packets, a.out, a.seed, sched,
                           validation=a.validation, log=log)


if __name__ == '__main__':
    main(sys.argv[1:])
