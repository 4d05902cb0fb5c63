"""ctypes front end of oracle/liboracle.so (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
use this module, as the checker or as the timed CPU baseline. It reuses the
ABI struct definitions of pyxopto_amd._native (plain ctypes; loading them
does not load libxomc.so)."""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_float, c_int, c_int32, c_uint32, c_uint64

import numpy as np

from pyxopto_amd import _native as N

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')
REF_RNG_PATH = os.path.join(HERE, '_ref', 'librng_ref.so')

MATH_LIBM = 0    # C library log/sin/cos/cbrt: independent of the product
MATH_KERNEL = 1  # the kernels' fixed-sequence functions: bit-exact runs

_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE, 'liboracle.so', 'ref'],
                   check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def use_library(path: str):
    """Load the oracle from another build of xo_oracle.c (the host-tuned
    CPU baseline of bench.py) for every later call."""
    global _lib
    _lib = _load(path)


def build_native(out_path: str) -> str:
    """Build xo_oracle.c for this host (gcc -O3 -march=native -fopenmp,
    the oracle/Makefile `native` recipe) into out_path."""
    subprocess.run(['make', '-s', '-C', HERE, 'native', 'OUT=' + out_path],
                   check=True)
    return out_path


def _load(path):
    l = ctypes.CDLL(path)
    l.oracle_mcml_run.restype = c_int
    l.oracle_mcml_run.argtypes = [
        POINTER(N.XoMcmlCfg), POINTER(c_float), c_uint32,
        POINTER(c_uint64), POINTER(c_uint32), c_uint64, c_int, c_int,
        POINTER(c_uint64), c_uint64, POINTER(c_uint64), POINTER(c_uint32)]
    l.oracle_mcml_run_trace.restype = c_int
    l.oracle_mcml_run_trace.argtypes = [
        POINTER(N.XoMcmlCfg), POINTER(c_float), c_uint32,
        POINTER(c_uint64), POINTER(c_uint32), c_uint64, c_int, c_int,
        POINTER(c_uint64), c_uint64, POINTER(c_uint64), POINTER(c_uint32),
        POINTER(N.XoTrace), POINTER(c_float), POINTER(c_int32)]
    l.oracle_mcvox_run_trace.restype = c_int
    l.oracle_mcvox_run_trace.argtypes = [
        POINTER(N.XoMcvoxCfg), POINTER(c_int32), POINTER(c_float),
        c_uint32, POINTER(c_uint64), POINTER(c_uint32), c_uint64, c_int,
        c_int, POINTER(c_uint64), c_uint64, POINTER(c_uint64),
        POINTER(c_uint32), POINTER(N.XoTrace), POINTER(c_float),
        POINTER(c_int32)]
    l.oracle_sampling_volume.restype = c_int
    l.oracle_sampling_volume.argtypes = [
        POINTER(N.XoSamplingVolume), c_int32, POINTER(c_float),
        POINTER(c_int32), c_uint64, POINTER(c_uint64), c_uint64,
        POINTER(c_uint64), POINTER(c_uint64)]
    l.oracle_mcvox_run.restype = c_int
    l.oracle_mcvox_run.argtypes = [
        POINTER(N.XoMcvoxCfg), POINTER(c_int32), POINTER(c_float),
        c_uint32, POINTER(c_uint64), POINTER(c_uint32), c_uint64, c_int,
        c_int, POINTER(c_uint64), c_uint64, POINTER(c_uint64),
        POINTER(c_uint32)]
    l.oracle_rng.restype = c_int
    l.oracle_rng.argtypes = [POINTER(c_uint64), c_uint32, c_uint32,
                             POINTER(c_float)]
    l.oracle_max_threads.restype = c_int
    l.oracle_batch_run.restype = c_int
    l.oracle_batch_run.argtypes = [
        POINTER(N.XoBatchInstance), c_uint32, POINTER(c_float),
        POINTER(c_uint32), c_float, c_float, c_int, c_int, c_int, c_int,
        POINTER(c_uint64), c_uint64, POINTER(c_uint64)]
    l.oracle_pf_samples.restype = c_int
    l.oracle_pf_samples.argtypes = [
        c_int, POINTER(N.XoPf), POINTER(c_float), c_uint64, c_uint32,
        c_uint32, c_int, POINTER(c_float), POINTER(c_float)]
    l.oracle_scatter.restype = c_int
    l.oracle_scatter.argtypes = [c_uint32, POINTER(c_float),
                                 POINTER(c_float), POINTER(c_float), c_int,
                                 POINTER(c_float)]
    l.oracle_reflectance.restype = c_int
    l.oracle_reflectance.argtypes = [c_float, c_float, c_uint32,
                                     POINTER(c_float), POINTER(c_float)]
    l.oracle_math.restype = c_int
    l.oracle_math.argtypes = [c_int, c_uint32, POINTER(c_float),
                              POINTER(c_float), POINTER(c_float)]
    return l


def _p(arr, ctype):
    return arr.ctypes.data_as(POINTER(ctype))


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def kernel_math(which: str, values):
    """The kernels' log ('log'), log of a uniform deviate ('log_unit'),
    sin/cos of turns ('sincos', returns a pair), cube root ('cbrt') or
    exponential ('exp') evaluated on the host."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    out = np.empty_like(v)
    out2 = np.empty_like(v)
    code = {'log': 0, 'sincos': 1, 'cbrt': 2, 'exp': 3, 'log_unit': 4}[which]
    rc = lib().oracle_math(c_int(code), c_uint32(v.size), _p(v, c_float),
                           _p(out, c_float), _p(out2, c_float))
    if rc != 0:
        raise RuntimeError('oracle_math failed')
    return (out, out2) if which == 'sincos' else out


def pf_samples(kind: int, pf, lut, x: int, a: int, n: int,
               math_mode: int = MATH_LIBM):
    """n (cos_theta, azimuth turns) draws of the phase-function sampler."""
    cos_out = np.empty(n, dtype=np.float32)
    fi_out = np.empty(n, dtype=np.float32)
    lut_p = None
    if lut is not None and len(lut):
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        lut_p = _p(lut, c_float)
    lib().oracle_pf_samples(c_int(kind), ctypes.byref(pf), lut_p,
                            c_uint64(int(x)), c_uint32(int(a)), c_uint32(n),
                            c_int(math_mode), _p(cos_out, c_float),
                            _p(fi_out, c_float))
    return cos_out, fi_out


def scatter(dirs, cos_theta, xi_fi, math_mode: int = MATH_LIBM):
    d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
    ct = np.ascontiguousarray(cos_theta, dtype=np.float32)
    fi = np.ascontiguousarray(xi_fi, dtype=np.float32)
    out = np.empty_like(d)
    lib().oracle_scatter(c_uint32(d.shape[0]), _p(d, c_float), _p(ct, c_float),
                         _p(fi, c_float), c_int(math_mode), _p(out, c_float))
    return out


def reflectance(n1: float, n2: float, cos1):
    c = np.ascontiguousarray(cos1, dtype=np.float32)
    out = np.empty_like(c)
    lib().oracle_reflectance(c_float(n1), c_float(n2), c_uint32(c.size),
                             _p(c, c_float), _p(out, c_float))
    return out


def rng(x: int, a: int, n: int):
    xs = c_uint64(int(x))
    out = np.empty(n, dtype=np.float32)
    lib().oracle_rng(ctypes.byref(xs), c_uint32(int(a)), c_uint32(n),
                     _p(out, c_float))
    return out, int(xs.value)


def _keyed_table(keyed):
    if not keyed:
        return None
    t = np.ascontiguousarray(N.safeprime_table(N.XO_NUM_PRIMES)[1:])
    return t


def mcml_run(cfg, lut, rng_x, rng_a, npackets: int, naccum: int,
             math_mode: int = MATH_LIBM, nthreads: int = 0,
             keyed: bool = False):
    """Static-schedule run on len(rng_x) lanes. Returns
    (accum, counts[steps, boundary, crossings, deposits], rng_x_after)."""
    x = np.ascontiguousarray(rng_x, dtype=np.uint64).copy()
    a = np.ascontiguousarray(rng_a, dtype=np.uint32)
    accum = np.zeros(naccum, dtype=np.uint64)
    counts = np.zeros(4, dtype=np.uint64)
    lut_p = None
    if lut is not None and len(lut):
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        lut_p = _p(lut, c_float)
    table = _keyed_table(keyed)
    rc = lib().oracle_mcml_run(ctypes.byref(cfg), lut_p, c_uint32(x.size),
                               _p(x, c_uint64), _p(a, c_uint32),
                               c_uint64(int(npackets)), c_int(math_mode),
                               c_int(nthreads), _p(accum, c_uint64),
                               c_uint64(naccum), _p(counts, c_uint64),
                               None if table is None else _p(table, c_uint32))
    if rc != 0:
        raise RuntimeError('oracle_mcml_run failed ({})'.format(rc))
    return accum, counts, x


def mcml_run_trace(cfg, lut, rng_x, rng_a, npackets: int, naccum: int,
                   trace, math_mode: int = MATH_LIBM, nthreads: int = 0,
                   keyed: bool = False):
    """mcml_run with the packet trace (xo_trace `trace`). Returns (accum,
    counts, rng_x_after, data float32[npackets][max_events][8],
    n int32[npackets])."""
    npackets = int(npackets)
    x = np.ascontiguousarray(rng_x, dtype=np.uint64).copy()
    a = np.ascontiguousarray(rng_a, dtype=np.uint32)
    accum = np.zeros(naccum, dtype=np.uint64)
    counts = np.zeros(4, dtype=np.uint64)
    data = np.zeros((npackets, trace.max_events, N.XO_TRACE_ENTRY_LEN),
                    dtype=np.float32)
    n = np.zeros(npackets, dtype=np.int32)
    lut_p = None
    if lut is not None and len(lut):
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        lut_p = _p(lut, c_float)
    table = _keyed_table(keyed)
    rc = lib().oracle_mcml_run_trace(
        ctypes.byref(cfg), lut_p, c_uint32(x.size), _p(x, c_uint64),
        _p(a, c_uint32), c_uint64(npackets), c_int(math_mode), c_int(nthreads),
        _p(accum, c_uint64), c_uint64(naccum), _p(counts, c_uint64),
        None if table is None else _p(table, c_uint32), ctypes.byref(trace),
        _p(data, c_float), _p(n, c_int32))
    if rc != 0:
        raise RuntimeError('oracle_mcml_run_trace failed ({})'.format(rc))
    return accum, counts, x, data, n


def sampling_volume(sv, data, n, naccum: int):
    """(accum, total_weight, [walk steps, truncated]) of the sampling volume
    of a trace."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    n = np.ascontiguousarray(n, dtype=np.int32)
    accum = np.zeros(naccum, dtype=np.uint64)
    total = c_uint64(0)
    counts = np.zeros(2, dtype=np.uint64)
    rc = lib().oracle_sampling_volume(
        ctypes.byref(sv), c_int32(data.shape[1]), _p(data, c_float),
        _p(n, c_int32), c_uint64(n.size), _p(accum, c_uint64),
        c_uint64(naccum), ctypes.byref(total), _p(counts, c_uint64))
    if rc != 0:
        raise RuntimeError('oracle_sampling_volume failed ({})'.format(rc))
    return accum, int(total.value), counts


def mcvox_run(cfg, voxels, lut, rng_x, rng_a, npackets: int, naccum: int,
              math_mode: int = MATH_LIBM, nthreads: int = 0,
              keyed: bool = False):
    x = np.ascontiguousarray(rng_x, dtype=np.uint64).copy()
    a = np.ascontiguousarray(rng_a, dtype=np.uint32)
    vox = np.ascontiguousarray(voxels, dtype=np.int32)
    accum = np.zeros(naccum, dtype=np.uint64)
    counts = np.zeros(4, dtype=np.uint64)
    lut_p = None
    if lut is not None and len(lut):
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        lut_p = _p(lut, c_float)
    table = _keyed_table(keyed)
    rc = lib().oracle_mcvox_run(ctypes.byref(cfg), _p(vox, c_int32), lut_p,
                                c_uint32(x.size), _p(x, c_uint64),
                                _p(a, c_uint32), c_uint64(int(npackets)),
                                c_int(math_mode), c_int(nthreads),
                                _p(accum, c_uint64), c_uint64(naccum),
                                _p(counts, c_uint64),
                                None if table is None else _p(table, c_uint32))
    if rc != 0:
        raise RuntimeError('oracle_mcvox_run failed ({})'.format(rc))
    return accum, counts, x


def mcvox_run_trace(cfg, voxels, lut, rng_x, rng_a, npackets: int,
                    naccum: int, trace, math_mode: int = MATH_LIBM,
                    nthreads: int = 0, keyed: bool = False):
    """mcvox_run with the packet trace; returns (accum, counts,
    rng_x_after, data, n) as mcml_run_trace."""
    npackets = int(npackets)
    x = np.ascontiguousarray(rng_x, dtype=np.uint64).copy()
    a = np.ascontiguousarray(rng_a, dtype=np.uint32)
    vox = np.ascontiguousarray(voxels, dtype=np.int32)
    accum = np.zeros(naccum, dtype=np.uint64)
    counts = np.zeros(4, dtype=np.uint64)
    data = np.zeros((npackets, trace.max_events, N.XO_TRACE_ENTRY_LEN),
                    dtype=np.float32)
    n = np.zeros(npackets, dtype=np.int32)
    lut_p = None
    if lut is not None and len(lut):
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        lut_p = _p(lut, c_float)
    table = _keyed_table(keyed)
    rc = lib().oracle_mcvox_run_trace(
        ctypes.byref(cfg), _p(vox, c_int32), lut_p, c_uint32(x.size),
        _p(x, c_uint64), _p(a, c_uint32), c_uint64(npackets), c_int(math_mode),
        c_int(nthreads), _p(accum, c_uint64), c_uint64(naccum),
        _p(counts, c_uint64), None if table is None else _p(table, c_uint32),
        ctypes.byref(trace), _p(data, c_float), _p(n, c_int32))
    if rc != 0:
        raise RuntimeError('oracle_mcvox_run_trace failed ({})'.format(rc))
    return accum, counts, x, data, n


def batch_run(instances, lut, naccum: int, w_min=1e-4, lottery_chance=0.1,
              use_lottery=True, math_mode: int = MATH_KERNEL,
              nthreads: int = 0, method: int = 0):
    """A batch of layered instances with per-packet generators
    (schedule independent), every instance with the stepping method
    `method` (xo_method). instances: ctypes array of XoBatchInstance."""
    table = np.ascontiguousarray(N.safeprime_table(N.XO_NUM_PRIMES)[1:])
    accum = np.zeros(naccum, dtype=np.uint64)
    counts = np.zeros(4, dtype=np.uint64)
    lut_p = None
    if lut is not None and len(lut):
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        lut_p = _p(lut, c_float)
    rc = lib().oracle_batch_run(instances, c_uint32(len(instances)), lut_p,
                                _p(table, c_uint32), c_float(w_min),
                                c_float(lottery_chance), c_int(int(use_lottery)),
                                c_int(int(method)), c_int(math_mode),
                                c_int(nthreads),
                                _p(accum, c_uint64), c_uint64(naccum),
                                _p(counts, c_uint64))
    if rc != 0:
        raise RuntimeError('oracle_batch_run failed ({})'.format(rc))
    return accum, counts


def ref_rng_lib():
    """The reference's own rng.cpp compiled into _ref/ (None if absent)."""
    if not os.path.isfile(REF_RNG_PATH):
        return None
    l = ctypes.CDLL(REF_RNG_PATH)
    l.init_RNG.restype = c_int
    l.init_RNG.argtypes = [POINTER(c_uint64), POINTER(c_uint32),
                           POINTER(c_uint32), c_uint32, c_uint64]
    for name in ('rand_cc', 'rand_co', 'rand_oc', 'rand_oo'):
        fn = getattr(l, name)
        fn.restype = c_int
        fn.argtypes = [POINTER(c_float), c_uint32, POINTER(c_uint64), c_uint32]
    return l
