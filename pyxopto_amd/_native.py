"""ctypes binding of libxomc.so (see include/xomc.h).

The structures below mirror the C structs field for field; their sizes are
checked against the library's own ``xo_struct_sizes`` when the library is
loaded (the analogue of the reference's ``sizeof_datatypes`` check,
xopto/mcml/mc.py:674-728). A missing library raises immediately: there is no
fallback implementation.
"""
import ctypes
import os
import threading
import weakref
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int32,
                    c_uint32, c_uint64, c_void_p)

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, 'libxomc.so')
DATA_DIR = os.path.join(PKG_DIR, 'data')
PRIME_CACHE = os.path.join(DATA_DIR, 'safeprimes_base32.u32')

XO_ACCU_K = 0x7FFFFF
XO_NUM_PRIMES = 500000
XO_MAX_SEEDS = XO_NUM_PRIMES - 1
XO_FIRST_PRIME_A = 4294967118
XO_MAX_PACKET_STEPS = 1 << 26
XO_MAX_LAYERS = 16
XO_MAX_MATERIALS = 32
XO_BATCH_MAX_LAYERS = 8
XO_NOALLOC_SCRATCH_BYTES = 16 << 20

XO_PF_HG, XO_PF_MHG, XO_PF_LUT = 0, 1, 2
XO_DET_NONE, XO_DET_TOTAL, XO_DET_RADIAL = 0, 1, 2
XO_SRC_LINE, XO_SRC_GAUSSIAN, XO_SRC_FIBER_NI, XO_SRC_FIBER = 0, 1, 2, 3
XO_METHOD_AW, XO_METHOD_AR, XO_METHOD_MBL = 0, 1, 2
XO_FLUENCE_NONE, XO_FLUENCE_XYZ, XO_FLUENCE_RZ = 0, 1, 2
XO_TRACE_NONE, XO_TRACE_START, XO_TRACE_END, XO_TRACE_ALL = 0, 1, 2, 7
XO_EVENT_REFLECTION, XO_EVENT_REFRACTION, XO_EVENT_BOUNDARY_HIT = 1, 2, 4
XO_EVENT_LAUNCH, XO_EVENT_ABSORPTION, XO_EVENT_SCATTERING = 8, 16, 32
XO_EVENT_TERMINATED, XO_EVENT_ESCAPED = 64, 128
XO_TRACE_ENTRY_LEN = 8
XO_FLAG_STATIC_SCHEDULE = 1 << 0
XO_FLAG_NO_LDS_BINS = 1 << 1
XO_FLAG_PACKET_KEYED = 1 << 2
XO_FLAG_FAST_MATH = 1 << 3
(XO_MATH_LOG, XO_MATH_SIN_TURNS, XO_MATH_COS_TURNS, XO_MATH_SQRT, XO_MATH_RCP,
 XO_MATH_RSQRT, XO_MATH_CBRT, XO_MATH_EXP) = range(8)

f3 = c_float * 3


class XoPf(ctypes.Structure):
    _fields_ = [('g', c_float), ('beta', c_float), ('a', c_float),
                ('b', c_float), ('c', c_float), ('offset', c_uint32),
                ('size', c_uint32), ('reserved', c_uint32)]


class XoLayer(ctypes.Structure):
    _fields_ = [('thickness', c_float), ('top', c_float), ('bottom', c_float),
                ('n', c_float), ('cos_critical_top', c_float),
                ('cos_critical_bottom', c_float), ('mus', c_float),
                ('mua', c_float), ('inv_mut', c_float),
                ('mua_inv_mut', c_float), ('reserved', c_float * 2),
                ('pf', XoPf)]


class XoSourceLine(ctypes.Structure):
    _fields_ = [('position', f3), ('direction_medium', f3),
                ('direction_sample', f3), ('direction_reflected', f3),
                ('reflectance', c_float), ('reserved', c_float * 3)]


class XoSourceGaussian(ctypes.Structure):
    _fields_ = [('transformation', c_float * 9), ('position', f3),
                ('direction', f3), ('sigma', c_float * 2), ('clip', c_float)]


class XoDetector(ctypes.Structure):
    _fields_ = [('kind', c_int32), ('direction', f3),
                ('position', c_float * 2), ('r_min', c_float),
                ('inv_dr', c_float), ('cos_min', c_float), ('n', c_uint32),
                ('offset', c_uint32), ('log_scale', c_int32)]


class XoFluence(ctypes.Structure):
    _fields_ = [('enabled', c_int32), ('inv_step', f3), ('top_left', f3),
                ('shape', c_uint32 * 3), ('offset', c_uint32), ('k', c_int32),
                ('rate', c_int32)]


class XoSourceFiber(ctypes.Structure):
    _fields_ = [('position', c_float * 3), ('radius', c_float),
                ('cos_min', c_float), ('n', c_float),
                ('reserved', c_float * 2)]


class XoFiberAxis(ctypes.Structure):
    _fields_ = [('transformation', c_float * 9), ('direction', c_float * 3)]


class XoMcmlCfg(ctypes.Structure):
    _fields_ = [('num_layers', c_uint32), ('pf_kind', c_int32),
                ('method', c_int32), ('rmax', c_float), ('w_min', c_float),
                ('lottery_chance', c_float), ('use_lottery', c_int32),
                ('source_kind', c_int32), ('source', XoSourceLine),
                ('top', XoDetector), ('bottom', XoDetector),
                ('specular', XoDetector),
                ('layers', XoLayer * XO_MAX_LAYERS),
                ('packet_seed', c_uint64), ('first_packet', c_uint64),
                ('fiber', XoSourceFiber), ('fluence', XoFluence),
                ('fiber_axis', XoFiberAxis)]


class XoMaterial(ctypes.Structure):
    _fields_ = [('n', c_float), ('mus', c_float), ('mua', c_float),
                ('inv_mut', c_float), ('mua_inv_mut', c_float),
                ('reserved', c_float * 3), ('pf', XoPf)]


class XoVoxelCfg(ctypes.Structure):
    _fields_ = [('top_left', f3), ('bottom_right', f3), ('size', f3),
                ('shape', c_int32 * 3)]


class XoMcvoxCfg(ctypes.Structure):
    _fields_ = [('num_materials', c_uint32), ('pf_kind', c_int32),
                ('method', c_int32), ('rmax', c_float), ('w_min', c_float),
                ('lottery_chance', c_float), ('use_lottery', c_int32),
                ('source_kind', c_int32), ('voxels', XoVoxelCfg),
                ('line', XoSourceLine), ('gaussian', XoSourceGaussian),
                ('fluence', XoFluence), ('top', XoDetector),
                ('bottom', XoDetector), ('specular', XoDetector),
                ('materials', XoMaterial * XO_MAX_MATERIALS),
                ('packet_seed', c_uint64), ('first_packet', c_uint64)]


class XoBatchInstance(ctypes.Structure):
    _fields_ = [('num_layers', c_uint32), ('pf_kind', c_int32),
                ('rmax', c_float), ('accum_base', c_uint32),
                ('accum_count', c_uint32), ('source_kind', c_int32),
                ('npackets', c_uint64), ('seed', c_uint64),
                ('source', XoSourceLine), ('top', XoDetector),
                ('bottom', XoDetector), ('specular', XoDetector),
                ('layers', XoLayer * XO_BATCH_MAX_LAYERS),
                ('fiber', XoSourceFiber), ('fiber_axis', XoFiberAxis),
                ('fluence', XoFluence)]


class XoBatchOpts(ctypes.Structure):
    _fields_ = [('chunk', c_uint32), ('block', c_uint32),
                ('blocks_per_cu', c_uint32), ('flags', c_uint32),
                ('w_min', c_float), ('lottery_chance', c_float),
                ('use_lottery', c_int32), ('refill_min', c_uint32),
                ('defer_min', c_uint32), ('method', c_int32)]


class XoServeOpts(ctypes.Structure):
    _fields_ = [('accum_capacity', c_uint64), ('lut_capacity', c_uint64),
                ('voxel_capacity', c_uint64), ('tick_packets', c_uint64),
                ('vox_tick_packets', c_uint64), ('instance_slots', c_uint32),
                ('max_items', c_uint32), ('vox_lanes', c_uint32),
                ('chunk', c_uint32), ('block', c_uint32),
                ('blocks_per_cu', c_uint32), ('w_min', c_float),
                ('lottery_chance', c_float), ('use_lottery', c_int32),
                ('use_graph', c_int32), ('fast_math', c_int32),
                ('reserved', c_int32)]


class XoServeStats(ctypes.Structure):
    _fields_ = [('admitted', c_uint64), ('completed', c_uint64),
                ('cancelled', c_uint64), ('ticks', c_uint64),
                ('packets', c_uint64), ('steps', c_uint64),
                ('accum_used', c_uint64), ('lut_used', c_uint64),
                ('voxel_used', c_uint64), ('busy_ms', c_double)]

    def todict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class XoTrace(ctypes.Structure):
    _fields_ = [('max_events', c_int32), ('options', c_uint32),
                ('event_mask', c_uint32), ('use_events', c_int32),
                ('plon', c_int32), ('reserved', c_int32 * 3)]


class XoSamplingVolume(ctypes.Structure):
    _fields_ = [('top_left', c_float * 3), ('voxel_size', c_float * 3),
                ('shape', c_uint32 * 3), ('multiplier', c_float),
                ('offset', c_uint32), ('k', c_int32), ('end_field', c_int32),
                ('reserved', c_int32 * 3)]


class XoRunReport(ctypes.Structure):
    _fields_ = [('packets', c_uint64), ('steps', c_uint64),
                ('boundary_steps', c_uint64), ('voxel_crossings', c_uint64),
                ('deposits', c_uint64), ('truncated', c_uint64),
                ('lanes', c_uint32),
                ('threads_used', c_uint32), ('kernel_ms', c_double),
                ('total_ms', c_double)]

    def todict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


STRUCTS_IN_ABI_ORDER = [XoPf, XoLayer, XoSourceLine, XoSourceGaussian,
                        XoDetector, XoFluence, XoMcmlCfg, XoMaterial,
                        XoVoxelCfg, XoMcvoxCfg, XoRunReport, XoBatchInstance,
                        XoBatchOpts, XoServeOpts, XoServeStats,
                        XoSourceFiber, XoTrace, XoSamplingVolume,
                        XoFiberAxis]

# Every entry point of include/xomc.h: name -> (restype, argtypes)
PROTOTYPES = {
    'xo_abi_version': (c_int, []),
    'xo_last_error': (c_char_p, []),
    'xo_device_count': (c_int, []),
    'xo_struct_sizes': (c_int, [POINTER(c_uint32), c_int]),
    'xo_safeprime_multipliers': (c_int, [c_uint32, POINTER(c_uint32)]),
    'xo_safeprime_preload': (c_int, [POINTER(c_uint32), c_uint32]),
    'xo_init_rng': (c_int, [c_uint32, c_uint64, POINTER(c_uint64),
                            POINTER(c_uint32)]),
    'xo_rng_host': (c_int, [POINTER(c_uint64), c_uint32, c_uint32,
                            POINTER(c_float)]),
    'xo_ctx_create': (c_void_p, [c_int, c_uint32, c_uint64]),
    'xo_ctx_destroy': (None, [c_void_p]),
    'xo_ctx_info': (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32),
                            POINTER(c_uint32)]),
    'xo_ctx_get_rng': (c_int, [c_void_p, c_uint32, POINTER(c_uint64),
                               POINTER(c_uint32)]),
    'xo_ctx_set_rng': (c_int, [c_void_p, c_uint32, POINTER(c_uint64),
                               POINTER(c_uint32)]),
    'xo_mcml_run': (c_int, [c_void_p, POINTER(XoMcmlCfg), POINTER(c_float),
                            c_uint32, c_uint64, c_uint32, c_uint32,
                            POINTER(c_uint64), c_uint64,
                            POINTER(XoRunReport)]),
    'xo_mcml_run_trace': (c_int, [c_void_p, POINTER(XoMcmlCfg),
                                  POINTER(c_float), c_uint32, c_uint64,
                                  c_uint32, c_uint32, POINTER(XoTrace),
                                  POINTER(c_float), POINTER(c_int32),
                                  POINTER(c_uint64), c_uint64,
                                  POINTER(XoRunReport)]),
    'xo_mcvox_run_trace': (c_int, [c_void_p, POINTER(XoMcvoxCfg),
                                   POINTER(c_int32), c_uint64,
                                   POINTER(c_float), c_uint32, c_uint64,
                                   c_uint32, c_uint32, POINTER(XoTrace),
                                   POINTER(c_float), POINTER(c_int32),
                                   POINTER(c_uint64), c_uint64,
                                   POINTER(XoRunReport)]),
    'xo_sampling_volume_run': (c_int, [c_void_p, POINTER(XoSamplingVolume),
                                       c_int32, POINTER(c_float),
                                       POINTER(c_int32), c_uint64,
                                       POINTER(c_uint64), c_uint64,
                                       POINTER(c_uint64),
                                       POINTER(XoRunReport)]),
    'xo_batch_run': (c_int, [c_void_p, POINTER(XoBatchInstance), c_uint32,
                             POINTER(c_float), c_uint32, POINTER(XoBatchOpts),
                             POINTER(c_uint64), c_uint64,
                             POINTER(XoRunReport)]),
    'xo_packet_seed': (c_int, [c_uint64, c_uint64, POINTER(c_uint64),
                               POINTER(c_uint32)]),
    'xo_batch_plan_create': (c_void_p, [c_void_p, POINTER(XoBatchInstance),
                                        c_uint32, POINTER(c_float), c_uint32,
                                        POINTER(XoBatchOpts), c_uint64]),
    'xo_batch_plan_run': (c_int, [c_void_p, POINTER(c_uint64),
                                  POINTER(XoRunReport)]),
    'xo_batch_plan_destroy': (None, [c_void_p]),
    'xo_batch_run_noalloc': (c_int, [c_void_p, POINTER(XoBatchInstance),
                                     c_uint32, c_void_p, c_uint32,
                                     POINTER(XoBatchOpts), c_void_p, c_uint64,
                                     POINTER(XoRunReport)]),
    'xo_mcvox_run': (c_int, [c_void_p, POINTER(XoMcvoxCfg), POINTER(c_int32),
                             c_uint64, POINTER(c_float), c_uint32, c_uint64,
                             c_uint32, c_uint32, POINTER(c_uint64), c_uint64,
                             POINTER(XoRunReport)]),
    'xo_serve_default_opts': (None, [POINTER(XoServeOpts)]),
    'xo_serve_create': (c_void_p, [c_void_p, POINTER(XoServeOpts)]),
    'xo_serve_destroy': (None, [c_void_p]),
    'xo_serve_admit_mcml': (c_int, [c_void_p, POINTER(XoBatchInstance),
                                    POINTER(c_float), c_uint32,
                                    POINTER(c_uint64)]),
    'xo_serve_admit_mcvox': (c_int, [c_void_p, POINTER(XoMcvoxCfg),
                                     POINTER(c_int32), c_uint64,
                                     POINTER(c_float), c_uint32, c_uint64,
                                     c_uint64, c_uint64, POINTER(c_uint64)]),
    'xo_serve_step': (c_int, [c_void_p, POINTER(c_uint32)]),
    'xo_serve_wait': (c_int, [c_void_p]),
    'xo_serve_poll': (c_int, [c_void_p, POINTER(c_uint64), c_uint32,
                              POINTER(c_uint32)]),
    'xo_serve_result': (c_int, [c_void_p, c_uint64, POINTER(c_uint64),
                                c_uint64]),
    'xo_serve_cancel': (c_int, [c_void_p, c_uint64]),
    'xo_serve_get_stats': (c_int, [c_void_p, POINTER(XoServeStats)]),
    'xo_mlp_create': (c_void_p, [c_void_p, POINTER(c_uint32), c_uint32,
                                 c_uint32]),
    'xo_mlp_destroy': (None, [c_void_p]),
    'xo_mlp_param_count': (c_uint64, [c_void_p]),
    'xo_mlp_set_params': (c_int, [c_void_p, POINTER(c_float), c_uint64]),
    'xo_mlp_get_params': (c_int, [c_void_p, POINTER(c_float), c_uint64]),
    'xo_mlp_forward': (c_int, [c_void_p, POINTER(c_float), c_uint64,
                               POINTER(c_float), POINTER(c_double)]),
    'xo_mlp_gradient': (c_int, [c_void_p, POINTER(c_float), POINTER(c_float),
                                c_uint32, POINTER(c_float), POINTER(c_float)]),
    'xo_mlp_set_data': (c_int, [c_void_p, POINTER(c_float), POINTER(c_float),
                                c_uint32]),
    'xo_mlp_train': (c_int, [c_void_p, c_uint32, c_uint32, c_float, c_uint64,
                             c_int, POINTER(c_float), POINTER(c_double)]),
    'xo_gemm_bf16': (c_int, [c_void_p, c_int, c_int, c_uint32, c_uint32,
                             c_uint32, POINTER(c_float), c_uint32,
                             POINTER(c_float), c_uint32, POINTER(c_float),
                             c_uint32]),
    'xo_probe_copy': (c_int, [c_void_p, c_uint64, c_uint32,
                              POINTER(c_double)]),
    'xo_probe_atomic_u64': (c_int, [c_void_p, c_uint64, c_uint64, c_uint64,
                                    POINTER(c_double), POINTER(c_uint64)]),
    'xo_probe_atomic_u64_xcd': (c_int, [c_void_p, c_uint64, c_uint64,
                                        c_uint64, c_uint64, c_int,
                                        POINTER(c_double),
                                        POINTER(c_uint64)]),
    'xo_probe_valu': (c_int, [c_void_p, c_uint32, POINTER(c_double),
                              POINTER(c_double)]),
    'xo_probe_cr_math': (c_int, [c_void_p, c_uint64, c_uint64,
                                 POINTER(c_uint64), POINTER(c_uint64),
                                 POINTER(c_uint64)]),
    'xo_probe_math': (c_int, [c_void_p, c_int, c_int, c_uint32,
                              POINTER(c_float), POINTER(c_float)]),
}

_lib = None
_lib_lock = threading.Lock()


class NativeError(RuntimeError):
    pass


def _load():
    if not os.path.isfile(LIB_PATH):
        raise ImportError(
            'libxomc.so is not built ({}); run '
            '"python -m pyxopto_amd.build_native" first'.format(LIB_PATH))
    lib = ctypes.CDLL(LIB_PATH)
    for name, (restype, argtypes) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    sizes = (c_uint32 * 32)()
    count = lib.xo_struct_sizes(sizes, 32)
    for i, struct in enumerate(STRUCTS_IN_ABI_ORDER[:count]):
        if ctypes.sizeof(struct) != sizes[i]:
            raise RuntimeError(
                'Size of {} differs between the host ({} B) and libxomc.so '
                '({} B)!'.format(struct.__name__, ctypes.sizeof(struct),
                                 sizes[i]))
    return lib


def lib():
    """The loaded library (loaded on first use)."""
    global _lib
    if _lib is None:
        with _lib_lock:
            if _lib is None:
                _lib = _load()
                _preload_primes(_lib)
    return _lib


def reload_library():
    global _lib
    with _lib_lock:
        _lib = None


def last_error() -> str:
    msg = lib().xo_last_error()
    return msg.decode() if msg else ''


def check(rc, what='libxomc'):
    if rc != 0:
        raise NativeError('{}: {}'.format(what, last_error()))


def _ptr(arr, ctype):
    return arr.ctypes.data_as(POINTER(ctype))


def _read_prime_cache():
    if not os.path.isfile(PRIME_CACHE):
        return None
    data = np.fromfile(PRIME_CACHE, dtype='<u4')
    if data.size == 0 or data[0] != XO_FIRST_PRIME_A:
        return None
    return data


def _preload_primes(l):
    data = _read_prime_cache()
    if data is not None:
        data = np.ascontiguousarray(data, dtype=np.uint32)
        l.xo_safeprime_preload(_ptr(data, c_uint32), c_uint32(data.size))


def safeprime_table(n: int = XO_NUM_PRIMES, ensure_cache: bool = False):
    """The first n multipliers of the safe-prime seed table (uint32)."""
    l = lib()
    out = np.empty(n, dtype=np.uint32)
    check(l.xo_safeprime_multipliers(c_uint32(n), _ptr(out, c_uint32)),
          'xo_safeprime_multipliers')
    if ensure_cache:
        cached = _read_prime_cache()
        if cached is None or cached.size < n:
            os.makedirs(DATA_DIR, exist_ok=True)
            tmp = PRIME_CACHE + '.tmp'
            out.astype('<u4').tofile(tmp)
            os.replace(tmp, PRIME_CACHE)
    return out


def init_rng(n: int, xinit: int):
    """init_RNG semantics: (x, a) seeds of n streams, or ValueError."""
    x = np.zeros(n, dtype=np.uint64)
    a = np.zeros(n, dtype=np.uint32)
    rc = lib().xo_init_rng(c_uint32(n), c_uint64(int(xinit)),
                           _ptr(x, c_uint64), _ptr(a, c_uint32))
    if rc == 1:
        raise ValueError('The provided initializer {} is not valid. Try a '
                         'different 64-bit unsigned integer!'.format(xinit))
    check(rc, 'xo_init_rng')
    return x, a


def rng_host(x: int, a: int, n: int):
    """n floats of one MWC stream; returns (values, new x)."""
    xs = c_uint64(int(x))
    out = np.empty(n, dtype=np.float32)
    check(lib().xo_rng_host(ctypes.byref(xs), c_uint32(int(a)), c_uint32(n),
                            _ptr(out, c_float)), 'xo_rng_host')
    return out, int(xs.value)


def packet_seed(seed: int, p: int):
    """(x, a) of packet p of a batch instance keyed by seed."""
    x, a = c_uint64(), c_uint32()
    check(lib().xo_packet_seed(c_uint64(int(seed)), c_uint64(int(p)),
                               ctypes.byref(x), ctypes.byref(a)),
          'xo_packet_seed')
    return int(x.value), int(a.value)


def device_count() -> int:
    return int(lib().xo_device_count())


class BatchPlan:
    """A batch whose tables live on the device and whose launch is a
    captured graph (xo_batch_plan_*): run() returns (accum, report) with the
    integers of Context.run_batch on the same arguments."""

    def __init__(self, ctx, instances, lut, opts, naccum: int):
        self._ctx = ctx            # the context must outlive the plan
        self.naccum = int(naccum)
        lut_p, nlut = None, 0
        if lut is not None and len(lut):
            lut = np.ascontiguousarray(lut, dtype=np.float32)
            lut_p, nlut = _ptr(lut, c_float), lut.size
        h = lib().xo_batch_plan_create(ctx._h, instances,
                                       c_uint32(len(instances)), lut_p,
                                       c_uint32(nlut), ctypes.byref(opts),
                                       c_uint64(self.naccum))
        if not h:
            raise NativeError('xo_batch_plan_create: {}'.format(last_error()))
        self._h = h
        ctx._dependents.add(self)

    def run(self):
        accum = np.zeros(self.naccum, dtype=np.uint64)
        rep = XoRunReport()
        check(lib().xo_batch_plan_run(self._h, _ptr(accum, c_uint64),
                                      ctypes.byref(rep)), 'xo_batch_plan_run')
        return accum, rep

    def close(self):
        if getattr(self, '_h', None):
            lib().xo_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """Owner of one xo_ctx: RNG streams and buffers on one device."""

    def __init__(self, device: int = 0, max_lanes: int = 0,
                 xinit: int = 0x5EED000000000001):
        l = lib()
        h = l.xo_ctx_create(c_int(device), c_uint32(max_lanes),
                            c_uint64(int(xinit)))
        if not h:
            msg = last_error()
            if 'not valid' in msg:
                raise ValueError(msg)
            raise NativeError('xo_ctx_create: ' + msg)
        self._h = h
        # objects holding this context's device, stream or buffers (batch
        # plans, networks): closed before the context is
        self._dependents = weakref.WeakSet()
        self.device = device
        p, m, c = c_uint32(), c_uint32(), c_uint32()
        check(l.xo_ctx_info(h, ctypes.byref(p), ctypes.byref(m),
                            ctypes.byref(c)), 'xo_ctx_info')
        self.persistent_lanes = p.value
        self.max_lanes = m.value
        self.compute_units = c.value

    def close(self):
        if getattr(self, '_h', None):
            for dep in list(getattr(self, '_dependents', ())):
                dep.close()
            lib().xo_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_rng(self, n: int):
        x = np.empty(n, dtype=np.uint64)
        a = np.empty(n, dtype=np.uint32)
        check(lib().xo_ctx_get_rng(self._h, c_uint32(n), _ptr(x, c_uint64),
                                   _ptr(a, c_uint32)), 'xo_ctx_get_rng')
        return x, a

    def set_rng(self, x, a=None):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        ap = None
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.uint32)
            ap = _ptr(a, c_uint32)
        check(lib().xo_ctx_set_rng(self._h, c_uint32(x.size),
                                   _ptr(x, c_uint64), ap), 'xo_ctx_set_rng')

    def run_mcml(self, cfg: XoMcmlCfg, lut, npackets: int, naccum: int,
                 lanes: int = 0, flags: int = 0):
        accum = np.zeros(naccum, dtype=np.uint64)
        lut_p, nlut = None, 0
        if lut is not None and len(lut):
            lut = np.ascontiguousarray(lut, dtype=np.float32)
            lut_p, nlut = _ptr(lut, c_float), lut.size
        rep = XoRunReport()
        check(lib().xo_mcml_run(self._h, ctypes.byref(cfg), lut_p,
                                c_uint32(nlut), c_uint64(int(npackets)),
                                c_uint32(lanes), c_uint32(flags),
                                _ptr(accum, c_uint64), c_uint64(naccum),
                                ctypes.byref(rep)), 'xo_mcml_run')
        return accum, rep

    def run_mcml_trace(self, cfg: XoMcmlCfg, lut, npackets: int, naccum: int,
                       trace: XoTrace, lanes: int = 0, flags: int = 0):
        """run_mcml with the packet trace: (accum, report, data, n) with
        data float32[npackets][max_events][8] and n int32[npackets]."""
        npackets = int(npackets)
        accum = np.zeros(naccum, dtype=np.uint64)
        data = np.zeros((npackets, trace.max_events, XO_TRACE_ENTRY_LEN),
                        dtype=np.float32)
        n = np.zeros(npackets, dtype=np.int32)
        lut_p, nlut = None, 0
        if lut is not None and len(lut):
            lut = np.ascontiguousarray(lut, dtype=np.float32)
            lut_p, nlut = _ptr(lut, c_float), lut.size
        rep = XoRunReport()
        check(lib().xo_mcml_run_trace(
            self._h, ctypes.byref(cfg), lut_p, c_uint32(nlut),
            c_uint64(npackets), c_uint32(lanes), c_uint32(flags),
            ctypes.byref(trace), _ptr(data, c_float), _ptr(n, c_int32),
            _ptr(accum, c_uint64), c_uint64(naccum), ctypes.byref(rep)),
            'xo_mcml_run_trace')
        return accum, rep, data, n

    def sampling_volume(self, sv: XoSamplingVolume, data, n, naccum: int):
        """(accum, total_weight, report) of xo_sampling_volume_run over a
        trace (data float32[packets][max_events][8], n int32[packets])."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        if data.ndim != 3 or data.shape[2] != XO_TRACE_ENTRY_LEN or \
                data.shape[0] != n.size:
            raise ValueError('trace data must be [packets][max_events][8] '
                             'with one count per packet')
        accum = np.zeros(naccum, dtype=np.uint64)
        total = c_uint64(0)
        rep = XoRunReport()
        check(lib().xo_sampling_volume_run(
            self._h, ctypes.byref(sv), c_int32(data.shape[1]),
            _ptr(data, c_float), _ptr(n, c_int32), c_uint64(n.size),
            _ptr(accum, c_uint64), c_uint64(naccum), ctypes.byref(total),
            ctypes.byref(rep)), 'xo_sampling_volume_run')
        return accum, int(total.value), rep

    def run_batch(self, instances, lut, opts: XoBatchOpts, naccum: int):
        """instances: ctypes array of XoBatchInstance."""
        accum = np.zeros(naccum, dtype=np.uint64)
        lut_p, nlut = None, 0
        if lut is not None and len(lut):
            lut = np.ascontiguousarray(lut, dtype=np.float32)
            lut_p, nlut = _ptr(lut, c_float), lut.size
        rep = XoRunReport()
        check(lib().xo_batch_run(self._h, instances, c_uint32(len(instances)),
                                 lut_p, c_uint32(nlut), ctypes.byref(opts),
                                 _ptr(accum, c_uint64), c_uint64(naccum),
                                 ctypes.byref(rep)), 'xo_batch_run')
        return accum, rep

    def batch_plan(self, instances, lut, opts: XoBatchOpts, naccum: int):
        """xo_batch_plan_create: the batch uploaded and its launch captured
        once; BatchPlan.run replays it."""
        return BatchPlan(self, instances, lut, opts, naccum)

    def run_batch_noalloc(self, instances, lut_dev: int, nlut: int,
                          opts: XoBatchOpts, accum_dev: int, naccum: int):
        """xo_batch_run_noalloc: lut_dev and accum_dev are device addresses
        (integers) the caller allocated; the accumulator is left on the
        device. Returns the run report."""
        rep = XoRunReport()
        check(lib().xo_batch_run_noalloc(
            self._h, instances, c_uint32(len(instances)),
            c_void_p(lut_dev or None), c_uint32(nlut), ctypes.byref(opts),
            c_void_p(accum_dev), c_uint64(naccum), ctypes.byref(rep)),
            'xo_batch_run_noalloc')
        return rep

    def probe_copy(self, nbytes: int, reps: int = 10) -> float:
        """Average ms of one device-to-device copy of nbytes."""
        ms = c_double()
        check(lib().xo_probe_copy(self._h, c_uint64(int(nbytes)),
                                  c_uint32(reps), ctypes.byref(ms)),
              'xo_probe_copy')
        return ms.value

    def probe_valu(self, iters: int = 20000):
        """(kernel ms, lane-operations) of the fp32 FMA throughput probe."""
        ms, ops = c_double(), c_double()
        check(lib().xo_probe_valu(self._h, c_uint32(iters), ctypes.byref(ms),
                                  ctypes.byref(ops)), 'xo_probe_valu')
        return ms.value, ops.value

    def probe_cr_math(self, ndiv: int = 1 << 30, seed: int = 1):
        """(square roots checked, their mismatches, division mismatches of
        ndiv random pairs) of the range-restricted correctly rounded
        operations against the general ones."""
        chk, bs, bd = c_uint64(), c_uint64(), c_uint64()
        check(lib().xo_probe_cr_math(self._h, c_uint64(int(ndiv)),
                                     c_uint64(int(seed)), ctypes.byref(chk),
                                     ctypes.byref(bs), ctypes.byref(bd)),
              'xo_probe_cr_math')
        return int(chk.value), int(bs.value), int(bd.value)

    def probe_math(self, which: int, values, fast: bool):
        """f(values) on the device for the function `which` (XO_MATH_*) in
        the exact or the fast arithmetic of the kernels."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        out = np.zeros_like(v)
        check(lib().xo_probe_math(self._h, c_int(int(bool(fast))),
                                  c_int(int(which)), c_uint32(v.size),
                                  _ptr(v, c_float), _ptr(out, c_float)),
              'xo_probe_math')
        return out

    def probe_atomic_u64(self, array_bytes: int, natomics: int,
                         seed: int = 1):
        """(kernel ms, array sum) of natomics random u64 atomic adds."""
        ms, total = c_double(), c_uint64()
        check(lib().xo_probe_atomic_u64(self._h, c_uint64(int(array_bytes)),
                                        c_uint64(int(natomics)),
                                        c_uint64(int(seed)), ctypes.byref(ms),
                                        ctypes.byref(total)),
              'xo_probe_atomic_u64')
        return ms.value, int(total.value)

    def probe_atomic_u64_xcd(self, array_bytes: int, hot_bytes: int,
                             natomics: int, per_xcd: bool, seed: int = 1):
        """(kernel ms, sum) of natomics random u64 atomic adds among the
        first hot_bytes, into one array per XCD when per_xcd."""
        ms, total = c_double(), c_uint64()
        check(lib().xo_probe_atomic_u64_xcd(
            self._h, c_uint64(int(array_bytes)), c_uint64(int(hot_bytes)),
            c_uint64(int(natomics)), c_uint64(int(seed)), c_int(int(per_xcd)),
            ctypes.byref(ms), ctypes.byref(total)), 'xo_probe_atomic_u64_xcd')
        return ms.value, int(total.value)

    def run_mcvox(self, cfg: XoMcvoxCfg, voxels, lut, npackets: int,
                  naccum: int, lanes: int = 0, flags: int = 0, trace=None):
        """(accum, report), or (accum, report, data, n) with a trace
        (XoTrace): data float32[npackets][max_events][8], n int32."""
        accum = np.zeros(naccum, dtype=np.uint64)
        vox_p, nvox = None, 0
        if voxels is not None:
            voxels = np.ascontiguousarray(voxels, dtype=np.int32)
            vox_p, nvox = _ptr(voxels, c_int32), voxels.size
        lut_p, nlut = None, 0
        if lut is not None and len(lut):
            lut = np.ascontiguousarray(lut, dtype=np.float32)
            lut_p, nlut = _ptr(lut, c_float), lut.size
        rep = XoRunReport()
        if trace is not None:
            npackets = int(npackets)
            data = np.zeros((npackets, trace.max_events, XO_TRACE_ENTRY_LEN),
                            dtype=np.float32)
            n = np.zeros(npackets, dtype=np.int32)
            check(lib().xo_mcvox_run_trace(
                self._h, ctypes.byref(cfg), vox_p, c_uint64(nvox), lut_p,
                c_uint32(nlut), c_uint64(npackets), c_uint32(lanes),
                c_uint32(flags), ctypes.byref(trace), _ptr(data, c_float),
                _ptr(n, c_int32), _ptr(accum, c_uint64), c_uint64(naccum),
                ctypes.byref(rep)), 'xo_mcvox_run_trace')
            return accum, rep, data, n
        check(lib().xo_mcvox_run(self._h, ctypes.byref(cfg), vox_p,
                                 c_uint64(nvox), lut_p, c_uint32(nlut),
                                 c_uint64(int(npackets)), c_uint32(lanes),
                                 c_uint32(flags), _ptr(accum, c_uint64),
                                 c_uint64(naccum), ctypes.byref(rep)),
              'xo_mcvox_run')
        return accum, rep
