"""Common host runtime of the simulators: the device context (RNG streams
and buffers, created on the first run), accumulator allocation and the
read-only lookup-table array. Replaces ClWorker + ClWorkerRngMixin +
BufferAllocator + LutManager (xopto/mcbase/mcworker.py:46-1248,
xopto/mcbase/mcutil/buffer.py:121-288, xopto/mcbase/mcutil/lut.py:408-579)."""
import numpy as np

from .. import _native as N


def random_initializer(rng=None) -> int:
    """A valid MWC initializer drawn at random (xopto/cl/clrng.py:447-462)."""
    rng = np.random.default_rng() if rng is None else rng
    while True:
        xinit = int(rng.integers(1, 2 ** 64 - 1, dtype=np.uint64))
        if not ((xinit >> 32) >= N.XO_FIRST_PRIME_A - 1 or
                (xinit & 0xFFFFFFFF) >= 0xFFFFFFFF or xinit == 0):
            return xinit


# OpenCL build options that allow approximate math (xopto/cl/cloptions.py:
# 23-30, 137-152): with any of them the kernels use the hardware
# approximations of log, exp, sin/cos, sqrt and division (XO_FLAG_FAST_MATH).
FAST_MATH_OPTIONS = ('-cl-fast-relaxed-math', '-cl-unsafe-math-optimizations')


def fast_math_requested(cl_build_options) -> bool:
    """True if a build option (a string, or an object whose str() is the
    option, as cloptions.ClBuildOption) asks for approximate math."""
    for item in cl_build_options or ():
        text = str(getattr(item, 'value', item))
        if any(opt in text for opt in FAST_MATH_OPTIONS):
            return True
    return False


def validate_initializer(xinit: int) -> int:
    """init_RNG's validity test (xopto/src/rng/rng.cpp:74-80)."""
    xinit = int(xinit)
    if (xinit <= 0 or xinit >= 2 ** 64 or
            (xinit >> 32) >= N.XO_FIRST_PRIME_A - 1 or
            (xinit & 0xFFFFFFFF) >= 0xFFFFFFFF):
        raise ValueError('The provided initializer {} is not valid. Try a '
                         'different 64-bit unsigned integer!'.format(xinit))
    return xinit


class McWorker:
    """Device context owner shared by mcml.Mc and mcvox.Mc."""

    def __init__(self, options=None, rnginit=None, cl_devices=None,
                 cl_build_options=None, cl_profiling: bool = False,
                 types=None):
        from . import mcoptions
        self._options = list(options or [])
        self._opts = mcoptions.resolve(self._options)
        if rnginit is None:
            rnginit = random_initializer()
        self._rnginit = validate_initializer(rnginit)
        self._device = cl_devices if isinstance(cl_devices, int) else 0
        self._cl_build_options = [str(getattr(item, 'value', item))
                                  for item in (cl_build_options or [])]
        self._fast_math = fast_math_requested(cl_build_options)
        self._ctx = None
        self._rng_seeds_x = None
        self._rng_seeds_a = None
        self._alloc_offset = 0
        self._allocations = {}
        self._luts = []
        self._lut_offset = 0
        self._lut_keys = {}
        self._run_report = {'upload': None, 'download': None,
                            'execution': None, 'threads': None, 'items': None}

    # ---------------- device context ----------------
    def _context(self) -> N.Context:
        if self._ctx is None:
            self._ctx = N.Context(self._device, 0, self._rnginit)
        return self._ctx

    context = property(_context, None, None, 'Device context (created lazily).')
    rnginit = property(lambda self: self._rnginit, None, None,
                       'RNG initializer.')
    cl_build_options = property(lambda self: tuple(self._cl_build_options),
                                None, None, 'Build options as given.')
    fast_math = property(lambda self: self._fast_math, None, None,
                         'True if the build options allow approximate '
                         'math (hardware log, exp, sin/cos, sqrt, division).')
    run_report = property(lambda self: dict(self._run_report), None, None,
                          'Timing and thread report of the last run.')

    def _get_rng_seeds(self):
        x, a = self._context().get_rng(self._context().max_lanes)
        return x, a

    rng_seeds = property(_get_rng_seeds, None, None,
                         'Current (x, a) state of all RNG streams.')

    def _host_seeds(self):
        if self._rng_seeds_x is None:
            self._rng_seeds_x, self._rng_seeds_a = N.init_rng(
                N.XO_MAX_SEEDS, self._rnginit)
        return self._rng_seeds_x, self._rng_seeds_a

    def _copy_seeds(self):
        """run(copyseeds=True): the mutable generator states of every lane
        downloaded into rng_seeds_x (xopto/mcml/mc.py:962-966)."""
        ctx = self._context()
        x, a = ctx.get_rng(ctx.max_lanes)
        self._host_seeds()
        self._rng_seeds_x[:len(x)] = x
        self._rng_seeds_a[:len(a)] = a

    rng_seeds_x = property(lambda self: self._host_seeds()[0], None, None,
                           'Mutable random number generator seeds X: the '
                           'initial states, or the device states after a '
                           'run with copyseeds=True '
                           '(xopto/mcbase/mcworker.py:1229-1232).')
    rng_seeds_a = property(lambda self: self._host_seeds()[1], None, None,
                           'Immutable random number generator seeds A '
                           '(xopto/mcbase/mcworker.py:1234-1237).')

    # ---------------- packing helpers ----------------
    def _reset_allocations(self):
        self._alloc_offset = 0
        self._allocations = {}
        self._luts = []
        self._lut_offset = 0
        self._lut_keys = {}

    def allocate_accumulator(self, owner, n: int) -> int:
        offset = self._alloc_offset
        self._allocations[id(owner)] = (offset, int(n))
        self._alloc_offset += int(n)
        return offset

    def allocation(self, owner):
        return self._allocations[id(owner)]

    def append_r_lut(self, data) -> int:
        data = np.ascontiguousarray(data, dtype=np.float32)
        key = (data.size, data.tobytes())
        if key in self._lut_keys:
            return self._lut_keys[key]
        offset = self._lut_offset
        self._luts.append(data)
        self._lut_offset += data.size
        self._lut_keys[key] = offset
        return offset

    def _lut_array(self):
        if not self._luts:
            return None
        return np.concatenate(self._luts)

    def _apply_options(self, cfg):
        cfg.method = int(self._opts['method'])
        cfg.w_min = float(self._opts['w_min'])
        cfg.lottery_chance = float(self._opts['lottery_chance'])
        cfg.use_lottery = int(bool(self._opts['use_lottery']))

    def _check_nphotons(self, nphotons: int) -> int:
        nphotons = int(nphotons)
        if nphotons < 0:
            raise ValueError('The number of photon packets must not be '
                             'negative!')
        return nphotons

    def _run_flags(self, static_schedule: bool) -> int:
        flags = N.XO_FLAG_STATIC_SCHEDULE if static_schedule else 0
        if self._fast_math:
            flags |= N.XO_FLAG_FAST_MATH
        return flags

    def _record_report(self, rep: N.XoRunReport, nphotons: int):
        self._run_report.update(
            upload=None, download=None, execution=rep.kernel_ms * 1e-3,
            total=rep.total_ms * 1e-3, items=nphotons,
            threads=int(rep.threads_used), lanes=int(rep.lanes),
            steps=int(rep.steps), boundary_steps=int(rep.boundary_steps),
            voxel_crossings=int(rep.voxel_crossings),
            deposits=int(rep.deposits))
