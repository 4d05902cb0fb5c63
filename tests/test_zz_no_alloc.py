"""xo_batch_run_noalloc through the C ABI: the caller's device LUT and
accumulator, the library's static scratch array, and the same integers as
xo_batch_run on two calls in a row (the second with a smaller batch, fewer
bins and a shorter LUT, so stale scratch or accumulator contents would
show), with both arithmetic builds."""
import ctypes

import numpy as np
import pytest

from pyxopto_amd import _native as N
from pyxopto_amd.mcml.batch import Batch

from batch_helpers import small_batch

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceMemory:
    """Device arrays of the caller, from the HIP runtime the library uses
    (allocated outside the entry point under test)."""

    def __init__(self):
        self.hip = ctypes.CDLL('libamdhip64.so.7')
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError('{} failed: hipError {}'.format(what, rc))

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._ok(self.hip.hipMalloc(ctypes.byref(p), max(int(nbytes), 8)),
                 'hipMalloc')
        self.ptrs.append(p.value)
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self._ok(self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D),
                 'hipMemcpy')
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        self._ok(self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H),
                 'hipMemcpy')
        return out

    def fill(self, p, n, dtype, value):
        a = np.full(n, value, dtype=dtype)
        self._ok(self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D),
                 'hipMemcpy')

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.mark.parametrize('fast', [False, True])
def test_noalloc_batch_equals_batch_run_twice(fast):
    opts_fast = ['-cl-fast-relaxed-math'] if fast else None
    big = Batch(small_batch(40, seed=3), rnginit=0x5EED000000000001,
                cl_build_options=opts_fast)
    small = Batch(small_batch(9, seed=4, with_mhg=False),
                  rnginit=0x5EED000000000001, cl_build_options=opts_fast)
    ctx = big.context
    mem = DeviceMemory()
    try:
        arr1, lut1, nacc1 = big.pack()
        arr2, lut2, nacc2 = small.pack()
        assert nacc2 < nacc1 and len(lut2) < len(lut1)
        opts = big.opts(chunk=1024)
        lut1_d = mem.upload(np.asarray(lut1, dtype=np.float32))
        lut2_d = mem.upload(np.asarray(lut2, dtype=np.float32))
        acc_d = mem.alloc(8 * nacc1)
        # garbage in the caller's accumulator: the call must zero it
        mem.fill(acc_d, nacc1, np.uint64, 0xDEADBEEFDEADBEEF)
        for name, arr, lut, lut_d, nacc in (
                ('first (40 instances)', arr1, lut1, lut1_d, nacc1),
                ('second (9 instances)', arr2, lut2, lut2_d, nacc2)):
            print('case', name, 'fast', fast, flush=True)
            ref, rep_ref = ctx.run_batch(arr, lut, opts, nacc)
            rep = ctx.run_batch_noalloc(arr, lut_d, len(lut), opts, acc_d,
                                        nacc)
            got = mem.download(acc_d, nacc, np.uint64)
            np.testing.assert_array_equal(got, ref)
            assert rep.packets == rep_ref.packets
            assert rep.steps == rep_ref.steps
            assert rep.boundary_steps == rep_ref.boundary_steps
        # the bins past the second batch's accumulator are the caller's:
        # the second call wrote only its naccum bins
        rest = mem.download(acc_d, nacc1, np.uint64)[nacc2:]
        assert np.array_equal(rest, ref_first_tail(ctx, arr1, lut1, opts,
                                                   nacc1, nacc2))
    finally:
        mem.close()
        big.context.close()


def ref_first_tail(ctx, arr, lut, opts, nacc, start):
    acc, _ = ctx.run_batch(arr, lut, opts, nacc)
    return acc[start:]


def test_noalloc_rejects_a_batch_beyond_the_scratch_array():
    b = Batch(small_batch(2, seed=5, budget=(3_000_000, 3_000_001)),
              rnginit=0x5EED000000000001)
    arr, lut, nacc = b.pack()
    mem = DeviceMemory()
    try:
        lut_d = mem.upload(np.asarray(lut, dtype=np.float32))
        acc_d = mem.alloc(8 * nacc)
        # 6e6 one-packet items of 16 bytes: beyond the static array
        with pytest.raises(RuntimeError, match='scratch'):
            b.context.run_batch_noalloc(arr, lut_d, len(lut), b.opts(chunk=1),
                                        acc_d, nacc)
        # and the entry point still works afterwards
        ref, _ = b.context.run_batch(arr, lut, b.opts(chunk=4096), nacc)
        b.context.run_batch_noalloc(arr, lut_d, len(lut), b.opts(chunk=4096),
                                    acc_d, nacc)
        np.testing.assert_array_equal(mem.download(acc_d, nacc, np.uint64),
                                      ref)
    finally:
        mem.close()
        b.context.close()
