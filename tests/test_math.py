"""Accuracy of the kernels' fixed-sequence math (pyxopto_amd/csrc/xo_math.h)
against float64 references, evaluated through the host build of the same
functions (the GPU build rounds identically, which the bit-exact GPU tests
confirm)."""
import numpy as np

from oracle import pyoracle as O


def _ulp(x):
    x = np.abs(np.asarray(x, dtype=np.float32))
    return (np.nextafter(x, np.float32(np.inf)) - x).astype(np.float64)


def _rng_floats(n, seed=1):
    u = np.random.default_rng(seed).integers(0, 2 ** 32, n, dtype=np.uint64)
    return (u.astype(np.uint32).astype(np.float32) *
            np.float32(2.0 ** -32))


def test_logf_accuracy():
    x = _rng_floats(2_000_000)
    x = x[x > 0]
    x = np.concatenate([x, np.float32([2.0 ** -32, 0.5, 0.70710677, 1.0,
                                       1.5, 2.0, 1e-38, 1e-45, 3e38])])
    y = O.kernel_math('log', x)
    ref = np.log(x.astype(np.float64))
    err = np.abs(y - ref) / _ulp(ref.astype(np.float32))
    assert err.max() <= 2.0, err.max()
    assert O.kernel_math('log', np.float32([0.0]))[0] == -np.inf
    assert np.isnan(O.kernel_math('log', np.float32([-1.0]))[0])
    assert O.kernel_math('log', np.float32([1.0]))[0] == 0.0


def test_sincos_turns_accuracy_and_identities():
    t = np.concatenate([_rng_floats(2_000_000, 2),
                        np.float32([0.0, 0.125, 0.25, 0.5, 0.75, 1.0])])
    s, c = O.kernel_math('sincos', t)
    a = 2.0 * np.pi * t.astype(np.float64)
    assert np.max(np.abs(s - np.sin(a))) < 2e-7
    assert np.max(np.abs(c - np.cos(a))) < 2e-7
    # quadrant points are exact
    s4, c4 = O.kernel_math('sincos', np.float32([0.0, 0.25, 0.5, 0.75]))
    np.testing.assert_array_equal(s4, np.float32([0, 1, 0, -1]))
    np.testing.assert_array_equal(c4, np.float32([1, 0, -1, 0]))
    assert np.max(np.abs(s * s + c * c - 1.0)) < 5e-7


def test_cbrtf_accuracy():
    u = _rng_floats(2_000_000, 3)
    x = (np.float32(2.0) * u - np.float32(1.0))
    x = np.concatenate([x, np.float32([1e-40, -1e-40, 1.0, -1.0, 8.0, 27.0])])
    y = O.kernel_math('cbrt', x)
    ref = np.cbrt(x.astype(np.float64))
    nz = ref != 0
    err = np.abs(y[nz] - ref[nz]) / _ulp(ref[nz].astype(np.float32))
    assert err.max() <= 2.0, err.max()
    assert O.kernel_math('cbrt', np.float32([0.0]))[0] == 0.0
    np.testing.assert_array_equal(O.kernel_math('cbrt', np.float32([8.0, 27.0])),
                                  np.float32([2.0, 3.0]))


def test_expf_accuracy():
    rng = np.random.default_rng(4)
    x = np.concatenate([
        rng.uniform(-103.0, 88.0, 1_000_000),
        -rng.exponential(3.0, 1_000_000),          # exp(-mua * step)
        [0.0, -0.0, 1.0, -1.0, 88.72, -87.3, -100.0, -103.9, 1e-8, -1e-8]
    ]).astype(np.float32)
    y = O.kernel_math('exp', x)
    ref = np.exp(x.astype(np.float64))
    normal = ref >= 2.0 ** -126
    err = np.abs(y[normal] - ref[normal]) / _ulp(ref[normal].astype(np.float32))
    assert err.max() <= 2.0, err.max()
    # subnormal results: within one subnormal step
    sub = ~normal
    assert np.all(np.abs(y[sub] - ref[sub]) <= 2.0 ** -149 * 1.01)
    e = O.kernel_math('exp', np.float32([0.0, np.inf, -np.inf, 89.0, -104.0]))
    np.testing.assert_array_equal(e, np.float32([1.0, np.inf, 0.0, np.inf, 0.0]))
    assert np.isnan(O.kernel_math('exp', np.float32([np.nan]))[0])


def test_log_unit_equals_log_on_its_whole_domain():
    """xo_logf_unit (the log of uniform deviates in the transport loops)
    drops xo_logf's infinity and subnormal branches: on +0 and every normal
    float <= 1 it returns the same bits."""
    lo, hi = 0x00800000, 0x3F800000          # 2^-126 .. 1.0
    step = 1 << 24
    for start in range(lo, hi + 1, step):
        u = np.arange(start, min(start + step, hi + 1), dtype=np.uint32)
        x = u.view(np.float32)
        a = O.kernel_math('log', x)
        b = O.kernel_math('log_unit', x)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), start
    z = np.float32([0.0])
    assert O.kernel_math('log_unit', z)[0] == O.kernel_math('log', z)[0]
