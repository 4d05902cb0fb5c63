"""Seed table, init_RNG and the MWC generator, pinned against the
reference's own data and code (tests/golden/seeds.npz, written by
tests/golden/make_fixtures.py from xopto/data/primes/ and from the
reference's rng.cpp compiled into oracle/_ref/)."""
import hashlib
import os

import numpy as np
import pytest

from pyxopto_amd import _native as N
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'seeds.npz')


@pytest.fixture(scope='module')
def gold():
    return np.load(GOLD, allow_pickle=False)


def test_safeprime_table_matches_reference(gold):
    table = N.safeprime_table(N.XO_NUM_PRIMES)
    head = gold['table_head']
    assert int(gold['table_rows']) == N.XO_NUM_PRIMES
    np.testing.assert_array_equal(table[:head.shape[0]].astype(np.uint64),
                                  head[:, 0])
    # second column of the reference table is a*2^32 - 1
    np.testing.assert_array_equal(
        (head[:, 0] << np.uint64(32)) - np.uint64(1), head[:, 1])
    sha = hashlib.sha256(table.astype('<u4').tobytes()).hexdigest()
    assert sha == str(gold['fora_sha256'])


def _is_prime(n: int) -> bool:
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = pow(x, 2, n)
            if x == n - 1:
                break
        else:
            return False
    return True


def test_safeprime_definition_brute_force():
    """Every multiplier between consecutive table entries is rejected and
    every entry is accepted (first 40 rows, exhaustive)."""
    table = N.safeprime_table(40)
    a = int(table[0])
    for expected in table:
        while True:
            n2 = (a << 32) - 1
            ok = _is_prime(n2) and _is_prime((n2 - 1) // 2)
            if ok:
                break
            a -= 1
        assert a == int(expected)
        a -= 1


def test_init_rng_matches_reference_rng_cpp(gold):
    for i, xinit in enumerate(gold['xinits']):
        x, a = N.init_rng(gold['x'].shape[1], int(xinit))
        np.testing.assert_array_equal(x, gold['x'][i])
        np.testing.assert_array_equal(a, gold['a'][i])


def test_init_rng_invalid_initializer(gold):
    assert int(gold['bad_rc']) == 1
    with pytest.raises(ValueError, match='not valid'):
        N.init_rng(1, int(gold['bad_xinit']))
    with pytest.raises(ValueError):
        N.init_rng(1, 0)


def test_init_rng_states_are_valid():
    x, a = N.init_rng(20000, 987654321987)
    c = x >> np.uint64(32)
    lo = x & np.uint64(0xFFFFFFFF)
    assert np.all(x != 0)
    assert np.all(c < (a.astype(np.uint64) - np.uint64(1)))
    assert np.all(lo < np.uint64(0xFFFFFFFF))


def test_init_rng_prefix_property():
    """Stream i does not depend on how many streams are requested, so the
    first L rows of a context equal init_rng(L)."""
    x1, a1 = N.init_rng(1000, 5555555555)
    x2, a2 = N.init_rng(3000, 5555555555)
    np.testing.assert_array_equal(x1, x2[:1000])
    np.testing.assert_array_equal(a1, a2[:1000])


def test_mwc_recurrence_matches_reference_rand_cc(gold):
    """rand_cc of the reference library performs the same MWC update; its
    output is the low 23 bits of x scaled by 1/0x7FFFFF."""
    x0, a0 = int(gold['x'][0][0]), int(gold['a'][0][0])
    vals, x_after = N.rng_host(x0, a0, 1000)
    # reconstruct the integer sequence and compare the low 23 bits
    xs = []
    x = x0
    for _ in range(1000):
        x = (x & 0xFFFFFFFF) * a0 + (x >> 32)
        xs.append(x)
    assert x_after == x == int(gold['rand_cc_x_after'])
    low = np.array([v & 0x7FFFFF for v in xs], dtype=np.float32)
    np.testing.assert_array_equal(low / np.float32(0x7FFFFF),
                                  gold['rand_cc'])
    # fp_random_single: float(uint32 x) / 2^32, an exact power-of-two scale
    expect = np.array([np.float32(np.uint32(v & 0xFFFFFFFF)) for v in xs],
                      dtype=np.float32) * np.float32(2.0 ** -32)
    np.testing.assert_array_equal(vals, expect)
    assert np.all((vals >= 0.0) & (vals <= 1.0))


def test_oracle_rng_equals_product_host_rng(gold):
    x0, a0 = int(gold['x'][1][7]), int(gold['a'][1][7])
    v1, x1 = N.rng_host(x0, a0, 5000)
    v2, x2 = O.rng(x0, a0, 5000)
    np.testing.assert_array_equal(v1, v2)
    assert x1 == x2


def test_reference_rng_library_agrees_when_built(gold):
    ref = O.ref_rng_lib()
    if ref is None:
        pytest.skip('oracle/_ref/librng_ref.so not built (reference tree '
                    'absent on this machine)')
    import ctypes
    fora = np.ascontiguousarray(N.safeprime_table(5001))
    x = np.zeros(5000, np.uint64)
    a = np.zeros(5000, np.uint32)
    rc = ref.init_RNG(x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                      a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                      fora.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                      5000, 31415926535)
    assert rc == 0
    x2, a2 = N.init_rng(5000, 31415926535)
    np.testing.assert_array_equal(x, x2)
    np.testing.assert_array_equal(a, a2)
