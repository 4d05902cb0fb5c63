"""Batched engine, host side: per-packet generator states and packing
(no GPU needed)."""
import numpy as np

from pyxopto_amd import _native as N
from pyxopto_amd.mcml.batch import Batch
from oracle import pyoracle as O

from batch_helpers import small_batch
from helpers import batch_mean_se, z_scores


def test_packet_seed_states_are_valid_and_distinct():
    table = N.safeprime_table(N.XO_NUM_PRIMES)
    valid_a = set(int(v) for v in table[1:])
    states = set()
    for p in range(3000):
        x, a = N.packet_seed(123456789, p)
        c, lo = x >> 32, x & 0xFFFFFFFF
        assert a in valid_a
        assert 0 <= c < a - 1 and 0 < lo < 0xFFFFFFFF
        states.add((x, a))
    assert len(states) == 3000
    # a different key gives different streams
    assert N.packet_seed(123456790, 0) != N.packet_seed(123456789, 0)


def test_packing_allocates_contiguous_instance_ranges():
    insts = small_batch(6)
    b = Batch(insts, rnginit=0x5EED000000000001)
    arr, lut, naccum = b.pack()
    assert naccum == 6 * 102
    for i in range(6):
        assert arr[i].accum_base == 102 * i and arr[i].accum_count == 102
        assert arr[i].top.offset == 102 * i
        assert arr[i].bottom.offset == 102 * i + 100
        assert arr[i].specular.offset == 102 * i + 101
        assert arr[i].npackets == insts[i].nphotons
        assert arr[i].seed == insts[i].seed
    assert lut is not None and lut.size % 500 == 0


def test_batch_oracle_agrees_with_layered_oracle():
    """Per-packet generators (batch) and per-lane streams (Mc.run) sample
    the same physics: batch means of total diffuse reflectance agree."""
    insts = small_batch(3, seed=11, budget=(20000, 20001), with_mhg=False)
    b = Batch(insts, rnginit=0x5EED000000000001)
    arr, lut, naccum = b.pack()
    k = 40
    rd_b, rd_l = [], []
    for rep in range(k):
        for i in range(3):
            arr[i].seed = 1000 * rep + i + 1
        acc, _ = O.batch_run(arr, lut, naccum, math_mode=O.MATH_LIBM)
        rd_b.append([acc[102 * i:102 * i + 100].sum() / N.XO_ACCU_K / 20000
                     for i in range(3)])
    from pyxopto_amd.mcml import mc
    x, a = N.init_rng(512 * k, 99999)
    for rep in range(k):
        row = []
        for i in range(3):
            sim = mc.Mc(insts[i].layers, insts[i].source, insts[i].detectors,
                        rnginit=99999)
            cfg, lut_i, nacc = sim.pack()
            acc, _, _ = O.mcml_run(cfg, lut_i, x[512 * rep:512 * (rep + 1)],
                                   a[512 * rep:512 * (rep + 1)], 20000, nacc)
            row.append(acc[:100].sum() / N.XO_ACCU_K / 20000)
        rd_l.append(row)
    m1, s1 = batch_mean_se(np.asarray(rd_b))
    m2, s2 = batch_mean_se(np.asarray(rd_l))
    assert np.all(np.abs(z_scores(m1, s1, m2, s2)) < 4.0)


def test_batch_oracle_fluence_equals_keyed_layered_oracle():
    """A batch instance's fluence grid (oracle_batch_run) holds the integers
    of the layered oracle run with the same packet-keyed generators, the
    same layers, detectors and grid, for each stepping method and both
    fluence modes: the batch path adds no arithmetic of its own."""
    import pytest
    from pyxopto_amd.mcml import mc
    from pyxopto_amd.mcml.batch import Instance
    A = mc.mcfluence.Axis
    inst0 = small_batch(1, seed=31, budget=(3000, 3001))[0]
    for method in (0, 1, 2):
        for mode in ('deposition', 'fluence'):
            flu = mc.mcfluence.FluenceRz(A(0.0, 2e-3, 12), A(0.0, 2e-3, 10),
                                         mode=mode)
            inst = Instance(inst0.layers, inst0.source, inst0.detectors,
                            nphotons=3000, seed=987654321, fluence=flu)
            b = Batch([inst], options=[mc.mcoptions.McMethod(method)],
                      rnginit=0x5EED000000000001)
            arr, lut, naccum = b.pack()
            acc, counts = O.batch_run(arr, lut, naccum, method=method,
                                      math_mode=O.MATH_KERNEL, nthreads=4)
            sim = mc.Mc(inst0.layers, inst0.source, inst0.detectors,
                        fluence=flu, options=[mc.mcoptions.McMethod(method)],
                        rnginit=0x5EED000000000001)
            cfg, lut2, naccum2 = sim.pack()
            assert naccum2 == naccum
            cfg.packet_seed = 987654321
            cfg.first_packet = 0
            x, a = N.init_rng(64, 0x5EED000000000001)
            ref, rc, _ = O.mcml_run(cfg, lut2, x, a, 3000, naccum,
                                    O.MATH_KERNEL, keyed=True)
            np.testing.assert_array_equal(acc, ref)
            assert int(counts[3]) == int(rc[3]) > 0
            assert acc[cfg.fluence.offset:].sum() > 0
