"""The atomic-rate probes behind the fluence accumulator layout: every
atomic lands (the sums are exact) for a 64 MiB target, a 256 KiB target and
one copy per XCD written with workgroup-scope atomics by the waves of that
XCD, and the three rates are within a factor of two of each other (the
device's u64 atomic rate depends on neither the footprint nor the scope)."""
import pytest

from pyxopto_amd import _native as N

pytestmark = pytest.mark.gpu


def test_atomic_probes_are_exact_and_rates_agree():
    ctx = N.Context(0, 1, 0x5EED000000000001)
    n = 1 << 26
    ms_all, s_all = ctx.probe_atomic_u64(64 << 20, n, seed=3)
    ms_hot, s_hot = ctx.probe_atomic_u64_xcd(64 << 20, 256 << 10, n, False,
                                             seed=3)
    ms_xcd, s_xcd = ctx.probe_atomic_u64_xcd(64 << 20, 64 << 20, n, True,
                                             seed=3)
    print('atomics/s: 64 MiB {:.3g}, 256 KiB {:.3g}, per-XCD copies '
          '{:.3g}'.format(n / ms_all * 1e3, n / ms_hot * 1e3,
                          n / ms_xcd * 1e3), flush=True)
    assert s_all == n and s_hot == n and s_xcd == n
    for ms in (ms_hot, ms_xcd):
        assert 0.5 < ms / ms_all < 2.0
    ctx.close()
