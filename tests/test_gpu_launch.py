"""GPU parity of the GaussianBeam launch and the log-scale radial axis on
the cases the oracle is pinned with by closed forms
(tests/test_oracle_launch.py): the traced voxel kernel's launch states of an
oblique beam with unequal sigmas equal the oracle's bit for bit (static
schedule, the kernels' math on the host), and every known-radius escape of
the mirror slab lands in the geometric bin that holds it."""
import numpy as np
import pytest

from pyxopto_amd import _native as N
from oracle import pyoracle as O

import test_oracle_launch as L
from helpers import XINIT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n_in', [1.0, 1.4])
def test_gaussian_launch_states_bit_exact(n_in):
    print('case gaussian oblique n_in', n_in, flush=True)
    theta, tilt = np.deg2rad(30.0), np.deg2rad(20.0)
    direction = (np.sin(theta) * np.cos(tilt), np.sin(theta) * np.sin(tilt),
                 np.cos(theta))
    sim = L._beam_sim((40e-6, 90e-6), 2.0, (2e-4, 1e-4, 3e-4), direction,
                      n_out=1.0, n_in=n_in)
    cfg, lut, naccum = sim.pack()
    lanes, npk = 1024, 20000
    tr = N.XoTrace()
    tr.max_events, tr.options, tr.plon = 1, N.XO_TRACE_START, 0
    ctx = N.Context(0, 0, XINIT)
    acc, rep, data, n = ctx.run_mcvox(cfg, sim.voxels.material, lut, npk,
                                      naccum, lanes, N.XO_FLAG_STATIC_SCHEDULE,
                                      trace=tr)
    x, a = N.init_rng(lanes, XINIT)
    racc, _, _, rdata, rn = O.mcvox_run_trace(
        cfg, sim.voxels.material, lut, x, a, npk, naccum, tr, O.MATH_KERNEL)
    assert np.array_equal(n, rn)
    assert np.array_equal(data.view(np.uint32), rdata.view(np.uint32))
    assert np.array_equal(acc, racc)


def test_log_scale_radial_axis_known_radii_on_gpu():
    start, stop, nbins = 1e-5, 5e-3, 50
    edges = np.geomspace(start, stop, nbins + 1)
    D = L.mcml.mcdetector
    raxis = D.RadialAxis(start, stop, nbins, logscale=True)
    theta = np.deg2rad(30.0)
    ctx = N.Context(0, 0, XINIT)
    for i in (0, 7, 19, 30, 44, 49):
        r = np.sqrt(edges[i] * edges[i + 1])
        print('case log-axis bin', i, flush=True)
        sim = L._mirror_slab_sim(r / (2.0 * np.tan(theta)), theta, raxis)
        cfg, lut, naccum = sim.pack()
        acc, _ = ctx.run_mcml(cfg, lut, 2000, naccum)
        off, nb = sim.allocation(sim.detectors.top)
        assert np.nonzero(acc[off:off + nb])[0].tolist() == [i]
