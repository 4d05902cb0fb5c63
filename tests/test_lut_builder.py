"""The Monte Carlo lookup-table builder (`PfBase.mclut`, the reference's
xopto/pf/pfbase.py:439-452 CDF and :518-567 fit and table) pinned by
closed forms:

- the Simpson CDF of the Henyey-Greenstein function against its closed
  form F(x) = (1 - g^2)/(2g) [(1 + g^2 - 2gx)^-1/2 - 1/(1 + g)];
- a phase function whose CDF is the LUT model itself, F(x) = a/(x + b) + c
  with a = 2c(1 - c), b = 2c - 1, on either branch (c > 1, c < 0): the fit
  must return that c and the table must be the identity on [-1, 1];
- the fitted c is the least-squares optimum of the model over both
  branches (a dense scan of c cannot do better);
- the table inverts the closed-form CDF at the model's quantiles.

A wrong sign of b, the other branch, a dropped normalization or swapped
interpolation axes each fail one of these. The reference's pfbase does not
import here (scipy 1.15 has no scipy.integrate.simps), so no vectors of its
own fit exist: these checks stand in for them.
"""
import numpy as np
import pytest

from pyxopto_amd.mcbase import pf as hostpf


def _hg_cdf(x, g):
    x = np.asarray(x, dtype=np.float64)
    return (1.0 - g * g) / (2.0 * g) * (1.0 / np.sqrt(1.0 + g * g - 2.0 * g * x)
                                        - 1.0 / (1.0 + g))


class _ModelPf(hostpf.PfBase):
    """pdf -a/(x + b)^2, whose CDF on [-1, 1] is a/(x + b) + c exactly."""

    def __init__(self, c):
        self.c = float(c)
        self.a = 2.0 * self.c * (1.0 - self.c)
        self.b = 2.0 * self.c - 1.0

    def __call__(self, costheta):
        return -self.a / (np.asarray(costheta, dtype=np.float64) + self.b) ** 2

    def g(self, n):
        x = np.linspace(-1.0, 1.0, 200001)
        return float(np.trapezoid(self(x) * x ** n, x))


@pytest.mark.parametrize('g,atol', [(0.8, 1e-9), (-0.3, 1e-12), (0.5, 1e-12),
                                    (0.95, 5e-5)])
def test_simpson_cdf_equals_the_hg_closed_form(g, atol):
    # Simpson's error at step 2e-4 grows as the peak at cos = 1 narrows
    # (width about (1 - g)^2 / 2g: 1e-5 at g = 0.95)
    # the 5001 Simpson nodes of the default 10001-point rule, where the
    # CDF is the cumulative rule itself (no interpolation)
    x = np.linspace(-1.0, 1.0, 5001)
    cdf = hostpf.Hg(g).cdf(x)
    assert cdf[0] == 0.0
    assert abs(cdf[-1] - 1.0) < 1e-15
    np.testing.assert_allclose(cdf, _hg_cdf(x, g), rtol=0, atol=atol)


@pytest.mark.parametrize('c0', [1.5, 1.05, -0.7, -3.0])
def test_fit_recovers_an_exact_model_and_the_table_is_the_identity(c0):
    pf = _ModelPf(c0)
    x = np.linspace(-1.0, 1.0, 11)
    np.testing.assert_allclose(pf.cdf(x), pf.a / (x + pf.b) + pf.c, atol=1e-9)
    (a, b, c), lut = pf.mclut(n=500)
    assert c == pytest.approx(c0, rel=1e-4)
    assert a == 2.0 * c * (1.0 - c) and b == 2.0 * c - 1.0
    np.testing.assert_allclose(lut, np.linspace(-1.0, 1.0, 500), atol=2e-5)


@pytest.mark.parametrize('pf', [hostpf.Hg(0.8), hostpf.Hg(-0.3),
                                hostpf.MHg(0.9, 0.5), hostpf.MHg(0.2, 0.3)],
                         ids=repr)
def test_fit_is_the_least_squares_optimum_over_both_branches(pf):
    (a, b, c), _ = pf.mclut(n=1000)
    xi = np.linspace(-1.0, 1.0, 10000)
    cdf = pf.cdf(xi)

    def cost(cc):
        return np.linalg.norm(hostpf.lut_function(xi, [cc]) - cdf)

    t = np.linspace(-12.0, 8.0, 4001)
    scan = [cost(1.0 + np.exp(v)) for v in t] + [cost(-np.exp(v)) for v in t]
    assert cost(c) <= min(scan) * (1.0 + 1e-6)
    # the model's end points: F(-1) = 0 and F(1) = 1
    assert abs(a / (b - 1.0) + c) < 1e-9 * max(1.0, abs(c))
    assert abs(a / (b + 1.0) + c - 1.0) < 1e-9 * max(1.0, abs(c))


@pytest.mark.parametrize('g', [0.8, -0.3, 0.9])
def test_table_inverts_the_closed_form_cdf(g):
    n = 1000
    (a, b, c), lut = hostpf.Hg(g).mclut(n=n)
    assert abs(lut[0] + 1.0) < 1e-12 and abs(lut[-1] - 1.0) < 1e-12
    assert np.all(np.diff(lut) >= 0.0)
    q = np.clip(hostpf.lut_function(np.linspace(-1.0, 1.0, n), [c]), 0.0, 1.0)
    # linear interpolation of the inverse CDF on 10000 points: largest
    # where the CDF is steepest (next to cos = 1 for g > 0)
    err = np.abs(_hg_cdf(lut, g) - q)
    assert err.max() < 1e-3
    assert np.median(err) < 1e-6
