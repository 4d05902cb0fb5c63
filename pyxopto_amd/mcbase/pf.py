"""Host-side scattering phase functions and the Monte Carlo lookup-table
builder (xopto/pf/pfbase.py:102-567, xopto/pf/hg.py, xopto/pf/mhg.py).

``mclut`` follows the construction of the reference (pfbase.py:518-567):
the CDF of the phase function is integrated with Simpson's rule on a fine
grid (pfbase.py:439-452), the model F(x) = a/(x + b) + c with
a = 2c(1 - c), b = 2c - 1 (so that F(-1) = 0 and F(1) = 1) is fitted to it
by least squares on both admissible branches (c > 1 and c < 0, the better
of the two kept; a bounded Brent search on each instead of the reference's
L-BFGS-B: the same objective and optimum at about 1/100 of the time, which
matters for batches of thousands of tables), and the table holds the inverse CDF evaluated at the model F of
N uniformly spaced cosines. The sampler then draws
index = (a/(xi - c) - b + 1)(N - 1)/2 and interpolates linearly.
"""
from typing import Tuple

import numpy as np
from scipy.interpolate import interp1d
from scipy.optimize import minimize_scalar


def lut_function(costheta, params):
    """F(x) = a/(x + b) + c with b = 2c - 1, a = 2c(1 - c)."""
    c = np.asarray(params, dtype=np.float64).ravel()[0]
    b = 2.0 * c - 1.0
    a = 2.0 * c * (1.0 - c)
    return a / (b + np.asarray(costheta)) + c


class PfBase:
    """Base of the host phase functions: pdf over cos(theta) in [-1, 1]."""

    def __call__(self, costheta):
        raise NotImplementedError

    def g(self, n: int) -> float:
        raise NotImplementedError

    def cdf(self, costheta, npts: int = 10000) -> np.ndarray:
        """CDF at the given cosines by composite Simpson integration."""
        costheta = np.asarray(costheta, dtype=np.float64)
        npts = int(max(npts, costheta.size) / 2) * 2 + 1
        x = np.linspace(-1.0, 1.0, npts)
        p = self(x)
        dx = 2.0 / (npts - 1.0)
        cum = np.zeros(p.size // 2 + 1)
        cum[1:] = dx / 3.0 * (p[:-2:2] + 4.0 * p[1:-1:2] + p[2::2])
        cum = np.cumsum(cum)
        out = np.interp(costheta, x[::2], cum)
        out *= 1.0 / cum[-1]
        return out

    def mclut(self, n: int = 1000, ncd: int = 10000) \
            -> Tuple[list, np.ndarray]:
        """([a, b, c], lut[n]) for the Lut sampler."""
        g1 = self.g(1)
        ncd = max(10 * int(n), int(ncd))
        if g1 > 0.95:
            ncd = max(50000, ncd)
        costheta = np.linspace(-1.0, 1.0, n)
        costhetai = np.linspace(-1.0, 1.0, ncd)
        cumpfi = self.cdf(costhetai)

        # the residual norm without BLAS: the tables are built by pools of
        # worker processes, where a threaded BLAS dot in each oversubscribes
        # the host cores
        def cost(c):
            diff = lut_function(costhetai, c) - cumpfi
            return float(np.sqrt(np.sum(diff * diff)))

        # one free parameter on two branches, c > 1 and c < 0: bounded
        # Brent on a logarithmic scale of |c - 1| and |c|, the lower
        # residual kept
        r1 = minimize_scalar(lambda t: cost(1.0 + np.exp(t)),
                             bounds=(-30.0, 15.0), method='bounded',
                             options={'xatol': 1e-10})
        r2 = minimize_scalar(lambda t: cost(-np.exp(t)),
                             bounds=(-30.0, 15.0), method='bounded',
                             options={'xatol': 1e-10})
        res_x = np.array([1.0 + np.exp(r1.x) if r1.fun < r2.fun else -np.exp(r2.x)])
        randlut = np.clip(lut_function(costheta, res_x), 0.0, 1.0)
        lut = interp1d(cumpfi, costhetai, bounds_error=False, fill_value=0.0,
                       kind='linear')(randlut)
        c = float(res_x[0])
        return [2.0 * c * (1.0 - c), 2.0 * c - 1.0, c], lut


class Hg(PfBase):
    """Henyey-Greenstein pdf (xopto/pf/hg.py:28-75)."""

    def __init__(self, g: float):
        eps = np.finfo(np.float64).eps
        self._g = max(min(float(g), 1.0 - eps), -1.0 + eps)

    def __call__(self, costheta):
        g = self._g
        return 0.5 * (1.0 - g * g) / (1.0 + g * g - 2.0 * g *
                                      np.asarray(costheta)) ** 1.5

    def g(self, n: int) -> float:
        return self._g ** n

    def __repr__(self):
        return 'Hg({})'.format(self._g)


class MHg(PfBase):
    """Modified HG: b*HG(g) + (1 - b)*3/2 cos^2 (xopto/pf/mhg.py:28-122)."""

    def __init__(self, g: float, b: float):
        self._b = float(np.clip(float(b), 0.0, 1.0))
        self._hg = Hg(g)

    def __call__(self, costheta):
        costheta = np.asarray(costheta)
        return self._b * self._hg(costheta) + \
            (1.0 - self._b) * 1.5 * costheta ** 2

    def g(self, n: int) -> float:
        rayleigh = {0: 1.0, 1: 0.0, 2: 0.4}.get(int(n), 0.0)
        return self._b * self._hg.g(n) + (1.0 - self._b) * rayleigh

    def __repr__(self):
        return 'MHg({}, {})'.format(self._hg._g, self._b)
