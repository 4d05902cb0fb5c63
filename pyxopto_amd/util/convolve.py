"""Reflectance collected by optical fibers from a radially symmetric
reflectance profile (xopto/util/convolve.py:27-131).

For a fiber of core radius r_c whose centre is at distance d from the
source, the collected reflectance is the integral of R(r) over the part of
each ring of radius r that lies inside the core:

    2 int_{max(d - r_c, 0)}^{d + r_c} arccos((r^2 + d^2 - r_c^2) / (2 r d)) R(r) r dr

(2 pi int_0^{r_c} R(r) r dr for d = 0), evaluated with Simpson's rule on an
odd number of evenly spaced points, R linearly interpolated (optionally in
log space) and extrapolated beyond the given points.
"""
import numpy as np
from scipy.integrate import simpson
from scipy.interpolate import interp1d


def fiber_reflectance(r, reflectance, sds, dcore: float, nsimps: int = 1000,
                      loginterp: bool = False, uneven: bool = False, **kwargs):
    """Reflectance through fibers at separations sds (one value or a
    sequence); reflectance is one profile over r or a 2D array of profiles
    in rows. Returns an array (profiles, len(sds))."""
    if isinstance(sds, float):
        sds = (sds,)
    rfib = dcore * 0.5
    nsimps = max(nsimps // 2 * 2 + 1, 3)
    reflectance = np.asarray(reflectance)
    nprof = reflectance.shape[0] if reflectance.ndim > 1 else 1
    values = np.log(reflectance) if loginterp else reflectance
    f = interp1d(r, values, assume_sorted=True, bounds_error=False,
                 fill_value='extrapolate', **kwargs)
    out = np.zeros((nprof, len(sds)))
    for index, d in enumerate(sds):
        rs = np.linspace(max(d - rfib, 0.0), max(d + rfib, 0.0), nsimps)
        x = rs if uneven else None
        dx = None if uneven else rs[1] - rs[0]
        fs = np.exp(f(rs)) if loginterp else f(rs)
        if d == 0.0:
            out[:, index] = 2.0 * np.pi * simpson(fs * rs, x=x, dx=dx if dx
                                                  is not None else 1.0)
        else:
            drs = rs * d
            drs[drs == 0.0] = np.finfo(np.float64).tiny
            kernel = np.arccos(np.clip((rs ** 2 + d ** 2 - rfib ** 2) /
                                       (2.0 * drs), -1.0, 1.0))
            out[:, index] = 2.0 * simpson(kernel * fs * rs, x=x,
                                          dx=dx if dx is not None else 1.0)
    return out
