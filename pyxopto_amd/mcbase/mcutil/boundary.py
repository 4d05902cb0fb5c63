"""Host-side boundary physics (xopto/mcbase/mcutil/boundary.py:25-150)."""
import numpy as np


def cos_critical(n1: float, n2: float) -> float:
    """Cosine of the critical angle for the boundary n1 -> n2 (0 if none)."""
    cc = 0.0
    if n1 > n2:
        cc = (1.0 - (n2 / n1) ** 2) ** 0.5
    return cc


def reflectance(n1: float, n2: float, costheta: float = 1.0) -> float:
    """Fresnel reflectance of unpolarized light (float64)."""
    if costheta < 0.0:
        raise ValueError('The incidence angle cosine must not be negative!')
    n1, n2 = float(n1), float(n2)
    sintheta = (1.0 - costheta ** 2) ** 0.5
    if n1 > n2 and sintheta >= n2 / n1:
        return 1.0
    a1 = n1 * costheta
    a2 = n2 * (1.0 - (n1 / n2 * sintheta) ** 2) ** 0.5
    rs = abs((a1 - a2) / (a1 + a2)) ** 2
    b1 = n1 * (1.0 - (n1 / n2 * sintheta) ** 2) ** 0.5
    b2 = n2 * costheta
    rp = abs((b1 - b2) / (b1 + b2)) ** 2
    return 0.5 * (rs + rp)


def refract(direction, normal, n1: float, n2: float) -> np.ndarray:
    """Refracted direction for a normal pointing either way
    (xopto/mcbase/mcutil/boundary.py:80-117)."""
    p = np.asarray(direction, dtype=np.float64)
    p = p / np.linalg.norm(p)
    n = np.asarray(normal, dtype=np.float64)
    n = n / np.linalg.norm(n)
    cos1 = float(np.dot(p, n))
    if abs(cos1) < cos_critical(n1, n2):
        raise ValueError('Cannot refract, the incidence angle exceeds the '
                         'critical angle!')
    k_n = n1 / n2
    sin2_sq = k_n * k_n * (1.0 - cos1 * cos1)
    k = k_n * abs(cos1) - np.sqrt(1.0 - sin2_sq)
    return k_n * p - np.sign(cos1) * k * n


def reflect(direction, normal) -> np.ndarray:
    """Mirror direction about the boundary (boundary.py:119-150)."""
    p = np.asarray(direction, dtype=np.float64)
    if np.linalg.norm(p) > 0.0:
        p = p / np.linalg.norm(p)
    n = np.asarray(normal, dtype=np.float64)
    n = n / np.linalg.norm(n)
    return p - 2.0 * np.dot(p, n) * n
