"""Photon packet sources of the voxelized simulator: Line
(xopto/mcvox/mcsource/line.py:30-255) and GaussianBeam
(xopto/mcvox/mcsource/gaussianbeam.py:33-353)."""
from typing import Tuple

import numpy as np

from .. import _native as N
from ..mcbase.mcutil import boundary, geometry


def _unit(v):
    v = np.asarray(v, dtype=np.float64).copy()
    norm = np.linalg.norm(v)
    if norm == 0.0:
        raise ValueError('The norm/length of the propagation direction '
                         'vector must not be 0!')
    return v / norm


def _pop_type(cls, data: dict) -> dict:
    data = dict(data)
    kind = data.pop('type')
    if kind != cls.__name__:
        raise TypeError('Expected "{}" type but got "{}"!'.format(
            cls.__name__, kind))
    return data


class Line:
    kind = N.XO_SRC_LINE

    def __init__(self, position: Tuple[float, float, float] = (0.0, 0.0, 0.0),
                 direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)):
        self._position = np.asarray(position, dtype=np.float64).copy()
        self._direction = _unit(direction)

    position = property(lambda self: self._position)
    direction = property(lambda self: self._direction)

    def pack(self, mc, cfg: N.XoMcvoxCfg):
        """Entry point on the box, specular reflectance and refraction
        computed on the host (line.py:181-237)."""
        voxels = mc.voxels
        search = -self._direction if voxels.contains(self._position) \
            else self._direction
        position, normal = voxels.intersect(self._position, search)
        if position is None:
            raise ValueError('The line source does not intersect the '
                             'voxelized sample!')
        n_medium = mc.materials[0].n
        n_sample = mc.material_at(position).n
        refl = boundary.reflectance(n_medium, n_sample,
                                    abs(float(np.dot(normal,
                                                     self._direction))))
        if refl >= 1.0:
            raise ValueError('The line source is fully reflected from the '
                             'surface of the voxelized sample!')
        ds = boundary.refract(self._direction, normal, n_medium, n_sample)
        dr = boundary.reflect(self._direction, normal)
        t = cfg.line
        for k in range(3):
            t.position[k] = position[k]
            t.direction_medium[k] = self._direction[k]
            t.direction_sample[k] = ds[k]
            t.direction_reflected[k] = dr[k]
        t.reflectance = refl
        cfg.source_kind = self.kind
        return cfg

    def todict(self):
        """mcvox/mcsource/line.py:244-250"""
        return {'position': self._position.tolist(),
                'direction': self._direction.tolist(), 'type': 'Line'}

    @classmethod
    def fromdict(cls, data: dict) -> 'Line':
        return cls(**_pop_type(cls, data))


class GaussianBeam:
    kind = N.XO_SRC_GAUSSIAN

    def __init__(self, sigma, clip: float = 5.0,
                 position: Tuple[float, float, float] = (0.0, 0.0, 0.0),
                 direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)):
        s = np.zeros(2)
        s[:] = sigma
        self._sigma = s
        self._clip = float(clip)
        if self._clip < 0.0:
            raise ValueError('Clip diameter/sigma must be greater than zero!.')
        self._position = np.asarray(position, dtype=np.float64).copy()
        self._direction = _unit(direction)

    sigma = property(lambda self: self._sigma)
    clip = property(lambda self: self._clip)
    position = property(lambda self: self._position)
    direction = property(lambda self: self._direction)

    def pack(self, mc, cfg: N.XoMcvoxCfg):
        """gaussianbeam.py:299-338"""
        T = geometry.transform_base((0.0, 0.0, 1.0), self._direction)
        g = cfg.gaussian
        for i in range(9):
            g.transformation[i] = T.flat[i]
        for k in range(3):
            g.position[k] = self._position[k]
            g.direction[k] = self._direction[k]
        g.sigma[0], g.sigma[1] = self._sigma
        g.clip = self._clip
        cfg.source_kind = self.kind
        return cfg

    def todict(self):
        """gaussianbeam.py:340-349"""
        return {'sigma': self._sigma.tolist(), 'clip': self._clip,
                'position': self._position.tolist(),
                'direction': self._direction.tolist(), 'type': 'GaussianBeam'}

    @classmethod
    def fromdict(cls, data: dict) -> 'GaussianBeam':
        return cls(**_pop_type(cls, data))


def fromdict(data: dict):
    """Line or GaussianBeam from its todict() data."""
    kinds = {'Line': Line, 'GaussianBeam': GaussianBeam}
    if data.get('type') not in kinds:
        raise TypeError('Unsupported source type "{}"!'.format(
            data.get('type')))
    return kinds[data['type']].fromdict(data)
