"""Photon packet sources of the layered simulator: Line
(xopto/mcml/mcsource/line.py:30-252), UniformFiberNI
(xopto/mcml/mcsource/fiberni.py:180-420) and UniformFiber
(xopto/mcml/mcsource/fiber.py:190-497)."""
from typing import Tuple

import numpy as np

from .. import _native as N
from ..mcbase.mcutil import boundary, geometry
from ..mcbase.mcutil.fiber import MultimodeFiber


class Line:
    kind = N.XO_SRC_LINE

    def __init__(self, position: Tuple[float, float, float] = (0.0, 0.0, 0.0),
                 direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)):
        self._position = np.zeros(3)
        self._direction = np.zeros(3)
        self.position = position
        self.direction = direction

    def _set_position(self, v):
        self._position[:] = v

    def _set_direction(self, v):
        self._direction[:] = v
        norm = np.linalg.norm(self._direction)
        if norm == 0.0:
            raise ValueError('The norm/length of the propagation direction '
                             'vector must not be 0!')
        self._direction /= norm

    position = property(lambda self: self._position, _set_position, None,
                        'Source position.')
    direction = property(lambda self: self._direction, _set_direction, None,
                         'Source direction.')

    def pack_cfg(self, mc, cfg):
        """Fill the source fields of an xo_mcml_cfg or xo_batch_instance."""
        cfg.source_kind = self.kind
        self.pack(mc, cfg.source)
        return cfg

    def pack(self, mc, target: N.XoSourceLine) -> N.XoSourceLine:
        """Project to the sample surface, refract, specular reflectance
        (line.py:181-237)."""
        d = self._direction
        t = -self._position[2] / d[2]
        position = self._position + d * t
        position[2] = 0.0
        n1 = mc.layer(0).n
        n2 = mc.layer(1).n
        reflectance = boundary.reflectance(n1, n2, d[2])
        if reflectance >= 1.0:
            raise ValueError('The line source is fully reflected from the '
                             'top surface of the sample!')
        sin1 = (1.0 - d[2] ** 2) ** 0.5
        sin2 = max(min(n1 * sin1 / n2, 1.0), -1.0)
        ds = (d[0] * n1 / n2, d[1] * n1 / n2,
              np.sign(d[2]) * (1.0 - sin2 ** 2) ** 0.5)
        for i in range(3):
            target.position[i] = position[i]
            target.direction_medium[i] = d[i]
            target.direction_sample[i] = ds[i]
        target.direction_reflected[0] = d[0]
        target.direction_reflected[1] = d[1]
        target.direction_reflected[2] = -d[2]
        target.reflectance = reflectance
        return target

    def todict(self):
        """mcml/mcsource/line.py:239-245"""
        return {'position': self._position.tolist(),
                'direction': self._direction.tolist(), 'type': 'Line'}

    @classmethod
    def fromdict(cls, data: dict) -> 'Line':
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError('Expected "{}" type but got "{}"!'.format(
                cls.__name__, kind))
        return cls(**data)

    def __repr__(self):
        return 'Line(position=({}, {}, {}), direction=({}, {}, {}))'.format(
            *self._position, *self._direction)


class UniformFiberNI:
    """A multimode fiber in contact with the sample (normal incidence):
    packets start uniformly over the core with cos(theta) uniform over the
    acceptance cone, refracted into the sample; the Fresnel-reflected part
    goes to the specular detector (xopto/mcml/mcsource/fiberni.py:180-420).
    """
    kind = N.XO_SRC_FIBER_NI

    def __init__(self, fiber: MultimodeFiber,
                 position: Tuple[float, float, float] = (0.0, 0.0, 0.0)):
        self._fiber = fiber
        self._position = np.zeros(3)
        self.position = position

    def _set_position(self, v):
        self._position[:] = v
        self._position[2] = 0.0

    def _set_fiber(self, fib):
        self._fiber = fib

    fiber = property(lambda self: self._fiber, _set_fiber, None,
                     'Multimode optical fiber.')
    position = property(lambda self: self._position, _set_position, None,
                        'Source position (z is always 0).')

    def pack_cfg(self, mc, cfg):
        """cl_pack of fiberni.py:341-375 into cfg.fiber."""
        cfg.source_kind = self.kind
        f = cfg.fiber
        f.n = self._fiber.ncore
        f.cos_min = (1.0 - self._fiber.na ** 2) ** 0.5
        for i in range(3):
            f.position[i] = self._position[i]
        f.radius = self._fiber.dcore * 0.5
        return cfg

    def update(self, other):
        if isinstance(other, UniformFiberNI):
            self.fiber = other.fiber
            self.position = other.position
        elif isinstance(other, dict):
            self.fiber = other.get('fiber', self.fiber)
            self.position = other.get('position', self.position)

    def todict(self):
        return {'fiber': self._fiber.todict(),
                'position': self._position.tolist(),
                'type': 'UniformFiberNI'}

    @classmethod
    def fromdict(cls, data: dict) -> 'UniformFiberNI':
        """fiberni.py:406-420"""
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError('Expected "{}" type but got "{}"!'.format(
                cls.__name__, kind))
        fiber = MultimodeFiber.fromdict(data.pop('fiber'))
        return cls(fiber, **data)

    def __repr__(self):
        return 'UniformFiberNI(fiber={}, position=({}, {}, {}))'.format(
            self._fiber, *self._position)


class UniformFiber(UniformFiberNI):
    """A multimode fiber in contact with the sample along any direction
    with z > 0, its tip cut parallel to the surface (an elliptical contact
    area); launch angles uniform in cos(theta) over the acceptance cone,
    inside the core (xopto/mcml/mcsource/fiber.py:190-497). As in the
    reference, the launch direction is not refracted into the sample (its
    refraction test compares the surface point's z, which is 0, with the
    critical cosine, fiber.py:305-309); the Fresnel reflectance at that
    direction goes to the specular detector."""
    kind = N.XO_SRC_FIBER

    def __init__(self, fiber: MultimodeFiber,
                 position: Tuple[float, float, float] = (0.0, 0.0, 0.0),
                 direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)):
        super().__init__(fiber, position)
        self._direction = np.array([0.0, 0.0, 1.0])
        self.direction = direction

    def _set_direction(self, v):
        v = np.asarray(v, dtype=np.float64)
        norm = float(np.linalg.norm(v))
        if norm == 0.0:
            raise ValueError('The direction vector is singular!')
        v = v / norm
        if v[2] <= 0.0:
            raise ValueError('Z component of the propagation direction '
                             'must be positive!')
        self._direction[:] = v

    direction = property(lambda self: self._direction, _set_direction, None,
                         'Fiber direction (z > 0).')

    def pack_cfg(self, mc, cfg):
        """cl_pack of fiber.py:429-465 into cfg.fiber and cfg.fiber_axis."""
        super().pack_cfg(mc, cfg)
        cfg.source_kind = self.kind
        T = geometry.transform_base((0.0, 0.0, 1.0), self._direction)
        ax = cfg.fiber_axis
        for i in range(9):
            ax.transformation[i] = T.flat[i]
        for i in range(3):
            ax.direction[i] = self._direction[i]
        return cfg

    def update(self, other):
        if isinstance(other, UniformFiber):
            self.fiber = other.fiber
            self.position = other.position
            self.direction = other.direction
        elif isinstance(other, dict):
            self.fiber = other.get('fiber', self.fiber)
            self.position = other.get('position', self.position)
            self.direction = other.get('direction', self.direction)

    def todict(self):
        """fiber.py:466-474"""
        return {'fiber': self._fiber.todict(),
                'position': self._position.tolist(),
                'direction': self._direction.tolist(),
                'type': 'UniformFiber'}

    def __repr__(self):
        return 'UniformFiber(fiber={}, position=({}, {}, {}), ' \
               'direction=({}, {}, {}))'.format(self._fiber, *self._position,
                                                *self._direction)


def fromdict(data: dict):
    """Line, UniformFiberNI or UniformFiber from its todict() data."""
    kinds = {'Line': Line, 'UniformFiberNI': UniformFiberNI,
             'UniformFiber': UniformFiber}
    if data.get('type') not in kinds:
        raise TypeError('Unsupported source type "{}"!'.format(
            data.get('type')))
    return kinds[data['type']].fromdict(data)
