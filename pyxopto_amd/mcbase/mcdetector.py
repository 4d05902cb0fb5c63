"""Surface detectors (xopto/mcml/mcdetector/{base,radial,total}.py).

Each detector owns integer accumulator bins; ``update_data`` adds
``accumulator / K`` (K = 0x7FFFFF) to ``raw`` and the packet count to
``nphotons`` (base.py:269-293). Normalization: Radial divides by the annulus
area and the packet count (radial.py:194-196, 242-244), Total by the packet
count (total.py:159-161).
"""
from typing import Tuple

import numpy as np

from .. import _native as N
from .mcutil import axis as _axis
from .mcutil.axis import Axis, RadialAxis


class Detector:
    kind = N.XO_DET_NONE

    def __init__(self, raw_data: np.ndarray, nphotons: int):
        self._raw_data = raw_data
        self._nphotons = int(nphotons)

    raw = property(lambda self: self._raw_data, None, None,
                   'Raw accumulator data (sum of packet weights).')
    nphotons = property(lambda self: self._nphotons, None, None,
                        'Packets that produced the raw data.')
    shape = property(lambda self: self._raw_data.shape, None, None,
                     'Shape of the raw data array.')
    total = property(lambda self: self._raw_data.sum(), None, None,
                     'Total collected weight.')

    def set_raw_data(self, data, nphotons: int):
        self._raw_data[:] = np.asarray(data, dtype=self._raw_data.dtype)
        self._nphotons = int(nphotons)

    def update_data(self, mc, accumulators, nphotons: int, **kwargs):
        new = np.reshape(np.asarray(accumulators[0]), self.shape)
        self._raw_data += new * (1.0 / N.XO_ACCU_K)
        self._nphotons += int(nphotons)

    def pack(self, mc, target: N.XoDetector) -> N.XoDetector:
        raise NotImplementedError


def _unit(v):
    v = np.asarray(v, dtype=np.float64).copy()
    norm = np.linalg.norm(v)
    if norm == 0.0:
        raise ValueError('Direction vector norm/length must not be 0!')
    return v / norm


class Radial(Detector):
    kind = N.XO_DET_RADIAL

    def __init__(self, raxis, position: Tuple[float, float] = (0.0, 0.0),
                 cosmin: float = 0.0,
                 direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)):
        if isinstance(raxis, Radial):
            other = raxis
            position, cosmin, direction = (other.position, other.cosmin,
                                           other.direction)
            raxis = type(other.raxis)(other.raxis)
            raw, nphotons = np.copy(other.raw), other.nphotons
        else:
            raw, nphotons = np.zeros((raxis.n,)), 0
        super().__init__(raw, nphotons)
        self._r_axis = raxis
        self._position = np.zeros(2)
        self._position[:] = position
        self._cosmin = min(max(float(cosmin), 0.0), 1.0)
        self._direction = _unit(direction)
        e = self._r_axis.edges
        self._inv_area = 1.0 / (np.pi * (e[1:] ** 2 - e[:-1] ** 2))

    raxis = property(lambda self: self._r_axis, None, None, 'Radial axis.')
    position = property(lambda self: self._position, None, None, 'Centre.')
    cosmin = property(lambda self: self._cosmin, None, None,
                      'Cosine of the acceptance angle.')
    direction = property(lambda self: self._direction, None, None,
                         'Reference direction.')
    r = property(lambda self: self._r_axis.centers, None, None, 'Bin centres.')
    edges = property(lambda self: self._r_axis.edges, None, None, 'Edges.')
    n = property(lambda self: self._r_axis.n, None, None, 'Number of bins.')
    logscale = property(lambda self: self._r_axis.logscale, None, None,
                        'Axis scale.')

    def _get_normalized(self):
        return self.raw * self._inv_area * (1.0 / max(self._nphotons, 1))
    normalized = property(_get_normalized, None, None, 'Normalized.')
    reflectance = property(_get_normalized, None, None, 'Reflectance.')
    transmittance = property(_get_normalized, None, None, 'Transmittance.')

    def pack(self, mc, target):
        target.kind = self.kind
        target.offset = mc.allocate_accumulator(self, self.n)
        target.position[0], target.position[1] = self._position
        target.r_min = self._r_axis.scaled_start
        target.inv_dr = 1.0 / self._r_axis.step if self._r_axis.step != 0.0 \
            else 0.0
        target.log_scale = int(self._r_axis.logscale)
        target.n = self._r_axis.n
        target.cos_min = self._cosmin
        for i in range(3):
            target.direction[i] = self._direction[i]
        return target

    def todict(self):
        """Configuration without the accumulator data (radial.py:292-309)."""
        return {'type': 'Radial', 'position': self._position.tolist(),
                'r_axis': self._r_axis.todict(), 'cosmin': self._cosmin,
                'direction': self._direction.tolist()}

    @staticmethod
    def fromdict(data: dict) -> 'Radial':
        data = dict(data)
        kind = data.pop('type')
        if kind != 'Radial':
            raise TypeError('Expected "Radial" type but got "{}"!'.format(kind))
        return Radial(_axis.fromdict(data.pop('r_axis')), **data)

    def __repr__(self):
        return 'Radial({}, position={}, cosmin={})'.format(
            self._r_axis, tuple(self._position), self._cosmin)


class Total(Detector):
    kind = N.XO_DET_TOTAL

    def __init__(self, cosmin: float = 0.0,
                 direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)):
        if isinstance(cosmin, Total):
            other = cosmin
            cosmin, direction = other.cosmin, other.direction
            raw, nphotons = np.copy(other.raw), other.nphotons
        else:
            raw, nphotons = np.zeros((1,)), 0
        super().__init__(raw, nphotons)
        self._cosmin = min(max(float(cosmin), 0.0), 1.0)
        self._direction = _unit(direction)

    cosmin = property(lambda self: self._cosmin, None, None,
                      'Cosine of the acceptance angle.')
    direction = property(lambda self: self._direction, None, None,
                         'Reference direction.')

    def _get_normalized(self):
        return self.raw * (1.0 / max(self._nphotons, 1.0))
    normalized = property(_get_normalized, None, None, 'Normalized.')
    reflectance = property(_get_normalized, None, None, 'Reflectance.')
    transmittance = property(_get_normalized, None, None, 'Transmittance.')

    def pack(self, mc, target):
        target.kind = self.kind
        target.offset = mc.allocate_accumulator(self, 1)
        target.n = 1
        target.cos_min = self._cosmin
        for i in range(3):
            target.direction[i] = self._direction[i]
        return target

    def todict(self):
        """Configuration without the accumulator data (total.py:197-212)."""
        return {'type': 'Total', 'cosmin': self._cosmin,
                'direction': self._direction.tolist()}

    @staticmethod
    def fromdict(data: dict) -> 'Total':
        data = dict(data)
        kind = data.pop('type')
        if kind != 'Total':
            raise TypeError('Expected "Total" type but got "{}"!'.format(kind))
        return Total(**data)

    def __repr__(self):
        return 'Total(cosmin={})'.format(self._cosmin)


class Detectors:
    """Top, bottom and specular detectors (xopto/mcml/mcdetector/base.py:295-560)."""

    def __init__(self, top: Detector = None, bottom: Detector = None,
                 specular: Detector = None):
        if isinstance(top, Detectors):
            other = top
            top = type(other.top)(other.top) if other.top is not None else None
            bottom = type(other.bottom)(other.bottom) \
                if other.bottom is not None else None
            specular = type(other.specular)(other.specular) \
                if other.specular is not None else None
        self._top, self._bottom, self._specular = top, bottom, specular

    top = property(lambda self: self._top, None, None, 'Top detector.')
    bottom = property(lambda self: self._bottom, None, None, 'Bottom detector.')
    specular = property(lambda self: self._specular, None, None,
                        'Specular detector.')

    def __iter__(self):
        return iter([d for d in (self._top, self._bottom, self._specular)
                     if d is not None])

    def items(self):
        return [('top', self._top), ('bottom', self._bottom),
                ('specular', self._specular)]

    def types(self):
        return tuple(type(d) for _, d in self.items())

    def pack(self, mc, cfg):
        for name, det in self.items():
            target = getattr(cfg, name)
            if det is None:
                target.kind = N.XO_DET_NONE
            else:
                det.pack(mc, target)

    def update_data(self, mc, accumulators_by_name, nphotons: int):
        for name, det in self.items():
            if det is not None and name in accumulators_by_name:
                det.update_data(mc, [accumulators_by_name[name]], nphotons)

    def todict(self) -> dict:
        """base.py:551-561; an unused location is {'type':
        'DetectorDefault'} (base.py:201-205)."""
        out = {'type': 'Detectors'}
        for name, det in self.items():
            out[name] = det.todict() if det is not None else \
                {'type': 'DetectorDefault'}
        return out

    @classmethod
    def fromdict(cls, data: dict) -> 'Detectors':
        data = dict(data)
        kind = data.pop('type')
        if kind != 'Detectors':
            raise TypeError('Expected "Detectors" type but got "{}"!'.format(
                kind))
        kinds = {'Radial': Radial, 'Total': Total}
        dets = {}
        for name in ('top', 'bottom', 'specular'):
            item = data.get(name)
            if item is None or item.get('type') == 'DetectorDefault':
                dets[name] = None
            elif item.get('type') in kinds:
                dets[name] = kinds[item['type']].fromdict(item)
            else:
                raise TypeError('Unsupported detector type "{}"!'.format(
                    item.get('type')))
        return cls(**dets)

    def __repr__(self):
        return 'Detectors(top={}, bottom={}, specular={})'.format(
            self._top, self._bottom, self._specular)
