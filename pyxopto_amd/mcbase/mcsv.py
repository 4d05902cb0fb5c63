"""Sampling volume of a photon packet trace (xopto/mcbase/mcsv.py:34-409).

SamplingVolume(xaxis, yaxis, zaxis) is a voxel grid; Mc.sampling_volume
walks every traced packet's path through it and adds, to each voxel the
path crosses, the packet's terminal weight times the length of path inside
the voxel. Integers: round(weight * length * multiplier * k) with k =
0x7FFFFF and multiplier = 1 / smallest voxel size (mcsv.py:296-305);
``data`` is the sum divided by k * multiplier, ``weight`` the summed
terminal weight.

Terminal weight: the reference's kernel reads the last float of a packet's
last trace entry (mcsv.template.c:278), which since the entry grew to eight
fields is the optical path length 'pl', not the weight 'w'. The default
``weight_field='pl'`` reproduces that; ``weight_field='w'`` weights by the
packet weight.
"""
from typing import Tuple

import numpy as np

from .. import _native as N
from .mcutil import axis as _axis
from .mcutil.axis import Axis

_FIELDS = {'w': 6, 'pl': 7}


class SamplingVolume:
    def __init__(self, xaxis, yaxis: Axis = None, zaxis: Axis = None,
                 weight_field: str = 'pl'):
        data, weight = None, 0.0
        if isinstance(xaxis, SamplingVolume):
            other = xaxis
            xaxis, yaxis, zaxis = (Axis(other.xaxis), Axis(other.yaxis),
                                   Axis(other.zaxis))
            weight_field = other.weight_field
            if other.data is not None:
                data = np.copy(other.data)
            weight = other.weight
        if weight_field not in _FIELDS:
            raise ValueError('weight_field must be "pl" or "w"!')
        self._x_axis, self._y_axis, self._z_axis = xaxis, yaxis, zaxis
        if xaxis.n * yaxis.n * zaxis.n <= 0:
            raise ValueError('Sampling volume accumulator array has one or '
                             'more dimensions equal to zero!')
        self._data = data
        self._weight = weight
        self._weight_field = weight_field
        self._k = N.XO_ACCU_K

    xaxis = property(lambda self: self._x_axis, None, None, 'x axis.')
    yaxis = property(lambda self: self._y_axis, None, None, 'y axis.')
    zaxis = property(lambda self: self._z_axis, None, None, 'z axis.')
    x = property(lambda self: self._x_axis.centers, None, None, 'x centres.')
    y = property(lambda self: self._y_axis.centers, None, None, 'y centres.')
    z = property(lambda self: self._z_axis.centers, None, None, 'z centres.')
    k = property(lambda self: self._k, None, None,
                 'Float to integer weight factor.')
    weight_field = property(lambda self: self._weight_field, None, None,
                            'Trace field read as the terminal weight.')
    shape = property(lambda self: (self._z_axis.n, self._y_axis.n,
                                   self._x_axis.n), None, None,
                     'Accumulator shape (z, y, x).')

    def _get_data(self):
        return self._data

    def _set_data(self, data):
        self._data = data
    data = property(_get_data, _set_data, None, 'Sampling volume data.')

    def _get_weight(self):
        return self._weight

    def _set_weight(self, w):
        self._weight = w
    weight = property(_get_weight, _set_weight, None,
                      'Total terminal weight of the processed packets.')

    def multiplier(self) -> float:
        """1 / smallest voxel size (mcsv.py:296-305)."""
        step = min(self._x_axis.step, self._y_axis.step, self._z_axis.step)
        if step <= 0.0:
            step = 1.0 / self._k
        return 1.0 / step

    def clear(self):
        if self._data is not None:
            self._data.fill(0)

    def pack(self, offset: int = 0) -> N.XoSamplingVolume:
        t = N.XoSamplingVolume()
        for i, ax in enumerate((self._x_axis, self._y_axis, self._z_axis)):
            t.top_left[i] = ax.start
            t.voxel_size[i] = ax.step
            t.shape[i] = ax.n
        t.multiplier = self.multiplier()
        t.offset = int(offset)
        t.k = self._k
        t.end_field = _FIELDS[self._weight_field]
        return t

    def update_data(self, mc, accumulators, total_weight: int, **kwargs):
        """mcsv.py:273-294"""
        m = self.multiplier()
        new = np.reshape(np.asarray(accumulators[0]), self.shape) * \
            (1.0 / (self._k * m))
        if self._data is not None:
            self._data += new
            self._weight += float(total_weight) / self._k
        else:
            self._data = new
            self._weight = float(total_weight) / self._k

    def todict(self) -> dict:
        """mcsv.py:150-163"""
        return {'type': 'SamplingVolume', 'xaxis': self._x_axis.todict(),
                'yaxis': self._y_axis.todict(),
                'zaxis': self._z_axis.todict()}

    @classmethod
    def fromdict(cls, data: dict) -> 'SamplingVolume':
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError('Expected data for type "{}" but got "{}"!'.format(
                cls.__name__, kind))
        return cls(_axis.fromdict(data.pop('xaxis')),
                   _axis.fromdict(data.pop('yaxis')),
                   _axis.fromdict(data.pop('zaxis')), **data)

    def __repr__(self):
        return 'SamplingVolume(xaxis={}, yaxis={}, zaxis={})'.format(
            self._x_axis, self._y_axis, self._z_axis)
