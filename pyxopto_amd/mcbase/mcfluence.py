"""Cartesian deposition accumulator (xopto/mcbase/mcfluence/fluence.py).

Deposition mode: each absorbed weight is added to the voxel that contains
the packet; ``data`` is raw/(nphotons*dx*dy*dz) (fluence.py:343-347)."""
import numpy as np

from .. import _native as N
from .mcutil.axis import Axis


class Fluence:
    def __init__(self, xaxis, yaxis: Axis = None, zaxis: Axis = None,
                 mode: str = 'deposition'):
        k = N.XO_ACCU_K
        data, nphotons = None, 0
        if isinstance(xaxis, Fluence):
            other = xaxis
            xaxis, yaxis, zaxis = (Axis(other.xaxis), Axis(other.yaxis),
                                   Axis(other.zaxis))
            mode, k, nphotons = other.mode, other.k, other.nphotons
            if other.raw is not None:
                data = np.copy(other.raw)
        if mode not in ('deposition',):
            raise NotImplementedError(
                'Only the "deposition" fluence mode is implemented!')
        self._x_axis, self._y_axis, self._z_axis = xaxis, yaxis, zaxis
        self._mode = mode
        self._k = int(k)
        self._data = data
        self._nphotons = int(nphotons)

    xaxis = property(lambda self: self._x_axis)
    yaxis = property(lambda self: self._y_axis)
    zaxis = property(lambda self: self._z_axis)
    mode = property(lambda self: self._mode)
    k = property(lambda self: self._k)
    nphotons = property(lambda self: self._nphotons)
    raw = property(lambda self: self._data)
    shape = property(lambda self: (self._z_axis.n, self._y_axis.n,
                                   self._x_axis.n))
    dx = property(lambda self: self._x_axis.step)
    dy = property(lambda self: self._y_axis.step)
    dz = property(lambda self: self._z_axis.step)

    def _get_data(self):
        k = 1.0 / (max(self._nphotons, 1) * self.dx * self.dy * self.dz)
        return self._data * k
    data = property(_get_data, None, None, 'Deposition (1/m^3).')

    def update_data(self, mc, accumulators, nphotons: int):
        new = np.asarray(accumulators[0]).reshape(self.shape) * (1.0 / self._k)
        if self._data is None:
            self._data = new
            self._nphotons = int(nphotons)
        else:
            self._data += new
            self._nphotons += int(nphotons)

    def pack(self, mc, target: N.XoFluence):
        n = self._x_axis.n * self._y_axis.n * self._z_axis.n
        target.enabled = 1
        target.offset = mc.allocate_accumulator(self, n)
        for i, ax in enumerate((self._x_axis, self._y_axis, self._z_axis)):
            target.top_left[i] = ax.start
            target.inv_step[i] = 1.0 / ax.step
            target.shape[i] = ax.n
        target.k = self._k
        return target

    def todict(self):
        return {'type': 'Fluence', 'xaxis': self._x_axis.todict(),
                'yaxis': self._y_axis.todict(), 'zaxis': self._z_axis.todict(),
                'mode': self._mode}
