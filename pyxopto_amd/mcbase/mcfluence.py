"""Fluence accumulators: Cartesian (xopto/mcbase/mcfluence/fluence.py)
and cylindrical r-z (xopto/mcbase/mcfluence/fluencerz.py).

Mode 'deposition': each absorbed weight is added to the voxel (or ring
cell) that contains the packet. Mode 'fluence' (fluence rate): the absorbed
weight divided by the absorption coefficient of the medium the packet is in
(nothing where that is 0), fluence.py:102-147 with MC_FLUENCE_MODE_RATE.
``data`` is raw/(nphotons*volume) in both modes, volume dx*dy*dz
(fluence.py:343-347) or pi (r1^2 - r0^2) dz (fluencerz.py:343-348): the
absorbed energy density (1/m^3) or the fluence rate (1/m^2)."""
import numpy as np

from .. import _native as N
from .mcutil import axis as _axis
from .mcutil.axis import Axis


def _pop_type(cls, data: dict) -> dict:
    data = dict(data)
    kind = data.pop('type')
    if kind != cls.__name__:
        raise TypeError('Expected "{}" type but got "{}"!'.format(
            cls.__name__, kind))
    return data


def _check_mode(mode):
    """fluence.py:207-210"""
    if mode not in ('fluence', 'deposition'):
        raise ValueError(
            'The value of mode parameter must be '
            '"fluence" or "deposition" but got {}!'.format(mode))


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
        _check_mode(mode)
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
    data = property(_get_data, None, None,
                    'Deposition (1/m^3) or fluence rate (1/m^2).')

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
        target.rate = int(self._mode == 'fluence')
        target.enabled = N.XO_FLUENCE_XYZ
        target.offset = mc.allocate_accumulator(self, n)
        for i, ax in enumerate((self._x_axis, self._y_axis, self._z_axis)):
            target.top_left[i] = ax.start
            target.inv_step[i] = 1.0 / ax.step
            target.shape[i] = ax.n
        target.k = self._k
        return target

    def todict(self):
        """Configuration without the accumulator data (fluence.py:247-262)."""
        return {'type': 'Fluence', 'mode': self._mode,
                'xaxis': self._x_axis.todict(), 'yaxis': self._y_axis.todict(),
                'zaxis': self._z_axis.todict()}

    @classmethod
    def fromdict(cls, data: dict) -> 'Fluence':
        """fluence.py:266-284"""
        data = _pop_type(cls, data)
        x = _axis.fromdict(data.pop('xaxis'))
        y = _axis.fromdict(data.pop('yaxis'))
        z = _axis.fromdict(data.pop('zaxis'))
        return cls(x, y, z, **data)


class FluenceRz:
    """Deposition on a cylindrical grid: rings of radius raxis around the
    axis through `center` (x, y), slices zaxis (fluencerz.py:169-450).
    The raw data are [z][r] (the kernel's flat index is iz * n_r + ir);
    the reference reshapes that flat array as (n_r, n_z), which only
    matches its own index when n_r == n_z — here the array is (n_z, n_r)."""

    def __init__(self, raxis, zaxis: Axis = None, center=(0.0, 0.0),
                 mode: str = 'deposition'):
        k = N.XO_ACCU_K
        data, nphotons = None, 0
        if isinstance(raxis, FluenceRz):
            other = raxis
            raxis, zaxis = Axis(other.raxis), Axis(other.zaxis)
            center, mode, k, nphotons = (other.center, other.mode, other.k,
                                         other.nphotons)
            if other.raw is not None:
                data = np.copy(other.raw)
        _check_mode(mode)
        self._r_axis, self._z_axis = raxis, zaxis
        self._center = np.array([float(center[0]), float(center[1])])
        self._mode = mode
        self._k = int(k)
        self._data = data
        self._nphotons = int(nphotons)

    raxis = property(lambda self: self._r_axis)
    zaxis = property(lambda self: self._z_axis)
    center = property(lambda self: self._center)
    mode = property(lambda self: self._mode)
    k = property(lambda self: self._k)
    nphotons = property(lambda self: self._nphotons)
    raw = property(lambda self: self._data)
    shape = property(lambda self: (self._z_axis.n, self._r_axis.n))
    dz = property(lambda self: self._z_axis.step)
    r = property(lambda self: self._r_axis.centers)
    z = property(lambda self: self._z_axis.centers)

    def _get_data(self):
        e = self._r_axis.edges
        area = np.pi * (e[1:] ** 2 - e[:-1] ** 2)
        k = 1.0 / (max(self._nphotons, 1) * area * self.dz)
        return self._data * k[np.newaxis, :]
    data = property(_get_data, None, None,
                    'Deposition (1/m^3) or fluence rate (1/m^2).')

    def update_data(self, mc, accumulators, nphotons: int):
        new = np.asarray(accumulators[0]).reshape(self.shape) * (1.0 / self._k)
        if self._data is None:
            self._data = new
            self._nphotons = int(nphotons)
        else:
            self._data += new
            self._nphotons += int(nphotons)

    def pack(self, mc, target: N.XoFluence):
        """cl_pack of fluencerz.py:425-450."""
        target.rate = int(self._mode == 'fluence')
        target.enabled = N.XO_FLUENCE_RZ
        target.offset = mc.allocate_accumulator(
            self, self._r_axis.n * self._z_axis.n)
        target.top_left[0] = self._center[0]
        target.top_left[1] = self._center[1]
        target.top_left[2] = self._z_axis.start
        target.inv_step[0] = (1.0 / self._r_axis.step
                              if self._r_axis.step != 0.0 else 0.0)
        target.inv_step[1] = 0.0
        target.inv_step[2] = 1.0 / self._z_axis.step
        target.shape[0] = self._r_axis.n
        target.shape[1] = 1
        target.shape[2] = self._z_axis.n
        target.k = self._k
        return target

    def todict(self):
        """Configuration without the accumulator data
        (fluencerz.py:254-269)."""
        return {'type': 'FluenceRz', 'mode': self._mode,
                'raxis': self._r_axis.todict(), 'zaxis': self._z_axis.todict(),
                'center': self._center.tolist()}

    @classmethod
    def fromdict(cls, data: dict) -> 'FluenceRz':
        """fluencerz.py:273-290"""
        data = _pop_type(cls, data)
        r = _axis.fromdict(data.pop('raxis'))
        z = _axis.fromdict(data.pop('zaxis'))
        return cls(r, z, **data)


def fromdict(data: dict):
    """Fluence or FluenceRz from its todict() data."""
    kinds = {'Fluence': Fluence, 'FluenceRz': FluenceRz}
    if data.get('type') not in kinds:
        raise TypeError('Unsupported fluence type "{}"!'.format(
            data.get('type')))
    return kinds[data['type']].fromdict(data)
