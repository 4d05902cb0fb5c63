"""Voxelized sample volume (xopto/mcvox/mcgeometry/voxel.py:89-585).

The material index of every voxel is an int32 array indexed [z, y, x]
(C order). Material 0 is the surrounding medium. Call ``update()`` after
changing ``material`` so the next run uploads the grid
(voxel.py:380-411)."""
from typing import Tuple

import numpy as np

from .. import _native as N
from ..mcbase.mcutil import axis as _axis
from ..mcbase.mcutil.axis import Axis


class Voxels:
    def __init__(self, xaxis: Axis, yaxis: Axis, zaxis: Axis):
        if xaxis.logscale or yaxis.logscale or zaxis.logscale:
            raise ValueError('Logarithmic scale is not supported!')
        self._x_axis, self._y_axis, self._z_axis = xaxis, yaxis, zaxis
        self._voxel_size = (abs(xaxis.step), abs(yaxis.step), abs(zaxis.step))
        self._top_left = (min(xaxis.span), min(yaxis.span), min(zaxis.span))
        self._bottom_right = (max(xaxis.span), max(yaxis.span),
                              max(zaxis.span))
        self._material = np.zeros((zaxis.n, yaxis.n, xaxis.n), dtype=np.int32)
        self._grid = None
        self._update = True

    xaxis = property(lambda self: self._x_axis)
    yaxis = property(lambda self: self._y_axis)
    zaxis = property(lambda self: self._z_axis)
    shape = property(lambda self: self._material.shape, None, None,
                     'Shape (nz, ny, nx).')
    size = property(lambda self: self._voxel_size, None, None,
                    'Voxel size (dx, dy, dz).')
    top_left = property(lambda self: self._top_left)
    bottom_right = property(lambda self: self._bottom_right)
    material = property(lambda self: self._material, None, None,
                        'Material index of each voxel, [z, y, x].')

    def data(self, mc=None) -> np.ndarray:
        return self._material

    def __getitem__(self, item):
        return self._material[item]

    def __setitem__(self, item, value):
        self._material[item] = value
        self._update = True

    def update(self):
        """Mark the grid for upload before the next run."""
        self._update = True

    def update_required(self, clear: bool = True) -> bool:
        value = self._update
        if clear:
            self._update = False
        return value

    def meshgrid(self):
        """(z, y, x) coordinates of the voxel centres, each of shape."""
        if self._grid is None:
            self._grid = np.meshgrid(self._z_axis.centers,
                                     self._y_axis.centers,
                                     self._x_axis.centers, indexing='ij')
        return self._grid

    def contains(self, position) -> bool:
        return (self._x_axis.start <= position[0] < self._x_axis.stop and
                self._y_axis.start <= position[1] < self._y_axis.stop and
                self._z_axis.start <= position[2] < self._z_axis.stop)

    def index(self, position) -> Tuple[int, int, int]:
        """(ix, iy, iz) of the voxel holding position (no range check)."""
        return tuple(int((position[k] - self._top_left[k]) /
                         self._voxel_size[k]) for k in range(3))

    def intersect(self, pos, direction):
        """First intersection of the ray with the sample box at a positive
        distance, and the box normal that points along the ray
        (voxel.py:317-378)."""
        inv = [np.inf if direction[k] == 0.0 else 1.0 / direction[k]
               for k in range(3)]
        tl, br = self._top_left, self._bottom_right
        t = [((tl[k] - pos[k]) * inv[k], (br[k] - pos[k]) * inv[k])
             for k in range(3)]
        tmins = [min(a, b) for a, b in t]
        tmaxs = [max(a, b) for a, b in t]
        tmin, tmax = max(tmins), min(tmaxs)
        if tmin > tmax or tmax < 0.0:
            return None, None
        normal = np.zeros(3)
        axis = int(np.argmax(tmins))
        normal[axis] = np.sign(direction[axis]) if direction[axis] != 0 else 1
        dist = tmin if tmin > 0.0 else tmax
        point = np.asarray(pos, dtype=np.float64) + dist * \
            np.asarray(direction, dtype=np.float64)
        return point, normal

    def pack(self, mc, target: N.XoVoxelCfg):
        for k in range(3):
            target.top_left[k] = self._top_left[k]
            target.bottom_right[k] = self._bottom_right[k]
            target.size[k] = self._voxel_size[k]
        target.shape[0] = self._x_axis.n
        target.shape[1] = self._y_axis.n
        target.shape[2] = self._z_axis.n
        return target

    def todict(self):
        """voxel.py:687-696; the material grid is not part of the dict."""
        return {'type': 'Voxels', 'xaxis': self._x_axis.todict(),
                'yaxis': self._y_axis.todict(),
                'zaxis': self._z_axis.todict()}

    @classmethod
    def fromdict(cls, data: dict) -> 'Voxels':
        """voxel.py:699-720"""
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError('Expected data of instance Voxels '
                            'but got "{}"!'.format(kind))
        return cls(_axis.fromdict(data.pop('xaxis')),
                   _axis.fromdict(data.pop('yaxis')),
                   _axis.fromdict(data.pop('zaxis')))
