"""Accumulator axes (xopto/mcbase/mcutil/axis.py:26-265).

Axis: linearly or logarithmically spaced bins; RadialAxis: the same bins with
area-corrected centres 2/3 (r_i^2 + r_i r_{i+1} + r_{i+1}^2)/(r_i + r_{i+1}).
"""
from typing import Tuple

import numpy as np


class Axis:
    def __init__(self, start=0.0, stop: float = 1.0, n: int = 1,
                 logscale: bool = False):
        if isinstance(start, Axis):
            other = start
            start, stop, n, logscale = (other.start, other.stop, other.n,
                                        other.logscale)
        start, stop, n = float(start), float(stop), int(n)
        logscale = bool(logscale)
        if n < 1:
            raise ValueError('An axis needs at least one bin!')
        self._logscale = logscale
        self._n = n
        if logscale:
            start = max(start, float(np.finfo(np.float64).eps))
            self._edges = np.logspace(np.log(start), np.log(stop), n + 1,
                                      base=np.e)
            self._step = (np.log(stop) - np.log(start)) / n
        else:
            self._edges = np.linspace(start, stop, n + 1)
            self._step = (stop - start) / n
        self._span = np.array((start, stop), dtype=np.float64)
        self._scaled_span = np.log(self._span) if logscale else self._span
        self._centers = 0.5 * (self._edges[:-1] + self._edges[1:])

    def todict(self) -> dict:
        return {'start': self._span[0], 'stop': self._span[1], 'n': self._n,
                'logscale': self._logscale, 'type': type(self).__name__}

    @classmethod
    def fromdict(cls, data: dict) -> 'Axis':
        data = dict(data)
        kind = data.pop('type', cls.__name__)
        if kind != cls.__name__:
            raise TypeError('Expected "{}" type but got "{}"!'.format(
                cls.__name__, kind))
        return cls(**data)

    logscale = property(lambda self: self._logscale, None, None,
                        'Accumulator scale.')
    step = property(lambda self: self._step, None, None, 'Accumulator step.')
    n = property(lambda self: self._n, None, None, 'Number of bins.')
    span = property(lambda self: self._span, None, None, 'Span (min, max).')
    start = property(lambda self: self._span[0], None, None, 'Start.')
    stop = property(lambda self: self._span[1], None, None, 'Stop.')
    scaled_span = property(lambda self: self._scaled_span, None, None,
                           'Scaled span.')
    scaled_start = property(lambda self: self._scaled_span[0], None, None,
                            'Scaled start.')
    scaled_stop = property(lambda self: self._scaled_span[1], None, None,
                           'Scaled stop.')
    edges = property(lambda self: self._edges, None, None, 'Bin edges.')

    def _get_centers(self) -> np.ndarray:
        return self._centers
    centers = property(_get_centers, None, None, 'Bin centres.')

    def __str__(self):
        return '{}(start={}, stop={}, n={}, logscale={})'.format(
            type(self).__name__, self._span[0], self._span[1], self._n,
            self._logscale)

    __repr__ = __str__


class RadialAxis(Axis):
    def __init__(self, start=0.0, stop: float = 1.0, n: int = 1,
                 logscale: bool = False):
        Axis.__init__(self, start, stop, n, logscale)
        e = self.edges
        self._true_centers = (2.0 / 3.0) * (e[:-1] ** 2 + e[:-1] * e[1:] +
                                            e[1:] ** 2) / (e[:-1] + e[1:])

    def _get_centers(self) -> np.ndarray:
        return self._true_centers
    centers = property(_get_centers, None, None,
                       'Area-corrected centres of the bins.')


def fromdict(data: dict) -> Axis:
    """Axis or RadialAxis from its todict() data (axis.py:125-146,
    251-264). RadialAxis.todict returns its dict here; the reference's
    returns None (axis.py:251-256 has no return statement)."""
    kinds = {'Axis': Axis, 'RadialAxis': RadialAxis}
    kind = data.get('type', 'Axis')
    if kind not in kinds:
        raise TypeError('Unsupported axis type "{}"!'.format(kind))
    return kinds[kind].fromdict(data)
