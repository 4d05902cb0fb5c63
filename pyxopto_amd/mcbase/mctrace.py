"""Photon packet trace and its filter (xopto/mcbase/mctrace.py:31-1234).

Trace(maxlen, options, filter, event_mask, plon) logs the state of every
packet of a run: entries (x, y, z, px, py, pz, w, pl) at the launch
(TRACE_START), at the end (TRACE_END) or at every step (TRACE_ALL); with an
event mask only steps whose events match it. ``data`` is a structured
float32 array [packet][maxlen], ``n`` the logged event count of each packet
(it may exceed maxlen: later events overwrite the last entry). The kernel
side is the traced build of the layered kernel (xo_mcml_run_trace).

Filter selects packets by their terminal state (position, direction, radial
distance, optical path length); its semantics follow the reference's
Filter.__call__ (mctrace.py:259-352), including that the path-length test
is skipped for a trace without path lengths and ORs into the mask left by
the previous test.
"""
from typing import Tuple

import numpy as np

from .. import _native as N


def _range(value) -> Tuple[float, float]:
    if not isinstance(value, (list, tuple)) or len(value) != 2:
        raise TypeError(
            'Range must be specified as a tuple of two float: (low, high)!')
    return float(value[0]), float(value[1])


def _origin2(value) -> Tuple[float, float]:
    if not isinstance(value, (list, tuple)) or len(value) != 2:
        raise TypeError(
            'Origin must be specified as a tuple of two float: (x, y)!')
    return float(value[0]), float(value[1])


def _unit3(p) -> Tuple[float, float, float]:
    if not isinstance(p, (list, tuple)) or len(p) != 3:
        raise TypeError(
            'Direction must be specified as a tuple of three float: (x, y, z)!')
    length = float(np.sqrt(p[0] ** 2 + p[1] ** 2 + p[2] ** 2))
    if length == 0.0:
        raise ValueError('Direction vector length must not be 0!')
    return float(p[0]) / length, float(p[1]) / length, float(p[2]) / length


class Filter:
    """Selection of traced packets by their final state
    (mctrace.py:31-257): x, y, z, pz, pl as (low, high) or a list of such
    intervals; r as (rmin, rmax, (x0, y0)) or a list; dir as
    (cosmin, cosmax, (px, py, pz)) or a list. Intervals of one quantity are
    ORed, quantities are ANDed, bounds are inclusive."""

    def __init__(self, x=None, y=None, z=None, pz=None, r=None, dir=None,
                 pl=None):
        if isinstance(x, Filter):
            other = x
            x, y, z, pz, r, dir, pl = (other.x, other.y, other.z, other.pz,
                                       other.r, other.dir, other.pl)
        self._x = self._ranges('x', x)
        self._y = self._ranges('y', y)
        self._z = self._ranges('z', z)
        self._pz = self._ranges('pz', pz)
        self._r = self._radii(r)
        self._dir = self._dirs(dir)
        self._pl = self._ranges('pl', pl)

    @staticmethod
    def _ranges(name, value):
        if value is None:
            return None
        if not isinstance(value, (list, tuple)):
            raise TypeError('Filter parameter {} must be a tuple of two float '
                            'values (low, high), or a list of such '
                            'tuples!'.format(name))
        if isinstance(value[0], (list, tuple)):
            return [_range(v) for v in value]
        return [_range(value)]

    @staticmethod
    def _radii(r):
        if r is None:
            return None
        if not isinstance(r, (list, tuple)):
            raise TypeError('Filter parameter r must be a tuple (rmin, rmax, '
                            '(x_origin, y_origin)) or a list of such tuples!')
        if isinstance(r[0], (list, tuple)):
            return [(float(v[0]), float(v[1]), _origin2(v[2])) for v in r]
        return [(float(r[0]), float(r[1]), _origin2(r[2]))]

    @staticmethod
    def _dirs(d):
        if d is None:
            return None
        if not isinstance(d, (list, tuple)):
            raise TypeError('Filter parameter dir must be a tuple (cosmin, '
                            'cosmax, (px, py, pz)) or a list of such tuples!')
        if isinstance(d[0], (list, tuple)):
            return [(float(v[0]), float(v[1]), _unit3(v[2])) for v in d]
        return [(float(d[0]), float(d[1]), _unit3(d[2]))]

    x = property(lambda self: self._x, None, None, 'x intervals.')
    y = property(lambda self: self._y, None, None, 'y intervals.')
    z = property(lambda self: self._z, None, None, 'z intervals.')
    pz = property(lambda self: self._pz, None, None, 'pz intervals.')
    r = property(lambda self: self._r, None, None,
                 '(rmin, rmax, (x0, y0)) tests.')
    dir = property(lambda self: self._dir, None, None,
                   '(cosmin, cosmax, direction) tests.')
    pl = property(lambda self: self._pl, None, None, 'Path length intervals.')

    def todict(self) -> dict:
        """mctrace.py:138-157"""
        return {'type': 'Filter', 'x': self._x, 'y': self._y, 'z': self._z,
                'pz': self._pz, 'dir': self._dir, 'r': self._r,
                'pl': self._pl}

    @classmethod
    def fromdict(cls, data: dict) -> 'Filter':
        data = dict(data)
        kind = data.pop('type')
        if kind != 'Filter':
            raise TypeError('Expected "{}" type but got "{}"!'.format(
                cls.__name__, kind))
        return cls(**data)

    def mask(self, trace: 'Trace') -> np.ndarray:
        """Packets whose terminal state passes every test (before the
        overflow rule)."""
        t = trace.terminal
        valid = np.ones(len(trace), dtype=bool)
        var = np.zeros(len(trace), dtype=bool)

        def apply(tests, one, reset=True):
            if tests is None:
                return
            if reset:
                var.fill(False)
            for test in tests:
                np.logical_or(one(test) & valid, var, out=var)
            np.logical_and(var, valid, out=valid)

        apply(self._x, lambda v: (t['x'] >= v[0]) & (t['x'] <= v[1]), False)
        apply(self._y, lambda v: (t['y'] >= v[0]) & (t['y'] <= v[1]))
        apply(self._z, lambda v: (t['z'] >= v[0]) & (t['z'] <= v[1]))
        apply(self._pz, lambda v: (t['pz'] >= v[0]) & (t['pz'] <= v[1]))

        def radius(v):
            x0, y0 = v[2] if len(v) > 2 else (0.0, 0.0)
            rr = (t['x'] - x0) ** 2 + (t['y'] - y0) ** 2
            return (rr >= v[0] ** 2) & (rr <= v[1] ** 2)
        apply(self._r, radius)

        def direction(v):
            c = t['px'] * v[2][0] + t['py'] * v[2][1] + t['pz'] * v[2][2]
            return (c >= v[0]) & (c <= v[1])
        apply(self._dir, direction)
        if trace.plon:
            # the reference does not clear the mask before this test
            # (mctrace.py:336-343)
            apply(self._pl, lambda v: (t['pl'] >= v[0]) & (t['pl'] <= v[1]),
                  False)
        return valid

    def __call__(self, trace: 'Trace', update: bool = True):
        """(filtered trace, packets that passed but overflowed the trace)
        (mctrace.py:259-352)."""
        if not isinstance(trace, Trace):
            raise TypeError('Trace filter can be applied only to Trace '
                            'objects!')
        valid = self.mask(trace)
        overflow = trace.overflow
        selected = valid & ~overflow
        dropped = int(np.count_nonzero(valid & overflow))
        out = trace if update else Trace(trace)
        out.data = trace.data[selected, :]
        out.n = trace.n[selected]
        return out, dropped

    def __repr__(self):
        return 'Filter(x={}, y={}, z={}, pz={}, dir={}, r={}, pl={})'.format(
            self._x, self._y, self._z, self._pz, self._dir, self._r, self._pl)


class Trace:
    """Photon packet trace (mctrace.py:444-1234)."""
    TRACE_NONE = N.XO_TRACE_NONE
    TRACE_START = N.XO_TRACE_START
    TRACE_END = N.XO_TRACE_END
    TRACE_ALL = N.XO_TRACE_ALL
    TRACE_ENTRY_LEN = N.XO_TRACE_ENTRY_LEN
    TRACE_EVENT_REFLECTION = N.XO_EVENT_REFLECTION
    TRACE_EVENT_REFRACTION = N.XO_EVENT_REFRACTION
    TRACE_EVENT_BOUNDARY_HIT = N.XO_EVENT_BOUNDARY_HIT
    TRACE_EVENT_LAUNCH = N.XO_EVENT_LAUNCH
    TRACE_EVENT_ABSORPTION = N.XO_EVENT_ABSORPTION
    TRACE_EVENT_SCATTERING = N.XO_EVENT_SCATTERING
    TRACE_EVENT_TERMINATION = N.XO_EVENT_TERMINATED
    TRACE_EVENT_ESCAPE = N.XO_EVENT_ESCAPED
    TRACE_EVENT_ALL = -1

    DTYPE = np.dtype([('x', np.float32), ('y', np.float32),
                      ('z', np.float32), ('px', np.float32),
                      ('py', np.float32), ('pz', np.float32),
                      ('w', np.float32), ('pl', np.float32)])

    def __init__(self, maxlen=500, options: int = TRACE_ALL,
                 filter: Filter = None, event_mask: int = None,
                 plon: bool = True):
        if isinstance(maxlen, Trace):
            other = maxlen
            self._maxlen = other.maxlen
            self._data, self._n = other.data, other.n
            self._options = other.options
            self._filter = other.filter
            self._n_dropped = other.dropped
            self._plon = other.plon
            self._event_mask = other.event_mask
        else:
            self._data = self._n = None
            self._options = int(options) & 7
            self._set_maxlen(maxlen)
            self._filter = filter
            self._n_dropped = 0
            self._plon = bool(plon)
            self._event_mask = None if event_mask is None else int(event_mask)
        self._terminal = None
        self._overflow = None

    def _set_maxlen(self, maxlen):
        """mctrace.py:848-859: 1 entry for START or END alone, 2 for both."""
        if self._options == Trace.TRACE_ALL:
            pass
        elif self._options == (Trace.TRACE_START | Trace.TRACE_END):
            maxlen = 2
        else:
            maxlen = 1
        if int(maxlen) < 1:
            raise ValueError('Maximum trace length must be at least 1!')
        self._maxlen = int(maxlen)

    maxlen = property(lambda self: self._maxlen, _set_maxlen, None,
                      'Maximum trace length.')
    options = property(lambda self: self._options, None, None,
                       'Trace options (TRACE_START/END/ALL).')
    plon = property(lambda self: self._plon, None, None,
                    'Optical path length tracking.')
    event_mask = property(lambda self: self._event_mask, None, None,
                          'Mask of traced events or None (all).')
    dropped = property(lambda self: self._n_dropped, None, None,
                       'Packets that passed the filter but overflowed.')

    def _get_filter(self):
        return self._filter

    def _set_filter(self, f):
        if not isinstance(f, Filter):
            raise TypeError('Expected a "Filter" instance but got '
                            '"{}"'.format(type(f)))
        self._filter = f
    filter = property(_get_filter, _set_filter, None, 'Trace filter.')

    def _get_data(self):
        return self._data

    def _set_data(self, data):
        self._data = data
        self._terminal = self._overflow = None
    data = property(_get_data, _set_data, None,
                    'Trace entries [packet][maxlen] (structured).')

    def _get_n(self):
        return self._n

    def _set_n(self, n):
        self._n = n
        self._terminal = self._overflow = None
    n = property(_get_n, _set_n, None, 'Logged events per packet.')

    nphotons = property(lambda self: 0 if self._n is None else self._n.size,
                        None, None, 'Packets in the trace.')

    def __len__(self):
        return 0 if self._n is None else self._n.size

    def _get_terminal(self):
        if self._terminal is None and self._data is not None:
            last = np.minimum(self._n - 1, self._maxlen - 1)
            self._terminal = self._data[np.arange(self._n.size), last]
        return self._terminal
    terminal = property(_get_terminal, None, None,
                        'Final state of every packet.')

    def _get_overflow(self):
        if self._overflow is None and self._n is not None:
            self._overflow = self._n >= self._maxlen
        return self._overflow
    overflow = property(_get_overflow, None, None,
                        'Packets with more events than maxlen.')

    def dtype(self, mc=None) -> np.dtype:
        return Trace.DTYPE

    def pack(self, mc=None) -> N.XoTrace:
        t = N.XoTrace()
        t.max_events = self._maxlen
        t.options = self._options
        t.use_events = int(self._event_mask is not None)
        t.event_mask = (self._event_mask if self._event_mask is not None
                        else 0) & 0xFFFFFFFF
        t.plon = int(self._plon)
        return t

    def raw(self) -> Tuple[np.ndarray, np.ndarray]:
        """(float32 [packet][maxlen][8], int32 [packet]) for the kernels."""
        data = np.ascontiguousarray(self._data, dtype=Trace.DTYPE)
        return (data.view(np.float32).reshape(len(self), self._maxlen,
                                              N.XO_TRACE_ENTRY_LEN),
                np.ascontiguousarray(self._n, dtype=np.int32))

    def update_data(self, mc, data: np.ndarray, n: np.ndarray,
                    nphotons: int = None, **kwargs):
        """Append a run's raw trace (float32 [packet][maxlen][8], counts);
        the filter applies to the new packets (mctrace.py:1120-1171)."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        new = data.view(Trace.DTYPE).reshape(data.shape[0], self._maxlen)
        n = np.asarray(n, dtype=np.int32)
        if self._data is None:
            self.data, self.n = new, n
            self.apply_filter()
        else:
            tmp = Trace(self)
            tmp.data, tmp.n = new, n
            tmp.apply_filter()
            self.n = np.hstack([self._n, tmp.n])
            self.data = np.vstack([self._data, tmp.data])

    def apply_filter(self):
        if self._filter is not None:
            self._n_dropped = self._filter(self, update=True)[1]
            self._terminal = None

    def update(self, obj: 'Trace'):
        """Append the packets of another trace (filtered by this one's
        filter)."""
        if self._maxlen != obj.maxlen:
            raise ValueError('Cannot update the trace with trace data of '
                             'different maximum length!')
        if self._data is None:
            self.data, self.n = obj.data, obj.n
            return
        if self._filter is not None:
            obj = self._filter(obj, update=False)[0]
        self.n = np.hstack([self._n, obj.n])
        self.data = np.vstack([self._data, obj.data])

    def todict(self) -> dict:
        """mctrace.py:686-705"""
        return {'type': 'Trace', 'maxlen': self._maxlen,
                'options': self._options, 'plon': self._plon,
                'filter': self._filter.todict() if self._filter else None}

    @classmethod
    def fromdict(cls, data: dict) -> 'Trace':
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError('Expected "{}" type but got "{}"!'.format(
                cls.__name__, kind))
        f = data.pop('filter', None)
        return cls(filter=Filter.fromdict(f) if f is not None else None,
                   **data)

    def __repr__(self):
        return 'Trace(maxlen={}, options={}, filter={}, plon={})'.format(
            self._maxlen, self._options, self._filter, self._plon)


TRACE_EVENT_REFLECTION = Trace.TRACE_EVENT_REFLECTION
TRACE_EVENT_REFRACTION = Trace.TRACE_EVENT_REFRACTION
TRACE_EVENT_BOUNDARY_HIT = Trace.TRACE_EVENT_BOUNDARY_HIT
TRACE_EVENT_LAUNCH = Trace.TRACE_EVENT_LAUNCH
TRACE_EVENT_ABSORPTION = Trace.TRACE_EVENT_ABSORPTION
TRACE_EVENT_SCATTERING = Trace.TRACE_EVENT_SCATTERING
TRACE_EVENT_TERMINATION = Trace.TRACE_EVENT_TERMINATION
TRACE_EVENT_ESCAPE = Trace.TRACE_EVENT_ESCAPE
TRACE_EVENT_ALL = Trace.TRACE_EVENT_ALL
