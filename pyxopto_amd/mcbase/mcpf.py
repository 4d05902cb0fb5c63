"""Device phase-function samplers as plugin objects: Hg
(xopto/mcbase/mcpf/hg.py:25-140), MHg (mcpf/mhg.py), Lut and LutEx
(mcpf/lut.py:30-300). Each packs its parameters into the xo_pf struct; the
kernels specialise on the kind. todict/fromdict keep the reference's dict
schema ({'type': class name, constructor arguments...}; hg.py:160-164,
mhg.py:186-190, lut.py:258-263, 317-323, pfbase.py:36-48)."""
import numpy as np

from .. import _native as N
from . import pf as hostpf


class PfBase:
    kind = None

    def pack(self, mc, target: N.XoPf) -> N.XoPf:
        raise NotImplementedError

    def todict(self) -> dict:
        raise NotImplementedError

    @classmethod
    def fromdict(cls, data: dict) -> 'PfBase':
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError(
                'Cannot initialize an instance of {:s} scattering phase '
                'function from the data of "{}"!'.format(cls.__name__, kind))
        return cls(**data)


class Hg(PfBase):
    kind = N.XO_PF_HG

    def __init__(self, g: float):
        self._g = 0.0
        self.g = g

    def _get_g(self):
        return self._g

    def _set_g(self, g):
        self._g = min(max(float(g), -1.0), 1.0)
    g = property(_get_g, _set_g, None, 'Anisotropy factor.')

    def pf(self):
        return hostpf.Hg(self._g)

    def pack(self, mc, target):
        target.g = self._g
        return target

    def todict(self):
        return {'g': self._g, 'type': 'Hg'}

    def __repr__(self):
        return 'Hg(g={})'.format(self._g)


class MHg(PfBase):
    kind = N.XO_PF_MHG

    def __init__(self, g: float, b: float):
        self._g = min(max(float(g), -1.0), 1.0)
        self._b = min(max(float(b), 0.0), 1.0)

    g = property(lambda self: self._g, None, None, 'Anisotropy of the HG term.')
    b = property(lambda self: self._b, None, None, 'Weight of the HG term.')

    def pf(self):
        return hostpf.MHg(self._g, self._b)

    def pack(self, mc, target):
        target.g = self._g
        target.beta = self._b
        return target

    def todict(self):
        return {'g': self._g, 'b': self._b, 'type': 'MHg'}

    def __repr__(self):
        return 'MHg(g={}, b={})'.format(self._g, self._b)


class Lut(PfBase):
    kind = N.XO_PF_LUT

    def __init__(self, params, lut):
        self._params = np.zeros(3)
        self._params[:] = params
        self._lut = np.asarray(lut, dtype=np.float64)
        self._offset = 0

    params = property(lambda self: self._params, None, None,
                      'Index transform parameters (a, b, c).')
    lut = property(lambda self: self._lut, None, None, 'Lookup table.')
    size = property(lambda self: self._lut.size, None, None, 'Table size.')
    offset = property(lambda self: self._offset, None, None,
                      'Offset of the table in the LUT array.')

    def pack(self, mc, target):
        self._offset = mc.append_r_lut(self._lut)
        target.a = self._params[0]
        target.b = self._params[1]
        target.c = self._params[2]
        target.offset = self._offset
        target.size = self._lut.size
        return target

    def todict(self):
        return {'params': self._params.tolist(), 'lut': self._lut.tolist(),
                'type': 'Lut'}

    def __repr__(self):
        return 'Lut(params=({}, {}, {}), size={})'.format(
            *self._params, self._lut.size)


class LutEx(Lut):
    """Lut built from a host phase function: LutEx(MHg, (g, b), 2000)."""

    def __init__(self, pftype, pfargs, lutsize: int = 2000, **kwargs):
        if isinstance(pftype, str):
            pftype = getattr(hostpf, pftype)
        self._pfargs = tuple(pfargs)
        self._kwargs = dict(kwargs)
        self._pf_obj = pftype(*pfargs)
        params, lut = self._pf_obj.mclut(lutsize, **kwargs)
        super().__init__(params, lut)

    pfargs = property(lambda self: self._pfargs, None, None,
                      'Phase function arguments.')

    def todict(self):
        return {'pftype': type(self._pf_obj).__name__,
                'pfargs': list(self._pfargs), 'lutsize': int(self._lut.size),
                'kwargs': dict(self._kwargs), 'type': 'LutEx'}

    @classmethod
    def fromdict(cls, data: dict) -> 'LutEx':
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError(
                'Cannot initialize an instance of {:s} scattering phase '
                'function from the data of "{}"!'.format(cls.__name__, kind))
        kwargs = dict(data.pop('kwargs', None) or {})
        return cls(data['pftype'], data['pfargs'], data['lutsize'], **kwargs)

    def pf(self):
        return self._pf_obj


def fromdict(data: dict) -> PfBase:
    """Phase function object of any kind from its todict() data."""
    kinds = {'Hg': Hg, 'MHg': MHg, 'Lut': Lut, 'LutEx': LutEx}
    kind = data.get('type')
    if kind not in kinds:
        raise TypeError('Scattering phase function "{}" not '
                        'implemented'.format(kind))
    return kinds[kind].fromdict(data)
