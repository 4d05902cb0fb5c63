"""Sample layers (xopto/mcml/mclayer/layer.py:35-1006)."""
from typing import List

import numpy as np

from .. import _native as N
from ..mcbase.mcmaterial import _mut_terms
from ..mcbase import mcpf
from ..mcbase.mcpf import PfBase
from ..mcbase.mcutil import boundary


class Layer:
    def __init__(self, d: float, n: float, mua: float, mus: float,
                 pf: PfBase):
        self._d = float(d)
        self._n = float(n)
        self._mua = float(mua)
        self._mus = float(mus)
        self._pf = pf

    def _set_pf(self, pf):
        if type(self._pf) != type(pf):
            raise ValueError('The scattering phase function type of the '
                             'layer must not change!')
        self._pf = pf

    d = property(lambda self: self._d,
                 lambda self, v: setattr(self, '_d', float(v)), None,
                 'Layer thickness (m).')
    n = property(lambda self: self._n,
                 lambda self, v: setattr(self, '_n', float(v)), None,
                 'Refractive index.')
    mua = property(lambda self: self._mua,
                   lambda self, v: setattr(self, '_mua', float(v)), None,
                   'Absorption coefficient (1/m).')
    mus = property(lambda self: self._mus,
                   lambda self, v: setattr(self, '_mus', float(v)), None,
                   'Scattering coefficient (1/m).')
    pf = property(lambda self: self._pf, _set_pf, None, 'Phase function.')

    def pack(self, mc, target: N.XoLayer) -> N.XoLayer:
        """Fields independent of the neighbours (layer.py:304-354)."""
        inv_mut, mua_inv_mut = _mut_terms(self._mua, self._mus)
        target.thickness = self._d
        target.n = self._n
        target.mua = self._mua
        target.mus = self._mus
        target.inv_mut = inv_mut
        target.mua_inv_mut = mua_inv_mut
        self._pf.pack(mc, target.pf)
        return target

    def todict(self):
        """layer.py:356-361"""
        return {'d': self._d, 'n': self._n, 'mua': self._mua,
                'mus': self._mus, 'pf': self._pf.todict(), 'type': 'Layer'}

    @classmethod
    def fromdict(cls, data: dict) -> 'Layer':
        """layer.py:364-385"""
        data = dict(data)
        if data.pop('type') != 'Layer':
            raise ValueError('Cannot create a Layer instance from the data!')
        pf = mcpf.fromdict(data.pop('pf'))
        return cls(pf=pf, **data)

    def __repr__(self):
        return 'Layer(d={}, n={}, mua={}, mus={}, pf={})'.format(
            self._d, self._n, self._mua, self._mus, self._pf)


class Layers:
    def __init__(self, layers):
        if isinstance(layers, Layers):
            layers = layers.tolist()
        self._layers = list(layers)
        self.check()

    def check(self):
        if len(self._layers) < 3:
            raise ValueError('At least three layers are required, but got '
                             'only {:d}!'.format(len(self._layers)))
        if len(self._layers) > N.XO_MAX_LAYERS:
            raise ValueError('At most {} layers are supported!'.format(
                N.XO_MAX_LAYERS))
        pf_type = type(self._layers[1].pf)
        for layer in self._layers:
            if not isinstance(layer, Layer):
                raise TypeError('All the sample layers must be instances of '
                                'Layer but found {}!'.format(
                                    type(layer).__name__))
            if type(layer.pf) != pf_type:
                raise TypeError(
                    'All the sample layers must use the same scattering phase '
                    'function model! Found {} and {}!'.format(
                        pf_type.__name__, type(layer.pf).__name__))

    def tolist(self) -> List[Layer]:
        return list(self._layers)

    def todict(self) -> dict:
        """layer.py:1009-1014"""
        return {'layers': [layer.todict() for layer in self._layers],
                'type': 'Layers'}

    @classmethod
    def fromdict(cls, data: dict) -> 'Layers':
        """layer.py:1028-1046"""
        data = dict(data)
        if data.pop('type') != 'Layers':
            raise ValueError(
                'Cannot create a Layers instance from the given data!')
        return cls([Layer.fromdict(item) for item in data.pop('layers')])

    def __len__(self):
        return len(self._layers)

    def __getitem__(self, i):
        return self._layers[i]

    def __iter__(self):
        return iter(self._layers)

    def layer(self, index: int) -> Layer:
        return self._layers[index]

    def layer_index(self, z: float) -> int:
        index = len(self._layers) - 1
        bottom = 0.0
        for pos, layer in enumerate(self._layers):
            if pos > 0:
                bottom += layer.d
            if z < bottom:
                index = pos
                break
        return index

    def thickness(self) -> float:
        return float(sum(layer.d for layer in self._layers[1:-1]))

    pf_kind = property(lambda self: self._layers[1].pf.kind)

    def pack(self, mc, cfg: N.XoMcmlCfg):
        """Stack fields (layer.py:944-1006): z bounds accumulate through the
        binary32 values already stored, critical cosines per interface."""
        num = len(self._layers)
        cfg.num_layers = num
        for i, layer in enumerate(self._layers):
            t = cfg.layers[i]
            layer.pack(mc, t)
            cc_top = cc_bottom = 0.0
            if i > 0:
                cc_top = boundary.cos_critical(layer.n, self._layers[i - 1].n)
            if i + 1 < num:
                cc_bottom = boundary.cos_critical(layer.n,
                                                  self._layers[i + 1].n)
            if i == 0:
                t.top = -np.inf
                t.bottom = 0.0
                t.thickness = np.inf
            elif i == num - 1:
                t.top = cfg.layers[i - 1].bottom
                t.bottom = np.inf
                t.thickness = np.inf
            else:
                t.top = cfg.layers[i - 1].bottom
                t.bottom = float(cfg.layers[i - 1].bottom) + layer.d
            t.cos_critical_top = cc_top
            t.cos_critical_bottom = cc_bottom
        return cfg

    def __repr__(self):
        return 'Layers([{}])'.format(', '.join(repr(l) for l in self._layers))
