"""Voxel materials (xopto/mcbase/mcmaterial.py:37-800)."""
from typing import List

from .. import _native as N
from . import mcpf
from .mcpf import PfBase


def _mut_terms(mua: float, mus: float):
    """inv_mut and mua_inv_mut as packed by the reference
    (xopto/mcbase/mcmaterial.py:255-266, xopto/mcml/mclayer/layer.py:330-341)."""
    mut = mua + mus
    inv_mut = 1.0 / mut if mut > 0.0 else float('inf')
    mua_inv_mut = 1.0 if mus == 0.0 else mua * inv_mut
    return inv_mut, mua_inv_mut


class Material:
    def __init__(self, n: float, mua: float, mus: float, pf: PfBase):
        self.n, self.mua, self.mus = float(n), float(mua), float(mus)
        self.pf = pf

    def pack(self, mc, target: N.XoMaterial):
        inv_mut, mua_inv_mut = _mut_terms(self.mua, self.mus)
        target.n = self.n
        target.mua = self.mua
        target.mus = self.mus
        target.inv_mut = inv_mut
        target.mua_inv_mut = mua_inv_mut
        self.pf.pack(mc, target.pf)
        return target

    def todict(self):
        """mcmaterial.py:278-283"""
        return {'n': self.n, 'mua': self.mua, 'mus': self.mus,
                'pf': self.pf.todict(), 'type': 'Material'}

    @classmethod
    def fromdict(cls, data: dict) -> 'Material':
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError('Expected "{}" type but got "{}"!'.format(
                cls.__name__, kind))
        pf = mcpf.fromdict(data.pop('pf'))
        return cls(pf=pf, **data)

    def __repr__(self):
        return 'Material(n={}, mua={}, mus={}, pf={})'.format(
            self.n, self.mua, self.mus, self.pf)


class Materials:
    def __init__(self, materials: List[Material]):
        if isinstance(materials, Materials):
            materials = list(materials)
        self._materials = list(materials)
        if not self._materials:
            raise ValueError('At least one material (the surrounding medium) '
                             'is required!')
        if len(self._materials) > N.XO_MAX_MATERIALS:
            raise ValueError('At most {} materials are supported!'.format(
                N.XO_MAX_MATERIALS))
        kinds = {type(m.pf) for m in self._materials}
        if len(kinds) != 1:
            raise TypeError('All the materials must use the same scattering '
                            'phase function model!')

    def __len__(self):
        return len(self._materials)

    def __getitem__(self, i):
        return self._materials[i]

    def __iter__(self):
        return iter(self._materials)

    pf_kind = property(lambda self: self._materials[0].pf.kind)

    def todict(self) -> dict:
        """mcmaterial.py:807-814"""
        return {'type': 'Materials',
                'materials': [m.todict() for m in self._materials]}

    @classmethod
    def fromdict(cls, data: dict) -> 'Materials':
        data = dict(data)
        kind = data.pop('type')
        if kind != cls.__name__:
            raise TypeError('Expected "{}" type but got "{}"!'.format(
                cls.__name__, kind))
        return cls([Material.fromdict(m) for m in data.pop('materials')])
