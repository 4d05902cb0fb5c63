"""Multimode step-index optical fiber: the geometry and optics a fiber
source packs (xopto/mcbase/mcutil/fiber.py:33-222)."""


class MultimodeFiber:
    """Core diameter (m), cladding diameter (m), core refractive index and
    numerical aperture of an ideal step-index fiber."""

    @staticmethod
    def compute_na(ncore: float, ncladding: float) -> float:
        return (ncore ** 2 - ncladding ** 2) ** 0.5

    @staticmethod
    def compute_ncladding(ncore: float, na: float) -> float:
        return (ncore ** 2 - na ** 2) ** 0.5

    def __init__(self, dcore: float, dcladding: float, ncore: float,
                 na: float):
        self._dcore = float(dcore)
        self._dcladding = float(dcladding)
        self._ncore = float(ncore)
        self._na = float(na)

    def validate(self):
        if self._dcore > self._dcladding:
            raise ValueError('Fiber core diameter must be smaller than the '
                             'fiber cladding diameter!')

    def _set_dcore(self, v):
        self._dcore = float(v)

    def _set_dcladding(self, v):
        self._dcladding = float(v)

    def _set_ncore(self, v):
        self._ncore = float(v)

    def _set_na(self, v):
        self._na = float(v)

    dcore = property(lambda self: self._dcore, _set_dcore, None,
                     'Diameter (m) of the fiber core.')
    dcladding = property(lambda self: self._dcladding, _set_dcladding, None,
                         'Diameter (m) of the fiber cladding.')
    ncore = property(lambda self: self._ncore, _set_ncore, None,
                     'Refractive index of the fiber core.')
    na = property(lambda self: self._na, _set_na, None,
                  'Numerical aperture of the fiber core.')
    ncladding = property(
        lambda self: MultimodeFiber.compute_ncladding(self._ncore, self._na),
        None, None, 'Refractive index of the cladding, sqrt(ncore^2 - na^2).')

    def todict(self) -> dict:
        return {'dcore': self._dcore, 'dcladding': self._dcladding,
                'ncore': self._ncore, 'na': self._na,
                'type': 'MultimodeFiber'}

    @classmethod
    def fromdict(cls, data: dict) -> 'MultimodeFiber':
        data = dict(data)
        data.pop('type', None)
        return cls(**data)

    def __repr__(self):
        return 'MultimodeFiber(dcore={}, dcladding={}, ncore={}, na={})'.format(
            self._dcore, self._dcladding, self._ncore, self._na)
