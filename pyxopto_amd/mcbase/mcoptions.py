"""Simulator options (xopto/mcbase/mcoptions.py:269-732). In the reference
these become compile-time #defines; here they are run-time fields of the
packed configuration. Options that only select memory spaces or math
variants of the OpenCL build are accepted and have no effect."""


class McOption:
    name = None

    def __init__(self, value):
        self.value = value

    def __repr__(self):
        return '{}({!r})'.format(type(self).__name__, self.value)


class McMethod(McOption):
    """Stepping method: albedo weight (0), albedo rejection (1) or
    microscopic Beer-Lambert (2) (xopto/mcbase/mcoptions.py:269-318).
    Layered and voxelized simulations take all three; batches of layered
    instances take albedo weight."""
    name = 'method'

    def __init__(self, value: int = 0):
        if isinstance(value, str) or value not in (0, 1, 2):
            raise ValueError('Allowed values are 0, 1 or 2!')
        super().__init__(int(value))


McMethod.albedo_weight = McMethod.aw = McMethod.default = McMethod(0)
McMethod.albedo_rejection = McMethod.ar = McMethod(1)
McMethod.microscopic_beer_lambert = McMethod.mbl = McMethod(2)


class McMinimumPacketWeight(McOption):
    name = 'w_min'

    def __init__(self, value: float = 1e-4):
        super().__init__(float(value))


class McPacketLotteryChance(McOption):
    name = 'lottery_chance'

    def __init__(self, value: float = 0.1):
        super().__init__(float(value))


class McUseLottery(McOption):
    name = 'use_lottery'

    def __init__(self, value: bool = True):
        super().__init__(bool(value))


McUseLottery.on = McUseLottery(True)
McUseLottery.off = McUseLottery(False)


class _NoEffect(McOption):
    """Accepted for API compatibility (OpenCL build detail)."""


class McUseNativeMath(_NoEffect):
    pass


McUseNativeMath.on = McUseNativeMath(True)
McUseNativeMath.off = McUseNativeMath(False)


class McFpLutMemory(_NoEffect):
    pass


McFpLutMemory.global_mem = McFpLutMemory('global')
McFpLutMemory.constant_mem = McFpLutMemory('constant')


class McMaterialMemory(_NoEffect):
    pass


McMaterialMemory.global_mem = McMaterialMemory('global')
McMaterialMemory.constant_mem = McMaterialMemory('constant')


DEFAULTS = {'method': 0, 'w_min': 1e-4, 'lottery_chance': 0.1,
            'use_lottery': True}


def resolve(options) -> dict:
    """Option values keyed by name; the last duplicate wins."""
    out = dict(DEFAULTS)
    for opt in options or []:
        if not isinstance(opt, McOption):
            raise TypeError('Options must be McOption instances!')
        if opt.name is not None:
            out[opt.name] = opt.value
    return out
