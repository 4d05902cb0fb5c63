"""Batched layered-media simulations: many Mc-like instances in one
persistent launch (xo_batch_run). The reference's batch workflow calls
Mc.run once per instance (xopto/mcml/mcrun/helper.py:640-686); here every
instance keeps its own layers, phase functions, detectors and packet budget,
and the GPU works through a device queue of (instance, packet chunk) items.

Packet p of an instance draws from a generator keyed on (instance seed, p),
so an instance's integer accumulators do not depend on what else runs in
the batch or on the launch shape.
"""
from typing import List, Optional

import numpy as np

from .. import _native as N
from ..mcbase import mcdetector
from ..mcbase.mcworker import McWorker
from . import mclayer, mcsource


class Instance:
    """One layered simulation: layers, source, detectors, optional fluence
    accumulator (Fluence or FluenceRz), packet budget."""

    def __init__(self, layers, source=None, detectors=None,
                 nphotons: int = 100000, seed: Optional[int] = None,
                 rmax: float = float('inf'), fluence=None):
        self.layers = mclayer.Layers(layers)
        if len(self.layers) > N.XO_BATCH_MAX_LAYERS:
            raise ValueError('A batch instance supports at most {} layers '
                             'including the surrounding media.'.format(
                                 N.XO_BATCH_MAX_LAYERS))
        self.source = source if source is not None else mcsource.Line()
        self.detectors = detectors
        self.nphotons = int(nphotons)
        self.seed = seed
        self.rmax = float(rmax)
        self.fluence = fluence

    def layer(self, index: int):
        return self.layers.layer(index)


class _InstancePacker:
    """The mc object an instance's plugins pack against: layer access from
    the instance, LUT array and accumulator allocation from the batch."""

    def __init__(self, batch, inst):
        self._batch, self._inst = batch, inst
        self.allocations = {}

    def layer(self, index):
        return self._inst.layer(index)

    def append_r_lut(self, data):
        return self._batch.append_r_lut(data)

    def allocate_accumulator(self, owner, n):
        offset = self._batch.allocate_accumulator(owner, n)
        self.allocations[id(owner)] = (offset, int(n))
        return offset


class Batch(McWorker):
    def __init__(self, instances: List[Instance], options=None, rnginit=None,
                 cl_devices=None, cl_build_options=None):
        super().__init__(options=options, rnginit=rnginit,
                         cl_devices=cl_devices,
                         cl_build_options=cl_build_options)
        self.instances = list(instances)
        if not self.instances:
            raise ValueError('A batch needs at least one instance.')

    def instance_seed(self, index: int) -> int:
        inst = self.instances[index]
        if inst.seed is not None:
            return int(inst.seed) & 0xFFFFFFFFFFFFFFFF
        return (self._rnginit * 0x9E3779B97F4A7C15 + index) & \
            0xFFFFFFFFFFFFFFFF

    def pack(self):
        """(XoBatchInstance array, lut, naccum)."""
        self._reset_allocations()
        arr = (N.XoBatchInstance * len(self.instances))()
        self._packers = []
        for i, inst in enumerate(self.instances):
            t = arr[i]
            packer = _InstancePacker(self, inst)
            self._packers.append(packer)
            cfg = N.XoMcmlCfg()
            inst.layers.pack(packer, cfg)
            t.num_layers = cfg.num_layers
            for l in range(cfg.num_layers):
                t.layers[l] = cfg.layers[l]
            t.pf_kind = inst.layers.pf_kind
            t.rmax = inst.rmax
            inst.source.pack_cfg(packer, t)
            base = self._alloc_offset
            if inst.detectors is not None:
                inst.detectors.pack(packer, t)
            if inst.fluence is not None:
                inst.fluence.pack(packer, t.fluence)
            t.accum_base = base
            t.accum_count = self._alloc_offset - base
            t.npackets = inst.nphotons
            t.seed = self.instance_seed(i)
        return arr, self._lut_array(), max(self._alloc_offset, 1)

    def opts(self, chunk: int = 2048, block: int = 256,
             blocks_per_cu: int = 0, refill_min: int = 0,
             defer_min: int = 0) -> N.XoBatchOpts:
        o = N.XoBatchOpts()
        o.chunk, o.block, o.blocks_per_cu = chunk, block, blocks_per_cu
        o.refill_min = refill_min
        o.defer_min = defer_min
        o.w_min = float(self._opts['w_min'])
        o.lottery_chance = float(self._opts['lottery_chance'])
        o.use_lottery = int(bool(self._opts['use_lottery']))
        o.method = int(self._opts['method'])
        o.flags = N.XO_FLAG_FAST_MATH if self._fast_math else 0
        return o

    def run(self, opts: N.XoBatchOpts = None, packed=None,
            fluence: bool = False):
        """Run every instance; returns (list of Detectors, run report), or
        with fluence=True (list of (fluence, detectors), run report) — the
        last two items of Mc.run's result per instance, fluence None for an
        instance without a fluence accumulator."""
        arr, lut, naccum = packed if packed is not None else self.pack()
        if opts is None:
            from .autotune import tuned_opts
            opts = tuned_opts(self, (arr, lut, naccum))
        accum, rep = self._context().run_batch(arr, lut, opts, naccum)
        self._record_report(rep, int(sum(i.nphotons for i in self.instances)))
        if fluence:
            return list(zip(self.fluence_results(accum),
                            self.results(accum))), rep
        return self.results(accum), rep

    def plan(self, opts: N.XoBatchOpts = None, packed=None) -> 'BatchRunPlan':
        """The batch uploaded once with its launch captured as a graph
        (xo_batch_plan_create); BatchRunPlan.run() replays it and returns
        what run() returns, the same integers."""
        arr, lut, naccum = packed if packed is not None else self.pack()
        if opts is None:
            from .autotune import tuned_opts
            opts = tuned_opts(self, (arr, lut, naccum))
        return BatchRunPlan(self, self._context().batch_plan(arr, lut, opts,
                                                             naccum))

    def fluence_results(self, accum) -> list:
        """Each instance's fluence object filled from the accumulators (None
        where the instance has none)."""
        out = []
        for inst, packer in zip(self.instances, self._packers):
            if inst.fluence is None:
                out.append(None)
                continue
            off, n = packer.allocations[id(inst.fluence)]
            res = type(inst.fluence)(inst.fluence)
            res.update_data(self, [accum[off:off + n]], inst.nphotons)
            out.append(res)
        return out

    def results(self, accum) -> List[mcdetector.Detectors]:
        out = []
        for inst, packer in zip(self.instances, self._packers):
            if inst.detectors is None:
                out.append(None)
                continue
            res = mcdetector.Detectors(inst.detectors)
            for name, det in inst.detectors.items():
                if det is None:
                    continue
                off, n = packer.allocations[id(det)]
                getattr(res, name).update_data(self, [accum[off:off + n]],
                                               inst.nphotons)
            out.append(res)
        return out


class BatchRunPlan:
    """A Batch whose device tables and captured launch are kept for repeated
    runs (Batch.plan)."""

    def __init__(self, batch: Batch, plan: N.BatchPlan):
        self._batch, self._plan = batch, plan
        self._npk = int(sum(i.nphotons for i in batch.instances))

    def run(self):
        """(list of Detectors, run report), as Batch.run."""
        accum, rep = self._plan.run()
        self._batch._record_report(rep, self._npk)
        return self._batch.results(accum), rep

    def run_raw(self):
        """(accumulator array, run report) without building result objects."""
        return self._plan.run()

    def close(self):
        self._plan.close()
