"""Layered-media Monte Carlo simulator: the drop-in for xopto.mcml.mc.Mc
(xopto/mcml/mc.py:77-1240).

    from pyxopto_amd.mcml import mc
    layers = mc.mclayer.Layers([...])
    mc_obj = mc.Mc(layers, mc.mcsource.Line(), mc.mcdetector.Detectors(...))
    mc_obj.rmax = 25e-3
    trace, fluence, detectors = mc_obj.run(1e6)

Each run packs the plugin objects into one xo_mcml_cfg and calls
xo_mcml_run of libxomc.so (one persistent gfx950 kernel launch).
"""
import numpy as np

from .. import _native as N
from ..mcbase import mcdetector, mcoptions, mcpf
from ..mcbase import mcfluence, mcsv, mctrace
from ..mcbase.mcutil import axis as _axis
from ..mcbase.mcutil import fiber as fiberutil
from ..mcbase.mcworker import McWorker
from . import mclayer, mcsource

mcfluence.Axis = _axis.Axis
mcsv.Axis = _axis.Axis

# Module layout of xopto.mcml.mc: the plugin modules are reachable from here.
mcdetector.Axis = _axis.Axis
mcdetector.RadialAxis = _axis.RadialAxis

__all__ = ['Mc', 'mclayer', 'mcsource', 'mcdetector', 'mcpf', 'mcoptions',
           'mcfluence', 'mctrace', 'mcsv', 'fiberutil']


class Mc(McWorker):
    def __init__(self, layers, source, detectors: mcdetector.Detectors = None,
                 trace=None, fluence=None, surface=None, types=None,
                 options=None, rnginit=None, cl_devices=None,
                 cl_build_options=None, cl_profiling: bool = False):
        if surface is not None:
            raise NotImplementedError(
                'Surface layouts are not part of this engine.')
        if trace is not None and not isinstance(trace, mctrace.Trace):
            raise TypeError('trace must be a Trace object')
        self._trace = trace
        if fluence is not None and not isinstance(
                fluence, (mcfluence.Fluence, mcfluence.FluenceRz)):
            raise TypeError('fluence must be a Fluence or FluenceRz object')
        self._fluence = fluence
        super().__init__(options=options, rnginit=rnginit,
                         cl_devices=cl_devices,
                         cl_build_options=cl_build_options,
                         cl_profiling=cl_profiling, types=types)
        self._layers = mclayer.Layers(layers)
        self._pf_type = type(self._layers[1].pf)
        self._source = source
        self._source_type = type(source)
        if not isinstance(source, (mcsource.Line, mcsource.UniformFiberNI)):
            raise NotImplementedError('Only the Line, UniformFiberNI and '
                                      'UniformFiber '
                                      'sources are implemented for layered '
                                      'media.')
        self._detectors = detectors
        self._detector_types = detectors.types() if detectors else None
        self._rmax = float('inf')

    # ---------------- properties (xopto/mcml/mc.py:1190-1290) ----------
    def _get_rmax(self):
        return self._rmax

    def _set_rmax(self, rmax):
        rmax = float(rmax)
        if rmax == 0.0:
            raise ValueError('Simulation radius must be > 0.0 m!')
        self._rmax = rmax
    rmax = property(_get_rmax, _set_rmax, None,
                    'Maximum simulation radius measured from the photon '
                    'packet source position (m).')

    layers = property(lambda self: self._layers, None, None, 'Layer stack.')
    source = property(lambda self: self._source, None, None, 'Source.')
    fluence = property(lambda self: self._fluence, None, None,
                       'Fluence/deposition accumulator or None.')
    trace = property(lambda self: self._trace, None, None,
                     'Trace object if available.')
    detectors = property(lambda self: self._detectors, None, None,
                         'Detectors.')
    options = property(lambda self: list(self._options), None, None,
                       'Simulator options.')

    def layer(self, index: int) -> mclayer.Layer:
        return self._layers.layer(index)

    def layer_index(self, z: float) -> int:
        return self._layers.layer_index(z)

    # ---------------- packing (xopto/mcml/mc.py:466-529) ----------------
    def pack(self) -> tuple:
        """Pack the configuration: (cfg, lut, naccum)."""
        self._reset_allocations()
        if type(self._layers[1].pf) != self._pf_type:
            raise ValueError('The scattering phase function kind/type must '
                             'not change between simulation calls!')
        if type(self._source) != self._source_type:
            raise ValueError('The photon packet source type/kind must not '
                             'change between simulation calls!')
        cfg = N.XoMcmlCfg()
        self._apply_options(cfg)
        cfg.pf_kind = self._layers.pf_kind
        cfg.rmax = self._rmax
        self._layers.pack(self, cfg)
        self._source.pack_cfg(self, cfg)
        if self._detectors is not None:
            if self._detectors.types() != self._detector_types:
                raise ValueError('Detector types must not change between '
                                 'simulation calls!')
            self._detectors.pack(self, cfg)
        if self._fluence is not None:
            self._fluence.pack(self, cfg.fluence)
        naccum = max(self._alloc_offset, 1)
        return cfg, self._lut_array(), naccum

    # ---------------- run (xopto/mcml/mc.py:730-1018) -------------------
    def run(self, nphotons: int, out=None, wgsize: int = None,
            maxthreads: int = None, copyseeds: bool = False,
            exportsrc=None, verbose: bool = False,
            static_schedule: bool = False, packet_seed: int = None):
        """Simulate nphotons packets; returns (trace, fluence, detectors).

        maxthreads caps the lanes (default: the persistent grid of the GPU).
        static_schedule=True makes lane g run packets g, g+L, ... so that the
        result is a deterministic function of the seeds and the lane count.
        packet_seed: draw packet p from the generator keyed on
        (packet_seed, p) instead of the per-lane streams, so the result does
        not depend on the lanes or schedule (the batch and serving engines
        key every packet this way).
        """
        if out is None:
            out = (None, None, None)
        nphotons = self._check_nphotons(nphotons)
        cfg, lut, naccum = self.pack()
        flags = self._run_flags(static_schedule)
        if packet_seed is not None:
            cfg.packet_seed = int(packet_seed) & 0xFFFFFFFFFFFFFFFF
            cfg.first_packet = 0
            flags |= N.XO_FLAG_PACKET_KEYED
        ctx = self._context()
        lanes = 0 if maxthreads is None else int(maxthreads)
        trace_res = out[0]
        if self._trace is not None:
            accum, rep, tdata, tn = ctx.run_mcml_trace(
                cfg, lut, nphotons, naccum, self._trace.pack(), lanes, flags)
            if trace_res is None:
                trace_res = mctrace.Trace(self._trace)
            trace_res.update_data(self, tdata, tn, nphotons)
        else:
            accum, rep = ctx.run_mcml(cfg, lut, nphotons, naccum, lanes,
                                      flags)
        self._record_report(rep, nphotons)
        if copyseeds:
            self._copy_seeds()
        self._last_accum = accum
        detectors_res = out[2]
        if self._detectors is not None:
            if detectors_res is None:
                detectors_res = mcdetector.Detectors(self._detectors)
            for name, det in self._detectors.items():
                if det is None:
                    continue
                off, n = self.allocation(det)
                getattr(detectors_res, name).update_data(
                    self, [accum[off:off + n]], nphotons)
        if verbose:
            print('McKernel processed {:,d} packets in {:,d} threads: '
                  '{:.3f} ms'.format(nphotons, rep.threads_used,
                                     rep.kernel_ms))
        fluence_res = out[1]
        if self._fluence is not None:
            if fluence_res is None:
                fluence_res = type(self._fluence)(self._fluence)
            off, n = self.allocation(self._fluence)
            fluence_res.update_data(self, [accum[off:off + n]], nphotons)
        return trace_res, fluence_res, detectors_res

    # ---------------- sampling volume (xopto/mcml/mc.py:1040-1200) ------
    def sampling_volume(self, trace: mctrace.Trace, sv: mcsv.SamplingVolume,
                        wgsize: int = None, maxthreads: int = None,
                        exportsrc=None, verbose: bool = False):
        """Add the sampling volume of `trace` to `sv` and return `sv`
        (one xo_sampling_volume_run launch)."""
        if self._trace is None:
            raise RuntimeError(
                'This Monte Carlo simulator was build without the Trace '
                'functionality (the trace argument of Mc.__init__ was None)! '
                'Sampling volume analysis requires trace functionality!')
        data, n = trace.raw()
        naccum = int(np.prod(sv.shape))
        accum, total, rep = self._context().sampling_volume(
            sv.pack(0), data, n, naccum)
        sv.update_data(self, [accum], total_weight=total)
        if verbose:
            print('SamplingVolume processed {:,d} packet traces: {:.3f} '
                  'ms'.format(len(n), rep.kernel_ms))
        return sv

    def __repr__(self):
        return 'Mc({}, {}, {})'.format(self._layers, self._source,
                                       self._detectors)
