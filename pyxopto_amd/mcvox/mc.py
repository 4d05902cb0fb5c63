"""Voxelized-media Monte Carlo simulator: the drop-in for xopto.mcvox.mc.Mc
(xopto/mcvox/mc.py:89-1240). Each run packs the plugin objects into one
xo_mcvox_cfg and calls xo_mcvox_run of libxomc.so; the voxel grid is
uploaded (as uint8 material indices) only when it changed."""
import numpy as np

from .. import _native as N
from ..mcbase import mcdetector, mcfluence, mcmaterial, mcoptions, mcpf
from ..mcbase import mcsv, mctrace
from ..mcbase.mcutil import axis as _axis
from ..mcbase.mcworker import McWorker
from . import mcgeometry, mcsource

mcdetector.Axis = _axis.Axis
mcdetector.RadialAxis = _axis.RadialAxis
mcfluence.Axis = _axis.Axis
mcgeometry.Axis = _axis.Axis
mcsv.Axis = _axis.Axis

__all__ = ['Mc', 'mcgeometry', 'mcsource', 'mcdetector', 'mcpf', 'mcoptions',
           'mcfluence', 'mcmaterial', 'mctrace', 'mcsv']


class Mc(McWorker):
    def __init__(self, voxels: mcgeometry.Voxels, materials, source,
                 detectors: mcdetector.Detectors = None, trace=None,
                 fluence: mcfluence.Fluence = None, surface=None, types=None,
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
        super().__init__(options=options, rnginit=rnginit,
                         cl_devices=cl_devices,
                         cl_build_options=cl_build_options,
                         cl_profiling=cl_profiling, types=types)
        self._voxels = voxels
        self._materials = mcmaterial.Materials(materials)
        self._pf_type = type(self._materials[0].pf)
        self._source = source
        self._source_type = type(source)
        self._detectors = detectors
        self._detector_types = detectors.types() if detectors else None
        self._fluence = fluence
        self._rmax = float('inf')
        self._uploaded_shape = None

    def _get_rmax(self):
        return self._rmax

    def _set_rmax(self, rmax):
        rmax = float(rmax)
        if rmax == 0.0:
            raise ValueError('Simulation radius must be > 0.0 m!')
        self._rmax = rmax
    rmax = property(_get_rmax, _set_rmax, None,
                    'Maximum simulation radius from the source (m).')

    voxels = property(lambda self: self._voxels, None, None, 'Voxels.')
    materials = property(lambda self: self._materials, None, None,
                         'Materials.')
    source = property(lambda self: self._source, None, None, 'Source.')
    detectors = property(lambda self: self._detectors, None, None,
                         'Detectors.')
    fluence = property(lambda self: self._fluence, None, None, 'Fluence.')
    trace = property(lambda self: self._trace, None, None,
                     'Trace object if available.')

    def material(self, index: int) -> mcmaterial.Material:
        return self._materials[index]

    def material_at(self, position) -> mcmaterial.Material:
        ix, iy, iz = self._voxels.index(position)
        nz, ny, nx = self._voxels.shape
        ix, iy, iz = (min(max(ix, 0), nx - 1), min(max(iy, 0), ny - 1),
                      min(max(iz, 0), nz - 1))
        return self._materials[int(self._voxels.material[iz, iy, ix])]

    def pack(self):
        """(cfg, lut, naccum): detector bins first, then the fluence bins."""
        self._reset_allocations()
        if type(self._materials[0].pf) != self._pf_type:
            raise ValueError('The scattering phase function kind/type must '
                             'not change between simulation calls!')
        if type(self._source) != self._source_type:
            raise ValueError('The photon packet source type/kind must not '
                             'change between simulation calls!')
        cfg = N.XoMcvoxCfg()
        self._apply_options(cfg)
        cfg.pf_kind = self._materials.pf_kind
        cfg.rmax = self._rmax
        cfg.num_materials = len(self._materials)
        for i, m in enumerate(self._materials):
            m.pack(self, cfg.materials[i])
        self._voxels.pack(self, cfg.voxels)
        self._source.pack(self, cfg)
        if self._detectors is not None:
            if self._detectors.types() != self._detector_types:
                raise ValueError('Detector types must not change between '
                                 'simulation calls!')
            self._detectors.pack(self, cfg)
        if self._fluence is not None:
            self._fluence.pack(self, cfg.fluence)
        naccum = max(self._alloc_offset, 1)
        return cfg, self._lut_array(), naccum

    def run(self, nphotons: int, out=None, wgsize: int = None,
            maxthreads: int = None, copyseeds: bool = False, exportsrc=None,
            verbose: bool = False, static_schedule: bool = False,
            packet_seed: int = None):
        """Simulate nphotons packets; returns (trace, fluence, detectors).
        packet_seed: draw packet p from the generator keyed on
        (packet_seed, p), independent of lanes and schedule."""
        if out is None:
            out = (None, None, None)
        nphotons = self._check_nphotons(nphotons)
        cfg, lut, naccum = self.pack()
        flags = self._run_flags(static_schedule)
        if packet_seed is not None:
            cfg.packet_seed = int(packet_seed) & 0xFFFFFFFFFFFFFFFF
            cfg.first_packet = 0
            flags |= N.XO_FLAG_PACKET_KEYED
        upload = self._voxels.update_required() or \
            self._uploaded_shape != self._voxels.shape
        voxels = self._voxels.material if upload else None
        ctx = self._context()
        lanes = 0 if maxthreads is None else int(maxthreads)
        trace_res = out[0]
        if self._trace is not None:
            accum, rep, tdata, tn = ctx.run_mcvox(
                cfg, voxels, lut, nphotons, naccum, lanes, flags,
                trace=self._trace.pack())
            if trace_res is None:
                trace_res = mctrace.Trace(self._trace)
            trace_res.update_data(self, tdata, tn, nphotons)
        else:
            accum, rep = ctx.run_mcvox(cfg, voxels, lut, nphotons, naccum,
                                       lanes, flags)
        self._uploaded_shape = self._voxels.shape
        self._record_report(rep, nphotons)
        if copyseeds:
            self._copy_seeds()
        _, fluence_res, detectors_res = self.unpack(accum, nphotons, out)
        if verbose:
            print('McKernel processed {:,d} packets: {:.3f} ms'.format(
                nphotons, rep.kernel_ms))
        return trace_res, fluence_res, detectors_res

    def sampling_volume(self, trace, sv, wgsize: int = None,
                        maxthreads: int = None, exportsrc=None,
                        verbose: bool = False):
        """Add the sampling volume of `trace` to `sv` and return `sv`
        (xopto/mcvox/mc.py:1089-1240)."""
        if self._trace is None:
            raise RuntimeError(
                'This Monte Carlo simulator was build without the Trace '
                'functionality (the trace argument of Mc.__init__ was None)! '
                'Sampling volume analysis requires trace functionality!')
        data, n = trace.raw()
        accum, total, rep = self._context().sampling_volume(
            sv.pack(0), data, n, int(np.prod(sv.shape)))
        sv.update_data(self, [accum], total_weight=total)
        return sv

    def unpack(self, accum, nphotons: int, out=None):
        """(None, fluence, detectors) from the integer accumulator array
        of a run packed by the last pack() call."""
        if out is None:
            out = (None, None, None)
        fluence_res = out[1]
        if self._fluence is not None:
            if fluence_res is None:
                fluence_res = type(self._fluence)(self._fluence)
            off, n = self.allocation(self._fluence)
            fluence_res.update_data(self, [accum[off:off + n]], nphotons)
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
        return None, fluence_res, detectors_res
