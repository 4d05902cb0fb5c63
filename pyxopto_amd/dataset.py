"""Simulation results as dataset files in the reference's layouts.

Three record layouts, each a dict that ``save`` writes with
``np.savez_compressed`` exactly as the reference's scripts do (plugin
configurations as their ``todict()`` dicts, which numpy stores as 0-d object
arrays; results as float arrays):

- layered sample, one file per (mua, musr) point
  (xopto/dataset/template/mcml/analytical_pf.template.py:226-246): layers,
  source, surface, detectors, reflectance, transmittance,
  total_reflectance, total_transmittance, method; file name
  ``mua-<mua 1/cm>-musr-<musr 1/cm>-invcm.npz`` with '.' -> '_'
  (:220-224);
- voxelized fluence (xopto/dataset/template/mcvox/fluence_skin_vessel.
  template.py:204-220): num_packets, batch_packets, voxels, voxel_data,
  materials, source, detectors, fluence, fluence_data, reflectance,
  transmittance, method; file name ``2-layer-skin-<d>um-vessel-<z>um-depth-
  deposition.npz`` (:181-182);
- one run with its full configuration (McRunHelper.collect_one,
  xopto/mcml/mcrun/helper.py:434-545, 588-624): epoch_timestamp, mc
  (source, surface, layers, detectors, fluence, trace, rmax, run_report),
  num_packets, detectors (reflectance, transmittance, specular), fluence
  (data), trace.

``load`` reads any of them back without executing anything from the file:
plain arrays through ``np.load(allow_pickle=False)``, object entries through
a static decoder of their pickle stream (``util.staticpickle``). ``objects``
rebuilds the plugin objects from a loaded record with the ``fromdict``
class methods.
"""
import os
import time
import zipfile
from typing import Dict, Optional

import numpy as np

from .util import staticpickle

METHOD_NAMES = {0: 'aw', 1: 'ar', 2: 'mbl'}


def _method_name(method) -> str:
    if isinstance(method, str):
        if method not in METHOD_NAMES.values():
            raise ValueError('Unknown method "{}"!'.format(method))
        return method
    value = getattr(method, 'value', method)
    return METHOD_NAMES[int(value)]


def _todict(obj):
    return obj.todict() if obj is not None else None


# ---------------------------------------------------------------- layered
def mcml_filename(mua: float, musr: float) -> str:
    """analytical_pf.template.py:220-224 (coefficients in 1/m)."""
    name = 'mua-{:.2f}-musr-{:.2f}-invcm'.format(mua * 1e-2, musr * 1e-2)
    return name.replace('.', '_') + '.npz'


def mcml_record(sim, detectors_res, method='aw') -> Dict:
    """Record of one layered simulation. `sim` is anything with layers,
    source and detectors (mcml.Mc, mcml.batch.Instance); `detectors_res`
    the Detectors result of its run."""
    top, bottom = detectors_res.top, detectors_res.bottom
    nphotons = max(top.nphotons if top is not None else 0, 1)
    return {
        'layers': sim.layers.todict(),
        'source': sim.source.todict(),
        'surface': None,
        'detectors': sim.detectors.todict(),
        'reflectance': top.reflectance if top is not None else None,
        'transmittance': bottom.transmittance if bottom is not None
        else None,
        # both totals divide by the top detector's packet count, as the
        # reference's script does (analytical_pf.template.py:238-241)
        'total_reflectance': (top.raw.sum() / nphotons)
        if top is not None else None,
        'total_transmittance': (bottom.raw.sum() / nphotons)
        if bottom is not None else None,
        'method': _method_name(method),
    }


# ---------------------------------------------------------------- voxels
def mcvox_fluence_filename(vessel_diameter: float, depth: float) -> str:
    """fluence_skin_vessel.template.py:181-182 (lengths in m)."""
    name = '2-layer-skin-{:.0f}um-vessel-{:.0f}um-depth-deposition'.format(
        vessel_diameter * 1e6, depth * 1e6)
    return name.replace('.', '_') + '.npz'


def mcvox_fluence_record(sim, fluence_res, detectors_res, num_packets: int,
                         batch_packets: Optional[int] = None,
                         method='aw') -> Dict:
    """Record of one voxelized fluence simulation (mcvox.Mc `sim`)."""
    return {
        'num_packets': int(num_packets),
        'batch_packets': int(batch_packets if batch_packets is not None
                             else num_packets),
        'voxels': sim.voxels.todict(),
        'voxel_data': np.asarray(sim.voxels.material),
        'materials': sim.materials.todict(),
        'source': sim.source.todict(),
        'detectors': sim.detectors.todict(),
        'fluence': sim.fluence.todict(),
        'fluence_data': fluence_res.data,
        'reflectance': detectors_res.top.reflectance,
        'transmittance': detectors_res.bottom.transmittance,
        'method': _method_name(method),
    }


# ---------------------------------------------------------------- runs
def run_record(sim, nphotons: int, fluence_res=None, detectors_res=None,
               run_report: Optional[Dict] = None) -> Dict:
    """McRunHelper.collect_one layout of one layered run."""
    dets = {'reflectance': None, 'transmittance': None, 'specular': None}
    if detectors_res is not None:
        if detectors_res.top is not None:
            dets['reflectance'] = detectors_res.top.reflectance
        if detectors_res.bottom is not None:
            # helper.py:497 reads the bottom detector's reflectance
            # property, which equals its transmittance
            dets['transmittance'] = detectors_res.bottom.transmittance
        if detectors_res.specular is not None:
            dets['specular'] = detectors_res.specular.reflectance
    return {
        'epoch_timestamp': time.time(),
        'mc': {
            'source': sim.source.todict(),
            'surface': None,
            'layers': sim.layers.todict(),
            'detectors': _todict(sim.detectors),
            'fluence': _todict(getattr(sim, 'fluence', None)),
            'trace': None,
            'rmax': float(sim.rmax),
            'run_report': dict(run_report or {}),
        },
        'num_packets': int(nphotons),
        'detectors': dets,
        'fluence': {'data': fluence_res.data if fluence_res is not None
                    else None},
        'trace': None,
    }


# ---------------------------------------------------------------- files
def save(filename: str, record: Dict) -> str:
    """Write `record` with np.savez_compressed (the reference's call);
    creates the parent folder. Returns the file name."""
    folder = os.path.dirname(os.path.abspath(filename))
    os.makedirs(folder, exist_ok=True)
    np.savez_compressed(filename, **record)
    return filename


def _read_member(zf: zipfile.ZipFile, member: str):
    with zf.open(member) as f:
        version = np.lib.format.read_magic(f)
        if version == (1, 0):
            shape, fortran, dtype = np.lib.format.read_array_header_1_0(f)
        elif version in ((2, 0), (3, 0)):
            shape, fortran, dtype = np.lib.format.read_array_header_2_0(f)
        else:
            raise ValueError('Unsupported .npy version {}'.format(version))
        if dtype.hasobject:
            arr = staticpickle.loads(f.read())
            if not isinstance(arr, np.ndarray):
                raise ValueError('{}: object entry is not an array'.format(
                    member))
            return arr
    with zf.open(member) as f:
        return np.lib.format.read_array(f, allow_pickle=False)


def load(filename: str) -> Dict:
    """Record of a dataset file; 0-d entries (dicts, None, strings,
    scalars) are returned as Python values, other arrays as arrays."""
    out = {}
    with zipfile.ZipFile(filename) as zf:
        for member in zf.namelist():
            if not member.endswith('.npy'):
                continue
            arr = _read_member(zf, member)
            if arr.shape == ():
                value = arr[()]
                if isinstance(value, np.str_):
                    value = str(value)
                elif isinstance(value, np.generic):
                    value = value.item()
            else:
                value = arr
            out[member[:-4]] = value
    return out


def objects(record: Dict) -> Dict:
    """Plugin objects of a loaded record (layers, materials, voxels,
    source, detectors, fluence), rebuilt with their fromdict methods."""
    from .mcbase import mcdetector, mcfluence, mcmaterial
    from .mcml import mclayer
    from .mcml import mcsource as mcml_source
    from .mcvox import mcgeometry
    from .mcvox import mcsource as mcvox_source
    cfg = record.get('mc', record)
    out = {}
    if cfg.get('layers') is not None:
        out['layers'] = mclayer.Layers.fromdict(cfg['layers'])
    if cfg.get('materials') is not None:
        out['materials'] = mcmaterial.Materials.fromdict(cfg['materials'])
    if cfg.get('voxels') is not None:
        out['voxels'] = mcgeometry.Voxels.fromdict(cfg['voxels'])
        data = record.get('voxel_data')
        if data is not None:
            out['voxels'].material[...] = data
    if cfg.get('source') is not None:
        mod = mcvox_source if 'voxels' in out else mcml_source
        out['source'] = mod.fromdict(cfg['source'])
    if cfg.get('detectors') is not None:
        out['detectors'] = mcdetector.Detectors.fromdict(cfg['detectors'])
    if cfg.get('fluence') is not None and 'type' in cfg['fluence']:
        out['fluence'] = mcfluence.fromdict(cfg['fluence'])
    return out
