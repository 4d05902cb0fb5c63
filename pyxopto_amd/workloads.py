"""Synthetic workloads used by the benchmark, the smoke test and the tests.

- ``slab_1mm``: one turbid slab 1 mm thick (n 1.37, mua 1/cm, mus 100/cm,
  HG g 0.8) between air (n 1.0) and glass (n 1.45), pencil beam, radial
  reflectance 0-5 mm in 100 bins, total transmittance and specular
  reflectance; 1e6 packets.
- ``skin_vessel_200``: 200^3 voxels of 10 um; epidermis (voxel centres at
  z <= 100 um: mua 16/cm, mus 400/cm, g 0.8, n 1.34) over dermis (mua
  0.5/cm, mus 250/cm, g 0.9, n 1.40) with a horizontal vessel along y,
  radius 100 um, axis at 400 um depth (mua 15/cm, mus 300/cm, g 0.98,
  n 1.40 as the dermis); Gaussian beam sigma 50 um at normal incidence;
  energy deposition in every voxel; 1e8 packets. The surrounding medium is
  n 1.0 with no absorption or scattering.
"""
import numpy as np

from .mcml import mc as mcml
from .mcvox import mc as mcvox

SLAB_PACKETS = 1_000_000
SKIN_VESSEL_PACKETS = 100_000_000


def slab_1mm(rnginit: int = 0x5EED000000000001):
    L = mcml.mclayer
    layers = L.Layers([
        L.Layer(d=np.inf, n=1.0, mua=0.0, mus=0.0, pf=mcml.mcpf.Hg(0.8)),
        L.Layer(d=1e-3, n=1.37, mua=100.0, mus=1e4, pf=mcml.mcpf.Hg(0.8)),
        L.Layer(d=np.inf, n=1.45, mua=0.0, mus=0.0, pf=mcml.mcpf.Hg(0.8))])
    D = mcml.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, 5e-3, 100)),
                      bottom=D.Total(), specular=D.Total())
    return mcml.Mc(layers, mcml.mcsource.Line(), det, rnginit=rnginit)


def skin_vessel(n: int = 200, voxel: float = 10e-6,
                rnginit: int = 0x5EED000000000001, epidermis: float = 100e-6,
                vessel_depth: float = 400e-6, vessel_radius: float = 100e-6):
    half = 0.5 * n * voxel
    M = mcvox.mcmaterial
    materials = M.Materials([
        M.Material(n=1.0, mua=0.0, mus=0.0, pf=mcvox.mcpf.Hg(0.8)),
        M.Material(n=1.34, mua=1600.0, mus=4e4, pf=mcvox.mcpf.Hg(0.8)),
        M.Material(n=1.40, mua=50.0, mus=2.5e4, pf=mcvox.mcpf.Hg(0.9)),
        M.Material(n=1.40, mua=1500.0, mus=3e4, pf=mcvox.mcpf.Hg(0.98))])
    G = mcvox.mcgeometry
    ax = (G.Axis(-half, half, n), G.Axis(-half, half, n),
          G.Axis(0.0, 2 * half, n))
    voxels = G.Voxels(*ax)
    fluence = mcvox.mcfluence.Fluence(G.Axis(-half, half, n),
                                      G.Axis(-half, half, n),
                                      G.Axis(0.0, 2 * half, n))
    D = mcvox.mcdetector
    det = D.Detectors(top=D.Radial(D.RadialAxis(0.0, half, 100)),
                      bottom=D.Total(), specular=D.Total())
    sim = mcvox.Mc(voxels, materials, mcvox.mcsource.GaussianBeam(50e-6), det,
                   fluence=fluence, rnginit=rnginit)
    z, _, x = sim.voxels.meshgrid()
    mat = sim.voxels.material
    mat[...] = 2
    mat[z <= epidermis] = 1
    mat[x ** 2 + (z - vessel_depth) ** 2 <= vessel_radius ** 2] = 3
    sim.voxels.update()
    return sim
