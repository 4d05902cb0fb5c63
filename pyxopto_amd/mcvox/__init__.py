"""Voxelized-media photon transport (the xopto.mcvox path)."""
from . import mc, mcgeometry, mcsource
from ..mcbase import mcdetector, mcfluence, mcmaterial, mcoptions, mcpf
