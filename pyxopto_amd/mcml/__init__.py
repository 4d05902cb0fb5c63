"""Layered-media photon transport (the xopto.mcml path)."""
from . import mc, mclayer, mcsource
from ..mcbase import mcdetector, mcoptions, mcpf
