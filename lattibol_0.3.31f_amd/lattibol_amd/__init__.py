"""lattibol_amd -- MI355X-native lattice-Boltzmann core.

Python host side of liblattibol_amd.so:
  refsim.Simulation   the reference driver (stream / halo / collide / probe)
  solver.Solver       fused pull-scheme sweep (FP32/FP16/BF16/FP64 storage)
  batch.Batch         many ragged D2Q9 channels in one launch per step
  scenes              seeded synthetic workloads
"""
from . import _lib
from ._lib import LatticeError, device_count
from .solver import Solver, make_config, decode_populations, W9, W19, BC, STORAGE
from .refsim import Simulation, SimulationD2Q9, SimulationD3Q19
from .batch import Batch
from . import scenes

__all__ = ["LatticeError", "device_count", "Solver", "make_config", "decode_populations", "W9", "W19",
           "BC", "STORAGE", "Simulation", "SimulationD2Q9", "SimulationD3Q19", "Batch", "scenes"]
