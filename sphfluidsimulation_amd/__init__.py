"""MI355X-native SPH step engine (a third engine beside the reference's OpenCL
and multithreaded CPU engines of Murk42/SPHFluidSimulation).

The compute path is libsph_hip.so: hand-written gfx950 HIP kernels behind the C
ABI of include/sph_hip.h. This package is the Python host over that ABI
(SimulationEngineHIP), the seeded scene generators and the build recipe.
"""
from ._lib import SPHError, load
from .engine import SimulationEngineHIP
from . import scenes
from . import scene_json

__all__ = ["SPHError", "SimulationEngineHIP", "load", "scenes", "scene_json"]
