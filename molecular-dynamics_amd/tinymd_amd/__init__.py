"""tinymd_amd: Python side of libtinymd_amd.so, the MI355X-native tinyMD
per-timestep path (cell binning, Verlet lists, LJ/DEM forces, velocity
Verlet, thermo) behind the reference's md_* C ABI."""
from ._lib import LIB_PATH, TinyMDError, check, load
from .sim import (KERNELS, Simulation, device_info, generate_fcc, generate_granular_gas,
                  world_aabb)
from .batch import Batch, replica_boxes, replica_set

__all__ = ["LIB_PATH", "TinyMDError", "check", "load", "Simulation", "device_info", "generate_fcc",
           "generate_granular_gas", "world_aabb", "KERNELS", "Batch", "replica_boxes", "replica_set"]
