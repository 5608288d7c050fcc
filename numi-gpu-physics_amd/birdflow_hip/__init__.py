"""birdflow_hip — MI355X (gfx950) D3Q19 moving-boundary step.

Python host layer over libbirdflow_hip.so (C ABI in include/birdflow_hip.h).
"""
from . import _abi
from ._abi import (ARITHMETIC_DEFAULT, ARITHMETIC_FUSED, ARITHMETIC_STRICT, BC_FARFIELD, BC_OUTLET, BC_PERIODIC, COLLISION_REGULARIZED, COLLISION_RR3, COLLISION_TRT,
                   COLLISION_TRT_SYMMETRIC_LIMITER, LIBRARY_PATH, STORAGE_BF16_OFFSET, STORAGE_FP16_OFFSET,
                   STORAGE_FP32, WALL_HALFWAY, WALL_INTERPOLATED_HALF, WALL_INTERPOLATED_LINKS, WALL_INTERPOLATED_RAY,
                   WALL_INTERPOLATED_SDF,
                   LibraryMissing)
from .simulation import (ArticulatedMesh, BirdFlowError, BirdFlowSimulation, BirdGeometry, BodyParameters, FlappingWing,
                         Joint, KinematicParameter, LatticeScaling, SolverConfig, copy_bandwidth, device_info)

__all__ = [
    "ArticulatedMesh", "Joint", "KinematicParameter", "BirdFlowError", "BirdFlowSimulation", "BirdGeometry", "FlappingWing", "BodyParameters", "LatticeScaling", "SolverConfig", "device_info", "copy_bandwidth", "LibraryMissing",
    "LIBRARY_PATH", "BC_PERIODIC", "BC_FARFIELD", "BC_OUTLET", "COLLISION_TRT", "COLLISION_TRT_SYMMETRIC_LIMITER",
    "COLLISION_REGULARIZED", "COLLISION_RR3", "WALL_HALFWAY", "WALL_INTERPOLATED_SDF", "WALL_INTERPOLATED_HALF", "WALL_INTERPOLATED_LINKS",
    "WALL_INTERPOLATED_RAY",
    "STORAGE_FP32", "STORAGE_FP16_OFFSET", "STORAGE_BF16_OFFSET", "ARITHMETIC_DEFAULT", "ARITHMETIC_FUSED", "ARITHMETIC_STRICT",
]
