"""Sphere collision model and the GPU collision checker."""

from ._collision import colldist_from_sdf
from ._geometry import Sphere
from ._hip_collision import (
    CUDARobotCollisionChecker,
    HIPRobotCollisionChecker,
    make_cuda_checker,
    make_hip_checker,
    world_sphere_array,
)
from ._robot_collision import RobotCollisionSpherized

__all__ = [
    "CUDARobotCollisionChecker",
    "HIPRobotCollisionChecker",
    "RobotCollisionSpherized",
    "Sphere",
    "colldist_from_sdf",
    "make_cuda_checker",
    "make_hip_checker",
    "world_sphere_array",
]
