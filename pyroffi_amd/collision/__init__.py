"""Collision models (spherized and capsule robots), world geometry and the GPU checker."""

from ._collision import colldist_from_sdf
from ._geometry import Box, Capsule, HalfSpace, Sphere
from ._hip_collision import (
    CUDARobotCollisionChecker,
    HIPRobotCollisionChecker,
    make_cuda_checker,
    make_hip_checker,
    world_arrays,
    world_sphere_array,
)
from ._robot_collision import RobotCollision, RobotCollisionSpherized

__all__ = [
    "Box",
    "CUDARobotCollisionChecker",
    "Capsule",
    "HIPRobotCollisionChecker",
    "HalfSpace",
    "RobotCollision",
    "RobotCollisionSpherized",
    "Sphere",
    "colldist_from_sdf",
    "make_cuda_checker",
    "make_hip_checker",
    "world_arrays",
    "world_sphere_array",
]
