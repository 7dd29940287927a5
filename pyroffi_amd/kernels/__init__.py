"""Device entry points: thin wrappers over the C ABI, one per reference FFI target."""

from ._lib import EXPORTED_SYMBOLS, PkError, PkInvalidArgument, lib, library_path
from .ops import (
    collision_self_sphere,
    collision_world_sphere_reduced,
    fk_cuda,
    fk_hip,
    fk_links,
    ls_ik_cuda,
    ls_ik_batched,
    ls_ik_hip,
    ls_ik_select,
    ls_ik_select_segments,
    pose_hinge_cost,
    sphere_sweep,
    world_grid,
)

__all__ = [
    "EXPORTED_SYMBOLS",
    "PkError",
    "PkInvalidArgument",
    "collision_self_sphere",
    "collision_world_sphere_reduced",
    "fk_cuda",
    "fk_hip",
    "fk_links",
    "lib",
    "library_path",
    "ls_ik_cuda",
    "ls_ik_batched",
    "ls_ik_hip",
    "ls_ik_select",
    "ls_ik_select_segments",
    "pose_hinge_cost",
    "sphere_sweep",
    "world_grid",
]
