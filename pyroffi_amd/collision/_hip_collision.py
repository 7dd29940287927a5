"""GPU collision checker with the reference's ``CUDARobotCollisionChecker`` API.

Reference: ``src/pyroffi/collision/_cuda_collision.py:98-724``.  The reference
runs FK, a vmapped sphere transform, two AoS/SoA transposes and one of two
distance kernels per call; here one FK launch writes link poses and one sweep
launch computes every requested distance from them (sphere centres are placed
in LDS inside the sweep, never in HBM).

Semantics are those of the reference CUDA kernels:
* world distance per (link, obstacle) = min over the link's spheres of
  ``|c - o| - r - R``; a link without spheres reports ``1e9``;
* self distance per active pair = min over *all* sphere pairs of the two
  links (``_collision_cuda_kernel.cu:425-459``), not the JAX path's
  same-index "diagonal" pairs (``_robot_collision.py:817-827``), and without
  the JAX path's ``+1e-6`` bias inside the norm (``_utils.py:45``).

Only sphere worlds are supported (capsule/box/half-space worlds and capsule
robots are outside this package's scope).
"""

from __future__ import annotations

import numpy as np
import torch

from ._geometry import Sphere
from ._robot_collision import RobotCollisionSpherized


def world_sphere_array(world_geom) -> np.ndarray:
    """World geometry -> float32 ``(M, 4)`` ``[x, y, z, r]`` (``_cuda_collision.py:98-168``).

    Accepts a ``Sphere`` with at most one batch axis, or an ``(M, 4)`` array.
    """
    if isinstance(world_geom, Sphere):
        axes = world_geom.get_batch_axes()
        if len(axes) > 1:
            raise ValueError(
                "the collision checker only supports world_geom without leading batch "
                f"dimensions; got shape {axes}"
            )
        return world_geom.as_array().reshape(-1, 4)
    if isinstance(world_geom, (list, tuple)):
        parts = [world_sphere_array(g) for g in world_geom]
        return np.concatenate(parts, axis=0) if parts else np.zeros((0, 4), np.float32)
    arr = np.asarray(world_geom, dtype=np.float32)
    if arr.ndim == 2 and arr.shape[1] == 4:
        return np.ascontiguousarray(arr)
    raise NotImplementedError(
        f"world geometry of type {type(world_geom).__name__} is not supported; only spheres are"
    )


class CUDARobotCollisionChecker:
    """Drop-in for the reference checker, backed by the HIP sweep kernel."""

    def __init__(self, inner: RobotCollisionSpherized, coarse_inner: RobotCollisionSpherized | None = None):
        from ..kernels._lib import lib

        lib()  # fail loudly now if the library is missing
        if not isinstance(inner, RobotCollisionSpherized):
            raise NotImplementedError("only spherized robot collision models are supported")
        self._inner = inner
        self._coarse_inner = coarse_inner
        self._ws: torch.Tensor | None = None
        self._cached_world_id = None

    @property
    def num_links(self) -> int:
        return self._inner.num_links

    @property
    def link_names(self) -> tuple[str, ...]:
        return self._inner.link_names

    @property
    def active_idx_i(self):
        return self._inner.active_idx_i

    @property
    def active_idx_j(self):
        return self._inner.active_idx_j

    def at_config(self, robot, cfg):
        return self._inner.at_config(robot, cfg)

    def set_world(self, world_geom) -> None:
        """Upload the world once (keyed by object identity like the reference)."""
        self._ws = torch.as_tensor(world_sphere_array(world_geom), device="cuda").contiguous()
        self._cached_world_id = id(world_geom)

    def _ensure_world(self, world_geom) -> torch.Tensor:
        if self._ws is None or id(world_geom) != self._cached_world_id:
            self.set_world(world_geom)
        return self._ws

    def compute_world_collision_distance(self, robot, cfg, world_geom):
        ws = self._ensure_world(world_geom)
        if self._coarse_inner is not None:
            batch, back, _, om, _ = self._coarse_inner._sweep(robot, cfg, ws, world_min=True)
            if bool((om > 0).all()):
                return back(torch.ones((*batch, self.num_links, ws.shape[0]), device=om.device))
        batch, back, ow, _, _ = self._inner._sweep(robot, cfg, ws, world=True)
        return back(ow.reshape(*batch, self.num_links, ws.shape[0]))

    def compute_self_collision_distance(self, robot, cfg):
        if self._coarse_inner is not None and self._coarse_inner.num_pairs > 0:
            batch, back, _, _, osf = self._coarse_inner._sweep(robot, cfg, None, self_pairs=True)
            if bool((osf > 0).all()):
                return back(torch.ones((*batch, self._inner.num_pairs), device=osf.device))
        return self._inner.compute_self_collision_distance(robot, cfg, semantics="cuda")

    def compute_min_world_distance(self, robot, cfg, world_geom):
        """``(*batch,)``: min over links and obstacles (the sweep's wave min-reduction)."""
        ws = self._ensure_world(world_geom)
        batch, back, _, om, _ = self._inner._sweep(robot, cfg, ws, world_min=True)
        return back(om.reshape(batch))

    def compute_all(self, robot, cfg, world_geom):
        """Link poses, per-config world minimum and self pair distances in two launches."""
        ws = self._ensure_world(world_geom)
        from .._robot import _to_cuda
        from ..kernels.ops import sphere_sweep

        x, back = _to_cuda(cfg)
        batch = tuple(x.shape[:-1])
        poses = robot.forward_kinematics(x.reshape(-1, x.shape[-1]))
        t = self._inner.device_tables(poses.device)
        _, om, osf = sphere_sweep(
            poses, t["spheres_local"], t["link_sphere_start"], ws, t["self_entries"],
            self._inner.num_pairs, world_min=True, self_pairs=True,
        )
        return (
            back(poses.reshape(*batch, self.num_links, 7)),
            back(om.reshape(batch)),
            back(osf.reshape(*batch, self._inner.num_pairs)),
        )


def make_cuda_checker(inner, coarse_inner=None) -> CUDARobotCollisionChecker:
    return CUDARobotCollisionChecker(inner, coarse_inner=coarse_inner)


HIPRobotCollisionChecker = CUDARobotCollisionChecker
make_hip_checker = make_cuda_checker
