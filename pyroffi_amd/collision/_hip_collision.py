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

Worlds may mix the four obstacle kinds of ``_extract_world_arrays``
(``_cuda_collision.py:98-168``).  A sphere-only world goes through the fused
sweep; a world with capsules, boxes or half-spaces goes through
``pk_sphere_world_links_f32`` (the ``CollisionWorldSphereReducedFfi`` path).
"""

from __future__ import annotations

import numpy as np
import torch

from ._geometry import Sphere
from ._robot_collision import RobotCollision, RobotCollisionSpherized


_WORLD_WIDTH = {"spheres": 4, "capsules": 7, "boxes": 15, "halfspaces": 6}


def world_arrays(world_geom) -> dict[str, np.ndarray]:
    """World geometry (one container, a list of them, or an ``(M, 4)`` sphere array) ->
    the packed float32 arrays by kind (``_cuda_collision.py:98-168``)."""
    from ._geometry import Box, Capsule, HalfSpace

    out = {k: [] for k in _WORLD_WIDTH}
    items = world_geom if isinstance(world_geom, (list, tuple)) else [world_geom]
    for g in items:
        kind = {Sphere: "spheres", Capsule: "capsules", Box: "boxes", HalfSpace: "halfspaces"}.get(type(g))
        if kind is None:
            out["spheres"].append(world_sphere_array(g))
            continue
        if len(g.get_batch_axes()) > 1:
            raise ValueError("world geometry must have at most one batch axis; got shape "
                             f"{g.get_batch_axes()}")
        out[kind].append(g.as_array().reshape(-1, _WORLD_WIDTH[kind]))
    return {k: (np.concatenate(v, 0) if v else np.zeros((0, _WORLD_WIDTH[k]), np.float32)) for k, v in out.items()}


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
    """Drop-in for the reference checker over spherized (``RobotCollisionSpherized``)
    and capsule (``RobotCollision``) models, with an optional coarse-first guard:
    when every coarse distance is positive the fine distances are replaced by ones
    (``_cuda_collision.py:544-659``; not differentiable, like the reference's)."""

    def __init__(self, inner, coarse_inner=None):
        from ..kernels._lib import lib

        lib()  # fail loudly now if the library is missing
        for m in (inner, coarse_inner):
            if m is not None and not isinstance(m, (RobotCollisionSpherized, RobotCollision)):
                raise NotImplementedError("collision models must be RobotCollisionSpherized or RobotCollision")
        self._inner = inner
        self._coarse_inner = coarse_inner
        self._ws: torch.Tensor | None = None
        self._world: dict[str, torch.Tensor] = {}
        self._n_obstacles = 0
        self._mixed = False
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
        arrs = world_arrays(world_geom)
        self._world = {k: torch.as_tensor(v, device="cuda").contiguous() for k, v in arrs.items()}
        self._ws = self._world["spheres"]
        self._n_obstacles = sum(v.shape[0] for v in arrs.values())
        self._mixed = self._n_obstacles > arrs["spheres"].shape[0]
        self._cached_world_id = id(world_geom)

    def _ensure_world(self, world_geom) -> torch.Tensor:
        if self._ws is None or id(world_geom) != self._cached_world_id:
            self.set_world(world_geom)
        return self._ws

    def _links_world(self, model, robot, cfg):
        """Mixed-kind world distances ``(batch, back, (B, L, M), (B,))`` of a spherized model."""
        from .._robot import _to_cuda
        from ..kernels.ops import sphere_world_links

        model._check_robot(robot)
        x, back = _to_cuda(cfg)
        batch = tuple(x.shape[:-1])
        poses = robot.forward_kinematics(x.reshape(-1, x.shape[-1]))
        t = model.device_tables(poses.device)
        ow, om = sphere_world_links(poses, t["spheres_padded"], self._world)
        return batch, back, ow, om

    def _world_of(self, model, robot, cfg, *, full: bool):
        """``(batch, back, (B, L, M) or None, (B,) minimum or None)`` for one model."""
        from ..kernels.ops import collision_world_capsule, row_min

        if isinstance(model, RobotCollision):
            batch, back, c, _ = model._placed(robot, cfg, soa=True, aos=False)
            w = self._world
            ow = collision_world_capsule(c, w["spheres"], w["capsules"], w["boxes"], w["halfspaces"])
            return batch, back, ow, (None if full else row_min(ow))
        if self._mixed:
            return self._links_world(model, robot, cfg)
        if full:
            batch, back, ow, _, _ = model._sweep(robot, cfg, self._ws, world=True)
            return batch, back, ow, None
        batch, back, _, om, _ = model._sweep(robot, cfg, self._ws, world_min=True)
        return batch, back, None, om

    @staticmethod
    def _self_of(model, robot, cfg):
        if isinstance(model, RobotCollision):
            return model.compute_self_collision_distance(robot, cfg)
        return model.compute_self_collision_distance(robot, cfg, semantics="cuda")

    def compute_world_collision_distance(self, robot, cfg, world_geom):
        self._ensure_world(world_geom)
        M = self._n_obstacles
        if self._coarse_inner is not None:
            batch, back, _, om = self._world_of(self._coarse_inner, robot, cfg, full=False)
            if bool((om > 0).all()):
                return back(torch.ones((*batch, self.num_links, M), device=om.device))
        batch, back, ow, _ = self._world_of(self._inner, robot, cfg, full=True)
        return back(ow.reshape(*batch, self.num_links, M))

    def compute_self_collision_distance(self, robot, cfg):
        if self._coarse_inner is not None and self._coarse_inner.num_pairs > 0:
            osf = self._self_of(self._coarse_inner, robot, cfg)
            if bool((osf > 0).all()):
                shape = (*osf.shape[:-1], self._inner.num_pairs)
                if isinstance(osf, torch.Tensor):
                    return torch.ones(shape, device=osf.device)
                return np.ones(shape, np.float32)
        return self._self_of(self._inner, robot, cfg)

    def compute_min_world_distance(self, robot, cfg, world_geom):
        """``(*batch,)``: min over links and obstacles (the sweep's wave min-reduction)."""
        self._ensure_world(world_geom)
        batch, back, _, om = self._world_of(self._inner, robot, cfg, full=False)
        return back(om.reshape(batch))

    def compute_all(self, robot, cfg, world_geom):
        """Link poses, per-config world minimum and self pair distances; a spherized
        model in a sphere world takes two launches (FK, then the fused sweep)."""
        from .._robot import _to_cuda
        from ..kernels.ops import sphere_sweep

        ws = self._ensure_world(world_geom)
        x, back = _to_cuda(cfg)
        batch = tuple(x.shape[:-1])
        if isinstance(self._inner, RobotCollision) or self._mixed:
            poses = robot.forward_kinematics(x.reshape(-1, x.shape[-1]))
            _, _, _, om = self._world_of(self._inner, robot, cfg, full=False)
            osf = self._self_of(self._inner, robot, cfg)
            return back(poses.reshape(*batch, self.num_links, 7)), back(om.reshape(batch)), osf
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
