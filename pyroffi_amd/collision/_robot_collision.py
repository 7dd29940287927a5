"""Spherized robot collision model (``RobotCollisionSpherized``).

Restates ``src/pyroffi/collision/_robot_collision.py:429-549`` (model build) and
``:583-695`` (active pair lists):

* each link's spheres are its URDF ``<collision><sphere>`` entries in document
  order, centred at the collision origin's translation (the reference fits a
  bounding sphere to an icosphere of the same radius at that origin, which
  reproduces radius and centre up to the fitter's tolerance,
  ``src/pyroffi/collision/_geometry.py:206-241``);
* ``coll`` pads every link to ``S = max spheres per link`` with radius
  ``-1e9`` (``:492-503``);
* active self pairs ``(i, j)`` with ``i > j`` come from ``tril_indices(k=-1)``
  minus the diagonal, the parent/child pairs of every joint (when
  ``ignore_immediate_adjacents``) and ``user_ignore_pairs`` -- or, when an SRDF
  is given, minus the SRDF's ``<disable_collisions>`` pairs only.

Distances follow the semantics of the reference class of the same name by
default (``semantics="jax"``): the JAX path's ``+1e-6`` inside the norm
(``src/pyroffi/collision/_utils.py:45``) and, for self distances, only the
same-slot ("diagonal") sphere pairs of the two links (``:796-828``).
``semantics="cuda"`` selects the reference CUDA kernels' semantics instead
(no bias, padding skipped, the minimum over all S_i x S_j sphere pairs,
``_collision_cuda_kernel.cu:169-201, 425-459``), which is also what
``CUDARobotCollisionChecker`` returns.  Both run on HIP kernels: the JAX
conventions on ``pk_sphere_distances_jax_f32``, the CUDA ones on the fused
sweep.
"""

from __future__ import annotations

import logging

import numpy as np
import torch

from .._robot_srdf_parser import read_disabled_collisions_from_srdf
from .._robot_urdf_parser import RobotURDFParser
from .._se3 import SE3
from .._urdf import URDF
from ._geometry import Sphere

logger = logging.getLogger(__name__)

PAD_RADIUS = -1e9


class RobotCollisionSpherized:
    """Per-link sphere sets of a robot plus its active self-collision pairs."""

    SEMANTICS = ("jax", "cuda")

    def __init__(self, num_links, link_names, coll: Sphere, active_idx_i, active_idx_j,
                 link_spheres: list[np.ndarray], semantics: str = "jax"):
        if semantics not in self.SEMANTICS:
            raise ValueError(f"semantics must be one of {self.SEMANTICS}, got {semantics!r}")
        self.semantics = semantics
        self.num_links = num_links
        self.link_names = tuple(link_names)
        self.coll = coll  # batch axes (N, S)
        self.active_idx_i = np.asarray(active_idx_i, dtype=np.int32)
        self.active_idx_j = np.asarray(active_idx_j, dtype=np.int32)
        # Compact tables for the kernels: spheres grouped by link, no padding.
        self.link_sphere_start = np.zeros(num_links + 1, dtype=np.int32)
        rows = []
        for n, sph in enumerate(link_spheres):
            self.link_sphere_start[n + 1] = self.link_sphere_start[n] + len(sph)
            rows.extend(sph)
        self.spheres_local = np.asarray(rows, dtype=np.float32).reshape(-1, 4)
        self.self_entries = self._build_self_entries()
        self._dev_cache: dict[str, dict[str, torch.Tensor]] = {}

    @property
    def num_pairs(self) -> int:
        return int(self.active_idx_i.shape[0])

    @property
    def num_spheres(self) -> int:
        return int(self.spheres_local.shape[0])

    def _build_self_entries(self) -> np.ndarray:
        """One uint32 per sphere pair to test: a | b << 10 | p << 20."""
        st = self.link_sphere_start
        ents = []
        for p, (i, j) in enumerate(zip(self.active_idx_i, self.active_idx_j)):
            a = np.arange(st[i], st[i + 1], dtype=np.int64)
            b = np.arange(st[j], st[j + 1], dtype=np.int64)
            if a.size and b.size:
                aa, bb = np.meshgrid(a, b, indexing="ij")
                ents.append((aa | (bb << 10) | (p << 20)).ravel())
        if not ents:
            return np.zeros(0, dtype=np.uint32)
        return np.concatenate(ents).astype(np.uint32)

    @staticmethod
    def from_urdf(urdf, user_ignore_pairs=(), srdf_path=None, ignore_immediate_adjacents=True,
                  semantics: str = "jax"):
        if not isinstance(urdf, URDF):
            urdf = URDF.load(urdf)
        _, link_info = RobotURDFParser.parse(urdf)
        names = link_info.names
        link_spheres = []
        for name in names:
            link = urdf.link_map[name]
            if link.n_non_sphere_collisions:
                logger.warning("Unsupported collision geometry type for link %r.", name)
            link_spheres.append(
                np.array([[*s.origin[:3, 3], s.radius] for s in link.spheres], dtype=np.float64).reshape(-1, 4)
            )
        n = len(names)
        s_max = max((len(s) for s in link_spheres), default=0)
        centers = np.zeros((n, s_max, 3))
        radii = np.full((n, s_max), PAD_RADIUS)
        for i, sph in enumerate(link_spheres):
            centers[i, : len(sph)] = sph[:, :3]
            radii[i, : len(sph)] = sph[:, 3]
        # The reference stores sphere centres and radii as float32.
        centers = centers.astype(np.float32).astype(np.float64)
        radii = radii.astype(np.float32).astype(np.float64)
        coll = Sphere(SE3.from_translation(centers), radii)
        if srdf_path:
            ai, aj = RobotCollisionSpherized._compute_active_pair_indices_from_srdf(names, srdf_path)
        else:
            ai, aj = RobotCollisionSpherized._compute_active_pair_indices(
                names, urdf, user_ignore_pairs, ignore_immediate_adjacents
            )
        # Kernel table: per link, spheres ordered by radius so that equal radii
        # are adjacent (the sweep kernel shares one square root per obstacle
        # across such a run).  The padded `coll` keeps the URDF order.
        compact = []
        for c, r, s in zip(centers, radii, link_spheres):
            rows = np.concatenate([c[: len(s)], r[: len(s), None]], axis=1)
            compact.append(rows[np.argsort(-rows[:, 3], kind="stable")])
        return RobotCollisionSpherized(n, names, coll, ai, aj, compact, semantics=semantics)

    @staticmethod
    def _pairs_from_ignore(ignore: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        n = ignore.shape[0]
        ii, jj = np.tril_indices(n, k=-1)
        keep = ~ignore[ii, jj]
        return ii[keep].astype(np.int32), jj[keep].astype(np.int32)

    @staticmethod
    def _compute_active_pair_indices_from_srdf(link_names, srdf_path):
        idx = {nm: i for i, nm in enumerate(link_names)}
        ignore = np.eye(len(link_names), dtype=bool)
        try:
            for pair in read_disabled_collisions_from_srdf(srdf_path):
                a, b = pair["link1"], pair["link2"]
                if a in idx and b in idx:
                    ignore[idx[a], idx[b]] = ignore[idx[b], idx[a]] = True
        except FileNotFoundError:
            logger.warning("SRDF file not found: %s. Using all collision pairs.", srdf_path)
        return RobotCollisionSpherized._pairs_from_ignore(ignore)

    @staticmethod
    def _compute_active_pair_indices(link_names, urdf, user_ignore_pairs, ignore_immediate_adjacents):
        idx = {nm: i for i, nm in enumerate(link_names)}
        ignore = np.eye(len(link_names), dtype=bool)
        if ignore_immediate_adjacents:
            for joint in urdf.joint_map.values():
                if joint.parent in idx and joint.child in idx:
                    p, c = idx[joint.parent], idx[joint.child]
                    ignore[p, c] = ignore[c, p] = True
        for a, b in user_ignore_pairs:
            if a in idx and b in idx:
                ignore[idx[a], idx[b]] = ignore[idx[b], idx[a]] = True
        return RobotCollisionSpherized._pairs_from_ignore(ignore)

    # ----------------------------------------------------------------- device
    def device_tables(self, device) -> dict[str, torch.Tensor]:
        key = str(device)
        if key not in self._dev_cache:
            self._dev_cache[key] = {
                "spheres_local": torch.as_tensor(self.spheres_local, device=device).contiguous(),
                "link_sphere_start": torch.as_tensor(self.link_sphere_start, device=device).contiguous(),
                "self_entries": torch.as_tensor(self.self_entries.view(np.int32), device=device).contiguous(),
                "pair_i": torch.as_tensor(self.active_idx_i, device=device).contiguous(),
                "pair_j": torch.as_tensor(self.active_idx_j, device=device).contiguous(),
                "spheres_padded": torch.as_tensor(self.coll.as_array(), device=device).contiguous(),
            }
        return self._dev_cache[key]

    def _check_robot(self, robot) -> None:
        assert self.link_names == robot.links.names, "Link name mismatch between RobotCollision and Robot kinematics."

    def at_config(self, robot, cfg) -> Sphere:
        """World-frame spheres, batch axes ``(*batch, S, N)`` (padding keeps r = -1e9)."""
        from .._robot import _to_cuda

        self._check_robot(robot)
        x, _ = _to_cuda(cfg)
        poses = robot.forward_kinematics(x)  # (*batch, N, 7) on the GPU
        batch = tuple(poses.shape[:-2])
        c = torch.as_tensor(self.coll.pose.translation(), dtype=torch.float32, device=poses.device)  # (N, S, 3)
        q = poses[..., None, :4]
        t = poses[..., None, 4:]
        u = 2.0 * torch.cross(q[..., 1:].expand(*batch, *c.shape), c.expand(*batch, *c.shape), dim=-1)
        w = c + q[..., :1] * u + torch.cross(q[..., 1:].expand_as(u), u, dim=-1) + t
        w = torch.swapaxes(w, -3, -2).cpu().numpy().astype(np.float64)  # (*batch, S, N, 3)
        r = np.broadcast_to(self.coll.radius.T, batch + self.coll.radius.T.shape)
        return Sphere.from_center_and_radius(w, r)

    def _sweep(self, robot, cfg, world_arr, *, world=False, world_min=False, self_pairs=False):
        from .._robot import _to_cuda
        from ..kernels.ops import sphere_sweep

        self._check_robot(robot)
        x, back = _to_cuda(cfg)
        batch = tuple(x.shape[:-1])
        poses = robot.forward_kinematics(x.reshape(-1, x.shape[-1]))
        t = self.device_tables(poses.device)
        ws = world_arr if world_arr is not None else torch.zeros((0, 4), device=poses.device)
        ow, om, osf = sphere_sweep(
            poses, t["spheres_local"], t["link_sphere_start"], ws, t["self_entries"], self.num_pairs,
            world=world, world_min=world_min, self_pairs=self_pairs,
        )
        return batch, back, ow, om, osf

    def _semantics(self, semantics):
        sem = self.semantics if semantics is None else semantics
        if sem not in self.SEMANTICS:
            raise ValueError(f"semantics must be one of {self.SEMANTICS}, got {sem!r}")
        return sem

    def _jax_distances(self, robot, cfg, world_arr, *, world: bool, self_pairs: bool):
        from .._robot import _to_cuda
        from ..kernels.ops import sphere_distances_jax

        self._check_robot(robot)
        x, back = _to_cuda(cfg)
        batch = tuple(x.shape[:-1])
        poses = robot.forward_kinematics(x.reshape(-1, x.shape[-1]))
        t = self.device_tables(poses.device)
        ws = world_arr if world_arr is not None else torch.zeros((0, 4), device=poses.device)
        ow, osf = sphere_distances_jax(poses, t["spheres_padded"], ws, t["pair_i"], t["pair_j"], world=world,
                                       self_pairs=self_pairs)
        return batch, back, ow, osf

    def compute_world_collision_distance(self, robot, cfg, world_geom, semantics: str | None = None):
        """``(*batch, N, M)``: per link, min over its spheres of the sphere-sphere distance."""
        from ._hip_collision import world_arrays, world_sphere_array

        sem = self._semantics(semantics)
        arrs = world_arrays(world_geom)
        if sum(v.shape[0] for k, v in arrs.items() if k != "spheres"):
            if sem == "jax":
                raise NotImplementedError("JAX-path distances cover sphere worlds; pass semantics='cuda' for "
                                          "capsule, box or half-space obstacles")
            from .._robot import _to_cuda
            from ..kernels.ops import sphere_world_links

            self._check_robot(robot)
            x, back = _to_cuda(cfg)
            poses = robot.forward_kinematics(x.reshape(-1, x.shape[-1]))
            ow, _ = sphere_world_links(poses, self.device_tables(poses.device)["spheres_padded"], arrs)
            return back(ow.reshape(*x.shape[:-1], self.num_links, ow.shape[-1]))
        ws = torch.as_tensor(world_sphere_array(world_geom), device="cuda")
        if sem == "jax":
            batch, back, ow, _ = self._jax_distances(robot, cfg, ws, world=True, self_pairs=False)
        else:
            batch, back, ow, _, _ = self._sweep(robot, cfg, ws, world=True)
        return back(ow.reshape(*batch, self.num_links, ws.shape[0]))

    def compute_self_collision_distance(self, robot, cfg, semantics: str | None = None):
        """``(*batch, P)``: per active pair, the minimum over the two links' same-slot
        sphere pairs (``"jax"``) or over all their sphere pairs (``"cuda"``)."""
        if self._semantics(semantics) == "jax":
            batch, back, _, osf = self._jax_distances(robot, cfg, None, world=False, self_pairs=True)
        else:
            batch, back, _, _, osf = self._sweep(robot, cfg, None, self_pairs=True)
        return back(osf.reshape(*batch, self.num_pairs))


class RobotCollision:
    """Capsule robot model: one capsule per link plus the active self-collision pairs
    (``src/pyroffi/collision/_robot_collision.py:25-427``).

    Distances are those of the reference CUDA capsule kernels
    (``_collision_cuda_kernel.cu:63-161, 203-407``), which the reference checker
    returns for this model: per link and obstacle, the capsule-sphere /
    capsule-capsule / capsule-box / capsule-half-space distance; per active pair,
    the capsule-capsule distance.  The capsules are placed by
    ``pk_capsules_place_f32`` and measured by ``pk_collision_world_capsule_f32`` /
    ``pk_collision_self_capsule_f32``.
    """

    def __init__(self, num_links, link_names, coll, active_idx_i, active_idx_j):
        from ._geometry import Capsule

        if not isinstance(coll, Capsule) or coll.get_batch_axes() != (num_links,):
            raise ValueError(f"coll must be a Capsule with batch axes ({num_links},)")
        self.num_links = num_links
        self.link_names = tuple(link_names)
        self.coll = coll
        self.active_idx_i = np.asarray(active_idx_i, dtype=np.int32)
        self.active_idx_j = np.asarray(active_idx_j, dtype=np.int32)
        self._dev_cache: dict[str, dict[str, torch.Tensor]] = {}

    @property
    def num_pairs(self) -> int:
        return int(self.active_idx_i.shape[0])

    @staticmethod
    def from_urdf(urdf, user_ignore_pairs=(), ignore_immediate_adjacents=True):
        """One capsule per link bounding its URDF sphere collisions
        (``Capsule.from_spheres``); the pairs as ``RobotCollision._compute_active_pair_indices``
        (``:100-156``: lower triangle minus diagonal, adjacent and user pairs)."""
        from ._geometry import Capsule

        if not isinstance(urdf, URDF):
            urdf = URDF.load(urdf)
        _, link_info = RobotURDFParser.parse(urdf)
        names = link_info.names
        caps = []
        for name in names:
            link = urdf.link_map[name]
            if link.n_non_sphere_collisions:
                logger.warning("Unsupported collision geometry type for link %r.", name)
            sph = np.array([[*s.origin[:3, 3], s.radius] for s in link.spheres], dtype=np.float64).reshape(-1, 4)
            caps.append(Capsule.from_spheres(sph[:, :3], sph[:, 3]))
        wxyz_xyz = np.stack([c.pose.wxyz_xyz for c in caps])
        coll = Capsule(SE3(wxyz_xyz), np.array([float(c.radius) for c in caps]),
                       np.array([float(c.height) for c in caps]))
        ai, aj = RobotCollisionSpherized._compute_active_pair_indices(
            names, urdf, user_ignore_pairs, ignore_immediate_adjacents
        )
        return RobotCollision(len(names), names, coll, ai, aj)

    def device_tables(self, device) -> dict[str, torch.Tensor]:
        key = str(device)
        if key not in self._dev_cache:
            self._dev_cache[key] = {
                "capsules_local": torch.as_tensor(self.coll.as_array(), device=device).contiguous(),
                "pair_i": torch.as_tensor(self.active_idx_i, device=device).contiguous(),
                "pair_j": torch.as_tensor(self.active_idx_j, device=device).contiguous(),
            }
        return self._dev_cache[key]

    def _check_robot(self, robot) -> None:
        assert self.link_names == robot.links.names, "Link name mismatch between RobotCollision and Robot kinematics."

    def _placed(self, robot, cfg, *, soa: bool, aos: bool):
        from .._robot import _to_cuda
        from ..kernels.ops import capsules_place

        self._check_robot(robot)
        x, back = _to_cuda(cfg)
        batch = tuple(x.shape[:-1])
        poses = robot.forward_kinematics(x.reshape(-1, x.shape[-1]))
        t = self.device_tables(poses.device)
        c_soa, c_aos = capsules_place(poses, t["capsules_local"], soa=soa, aos=aos)
        return batch, back, c_soa, c_aos

    def placed_capsules(self, robot, cfg) -> torch.Tensor:
        """``(*batch, N, 7)`` world capsules ``[a, b, r]`` on the GPU."""
        batch, _, _, c = self._placed(robot, cfg, soa=False, aos=True)
        return c.reshape(*batch, self.num_links, 7)

    def at_config(self, robot, cfg):
        """World-frame capsules, batch axes ``(*batch, N)`` (``:323-345``)."""
        from ._geometry import Capsule

        c = self.placed_capsules(robot, cfg).cpu().numpy().astype(np.float64)
        a, b, r = c[..., :3], c[..., 3:6], c[..., 6]
        d = b - a
        h = np.linalg.norm(d, axis=-1)
        z = np.where(h[..., None] > 0, d / np.maximum(h, 1e-300)[..., None], np.array([0.0, 0.0, 1.0]))
        from ._geometry import _frame_with_z

        m = np.zeros(c.shape[:-1] + (4, 4))
        m[..., :3, :3] = _frame_with_z(z)
        m[..., :3, 3] = 0.5 * (a + b)
        m[..., 3, 3] = 1.0
        return Capsule(SE3.from_matrix(m), r, h)

    def compute_self_collision_distance(self, robot, cfg):
        """``(*batch, P)``: capsule-capsule distance of every active pair (``:347-386``)."""
        from ..kernels.ops import collision_self_capsule

        batch, back, _, c = self._placed(robot, cfg, soa=False, aos=True)
        t = self.device_tables(c.device)
        out = collision_self_capsule(c, t["pair_i"], t["pair_j"])
        return back(out.reshape(*batch, self.num_pairs))

    def compute_world_collision_distance(self, robot, cfg, world_geom):
        """``(*batch, N, M)``: every link capsule against every obstacle (``:388-427``)."""
        from ..kernels.ops import collision_world_capsule
        from ._hip_collision import world_arrays

        w = world_arrays(world_geom)
        batch, back, c, _ = self._placed(robot, cfg, soa=True, aos=False)
        out = collision_world_capsule(c, w["spheres"], w["capsules"], w["boxes"], w["halfspaces"])
        return back(out.reshape(*batch, self.num_links, out.shape[-1]))
