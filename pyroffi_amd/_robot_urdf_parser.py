"""URDF -> kinematic tables (``JointInfo`` / ``LinkInfo``).

Restates ``RobotURDFParser.parse`` (``src/pyroffi/_robot_urdf_parser.py:269-412``)
table for table, with numpy arrays in place of jax arrays:

* twists ``(J, 6)`` in ``[v; omega]`` order: revolute/continuous ``[0; axis]``,
  prismatic ``[axis; 0]``, everything else zero (``:454-471``);
* parent transforms ``(J, 7)`` wxyz_xyz from the joint origin (``:473-500``);
* parent joint index (``-1`` when the parent link is the root), actuated index,
  mimic index/multiplier/offset (``:414-452``);
* the topological order ``_topo_sort_inv`` (``:163-231``) and the depth-level
  schedule ``fk_level_starts`` / ``fk_level_joints`` (``:233-267``).

Device kernels read float32/int32 copies of exactly these tables, so a model
built here is interchangeable with the reference's ``JointInfo``.
"""

from __future__ import annotations

import dataclasses
import logging

import numpy as np

from ._se3 import se3_from_matrix
from ._urdf import URDF, Joint

logger = logging.getLogger(__name__)


@dataclasses.dataclass(frozen=True, eq=False)
class JointInfo:
    """Joint tables of a robot (field names follow the reference, ``:15-68``)."""

    num_joints: int
    num_actuated_joints: int
    names: tuple[str, ...]
    actuated_names: tuple[str, ...]
    twists: np.ndarray  # (J, 6) float64
    parent_transforms: np.ndarray  # (J, 7) float64 wxyz_xyz
    parent_indices: np.ndarray  # (J,) int32
    actuated_indices: np.ndarray  # (J,) int32
    lower_limits: np.ndarray  # (A,)
    upper_limits: np.ndarray  # (A,)
    velocity_limits: np.ndarray  # (A,)
    lower_limits_all: np.ndarray  # (J,)
    upper_limits_all: np.ndarray  # (J,)
    velocity_limits_all: np.ndarray  # (J,)
    mimic_multiplier: np.ndarray  # (J,)
    mimic_offset: np.ndarray  # (J,)
    mimic_act_indices: np.ndarray  # (J,) int32
    _topo_sort_inv: np.ndarray  # (J,) int32: sorted position -> joint index
    fk_level_starts: np.ndarray  # (levels + 1,) int32
    fk_level_joints: np.ndarray  # (J,) int32
    num_fk_levels: int

    def _map_to_full_joint_space(self, value_actuated: np.ndarray, apply_offset: bool) -> np.ndarray:
        """Actuated values -> all joints (``:70-111``): fixed joints read 0,
        mimic joints read their source times the multiplier (+ offset)."""
        value_actuated = np.asarray(value_actuated)
        batch = value_actuated.shape[:-1]
        padded = np.concatenate([value_actuated, np.zeros(batch + (1,))], axis=-1)
        src = np.where(self.mimic_act_indices == -1, self.actuated_indices, self.mimic_act_indices)
        src = np.where(src == -1, self.num_actuated_joints, src)
        out = padded[..., src] * self.mimic_multiplier
        return out + self.mimic_offset if apply_offset else out

    def get_full_config(self, cfg_actuated: np.ndarray) -> np.ndarray:
        return self._map_to_full_joint_space(cfg_actuated, apply_offset=True)


@dataclasses.dataclass(frozen=True, eq=False)
class LinkInfo:
    """Link names and the joint whose child each link is (``:131-157``)."""

    num_links: int
    names: tuple[str, ...]
    parent_joint_indices: np.ndarray  # (L,) int32, -1 for the root link

    def get_link_indices(self, link_names: tuple[str, ...]) -> np.ndarray:
        return np.array([i for i, n in enumerate(self.names) if n in link_names], dtype=np.int32)


class RobotURDFParser:
    """Builds ``JointInfo``/``LinkInfo`` from a parsed URDF."""

    @staticmethod
    def _topologically_sort_joints(urdf: URDF) -> np.ndarray:
        """Parent-first joint order honouring mimic dependencies (``:163-231``).

        Repeatedly takes the first remaining joint (document order) whose parent
        link is the root or already reached and whose mimicked joint is already
        placed.
        """
        joints = list(urdf.joint_map.values())
        idx_of = {j.name: i for i, j in enumerate(joints)}
        remaining = list(joints)
        placed_links: set[str] = set()
        placed_joints: set[str] = set()
        order: list[int] = []
        while remaining:
            for k, j in enumerate(remaining):
                parent_ok = j.parent == urdf.base_link or j.parent in placed_links
                mimic_ok = j.mimic is None or j.mimic.joint in placed_joints
                if parent_ok and mimic_ok:
                    order.append(idx_of[j.name])
                    placed_links.add(j.child)
                    placed_joints.add(j.name)
                    remaining.pop(k)
                    break
            else:
                raise ValueError(
                    f"Topological sort failed. Remaining: {[j.name for j in remaining]}"
                )
        return np.array(order, dtype=np.int32)

    @staticmethod
    def _compute_fk_depth_schedule(parent_idx: list[int], topo_sort_inv: np.ndarray):
        """Joints grouped by tree depth, in topological order (``:233-267``)."""
        n = int(topo_sort_inv.shape[0])
        depth = [0] * n
        for j in topo_sort_inv:
            p = parent_idx[int(j)]
            depth[int(j)] = 0 if p == -1 else depth[p] + 1
        max_depth = max(depth) if depth else 0
        starts = [0]
        level_joints: list[int] = []
        for d in range(max_depth + 1):
            level_joints.extend(int(j) for j in topo_sort_inv if depth[int(j)] == d)
            starts.append(len(level_joints))
        return (
            np.array(starts, dtype=np.int32),
            np.array(level_joints, dtype=np.int32),
            max_depth + 1,
        )

    @staticmethod
    def _get_joint_twist(joint: Joint) -> np.ndarray:
        if joint.type in ("revolute", "continuous"):
            return np.concatenate([np.zeros(3), joint.axis])
        if joint.type == "prismatic":
            return np.concatenate([joint.axis, np.zeros(3)])
        if joint.type != "fixed":
            logger.warning(
                "Unsupported joint type %s for joint %r; treating as fixed.", joint.type, joint.name
            )
        return np.zeros(6)

    @staticmethod
    def _get_act_joint_idx_and_mimic(urdf: URDF, joint: Joint) -> tuple[int, int, float, float]:
        """(actuated index, mimicked actuated index, multiplier, offset) (``:414-452``)."""
        mimic_act_idx, mult, off, act_idx = -1, 1.0, 0.0, -1
        act_names = [j.name for j in urdf.actuated_joints]
        if joint.mimic is not None:
            if joint.mimic.multiplier is not None:
                mult = joint.mimic.multiplier
            if joint.mimic.offset is not None:
                off = joint.mimic.offset
            if joint.mimic.joint in act_names:
                mimic_act_idx = act_names.index(joint.mimic.joint)
            else:
                logger.warning(
                    "Joint %r mimics %r which is not actuated; treating as fixed.",
                    joint.name,
                    joint.mimic.joint,
                )
        elif joint.name in act_names:
            act_idx = act_names.index(joint.name)
        return act_idx, mimic_act_idx, mult, off

    @staticmethod
    def _get_joint_limits(joint: Joint) -> tuple[float, float]:
        lim = joint.limit
        if joint.type == "fixed":
            return 0.0, 0.0
        if joint.type == "continuous":
            if lim is not None and lim.lower is not None and lim.upper is not None:
                return lim.lower, lim.upper
            return -np.pi, np.pi
        if lim is not None and lim.lower is not None and lim.upper is not None:
            return lim.lower, lim.upper
        raise ValueError(f"Joint {joint.name!r} ({joint.type}) has no limits.")

    @staticmethod
    def _get_joint_limit_vel(joint: Joint) -> float:
        if joint.type == "fixed":
            return 0.0
        if joint.limit is not None and joint.limit.velocity is not None:
            return joint.limit.velocity
        raise ValueError(f"Joint {joint.name!r} of type {joint.type!r} has no velocity limits.")

    @staticmethod
    def parse(urdf: URDF) -> tuple[JointInfo, LinkInfo]:
        joints = list(urdf.joint_map.values())
        act_names = {j.name for j in urdf.actuated_joints}
        joint_from_child = {j.child: j for j in joints}
        joint_idx = {j.name: i for i, j in enumerate(joints)}

        twists, ptf, pidx, aidx = [], [], [], []
        mmul, moff, midx = [], [], []
        lo_all, hi_all, vel_all = [], [], []
        lo_act, hi_act, vel_act = [], [], []
        names, act_list = [], []
        for joint in joints:
            names.append(joint.name)
            twists.append(RobotURDFParser._get_joint_twist(joint))
            a, m, mul, off = RobotURDFParser._get_act_joint_idx_and_mimic(urdf, joint)
            aidx.append(a)
            midx.append(m)
            mmul.append(mul)
            moff.append(off)
            if a != -1:
                act_list.append(joint.name)
            lo, hi = RobotURDFParser._get_joint_limits(joint)
            vel = RobotURDFParser._get_joint_limit_vel(joint)
            lo_all.append(lo)
            hi_all.append(hi)
            vel_all.append(vel)
            if joint.name in act_names and m == -1:
                lo_act.append(lo)
                hi_act.append(hi)
                vel_act.append(vel)
            origin = np.eye(4) if joint.origin is None else joint.origin
            ptf.append(se3_from_matrix(origin))
            parent_joint = joint_from_child.get(joint.parent)
            pidx.append(-1 if parent_joint is None else joint_idx[parent_joint.name])

        link_names, link_parent = [], []
        for link_name in urdf.link_map:
            link_names.append(link_name)
            pj = joint_from_child.get(link_name)
            link_parent.append(-1 if pj is None else joint_idx[pj.name])

        topo = RobotURDFParser._topologically_sort_joints(urdf)
        starts, level_joints, n_levels = RobotURDFParser._compute_fk_depth_schedule(pidx, topo)
        n_j = len(joints)
        ji = JointInfo(
            num_joints=n_j,
            num_actuated_joints=len(urdf.actuated_joints),
            names=tuple(names),
            actuated_names=tuple(act_list),
            twists=np.array(twists, dtype=np.float64).reshape(n_j, 6),
            parent_transforms=np.array(ptf, dtype=np.float64).reshape(n_j, 7),
            parent_indices=np.array(pidx, dtype=np.int32),
            actuated_indices=np.array(aidx, dtype=np.int32),
            lower_limits=np.array(lo_act, dtype=np.float64),
            upper_limits=np.array(hi_act, dtype=np.float64),
            velocity_limits=np.array(vel_act, dtype=np.float64),
            lower_limits_all=np.array(lo_all, dtype=np.float64),
            upper_limits_all=np.array(hi_all, dtype=np.float64),
            velocity_limits_all=np.array(vel_all, dtype=np.float64),
            mimic_multiplier=np.array(mmul, dtype=np.float64),
            mimic_offset=np.array(moff, dtype=np.float64),
            mimic_act_indices=np.array(midx, dtype=np.int32),
            _topo_sort_inv=topo,
            fk_level_starts=starts,
            fk_level_joints=level_joints,
            num_fk_levels=n_levels,
        )
        li = LinkInfo(
            num_links=len(link_names),
            names=tuple(link_names),
            parent_joint_indices=np.array(link_parent, dtype=np.int32),
        )
        return ji, li
