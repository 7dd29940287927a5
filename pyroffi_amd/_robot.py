"""``Robot``: the kinematic tree and its forward kinematics on the GPU.

Mirrors ``pyroffi.Robot`` (``src/pyroffi/_robot.py:16-109``): ``from_urdf``
builds the joint/link tables, ``forward_kinematics(cfg)`` returns the world
pose of every link, ``(*batch, n_links, 7)`` in wxyz_xyz order.  The reference
picks between a JAX implementation and its CUDA kernel with ``use_cuda``; this
package has exactly one implementation, the HIP kernel, which both values of
the flag run.

Array handling: ``cfg`` may be a torch tensor on the GPU (the result stays on
the GPU, nothing is copied), a CPU torch tensor, or anything numpy accepts (the
input is uploaded and the result is returned the same way it came in).
"""

from __future__ import annotations

import os

import numpy as np
import torch

from ._robot_urdf_parser import JointInfo, LinkInfo, RobotURDFParser
from ._urdf import URDF


class _DeviceTables:
    """float32/int32 copies of the joint tables on one GPU."""

    def __init__(self, joints: JointInfo, links: LinkInfo, device: torch.device):
        f = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float32), device=device).contiguous()
        i = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int32), device=device).contiguous()
        self.twists = f(joints.twists)
        self.parent_tf = f(joints.parent_transforms)
        self.parent_idx = i(joints.parent_indices)
        self.act_idx = i(joints.actuated_indices)
        self.mimic_mul = f(joints.mimic_multiplier)
        self.mimic_off = f(joints.mimic_offset)
        self.mimic_act_idx = i(joints.mimic_act_indices)
        self.topo_inv = i(joints._topo_sort_inv)
        self.fk_level_starts = i(joints.fk_level_starts)
        self.fk_level_joints = i(joints.fk_level_joints)
        self.link_parent_joint = i(links.parent_joint_indices)
        self.lower = f(joints.lower_limits)
        self.upper = f(joints.upper_limits)
        self._derived: dict = {}

    def derived(self, key, make):
        """Device tensors derived from host data, built once per ``key``.

        Solver calls reuse them instead of uploading small tables every call:
        an upload from pageable memory waits for the stream's queued work,
        which would keep the host from preparing the next call while the GPU
        runs the current one.
        """
        v = self._derived.get(key)
        if v is None:
            if len(self._derived) >= 256:
                self._derived.clear()
            v = self._derived[key] = make()
        return v

    def robot_buffers(self) -> tuple[torch.Tensor, ...]:
        """The 8 robot buffers of the reference FFI calls, in order."""
        return (
            self.twists,
            self.parent_tf,
            self.parent_idx,
            self.act_idx,
            self.mimic_mul,
            self.mimic_off,
            self.mimic_act_idx,
            self.topo_inv,
        )


def _to_cuda(x, dtype=torch.float32):
    """(device tensor, function mapping a device result back to x's kind)."""
    if isinstance(x, torch.Tensor):
        if x.is_cuda:
            return x.to(dtype).contiguous(), lambda y: y
        return x.to("cuda", dtype).contiguous(), lambda y: y.cpu()
    arr = np.asarray(x)
    return torch.as_tensor(arr, dtype=dtype, device="cuda").contiguous(), lambda y: y.cpu().numpy()


class Robot:
    """A robot kinematic tree (joint/link tables + GPU forward kinematics)."""

    def __init__(self, joints: JointInfo, links: LinkInfo, default_joint_cfg: np.ndarray):
        self.joints = joints
        self.links = links
        self.default_joint_cfg = default_joint_cfg
        self._dev_cache: dict[str, _DeviceTables] = {}

    @staticmethod
    def from_urdf(urdf, default_joint_cfg=None) -> "Robot":
        """Build from a parsed ``URDF`` (or a path to a ``.urdf`` file)."""
        if isinstance(urdf, (str, os.PathLike)):
            urdf = URDF.load(urdf)
        joints, links = RobotURDFParser.parse(urdf)
        if default_joint_cfg is None:
            default_joint_cfg = (joints.lower_limits + joints.upper_limits) / 2
        default_joint_cfg = np.asarray(default_joint_cfg, dtype=np.float64)
        assert default_joint_cfg.shape == (joints.num_actuated_joints,)
        return Robot(joints, links, default_joint_cfg)

    def device_tables(self, device=None) -> _DeviceTables:
        device = torch.device(device if device is not None else "cuda")
        if device.type == "cuda" and device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        key = str(device)
        if key not in self._dev_cache:
            self._dev_cache[key] = _DeviceTables(self.joints, self.links, device)
        return self._dev_cache[key]

    def forward_kinematics(self, cfg, unroll_fk: bool = False, use_cuda: bool = False):
        """World pose of each link, ``(*batch, n_links, 7)`` wxyz_xyz.

        ``unroll_fk`` and ``use_cuda`` are accepted for signature compatibility
        with the reference; both paths run the HIP kernel.
        """
        del unroll_fk, use_cuda
        from .kernels.ops import fk_links

        x, back = _to_cuda(cfg)
        assert x.shape[-1] == self.joints.num_actuated_joints, (
            f"cfg has {x.shape[-1]} values, robot has {self.joints.num_actuated_joints} actuated joints"
        )
        t = self.device_tables(x.device)
        out = fk_links(x, *t.robot_buffers(), t.link_parent_joint)
        return back(out)

    def forward_kinematics_joints(self, cfg):
        """World pose of each joint, ``(*batch, n_joints, 7)`` (the FK kernel's raw output)."""
        from .kernels.ops import fk_hip

        x, back = _to_cuda(cfg)
        t = self.device_tables(x.device)
        out = fk_hip(x, *t.robot_buffers(), t.fk_level_starts, t.fk_level_joints)
        return back(out)

    def inverse_kinematics(self, *args, **kwargs):
        raise NotImplementedError(
            "Robot.inverse_kinematics (HJCD-IK) is not part of this package yet; "
            "use pyroffi_amd.ls_ik_solve_cuda"
        )
