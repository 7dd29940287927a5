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


class _LinkPosesFn(torch.autograd.Function):
    """Link poses with the hand-written FK adjoint as backward (the torch
    counterpart of a ``jax.custom_vjp`` over the FK FFI call).  Only the
    configuration is kept for the backward, which recomputes FK."""

    @staticmethod
    def forward(ctx, x, tables):
        from .kernels.ops import fk_links

        ctx.tables = tables
        ctx.save_for_backward(x)
        return fk_links(x.detach(), *tables.robot_buffers(), tables.link_parent_joint)

    @staticmethod
    def backward(ctx, g):
        from .kernels.ops import fk_vjp

        (x,) = ctx.saved_tensors
        t = ctx.tables
        return fk_vjp(x.detach(), g.contiguous(), *t.robot_buffers(), t.link_parent_joint), None


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
        with the reference; both paths run the HIP kernel.  A GPU tensor that
        requires grad gets a differentiable result: its backward is the FK
        adjoint kernel (``pk_fk_vjp_f32``).
        """
        del unroll_fk, use_cuda
        from .kernels.ops import fk_links

        x, back = _to_cuda(cfg)
        assert x.shape[-1] == self.joints.num_actuated_joints, (
            f"cfg has {x.shape[-1]} values, robot has {self.joints.num_actuated_joints} actuated joints"
        )
        t = self.device_tables(x.device)
        if isinstance(cfg, torch.Tensor) and cfg.requires_grad and torch.is_grad_enabled():
            return back(_LinkPosesFn.apply(x, t))
        out = fk_links(x, *t.robot_buffers(), t.link_parent_joint)
        return back(out)

    def forward_kinematics_joints(self, cfg):
        """World pose of each joint, ``(*batch, n_joints, 7)`` (the FK kernel's raw output)."""
        from .kernels.ops import fk_hip

        x, back = _to_cuda(cfg)
        t = self.device_tables(x.device)
        out = fk_hip(x, *t.robot_buffers(), t.fk_level_starts, t.fk_level_joints)
        return back(out)

    def inverse_kinematics(self, target_link_name: str, target_pose, rng_key=None, previous_cfg=None,
                           num_seeds: int = 32, coarse_max_iter: int = 20, lm_max_iter: int = 40,
                           epsilon: float = 0.02, nu: float = float(np.pi / 2), lambda_init: float = 5e-3,
                           continuity_weight: float = 1e-3, fixed_joint_mask=None):
        """HJCD-IK for one link pose (``src/pyroffi/_robot.py:111-178``); returns ``(n_act,)``.

        ``previous_cfg`` defaults to the joint-range midpoint and ``rng_key`` to
        seed 0, as in the reference.
        """
        from .optimization_engines._hjcd_ik import hjcd_solve

        if rng_key is None:
            rng_key = 0
        if previous_cfg is None:
            previous_cfg = (self.joints.lower_limits + self.joints.upper_limits) / 2
        link = self.links.names.index(target_link_name)
        return hjcd_solve(self, (link,), (target_pose,), rng_key, previous_cfg, num_seeds=num_seeds,
                          coarse_max_iter=coarse_max_iter, lm_max_iter=lm_max_iter, epsilon=epsilon, nu=nu,
                          lambda_init=lambda_init, continuity_weight=continuity_weight,
                          fixed_joint_mask=fixed_joint_mask)
