"""FK + pose + hinge collision cost with a hand-written backward.

The reference gets gradients of its collision costs from JAX AD through FK
and the sphere distances (``src/pyroffi/optimization_engines/_chomp_optimization.py:263``)
and has no custom VJP.  Here the forward and the adjoint are HIP kernels
(``pk_fk_links_f32`` then ``pk_pose_hinge_cost_f32``), wrapped in a
``torch.autograd.Function`` -- the host-side counterpart of ``jax.custom_vjp``:
``forward`` saves dcost/dq computed by the adjoint kernel in the same launch
(rematerialising nothing and storing no per-link Jacobian), ``backward``
scales it by the incoming cotangent.
"""

from __future__ import annotations

import numpy as np
import torch

from ..kernels.ops import fk_links, pose_hinge_cost as _pose_hinge_cost


class PoseHingeCost(torch.autograd.Function):
    """cost(q) = |W r_pose(q)|^2 + sum max(0, margin - d(q))^2, per configuration."""

    @staticmethod
    def forward(ctx, q, robot, coll, targets, world, margin, ee_link, pos_weight, ori_weight):
        t = robot.device_tables(q.device)
        ct = coll.device_tables(q.device)
        poses = fk_links(q.detach(), *t.robot_buffers(), t.link_parent_joint)
        cost, grad = _pose_hinge_cost(
            poses, ct["spheres_local"], ct["link_sphere_start"], world, margin, ee_link, targets,
            t.twists, t.parent_idx, t.act_idx, t.mimic_act_idx, t.mimic_mul, t.topo_inv,
            t.link_parent_joint, robot.joints.num_actuated_joints, pos_weight=pos_weight,
            ori_weight=ori_weight, want_grad=q.requires_grad,
        )
        if grad is not None:
            ctx.save_for_backward(grad)
        return cost

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g[:, None] * grad, None, None, None, None, None, None, None, None


def pose_hinge_cost(robot, coll, q: torch.Tensor, targets, world, margin: float = 0.05,
                    ee_link: int | None = None, pos_weight: float = 1.0, ori_weight: float = 1.0):
    """Differentiable per-configuration cost for ``q`` of shape ``(B, n_act)`` on the GPU.

    ``targets`` is ``(B, 7)`` wxyz_xyz for link ``ee_link`` (default: the last
    link); ``world`` is ``(M, 4)`` spheres.
    """
    if ee_link is None:
        ee_link = robot.links.num_links - 1
    dev = q.device
    targets = torch.as_tensor(np.asarray(targets, np.float32) if not isinstance(targets, torch.Tensor) else targets,
                              device=dev, dtype=torch.float32)
    world = torch.as_tensor(np.asarray(world, np.float32) if not isinstance(world, torch.Tensor) else world,
                            device=dev, dtype=torch.float32)
    return PoseHingeCost.apply(q, robot, coll, targets, world, float(margin), int(ee_link), float(pos_weight),
                               float(ori_weight))
