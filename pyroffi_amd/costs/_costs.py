"""Residual factories of the reference's cost module, on the GPU.

``pose_cost``, ``self_collision_cost`` and ``world_collision_cost`` follow
``src/pyroffi/costs/_costs.py:11-29, 169-199``: each reads the joint
configuration ``vals[joint_var]`` and returns a flat residual vector.  In the
reference they are jaxls ``Cost`` factories; jaxls (and JAX) are not part of
this build, so here they are plain functions whose ``vals`` is any mapping
from ``joint_var`` to a configuration (a dict, for instance), or the
configuration itself when ``joint_var`` is None.  When jaxls is importable the
same functions are also exposed wrapped by ``jaxls.Cost.create_factory``
(``JAXLS_FACTORIES``).

The kinematics and distances run on the HIP kernels (``Robot.forward_kinematics``,
``compute_self_collision_distance``, ``compute_world_collision_distance``); the
pose residual is differentiable through the FK adjoint kernel when the
configuration tensor requires grad.

The pose residual is ``log(T_actual^-1 T_target)`` in jaxlie's tangent
convention ``[v; omega]`` (jaxlie 1.x, the reference's dependency, absent
here): ``omega = SO3.log`` of the relative rotation and
``v = V(omega)^-1 t`` with ``V^-1 = I - [omega]x / 2 + c(theta) [omega]x^2``,
``c = (1 - theta cos(theta/2) / (2 sin(theta/2))) / theta^2`` (``1/12`` near
zero).  ``pyroffi_amd._se3.se3_log`` is the FP64 host restatement of the same
map (pinned by an exp/log round trip); the device version below is checked
against it.
"""

from __future__ import annotations

import numpy as np
import torch

from ..collision._collision import colldist_from_sdf

__all__ = ["pose_cost", "self_collision_cost", "world_collision_cost", "se3_log_t", "JAXLS_FACTORIES"]


def _cfg(vals, joint_var):
    return vals if joint_var is None else vals[joint_var]


def _as_pose_tensor(pose, like: torch.Tensor) -> torch.Tensor:
    """wxyz_xyz tensor from an ``SE3``-like object (``.wxyz_xyz``) or an array."""
    p = getattr(pose, "wxyz_xyz", pose)
    return torch.as_tensor(np.asarray(p) if not isinstance(p, torch.Tensor) else p, dtype=like.dtype,
                           device=like.device)


def _quat_mul(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    aw, ax, ay, az = a.unbind(-1)
    bw, bx, by, bz = b.unbind(-1)
    return torch.stack([aw * bw - ax * bx - ay * by - az * bz,
                        aw * bx + ax * bw + ay * bz - az * by,
                        aw * by - ax * bz + ay * bw + az * bx,
                        aw * bz + ax * by - ay * bx + az * bw], dim=-1)


def _quat_rotate(q: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    u = q[..., 1:]
    t = 2.0 * torch.cross(u, v, dim=-1)
    return v + q[..., :1] * t + torch.cross(u, t, dim=-1)


def _inv_compose(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """a^-1 o b for wxyz_xyz poses."""
    qa = a[..., :4] * a.new_tensor([1.0, -1.0, -1.0, -1.0])
    return torch.cat([_quat_mul(qa, b[..., :4]), _quat_rotate(qa, b[..., 4:] - a[..., 4:])], dim=-1)


def _so3_log(q: torch.Tensor) -> torch.Tensor:
    """wxyz -> axis-angle, the shorter rotation (jaxlie ``SO3.log``)."""
    eps = torch.finfo(q.dtype).eps
    w = q[..., 0]
    v = q[..., 1:]
    nv2 = (v * v).sum(-1)
    small = nv2 < eps
    nv = torch.sqrt(torch.where(small, torch.ones_like(nv2), nv2))
    w_safe = torch.where(small, w, torch.ones_like(w))
    ang = torch.atan2(torch.where(w < 0, -nv, nv), w.abs())
    scale = torch.where(small, 2.0 / w_safe - 2.0 / 3.0 * nv2 / w_safe ** 3, 2.0 * ang / nv)
    return scale[..., None] * v


def se3_log_t(T: torch.Tensor) -> torch.Tensor:
    """wxyz_xyz -> ``[v; omega]`` tangent (jaxlie ``SE3.log``), on T's device."""
    omega = _so3_log(T[..., :4])
    eps = torch.finfo(T.dtype).eps
    th2 = (omega * omega).sum(-1)
    small = th2 < eps
    th2s = torch.where(small, torch.ones_like(th2), th2)
    th = torch.sqrt(th2s)
    half = 0.5 * th
    coef = torch.where(small, torch.full_like(th2, 1.0 / 12.0),
                       (1.0 - th * torch.cos(half) / (2.0 * torch.sin(half))) / th2s)
    t = T[..., 4:]
    wxt = torch.cross(omega, t, dim=-1)
    v = t - 0.5 * wxt + coef[..., None] * torch.cross(omega, wxt, dim=-1)
    return torch.cat([v, omega], dim=-1)


def pose_cost(vals, robot, joint_var, target_pose, target_link_index, pos_weight, ori_weight) -> torch.Tensor:
    """Weighted ``log(T_link^-1 T_target)`` of the target link(s)
    (``src/pyroffi/costs/_costs.py:11-29``): positions of every target first,
    then orientations, flattened, as the reference's concatenation."""
    cfg = _cfg(vals, joint_var)
    if not isinstance(cfg, torch.Tensor):
        cfg = torch.as_tensor(np.asarray(cfg, dtype=np.float32), device="cuda")
    idx = torch.as_tensor(np.asarray(target_link_index), device=cfg.device)
    assert idx.dtype == torch.int32, "target_link_index must be int32, as in the reference"
    links = robot.forward_kinematics(cfg)
    actual = links[..., idx.long(), :]
    residual = se3_log_t(_inv_compose(actual, _as_pose_tensor(target_pose, actual)))
    pw = torch.as_tensor(pos_weight, dtype=residual.dtype, device=residual.device)
    ow = torch.as_tensor(ori_weight, dtype=residual.dtype, device=residual.device)
    return torch.cat([residual[..., :3] * pw, residual[..., 3:] * ow]).flatten()


def self_collision_cost(vals, robot, robot_coll, joint_var, margin: float, weight) -> torch.Tensor:
    """``colldist_from_sdf(self distances, margin) * weight``, flattened
    (``src/pyroffi/costs/_costs.py:169-182``)."""
    d = robot_coll.compute_self_collision_distance(robot, _cfg(vals, joint_var))
    d = d if isinstance(d, torch.Tensor) else torch.as_tensor(np.asarray(d), device="cuda")
    return (colldist_from_sdf(d, margin) * weight).flatten()


def world_collision_cost(vals, robot, robot_coll, joint_var, world_geom, margin: float, weight) -> torch.Tensor:
    """``colldist_from_sdf(world distances, margin) * weight``, flattened
    (``src/pyroffi/costs/_costs.py:185-199``)."""
    d = robot_coll.compute_world_collision_distance(robot, _cfg(vals, joint_var), world_geom)
    d = d if isinstance(d, torch.Tensor) else torch.as_tensor(np.asarray(d), device="cuda")
    return (colldist_from_sdf(d, margin) * weight).flatten()


def _jaxls_factories():
    try:
        import jaxls  # noqa: F401
    except ImportError:
        return None
    return {f.__name__: jaxls.Cost.create_factory(f) for f in (pose_cost, self_collision_cost, world_collision_cost)}


JAXLS_FACTORIES = _jaxls_factories()
