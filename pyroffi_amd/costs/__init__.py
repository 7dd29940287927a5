"""Costs on the GPU: the reference's residual factories (pose, self- and
world-collision) and the differentiable pose + hinge cost."""

from ._costs import JAXLS_FACTORIES, pose_cost, se3_log_t, self_collision_cost, world_collision_cost
from ._differentiable import PoseHingeCost, pose_hinge_cost

__all__ = ["JAXLS_FACTORIES", "PoseHingeCost", "pose_cost", "pose_hinge_cost", "se3_log_t", "self_collision_cost",
           "world_collision_cost"]
