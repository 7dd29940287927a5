"""Differentiable costs on the GPU."""

from ._differentiable import PoseHingeCost, pose_hinge_cost

__all__ = ["PoseHingeCost", "pose_hinge_cost"]
