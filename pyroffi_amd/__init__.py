"""pyroffi_amd: pyroffi's batched kinematics hot path, native to MI355X (gfx950).

Public API (same names and arguments as pyroffi):

* ``Robot.from_urdf`` / ``Robot.forward_kinematics`` (differentiable: its
  backward is the FK adjoint kernel) / ``Robot.inverse_kinematics`` (HJCD-IK)
* ``collision.RobotCollisionSpherized``, ``collision.CUDARobotCollisionChecker``,
  ``collision.colldist_from_sdf``
* ``ls_ik_solve`` (the JAX-convention LM), ``hjcd_solve`` /
  ``hjcd_solve_batch``, ``ls_ik_solve_cuda`` / ``ls_ik_solve_cuda_batch``;
  ``UnevenLsIkSolver`` for uneven multi-robot batches in one graph
* ``costs.pose_cost`` / ``costs.self_collision_cost`` /
  ``costs.world_collision_cost`` (the residual factories; jaxls is optional) and
  ``costs.pose_hinge_cost`` (differentiable, hand-written adjoint)
* ``lbfgs_trajopt`` / ``LbfgsTrajOptConfig`` (direct L-BFGS trajectory
  optimisation) and ``lbfgs_trajopt_uneven``; ``chomp_trajopt`` /
  ``ChompTrajOptConfig`` (CHOMP with the covariant update);
  ``sco_trajopt`` / ``ScoTrajOptConfig`` / ``make_init_trajs`` (sequential
  convex optimisation)

``save_model`` / ``load_model`` store a robot's tables and collision model on
disk (no URDF needed to load them).

All computation runs in the HIP library ``libpyroffi_amd.so`` (C ABI in
``include/pyroffi_amd.h``); torch supplies device memory and the stream.
"""

from . import collision, costs, kernels
from ._model_file import load_model, save_model
from ._robot import Robot
from ._robot_urdf_parser import JointInfo, LinkInfo, RobotURDFParser
from ._se3 import SE3, SO3
from ._urdf import URDF, bundled_robot_names, bundled_srdf_path, bundled_urdf_path, load_bundled_urdf
from .optimization_engines import (
    ChompTrajOptConfig,
    LbfgsTrajOptConfig,
    ScoTrajOptConfig,
    UnevenLsIkSolver,
    chomp_trajopt,
    hjcd_solve,
    hjcd_solve_batch,
    lbfgs_trajopt,
    lbfgs_trajopt_uneven,
    ls_ik_solve,
    ls_ik_solve_cuda,
    ls_ik_solve_cuda_batch,
    make_init_trajs,
    sco_trajopt,
)

__all__ = [
    "ChompTrajOptConfig",
    "LbfgsTrajOptConfig",
    "ScoTrajOptConfig",
    "chomp_trajopt",
    "make_init_trajs",
    "sco_trajopt",
    "UnevenLsIkSolver",
    "lbfgs_trajopt",
    "lbfgs_trajopt_uneven",
    "SE3",
    "SO3",
    "URDF",
    "JointInfo",
    "LinkInfo",
    "Robot",
    "RobotURDFParser",
    "bundled_robot_names",
    "bundled_srdf_path",
    "bundled_urdf_path",
    "collision",
    "costs",
    "kernels",
    "hjcd_solve",
    "hjcd_solve_batch",
    "load_bundled_urdf",
    "load_model",
    "save_model",
    "ls_ik_solve",
    "ls_ik_solve_cuda",
    "ls_ik_solve_cuda_batch",
]
