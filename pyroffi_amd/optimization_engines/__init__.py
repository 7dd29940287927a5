"""IK and trajectory solvers."""

from ._lbfgs_trajopt_optimization import LbfgsTrajOptConfig, lbfgs_trajopt, lbfgs_trajopt_uneven
from ._hjcd_ik import hjcd_solve, hjcd_solve_batch
from ._ls_ik import ls_ik_solve, ls_ik_solve_cuda, ls_ik_solve_cuda_batch
from ._ls_ik_batched import UnevenLsIkSolver

__all__ = [
    "LbfgsTrajOptConfig",
    "hjcd_solve",
    "hjcd_solve_batch",
    "ls_ik_solve",
    "UnevenLsIkSolver",
    "lbfgs_trajopt",
    "lbfgs_trajopt_uneven",
    "ls_ik_solve_cuda",
    "ls_ik_solve_cuda_batch",
]
