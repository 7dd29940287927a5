"""IK and trajectory solvers."""

from ._chomp_optimization import ChompTrajOptConfig, chomp_trajopt, chomp_trajopt_uneven
from ._lbfgs_trajopt_optimization import LbfgsTrajOptConfig, lbfgs_trajopt, lbfgs_trajopt_uneven
from ._hjcd_ik import hjcd_solve, hjcd_solve_batch
from ._ls_ik import ls_ik_solve, ls_ik_solve_cuda, ls_ik_solve_cuda_batch
from ._ls_ik_batched import UnevenLsIkSolver
from ._sco_optimization import ScoTrajOptConfig, TrajOptConfig, make_init_trajs, sco_trajopt

__all__ = [
    "ChompTrajOptConfig",
    "LbfgsTrajOptConfig",
    "chomp_trajopt",
    "chomp_trajopt_uneven",
    "hjcd_solve",
    "hjcd_solve_batch",
    "ls_ik_solve",
    "ScoTrajOptConfig",
    "TrajOptConfig",
    "UnevenLsIkSolver",
    "make_init_trajs",
    "sco_trajopt",
    "lbfgs_trajopt",
    "lbfgs_trajopt_uneven",
    "ls_ik_solve_cuda",
    "ls_ik_solve_cuda_batch",
]
