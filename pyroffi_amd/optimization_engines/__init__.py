"""IK solvers."""

from ._ls_ik import ls_ik_solve_cuda, ls_ik_solve_cuda_batch

__all__ = ["ls_ik_solve_cuda", "ls_ik_solve_cuda_batch"]
