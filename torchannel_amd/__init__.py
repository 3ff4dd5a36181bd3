"""MI355X-native (gfx950) IMEX time step of the torChannel channel-flow DNS.

Modules mirror the reference's (operators, projection, projection_fft, tridiag, utils, solver,
initflow, turbstats) with the same function names and call signatures on torch tensors; every
hot operation is a hand-written HIP kernel in libtorchannel_hip.so behind a C ABI
(include/torchannel.h).  There is no CPU path.
"""
from . import _lib
from .adjoint import WallShearAdjoint, wall_shear_objective
from .autotune import tune
from .ensemble import ChannelEnsemble, ensemble_configs, make_ensemble
from .initflow import initialize_flow
from .operators import (compute_cfl_fused, compute_momentum_rhs_fused_imex, compute_momentum_rhs_fused_v2,
                        compute_momentum_rhs_separate, solve_implicit_diffusion_u, solve_implicit_diffusion_v, solve_implicit_diffusion_w,
                        solve_tridiagonal_batch)
from .projection import project_velocity
from .projection_fft import initialize_fft_solver, solve_poisson_fft, solve_tridiagonal
from .solver import ChannelFlow, apply_bc_all
from .tridiag import pcr_solve
from .turbstats import TurbulenceStats
from .utils import compute_bulk_velocity, compute_divergence, compute_u_tau, generate_grid, generate_hybrid_grid

__all__ = [
    "ChannelFlow", "apply_bc_all", "compute_bulk_velocity", "compute_cfl_fused", "compute_divergence",
    "compute_momentum_rhs_fused_imex", "compute_momentum_rhs_fused_v2", "compute_momentum_rhs_separate",
    "compute_u_tau", "generate_grid", "generate_hybrid_grid",
    "initialize_fft_solver", "initialize_flow", "pcr_solve", "project_velocity", "solve_implicit_diffusion_u",
    "solve_implicit_diffusion_v", "solve_implicit_diffusion_w", "solve_poisson_fft", "solve_tridiagonal",
    "solve_tridiagonal_batch", "TurbulenceStats", "WallShearAdjoint", "wall_shear_objective", "tune", "ChannelEnsemble",
    "ensemble_configs", "make_ensemble",
]
