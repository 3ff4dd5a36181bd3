"""Explicit right-hand side, Crank-Nicolson wall-normal diffusion and CFL rate (operators.py)
on the HIP library."""
import torch

from . import _lib as L
from .tridiag import pcr_solve


def _grid_from_u(u, nx, ny, nz, dx, dy, dz_c, dz_f, nu, top="dirichlet", others=()):
    """The grid of a call on field arrays u (and `others`), for their dtype and layout."""
    dt, pitch = L.require_fields(u, *others)
    return L.get_grid(nx, ny, nz, dx, dy, float(nu), L.top_code(top), dt, u.device, dz_f=dz_f, dz_c=dz_c,
                      pitch=pitch)


def _rhs(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu, branch):
    g = _grid_from_u(u, nx, ny, nz, dx, dy, dz_c, dz_f, nu, others=(v, w))
    ru, rv, rw = L.field_like(u), L.field_like(v), L.field_like(w)
    L.call("tc_rhs_imex", g.handle, branch, L.ptr(u), L.ptr(v), L.ptr(w), L.ptr(ru), L.ptr(rv),
           L.ptr(rw), L.stream_of(u))
    return ru, rv, rw


def compute_momentum_rhs_fused_imex(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu):
    """operators.py:1037-1225: (rhs_u, rhs_v, rhs_w) = nu*lap_xy(q) - div(q u), zero outside the
    interior (and on u-face nx / v-face ny, operators.py:1122, 1173)."""
    return _rhs(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu, L.TC_BRANCH_GPU)


def compute_momentum_rhs_separate(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu):
    """The CPU branch of step_imex (solver.py:601-622): diffusion_xy_* - advection_*, which also
    gives u-face nx and v-face ny their xy-Laplacian (operators.py:495, 521)."""
    return _rhs(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu, L.TC_BRANCH_CPU)


def compute_momentum_rhs_fused_v2(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu):
    """operators.py:817-1034: diffusion (x, y and z, all explicit) minus advection for the FE
    scheme's right-hand side, zero outside the interiors (u-face nx and v-face ny included)."""
    g = _grid_from_u(u, nx, ny, nz, dx, dy, dz_c, dz_f, nu, others=(v, w))
    ru, rv, rw = L.field_like(u), L.field_like(v), L.field_like(w)
    L.call("tc_rhs_fused_v2", g.handle, L.ptr(u), L.ptr(v), L.ptr(w), L.ptr(ru), L.ptr(rv), L.ptr(rw),
           L.stream_of(u))
    return ru, rv, rw


def _cn(q, comp, dt, nx, ny, nz, dz_c, dz_f, nu, theta, top):
    g = _grid_from_u(q, nx, ny, nz, 1.0, 1.0, dz_c, dz_f, nu, top)
    out = L.clone_field(q)
    if isinstance(dt, torch.Tensor):
        dtd = dt.to(device=q.device, dtype=torch.float64).reshape(()).contiguous()
        L.call("tc_cn_solve", g.handle, comp, 0.0, L.ptr(dtd), float(theta), L.ptr(q), L.ptr(out),
               L.stream_of(q))
    else:
        L.call("tc_cn_solve", g.handle, comp, float(dt), None, float(theta), L.ptr(q), L.ptr(out),
               L.stream_of(q))
    return out


def solve_implicit_diffusion_u(u, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5, top_wall_bc_type="dirichlet"):
    """operators.py:563-654: theta-scheme z-diffusion of u; returns a clone with the interior
    replaced."""
    return _cn(u, 0, dt, nx, ny, nz, dz_c, dz_f, nu, theta, top_wall_bc_type)


def solve_implicit_diffusion_v(v, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5, top_wall_bc_type="dirichlet"):
    """operators.py:657-731."""
    return _cn(v, 1, dt, nx, ny, nz, dz_c, dz_f, nu, theta, top_wall_bc_type)


def solve_implicit_diffusion_w(w, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5):
    """operators.py:734-811, including the reference's shifted explicit half for w."""
    return _cn(w, 2, dt, nx, ny, nz, dz_c, dz_f, nu, theta, "dirichlet")


def solve_tridiagonal_batch(a, b, c, d):
    """operators.py:551-560 (PCR, shared or batched diagonals)."""
    return pcr_solve(a, b, c, d)


def compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c):
    """operators.py:1232-1333: largest convective rate as a Python float (one device sync)."""
    g = _grid_from_u(u, nx, ny, nz, dx, dy, dz_c, dz_f, 0.0, others=(v, w))
    out = torch.empty((), dtype=torch.float64, device=u.device)
    L.call("tc_cfl_rate", g.handle, L.ptr(u), L.ptr(v), L.ptr(w), L.ptr(out), L.stream_of(u))
    return float(out.item())
