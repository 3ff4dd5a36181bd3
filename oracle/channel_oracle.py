"""CPU FP64 restatement of torChannel's IMEX time step.

THIS MODULE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker (or, for
``cpu_baseline``, as the timed CPU port).  The product path in ``torchannel_amd`` never calls
it: the product runs hand-written HIP kernels and raises when they are missing.

Everything here runs on CPU tensors in float64 and follows the reference's arithmetic on the
ghosted staggered arrays:

    u (nx+1, ny+2, nz+2)   v (nx+2, ny+1, nz+2)   w (nx+2, ny+2, nz+1)   p (nx+2, ny+2, nz+2)

Each function cites the reference lines it restates.  Pins (what checks each function against
something other than itself) are listed in ``PINS`` at the bottom of the file; the tests in
``tests/test_oracle_*.py`` exercise every entry.
"""
from __future__ import annotations

import math

import numpy as np
import torch

F64 = torch.float64


# ----------------------------------------------------------------------------------------------
# Grid (utils.py:5-45) and initial field (initflow.py:4-104, solver.py:268-276)
# ----------------------------------------------------------------------------------------------

def generate_grid(gamma: float, nz: int, Lz: float, stretching_type: str = "symmetric"):
    """tanh-stretched wall-normal grid; utils.py:22-45.

    Returns (z_f[nz+1], z_c[nz+2], dz_f[nz], dz_c[nz+1]).  z_c carries one mirrored ghost
    centre below the bottom wall and one above the top wall.
    """
    kk = torch.linspace(0, nz, nz + 1, dtype=F64)
    g = torch.tensor(gamma, dtype=F64)
    if stretching_type == "bottom":                              # utils.py:24-30
        s = kk / nz
        z_f = Lz * (1.0 - torch.tanh(gamma * (1.0 - s)) / torch.tanh(g))
    else:                                                        # utils.py:31-35
        s = (2 * kk / nz) - 1
        z_f = 0.5 * Lz * (1 + torch.tanh(gamma * s) / torch.tanh(g))
    mid = 0.5 * (z_f[:-1] + z_f[1:])                             # utils.py:37-39
    lo = torch.tensor([-mid[0]], dtype=F64)
    hi = torch.tensor([2 * z_f[-1] - mid[-1]], dtype=F64)
    z_c = torch.cat([lo, mid, hi])
    return z_f, z_c, z_f[1:] - z_f[:-1], z_c[1:] - z_c[:-1]      # utils.py:42-43


def generate_hybrid_grid(nz_uniform: int, nz_stretched: int, z_transition: float, Lz: float, gamma: float):
    """Uniform region [0, z_transition] plus one-sided tanh region above; utils.py:47-113.

    Faces: linspace over the uniform part (utils.py:68-69), then
    z_transition + H (1 - tanh(gamma (1 - xi)) / tanh(gamma)), xi = k/nz_stretched, k >= 1
    (utils.py:71-82); centres and spacings as in generate_grid (utils.py:84-90).
    """
    z_lo = torch.linspace(0, z_transition, nz_uniform + 1, dtype=F64)
    H = Lz - z_transition
    xi = torch.linspace(0, nz_stretched, nz_stretched + 1, dtype=F64) / nz_stretched
    z_hi = H * (1.0 - torch.tanh(gamma * (1.0 - xi)) / torch.tanh(torch.tensor(gamma, dtype=F64)))
    z_f = torch.cat([z_lo, z_transition + z_hi[1:]])
    mid = 0.5 * (z_f[:-1] + z_f[1:])
    z_c = torch.cat([torch.tensor([-mid[0]], dtype=F64), mid, torch.tensor([2 * z_f[-1] - mid[-1]], dtype=F64)])
    dz_f = z_f[1:] - z_f[:-1]
    if bool(torch.any(dz_f <= 0)) or bool(torch.any(z_f[1:] <= z_f[:-1])):   # utils.py:106-111
        raise ValueError("non-monotone hybrid grid")
    return z_f, z_c, dz_f, z_c[1:] - z_c[:-1]


def initialize_flow(nx, ny, nz, z_c, Lz, U_bulk=1.0, perturbation_intensity=0.0,
                    top_wall_bc_type="dirichlet", seed=None):
    """Laminar parabola plus uniform noise; initflow.py:6-26, 84-102.

    The reference draws ``torch.rand_like`` from the global CPU generator in the order u, v, w
    and never seeds it (initflow.py:90-92).  ``seed`` (if given) calls ``torch.manual_seed``
    first, which defines the seeded start used by the tests and the benchmark.
    """
    u = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=F64)
    v = torch.zeros(nx + 2, ny + 1, nz + 2, dtype=F64)
    w = torch.zeros(nx + 2, ny + 2, nz + 1, dtype=F64)
    p = torch.zeros(nx + 2, ny + 2, nz + 2, dtype=F64)
    peak = 1.5 * U_bulk
    if top_wall_bc_type == "neumann":                            # initflow.py:13-19
        zn = z_c / Lz
        prof = peak * zn * (2 - zn)
    else:                                                        # initflow.py:20-24
        zn = 2 * z_c / Lz - 1
        prof = peak * (1 - zn ** 2)
    u[:, :, :] = prof[None, None, :]
    if perturbation_intensity > 0.0:                             # initflow.py:84-92
        if seed is not None:
            torch.manual_seed(seed)
        amp = perturbation_intensity * U_bulk
        for q in (u, v, w):
            box = q[1:-1, 1:-1, 2:-2]
            box += (torch.rand_like(box) - 0.5) * 2 * amp
    u[:, :, 0] = -u[:, :, 1]                                     # initflow.py:95-102
    u[:, :, -1] = -u[:, :, -2]
    v[:, :, 0] = -v[:, :, 1]
    v[:, :, -1] = -v[:, :, -2]
    w[:, :, 0] = 0.0
    w[:, :, -1] = 0.0
    return u, v, w, p


# ----------------------------------------------------------------------------------------------
# Boundary conditions (solver.py:23-79)
# ----------------------------------------------------------------------------------------------

def apply_bc_all(u, v, w, top_wall_bc_type="dirichlet"):
    """Periodic x/y copies, then wall reflections in z, in the reference's order (x, y, z)."""
    top = 1.0 if top_wall_bc_type == "neumann" else -1.0
    u[0] = u[-1]                         # solver.py:45
    u[:, 0] = u[:, -2]                   # solver.py:47
    u[:, -1] = u[:, 1]                   # solver.py:48
    u[:, :, 0] = -u[:, :, 1]             # solver.py:50
    u[:, :, -1] = top * u[:, :, -2]      # solver.py:51-54
    v[0] = v[-2]                         # solver.py:58
    v[-1] = v[1]                         # solver.py:59
    v[:, 0] = v[:, -1]                   # solver.py:61
    v[:, :, 0] = -v[:, :, 1]             # solver.py:63
    v[:, :, -1] = top * v[:, :, -2]      # solver.py:64-67
    w[0] = w[-2]                         # solver.py:71
    w[-1] = w[1]                         # solver.py:72
    w[:, 0] = w[:, -2]                   # solver.py:74
    w[:, -1] = w[:, 1]                   # solver.py:75
    w[:, :, 0] = 0.0                     # solver.py:78
    w[:, :, -1] = 0.0                    # solver.py:79


# ----------------------------------------------------------------------------------------------
# Explicit right-hand side: advection + xy-diffusion
# ----------------------------------------------------------------------------------------------

def _h(a, b):
    return 0.5 * (a + b)


def rhs_fused_imex(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu):
    """GPU-branch explicit RHS, compute_momentum_rhs_fused_imex (operators.py:1037-1225).

    rhs = nu*(d2/dx2 + d2/dy2) q - div(q u) on i=1..nx-1 (u), j=1..ny-1 (v), k=1..nz-1 (w);
    zero everywhere else, including u-face nx and v-face ny (operators.py:1122, 1173).
    """
    ru = torch.zeros_like(u)
    rv = torch.zeros_like(v)
    rw = torch.zeros_like(w)
    rdx, rdy = 1.0 / dx, 1.0 / dy
    kxx = nu * (1.0 / (dx * dx))
    kyy = nu * (1.0 / (dy * dy))
    rdzf = (1.0 / dz_f).view(1, 1, -1)
    I, J, K = slice(1, nx), slice(1, ny + 1), slice(1, nz + 1)
    # --- u on faces 1..nx-1 (operators.py:1078-1122)
    fx = (_h(u[1:nx, J, K], u[2:nx + 1, J, K]) ** 2 - _h(u[0:nx - 1, J, K], u[1:nx, J, K]) ** 2) * rdx
    fy = (_h(v[1:nx, 1:ny + 1, K], v[2:nx + 1, 1:ny + 1, K]) * _h(u[I, 1:ny + 1, K], u[I, 2:ny + 2, K])
          - _h(v[1:nx, 0:ny, K], v[2:nx + 1, 0:ny, K]) * _h(u[I, 0:ny, K], u[I, 1:ny + 1, K])) * rdy
    fz = (_h(w[1:nx, J, 1:nz + 1], w[2:nx + 1, J, 1:nz + 1]) * _h(u[I, J, 1:nz + 1], u[I, J, 2:nz + 2])
          - _h(w[1:nx, J, 0:nz], w[2:nx + 1, J, 0:nz]) * _h(u[I, J, 0:nz], u[I, J, 1:nz + 1])) * rdzf[:, :, 0:nz]
    lap = ((u[2:nx + 1, J, K] - 2.0 * u[1:nx, J, K] + u[0:nx - 1, J, K]) * kxx
           + (u[I, 2:ny + 2, K] - 2.0 * u[I, J, K] + u[I, 0:ny, K]) * kyy)
    ru[1:nx, J, K] = lap - (fx + fy + fz)
    # --- v on faces 1..ny-1 (operators.py:1129-1173)
    I2, J2 = slice(1, nx + 1), slice(1, ny)
    fx = (_h(u[1:nx + 1, 1:ny, K], u[1:nx + 1, 2:ny + 1, K]) * _h(v[1:nx + 1, J2, K], v[2:nx + 2, J2, K])
          - _h(u[0:nx, 1:ny, K], u[0:nx, 2:ny + 1, K]) * _h(v[0:nx, J2, K], v[1:nx + 1, J2, K])) * rdx
    fy = (_h(v[I2, 1:ny, K], v[I2, 2:ny + 1, K]) ** 2 - _h(v[I2, 0:ny - 1, K], v[I2, 1:ny, K]) ** 2) * rdy
    fz = (_h(w[I2, 1:ny, 1:nz + 1], w[I2, 2:ny + 1, 1:nz + 1]) * _h(v[I2, J2, 1:nz + 1], v[I2, J2, 2:nz + 2])
          - _h(w[I2, 1:ny, 0:nz], w[I2, 2:ny + 1, 0:nz]) * _h(v[I2, J2, 0:nz], v[I2, J2, 1:nz + 1])) * rdzf[:, :, 0:nz]
    lap = ((v[2:nx + 2, J2, K] - 2.0 * v[I2, J2, K] + v[0:nx, J2, K]) * kxx
           + (v[I2, 2:ny + 1, K] - 2.0 * v[I2, J2, K] + v[I2, 0:ny - 1, K]) * kyy)
    rv[I2, J2, K] = lap - (fx + fy + fz)
    # --- w on faces 1..nz-1 (operators.py:1179-1223); the ww flux divides by dz_c
    Kw = slice(1, nz)
    fx = (_h(u[1:nx + 1, J, 1:nz], u[1:nx + 1, J, 2:nz + 1]) * _h(w[I2, J, Kw], w[2:nx + 2, J, Kw])
          - _h(u[0:nx, J, 1:nz], u[0:nx, J, 2:nz + 1]) * _h(w[0:nx, J, Kw], w[I2, J, Kw])) * rdx
    fy = (_h(v[I2, 1:ny + 1, 1:nz], v[I2, 1:ny + 1, 2:nz + 1]) * _h(w[I2, J, Kw], w[I2, 2:ny + 2, Kw])
          - _h(v[I2, 0:ny, 1:nz], v[I2, 0:ny, 2:nz + 1]) * _h(w[I2, 0:ny, Kw], w[I2, J, Kw])) * rdy
    fz = (_h(w[I2, J, 1:nz], w[I2, J, 2:nz + 1]) ** 2 - _h(w[I2, J, 0:nz - 1], w[I2, J, 1:nz]) ** 2) \
        / dz_c[1:nz].view(1, 1, -1)
    lap = ((w[2:nx + 2, J, Kw] - 2.0 * w[I2, J, Kw] + w[0:nx, J, Kw]) * kxx
           + (w[I2, 2:ny + 2, Kw] - 2.0 * w[I2, J, Kw] + w[I2, 0:ny, Kw]) * kyy)
    rw[I2, J, Kw] = lap - (fx + fy + fz)
    return ru, rv, rw


def rhs_fused_v2(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu):
    """compute_momentum_rhs_fused_v2 (operators.py:817-1034), the FE scheme's explicit RHS.

    The advection and xy-diffusion of rhs_fused_imex on the same interiors, plus the explicit z
    second derivative: for u and v (operators.py:904-909, 962-963) over 1/dz_c[m], 1/dz_c[m+1]
    and 1/dz_f[m] at cell m+1; for w (operators.py:1021-1026) dividing by dz_f[k-1], dz_f[k] and
    the control volume dz_c[k-1] (not dz_c[k]).
    """
    ru, rv, rw = rhs_fused_imex(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu)
    J, K = slice(1, ny + 1), slice(1, nz + 1)
    rzc = (1.0 / dz_c).view(1, 1, -1)
    rzf = (1.0 / dz_f).view(1, 1, -1)

    def dzz(q):
        return nu * ((q[:, :, 2:nz + 2] - q[:, :, K]) * rzc[:, :, 1:nz + 1]
                     - (q[:, :, K] - q[:, :, 0:nz]) * rzc[:, :, 0:nz]) * rzf

    ru[1:nx, J, K] += dzz(u[1:nx, J, :])
    rv[1:nx + 1, 1:ny, K] += dzz(v[1:nx + 1, 1:ny, :])
    wc = w[1:nx + 1, J, :]
    dzw = nu * ((wc[:, :, 2:nz + 1] - wc[:, :, 1:nz]) / dz_f[1:nz].view(1, 1, -1)
                - (wc[:, :, 1:nz] - wc[:, :, 0:nz - 1]) / dz_f[0:nz - 1].view(1, 1, -1)) / dz_c[0:nz - 1].view(1, 1, -1)
    rw[1:nx + 1, J, 1:nz] += dzw
    return ru, rv, rw


def rhs_separate(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu):
    """CPU-branch explicit RHS: diffusion_xy_* - advection_* (solver.py:601-622).

    advection_u/v/w (operators.py:126-247) divide by the spacings and fill i=1..nx-1 (u),
    j=1..ny-1 (v); diffusion_xy_u/v (operators.py:474-523) fill i=1..nx / j=1..ny, so u-face nx
    and v-face ny carry the xy-Laplacian only (periodic wrap through the face-1 copy).
    """
    J, K = slice(1, ny + 1), slice(1, nz + 1)
    I2 = slice(1, nx + 1)
    dzf = dz_f.view(1, 1, -1)
    # advection_u (operators.py:134-165)
    a_u = torch.zeros_like(u)
    fx = (_h(u[1:nx, J, K], u[2:nx + 1, J, K]) ** 2 - _h(u[0:nx - 1, J, K], u[1:nx, J, K]) ** 2) / dx
    fy = (_h(v[1:nx, 1:ny + 1, K], v[2:nx + 1, 1:ny + 1, K]) * _h(u[1:nx, 1:ny + 1, K], u[1:nx, 2:ny + 2, K])
          - _h(v[1:nx, 0:ny, K], v[2:nx + 1, 0:ny, K]) * _h(u[1:nx, 0:ny, K], u[1:nx, 1:ny + 1, K])) / dy
    fz = (_h(w[1:nx, J, 1:nz + 1], w[2:nx + 1, J, 1:nz + 1]) * _h(u[1:nx, J, 1:nz + 1], u[1:nx, J, 2:nz + 2])
          - _h(w[1:nx, J, 0:nz], w[2:nx + 1, J, 0:nz]) * _h(u[1:nx, J, 0:nz], u[1:nx, J, 1:nz + 1])) / dzf[:, :, 0:nz]
    a_u[1:nx, J, K] = fx + fy + fz
    # advection_v (operators.py:175-204)
    a_v = torch.zeros_like(v)
    J2 = slice(1, ny)
    fx = (_h(u[1:nx + 1, 1:ny, K], u[1:nx + 1, 2:ny + 1, K]) * _h(v[1:nx + 1, J2, K], v[2:nx + 2, J2, K])
          - _h(u[0:nx, 1:ny, K], u[0:nx, 2:ny + 1, K]) * _h(v[0:nx, J2, K], v[1:nx + 1, J2, K])) / dx
    fy = (_h(v[I2, 1:ny, K], v[I2, 2:ny + 1, K]) ** 2 - _h(v[I2, 0:ny - 1, K], v[I2, 1:ny, K]) ** 2) / dy
    fz = (_h(w[I2, 1:ny, 1:nz + 1], w[I2, 2:ny + 1, 1:nz + 1]) * _h(v[I2, J2, 1:nz + 1], v[I2, J2, 2:nz + 2])
          - _h(w[I2, 1:ny, 0:nz], w[I2, 2:ny + 1, 0:nz]) * _h(v[I2, J2, 0:nz], v[I2, J2, 1:nz + 1])) / dzf[:, :, 0:nz]
    a_v[I2, J2, K] = fx + fy + fz
    # advection_w (operators.py:216-245)
    a_w = torch.zeros_like(w)
    Kw = slice(1, nz)
    fx = (_h(u[1:nx + 1, J, 1:nz], u[1:nx + 1, J, 2:nz + 1]) * _h(w[I2, J, Kw], w[2:nx + 2, J, Kw])
          - _h(u[0:nx, J, 1:nz], u[0:nx, J, 2:nz + 1]) * _h(w[0:nx, J, Kw], w[I2, J, Kw])) / dx
    fy = (_h(v[I2, 1:ny + 1, 1:nz], v[I2, 1:ny + 1, 2:nz + 1]) * _h(w[I2, J, Kw], w[I2, 2:ny + 2, Kw])
          - _h(v[I2, 0:ny, 1:nz], v[I2, 0:ny, 2:nz + 1]) * _h(w[I2, 0:ny, Kw], w[I2, J, Kw])) / dy
    fz = (_h(w[I2, J, 1:nz], w[I2, J, 2:nz + 1]) ** 2 - _h(w[I2, J, 0:nz - 1], w[I2, J, 1:nz]) ** 2) \
        / dz_c[1:nz].view(1, 1, -1)
    a_w[I2, J, Kw] = fx + fy + fz
    # diffusion_xy_u (operators.py:482-495): x-periodic through an appended copy of face 1
    d_u = torch.zeros_like(u)
    ue = torch.cat([u, u[1:2]], dim=0)
    d_u[I2, J, K] = nu * ((ue[2:nx + 2, J, K] - 2 * ue[I2, J, K] + ue[0:nx, J, K]) / dx ** 2
                          + (u[I2, 2:ny + 2, K] - 2 * u[I2, J, K] + u[I2, 0:ny, K]) / dy ** 2)
    # diffusion_xy_v (operators.py:508-521): y-periodic through an appended copy of face 1
    d_v = torch.zeros_like(v)
    ve = torch.cat([v, v[:, 1:2]], dim=1)
    d_v[I2, J, K] = nu * ((v[2:nx + 2, J, K] - 2 * v[I2, J, K] + v[0:nx, J, K]) / dx ** 2
                          + (ve[I2, 2:ny + 2, K] - 2 * ve[I2, J, K] + ve[I2, 0:ny, K]) / dy ** 2)
    # diffusion_xy_w (operators.py:534-546)
    d_w = torch.zeros_like(w)
    d_w[I2, J, Kw] = nu * ((w[2:nx + 2, J, Kw] - 2 * w[I2, J, Kw] + w[0:nx, J, Kw]) / dx ** 2
                           + (w[I2, 2:ny + 2, Kw] - 2 * w[I2, J, Kw] + w[I2, 0:ny, Kw]) / dy ** 2)
    return d_u - a_u, d_v - a_v, d_w - a_w


# ----------------------------------------------------------------------------------------------
# Tridiagonal solvers: PCR (tridiag.py:20-58) and Thomas (projection_fft.py:167-215)
# ----------------------------------------------------------------------------------------------

def _shifted(x, s, down, pad):
    """y[m] = x[m-s] (down) or x[m+s] (up) along the last axis; out-of-range rows = pad."""
    n = x.shape[-1]
    y = torch.full_like(x, pad)
    if s < n:
        if down:
            y[..., s:] = x[..., :n - s]
        else:
            y[..., :n - s] = x[..., s:]
    return y


def pcr_solve(a, b, c, d):
    """Parallel cyclic reduction; tridiag.py:33-58.  a[0] and c[n-1] are ignored by the
    recurrence only through the pads (b->1, a/c/d->0); callers zero them as the reference does."""
    n = d.shape[-1]
    a, b, c, d = a.clone(), b.clone(), c.clone(), d.clone()
    s = 1
    while s < n:
        al = -a / _shifted(b, s, True, 1.0)
        ga = -c / _shifted(b, s, False, 1.0)
        b = b + al * _shifted(c, s, True, 0.0) + ga * _shifted(a, s, False, 0.0)
        d = d + al * _shifted(d, s, True, 0.0) + ga * _shifted(d, s, False, 0.0)
        a = al * _shifted(a, s, True, 0.0)
        c = ga * _shifted(c, s, False, 0.0)
        s *= 2
    return d / b


def thomas_solve(a, b, c, d):
    """Serial Thomas sweep over the last axis; projection_fft.py:184-215 (a, b, c, d all 2-D)."""
    nb, n = d.shape
    cp = torch.zeros((nb, n), dtype=d.dtype)
    dp = torch.zeros((nb, n), dtype=d.dtype)
    x = torch.zeros((nb, n), dtype=d.dtype)
    cp[:, 0] = c[:, 0] / b[:, 0]
    dp[:, 0] = d[:, 0] / b[:, 0]
    for m in range(1, n):
        den = b[:, m] - a[:, m] * cp[:, m - 1]
        if m < n - 1:
            cp[:, m] = c[:, m] / den
        dp[:, m] = (d[:, m] - a[:, m] * dp[:, m - 1]) / den
    x[:, n - 1] = dp[:, n - 1]
    for m in range(n - 2, -1, -1):
        x[:, m] = dp[:, m] - cp[:, m] * x[:, m + 1]
    return x


# ----------------------------------------------------------------------------------------------
# Crank-Nicolson wall-normal diffusion (operators.py:563-811)
# ----------------------------------------------------------------------------------------------

def cn_matrix_cell(dt, nz, dz_c, dz_f, nu, theta=0.5, top_wall_bc_type="dirichlet"):
    """Shared tridiagonal rows for u and v; operators.py:581-620."""
    alpha = theta * dt * nu
    dl, dr, dc = dz_c[0:nz], dz_c[1:nz + 1], dz_f[0:nz]
    lo = -alpha / (dl * dc)
    ce = 1.0 + alpha * (1.0 / dl + 1.0 / dr) / dc
    up = -alpha / (dr * dc)
    a, b, c = lo.clone(), ce.clone(), up.clone()
    a[0] = 0.0
    b[0] = ce[0] - lo[0]
    if top_wall_bc_type == "neumann":
        b[nz - 1] = ce[nz - 1] + up[nz - 1]
    else:
        b[nz - 1] = ce[nz - 1] - up[nz - 1]
    c[nz - 1] = 0.0
    return a, b, c


def cn_matrix_face(dt, nz, dz_c, dz_f, nu, theta=0.5):
    """Shared tridiagonal rows for w (nz-1 interior faces); operators.py:754-781."""
    alpha = theta * dt * nu
    ni = nz - 1
    dl, dr, dv = dz_f[0:ni], dz_f[1:nz], dz_c[1:nz]
    a = (-alpha / (dl * dv)).clone()
    b = (1.0 + alpha * (1.0 / dl + 1.0 / dr) / dv).clone()
    c = (-alpha / (dr * dv)).clone()
    a[0] = 0.0
    c[ni - 1] = 0.0
    return a, b, c


def solve_implicit_diffusion_uv(q, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5,
                                top_wall_bc_type="dirichlet"):
    """CN solve for u or v (identical code paths, operators.py:563-731); returns a clone."""
    dt = float(dt)
    out = q.clone()
    a, b, c = cn_matrix_cell(dt, nz, dz_c, dz_f, nu, theta, top_wall_bc_type)
    rhs = q[1:nx + 1, 1:ny + 1, 1:nz + 1].clone()
    if theta < 1.0:                                              # operators.py:627-644
        beta = (1.0 - theta) * dt * nu
        dl, dr, dc = dz_c[0:nz], dz_c[1:nz + 1], dz_f[0:nz]
        col = q[1:nx + 1, 1:ny + 1, :]
        lap = ((col[:, :, 2:nz + 2] - col[:, :, 1:nz + 1]) / dr
               - (col[:, :, 1:nz + 1] - col[:, :, 0:nz]) / dl) / dc
        lap[:, :, 0] = ((col[:, :, 2] - col[:, :, 1]) / dr[0] - 2.0 * col[:, :, 1] / dl[0]) / dc[0]
        if top_wall_bc_type == "neumann":
            lap[:, :, nz - 1] = (-(col[:, :, nz] - col[:, :, nz - 1]) / dl[nz - 1]) / dc[nz - 1]
        else:
            lap[:, :, nz - 1] = (-2.0 * col[:, :, nz] / dr[nz - 1]
                                 - (col[:, :, nz] - col[:, :, nz - 1]) / dl[nz - 1]) / dc[nz - 1]
        rhs += beta * lap
    x = pcr_solve(a, b, c, rhs.reshape(nx * ny, nz))             # operators.py:648-652
    out[1:nx + 1, 1:ny + 1, 1:nz + 1] = x.reshape(nx, ny, nz)
    return out


def solve_implicit_diffusion_w(w, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5):
    """CN solve for w with the reference's shifted explicit half; operators.py:734-811."""
    dt = float(dt)
    out = w.clone()
    ni = nz - 1
    a, b, c = cn_matrix_face(dt, nz, dz_c, dz_f, nu, theta)
    rhs = w[1:nx + 1, 1:ny + 1, 1:nz].clone()
    if theta < 1.0:                                              # operators.py:788-803
        beta = (1.0 - theta) * dt * nu
        dl, dr, dv = dz_f[0:ni], dz_f[1:nz], dz_c[1:nz]
        col = w[1:nx + 1, 1:ny + 1, :]
        lap = torch.empty((nx, ny, ni), dtype=w.dtype)
        lap[:, :, 1:ni - 1] = ((col[:, :, 2:ni] - col[:, :, 1:ni - 1]) / dr[1:ni - 1]
                               - (col[:, :, 1:ni - 1] - col[:, :, 0:ni - 2]) / dl[1:ni - 1]) / dv[1:ni - 1]
        lap[:, :, 0] = ((col[:, :, 1] - col[:, :, 0]) / dr[0] - col[:, :, 0] / dl[0]) / dv[0]
        lap[:, :, ni - 1] = (-col[:, :, ni - 1] / dr[ni - 1]
                             - (col[:, :, ni - 1] - col[:, :, ni - 2]) / dl[ni - 1]) / dv[ni - 1]
        rhs += beta * lap
    x = pcr_solve(a, b, c, rhs.reshape(nx * ny, ni))
    out[1:nx + 1, 1:ny + 1, 1:nz] = x.reshape(nx, ny, ni)
    return out


# ----------------------------------------------------------------------------------------------
# Divergence, bulk velocity, wall shear (utils.py:176-240), projection (projection.py:143-165)
# ----------------------------------------------------------------------------------------------

def compute_divergence(u, v, w, nx, ny, nz, dx, dy, dz_f):
    """utils.py:235-240."""
    return ((u[1:nx + 1, 1:ny + 1, 1:nz + 1] - u[0:nx, 1:ny + 1, 1:nz + 1]) / dx
            + (v[1:nx + 1, 1:ny + 1, 1:nz + 1] - v[1:nx + 1, 0:ny, 1:nz + 1]) / dy
            + (w[1:nx + 1, 1:ny + 1, 1:nz + 1] - w[1:nx + 1, 1:ny + 1, 0:nz]) / dz_f[0:nz].view(1, 1, -1))


def compute_bulk_velocity(u, cell_vol, total_volume):
    """utils.py:220-223: sum over faces 1..nx weighted by dx*dy*dz_f, / (Lx*Ly*Lz)."""
    nx, ny, nz = cell_vol.shape
    return torch.sum(u[1:nx + 1, 1:ny + 1, 1:nz + 1] * cell_vol) / total_volume


def compute_u_tau(u, z_c, nu, top_wall_bc_type="dirichlet"):
    """utils.py:187-204: means over the whole k=1 and k=nz planes, ghosts included."""
    d = z_c[1]
    ub = torch.sqrt(torch.abs(nu * torch.mean(u[:, :, 1]) / d))
    if top_wall_bc_type == "neumann":
        return ub
    ut = torch.sqrt(torch.abs(nu * torch.mean(u[:, :, -2]) / d))
    return 0.5 * (ub + ut)


def project_velocity(u, v, w, p, nx, ny, nz, dx, dy, dz_c, dz_f, dt):
    """In-place pressure correction q -= dt*grad(p); projection.py:154-165."""
    dt = float(dt)
    u[1:nx + 1, 1:ny + 1, 1:nz + 1] -= dt * ((p[2:nx + 2, 1:ny + 1, 1:nz + 1] - p[1:nx + 1, 1:ny + 1, 1:nz + 1]) / dx)
    v[1:nx + 1, 1:ny + 1, 1:nz + 1] -= dt * ((p[1:nx + 1, 2:ny + 2, 1:nz + 1] - p[1:nx + 1, 1:ny + 1, 1:nz + 1]) / dy)
    w[1:nx + 1, 1:ny + 1, 1:nz] -= dt * ((p[1:nx + 1, 1:ny + 1, 2:nz + 1] - p[1:nx + 1, 1:ny + 1, 1:nz])
                                         / dz_c[1:nz].view(1, 1, -1))
    return u, v, w


# ----------------------------------------------------------------------------------------------
# FFT + tridiagonal Poisson solver (projection_fft.py:3-165)
# ----------------------------------------------------------------------------------------------

def initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f, top_wall_bc_type="dirichlet"):
    """Modified wavenumbers and per-mode tridiagonal rows; projection_fft.py:24-99."""
    kx = torch.fft.fftfreq(nx, d=dx / (2 * math.pi), dtype=F64)
    ky = torch.fft.rfftfreq(ny, d=dy / (2 * math.pi), dtype=F64)
    kxm = (2.0 / dx) * torch.sin(kx * dx / 2.0)
    kym = (2.0 / dy) * torch.sin(ky * dy / 2.0)
    gx, gy = torch.meshgrid(kxm, kym, indexing="ij")
    k2 = gx ** 2 + gy ** 2
    cl = 1.0 / (dz_c[:-1] * dz_f)
    cr = 1.0 / (dz_c[1:] * dz_f)
    nkx, nky = len(kx), len(ky)
    ta = torch.zeros(nkx, nky, nz, dtype=F64)
    ta[:, :, 1:] = cl[1:].view(1, 1, -1)
    tc = torch.zeros(nkx, nky, nz, dtype=F64)
    tc[:, :, :-1] = cr[:-1].view(1, 1, -1)
    tb = -((cl + cr).view(1, 1, -1) + k2.unsqueeze(-1).expand(nkx, nky, nz))
    tb[:, :, 0] += cl[0]
    tb[:, :, -1] += cr[-1]
    return {"kx": kx, "ky": ky, "tri_a": ta, "tri_b": tb, "tri_c": tc, "nx": nx, "ny": ny,
            "nz": nz, "workspace_p": torch.zeros(nx + 2, ny + 2, nz + 2, dtype=F64),
            "top_wall_bc_type": top_wall_bc_type}


def build_poisson_matrix(nx, ny, nz, dx, dy, dz_c, dz_f, pin_pressure=True):
    """Dense 7-point Laplacian of the direct solver; projection.py:3-84.

    Unknown order i fastest: idx = i + j*nx + k*nx*ny (0-based).  Periodic in x and y; in z the
    rows use 1/(dz_f[k] dz_c[k]) below and 1/(dz_f[k] dz_c[k+1]) above, and a wall drops its
    neighbour (Neumann at both walls, whatever the top-wall type).  pin_pressure replaces row 0
    by p_0 (projection.py:80-82).  Small grids only: N^2 entries.
    """
    N = nx * ny * nz
    A = torch.zeros(N, N, dtype=F64)
    ii, jj, kk = torch.meshgrid(torch.arange(nx), torch.arange(ny), torch.arange(nz), indexing="ij")
    ii, jj, kk = ii.reshape(-1), jj.reshape(-1), kk.reshape(-1)
    row = ii + jj * nx + kk * nx * ny
    diag = torch.zeros(N, dtype=F64)
    for di, dj, cf in ((-1, 0, 1.0 / dx ** 2), (1, 0, 1.0 / dx ** 2), (0, -1, 1.0 / dy ** 2), (0, 1, 1.0 / dy ** 2)):
        col = (ii + di) % nx + ((jj + dj) % ny) * nx + kk * nx * ny
        A[row, col] += cf
        diag -= cf
    dzf, dzc = dz_f.to(F64), dz_c.to(F64)
    lo = kk > 0
    cf = 1.0 / (dzf[kk] * dzc[kk])
    A[row[lo], row[lo] - nx * ny] = cf[lo]
    diag[lo] -= cf[lo]
    hi = kk < nz - 1
    cf = 1.0 / (dzf[kk] * dzc[kk + 1])
    A[row[hi], row[hi] + nx * ny] = cf[hi]
    diag[hi] -= cf[hi]
    A[row, row] += diag
    if pin_pressure:
        A[0, :] = 0.0
        A[0, 0] = 1.0
    return A


def solve_poisson(A, div, nx, ny, nz, top_wall_bc_type="dirichlet"):
    """Direct solve of projection.py:87-140: least squares on A (the right-hand side is not
    zeroed at a pinned row, so the pinned equation reads p_0 = b_0), then the ghosts: periodic
    x, y; bottom Neumann; top Neumann for a no-slip wall, p = 0 for a free-slip one."""
    b = div.to(F64).permute(2, 1, 0).reshape(-1)
    sol = torch.linalg.lstsq(A, b.unsqueeze(1), driver="gelsd").solution.reshape(-1)
    p = torch.zeros(nx + 2, ny + 2, nz + 2, dtype=F64)
    p[1:nx + 1, 1:ny + 1, 1:nz + 1] = sol.reshape(nz, ny, nx).permute(2, 1, 0)
    p[0] = p[nx]
    p[nx + 1] = p[1]
    p[:, 0] = p[:, ny]
    p[:, ny + 1] = p[:, 1]
    p[:, :, 0] = p[:, :, 1]
    p[:, :, nz + 1] = 0.0 if top_wall_bc_type == "neumann" else p[:, :, nz]
    return p


def solve_poisson_fft(div, fft_data, gauge_zero_mode=False):
    """rfft2 -> per-mode Thomas -> irfft2 -> ghost fill; projection_fft.py:108-165.

    ``gauge_zero_mode`` replaces the singular (kx,ky)=(0,0) system by one with its first
    unknown pinned to zero (the reference solves it as is and lands on a round-off-driven
    additive constant).  Either way p is defined only up to a constant.
    """
    nx, ny, nz = fft_data["nx"], fft_data["ny"], fft_data["nz"]
    spec = torch.fft.rfft2(div, dim=(0, 1))
    nkx, nky = spec.shape[0], spec.shape[1]
    a = fft_data["tri_a"].reshape(-1, nz).to(spec.dtype)
    b = fft_data["tri_b"].reshape(-1, nz).to(spec.dtype)
    c = fft_data["tri_c"].reshape(-1, nz).to(spec.dtype)
    rhs = spec.reshape(-1, nz)
    if gauge_zero_mode:
        a, b, c, rhs = a.clone(), b.clone(), c.clone(), rhs.clone()
        b[0, 0], c[0, 0], rhs[0, 0] = 1.0, 0.0, 0.0
        a[0, 1] = 0.0
    sol = thomas_solve(a, b, c, rhs).reshape(nkx, nky, nz)
    interior = torch.fft.irfft2(sol, s=(nx, ny), dim=(0, 1))
    p = fft_data["workspace_p"]
    p.zero_()
    p[1:nx + 1, 1:ny + 1, 1:nz + 1] = interior
    p[0] = p[nx]
    p[nx + 1] = p[1]
    p[:, 0] = p[:, ny]
    p[:, ny + 1] = p[:, 1]
    p[:, :, 0] = p[:, :, 1]
    p[:, :, nz + 1] = p[:, :, nz]
    return p


# ----------------------------------------------------------------------------------------------
# CFL rate (operators.py:1232-1333) and adaptive dt (solver.py:465-496, 807-812)
# ----------------------------------------------------------------------------------------------

def compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c):
    """Largest convective rate over the three staggered stencils; returns a Python float."""
    rx, ry = 1.0 / dx, 1.0 / dy
    rzf, rzc = 1.0 / dz_f, 1.0 / dz_c
    I, J, K = slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1)
    ex = (torch.abs(u[I, J, K]) * rx
          + 0.25 * torch.abs(v[I, J, K] + v[I, 0:ny, K] + v[2:nx + 2, J, K] + v[2:nx + 2, 0:ny, K]) * ry
          + 0.25 * torch.abs(w[I, J, 1:nz + 1] + w[I, J, 0:nz] + w[2:nx + 2, J, 1:nz + 1] + w[2:nx + 2, J, 0:nz])
          * rzf[0:nz].view(1, 1, -1))
    ey = (0.25 * torch.abs(u[I, J, K] + u[I, 2:ny + 2, K] + u[0:nx, 2:ny + 2, K] + u[0:nx, J, K]) * rx
          + torch.abs(v[I, J, K]) * ry
          + 0.25 * torch.abs(w[I, J, 1:nz + 1] + w[I, 2:ny + 2, 1:nz + 1] + w[I, 2:ny + 2, 0:nz] + w[I, J, 0:nz])
          * rzf[0:nz].view(1, 1, -1))
    ez = (0.25 * torch.abs(u[I, J, 1:nz] + u[0:nx, J, 1:nz] + u[0:nx, J, 2:nz + 1] + u[I, J, 2:nz + 1]) * rx
          + 0.25 * torch.abs(v[I, J, 1:nz] + v[I, 0:ny, 1:nz] + v[I, 0:ny, 2:nz + 1] + v[I, J, 2:nz + 1]) * ry
          + torch.abs(w[I, J, 1:nz]) * rzc[1:nz].view(1, 1, -1))
    return float(torch.max(torch.stack([ex.max(), ey.max(), ez.max()])))


def cfl_dt(rate, cfl_target, dt_min, dt_max):
    """solver.py:483-496."""
    if rate < 1e-10:
        rate = 1.0
    return min(max(cfl_target * (1.0 / rate), dt_min), dt_max)


# ----------------------------------------------------------------------------------------------
# Turbulence statistics (turbstats.py:20-303)
# ----------------------------------------------------------------------------------------------

class TurbulenceStats:
    """Running sums of U(z), u'u', v'v', w'w', u'w' and wall-averaged 2-D spectra."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, z_c, dx, dy, nu, Re_tau_target, z_plus_target=15.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lz, self.z_c, self.nu = Lz, z_c, nu
        self.u_tau_target = Re_tau_target * nu / (Lz / 2.0)               # turbstats.py:52-53
        zt = z_plus_target * nu / self.u_tau_target                       # turbstats.py:58
        zin = z_c[1:nz + 1]
        self.k_bot = int(torch.argmin(torch.abs(zin - zt))) + 1            # turbstats.py:64-65
        self.k_top = int(torch.argmin(torch.abs(zin - (Lz - zt)))) + 1     # turbstats.py:68-69
        self.n_samples = 0
        self.sums = {k: torch.zeros(nz, dtype=F64) for k in ("U", "uu", "vv", "ww", "uw")}
        self.spec = {k: torch.zeros(nx // 2, ny // 2, dtype=F64) for k in ("uu", "vv", "ww", "uw")}
        self.kx = 2 * np.pi * np.fft.rfftfreq(nx, d=dx)[1:]
        self.ky = 2 * np.pi * np.fft.rfftfreq(ny, d=dy)[1:]

    def _fold(self, E):
        """E[kx, ky] + E[-kx, ky] for kx = 1..nx/2 (Nyquist counted twice), ky = 1..ny/2;
        turbstats.py:201-214."""
        hx, hy = self.nx // 2, self.ny // 2
        return E[1:hx + 1, 1:hy + 1] + torch.flip(E[self.nx - hx:self.nx, 1:hy + 1], dims=[0])

    def accumulate_statistics(self, u, v, w, u_tau_current=None):
        nx, ny, nz = self.nx, self.ny, self.nz
        uc = 0.5 * (u[0:nx, 1:ny + 1, 1:nz + 1] + u[1:nx + 1, 1:ny + 1, 1:nz + 1])   # turbstats.py:114-120
        vc = 0.5 * (v[1:nx + 1, 0:ny, 1:nz + 1] + v[1:nx + 1, 1:ny + 1, 1:nz + 1])   # :130
        wc = 0.5 * (w[1:nx + 1, 1:ny + 1, 0:nz] + w[1:nx + 1, 1:ny + 1, 1:nz + 1])   # :135
        U = torch.mean(uc, dim=(0, 1))
        uf = uc - U.view(1, 1, -1)
        vf = vc - torch.mean(vc, dim=(0, 1)).view(1, 1, -1)
        wf = wc - torch.mean(wc, dim=(0, 1)).view(1, 1, -1)
        self.sums["U"] += U
        self.sums["uu"] += torch.mean(uf * uf, dim=(0, 1))
        self.sums["vv"] += torch.mean(vf * vf, dim=(0, 1))
        self.sums["ww"] += torch.mean(wf * wf, dim=(0, 1))
        self.sums["uw"] += torch.mean(uf * wf, dim=(0, 1))
        norm = (nx * ny) ** 2                                                         # :176-185
        kb, kt = self.k_bot - 1, self.k_top - 1
        fb = [torch.fft.rfft2(q[:, :, kb]) for q in (uf, vf, wf)]
        ft = [torch.fft.rfft2(uf[:, :, kt]), torch.fft.rfft2(vf[:, :, kt]), torch.fft.rfft2(-wf[:, :, kt])]
        for key, fn in (("uu", lambda f: torch.abs(f[0]) ** 2 / norm),
                        ("vv", lambda f: torch.abs(f[1]) ** 2 / norm),
                        ("ww", lambda f: torch.abs(f[2]) ** 2 / norm),
                        ("uw", lambda f: (f[0] * torch.conj(f[2])).real / norm)):
            E = 0.5 * (fn(fb) + fn(ft))                                               # :188-191
            self.spec[key] += self._fold(E)[: nx // 2, : ny // 2]                     # :216-225
        self.n_samples += 1

    def finalize_statistics(self):
        """turbstats.py:237-303 (numbers only; no printing)."""
        if self.n_samples == 0:
            raise RuntimeError("No samples accumulated - cannot finalize statistics")
        n = self.n_samples
        out = {"n_samples": n, "z_c": self.z_c[1:self.nz + 1].numpy()}
        for k in ("U", "uu", "vv", "ww", "uw"):
            out[f"{k}_mean"] = (self.sums[k] / n).numpy()
        for k in ("uu", "vv", "ww", "uw"):
            out[f"E_{k}_2d"] = (self.spec[k] / n).numpy()
        U = out["U_mean"]
        dudz = float(0.5 * (U[0] + U[-1]) / float(self.z_c[1]))
        out["u_tau"] = float(np.sqrt(self.nu * dudz))
        out["kx"], out["ky"] = self.kx[: self.nx // 2], self.ky[: self.ny // 2]
        return out


# ----------------------------------------------------------------------------------------------
# Composition: ChannelFlow.__init__ (solver.py:215-383) and step_imex (solver.py:580-710)
# ----------------------------------------------------------------------------------------------

class Channel:
    """Stateful FP64 channel following ChannelFlow's IMEX step.

    ``branch='gpu'`` uses the fused RHS (what the reference runs on any 'cuda' device,
    solver.py:593-600); ``branch='cpu'`` uses the separate kernels (solver.py:601-622).
    """

    def __init__(self, nx, ny, nz, Lx, Ly, Lz=2.0, gamma=2.6, Re=2870.0, U_bulk=1.0,
                 perturbation=0.1, seed=0, top_wall_bc_type="dirichlet", branch="gpu",
                 gauge_zero_mode=False, grid=None, restart=None):
        """``restart`` = (u, v, w, step, time) starts from saved fields as initflow.py:106-167 and
        solver.py:246-249 do: wall-normal ghosts re-applied (no-slip reflections, zero wall w,
        whatever the top wall), no rescale (solver.py:271), step and time kept, forcing and AB2
        history from zero; the initial projection runs either way."""
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.nu = 1.0 / Re
        self.U_bulk = U_bulk
        self.top = top_wall_bc_type
        self.branch = branch
        self.gauge = gauge_zero_mode
        # grid: (z_f, z_c, dz_f, dz_c) of another generator (the hybrid grid), else the tanh grid
        self.z_f, self.z_c, self.dz_f, self.dz_c = grid if grid is not None else generate_grid(gamma, nz, Lz)
        self.dx, self.dy = Lx / nx, Ly / ny
        self.cell_vol = (self.dx * self.dy * self.dz_f.view(1, 1, -1)).expand(nx, ny, nz)
        self.total_volume = Lx * Ly * Lz
        step0, time0 = 0, 0.0
        if restart is not None:
            u, v, w, step0, time0 = restart
            self.u, self.v, self.w = (torch.as_tensor(q, dtype=F64).clone() for q in (u, v, w))
            self.p = torch.zeros(nx + 2, ny + 2, nz + 2, dtype=F64)
            self.u[:, :, 0] = -self.u[:, :, 1]                            # initflow.py:156-165
            self.u[:, :, -1] = -self.u[:, :, -2]
            self.v[:, :, 0] = -self.v[:, :, 1]
            self.v[:, :, -1] = -self.v[:, :, -2]
            self.w[:, :, 0] = 0.0
            self.w[:, :, -1] = 0.0
        else:
            self.u, self.v, self.w, self.p = initialize_flow(nx, ny, nz, self.z_c, Lz, U_bulk,
                                                             perturbation, top_wall_bc_type, seed)
            ub = compute_bulk_velocity(self.u, self.cell_vol, self.total_volume)  # solver.py:271-276
            if abs(ub) > 1e-9:
                self.u *= (U_bulk / ub)
        # the state fields_init.npz holds (solver.py:337-340): before the initial projection
        self.init_state = (self.u.clone(), self.v.clone(), self.w.clone(), self.p.clone())
        self.forcing = 0.0
        self.fft_data = initialize_fft_solver(nx, ny, nz, self.dx, self.dy, self.dz_c, self.dz_f,
                                              top_wall_bc_type)
        div = compute_divergence(self.u, self.v, self.w, nx, ny, nz, self.dx, self.dy, self.dz_f)
        self.p = solve_poisson_fft(div, self.fft_data, self.gauge)             # solver.py:345-357
        project_velocity(self.u, self.v, self.w, self.p, nx, ny, nz, self.dx, self.dy,
                         self.dz_c, self.dz_f, 1.0)                             # solver.py:364-366
        apply_bc_all(self.u, self.v, self.w, self.top)
        self.rhs_prev = None
        self.rhs_curr = None
        self.time = time0
        self.current_step = step0

    def rhs(self):
        f = rhs_fused_imex if self.branch == "gpu" else rhs_separate
        return f(self.u, self.v, self.w, self.nx, self.ny, self.nz, self.dx, self.dy,
                 self.dz_c, self.dz_f, self.nu)

    def step_imex(self, dt, body_force=None):
        """One IMEX step; solver.py:588-710.  ``body_force`` (shaped like u, optional) is an
        additive term next to the scalar forcing (no counterpart in the reference)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        apply_bc_all(self.u, self.v, self.w, self.top)                          # :588
        ru, rv, rw = self.rhs()                                                 # :593-622
        ru += self.forcing                                                      # :626
        if body_force is not None:
            ru += body_force
        if self.rhs_curr is None:                                               # :629-636
            self.rhs_curr = [ru, rv, rw]
        else:
            for dst, src in zip(self.rhs_curr, (ru, rv, rw)):
                dst[:] = src
        if self.rhs_prev is None:                                               # :639-646
            for q, r in zip((self.u, self.v, self.w), self.rhs_curr):
                q += dt * r
        else:
            for q, r, rp in zip((self.u, self.v, self.w), self.rhs_curr, self.rhs_prev):
                q += dt * (1.5 * r - 0.5 * rp)
        self.rhs_prev, self.rhs_curr = self.rhs_curr, self.rhs_prev             # :649-651
        apply_bc_all(self.u, self.v, self.w, self.top)                          # :654
        self.u = solve_implicit_diffusion_uv(self.u, dt, nx, ny, nz, self.dz_c, self.dz_f, self.nu,
                                             top_wall_bc_type=self.top)
        self.v = solve_implicit_diffusion_uv(self.v, dt, nx, ny, nz, self.dz_c, self.dz_f, self.nu,
                                             top_wall_bc_type=self.top)
        self.w = solve_implicit_diffusion_w(self.w, dt, nx, ny, nz, self.dz_c, self.dz_f, self.nu)
        apply_bc_all(self.u, self.v, self.w, self.top)                          # :668
        div = compute_divergence(self.u, self.v, self.w, nx, ny, nz, self.dx, self.dy, self.dz_f)
        self.p = solve_poisson_fft(div / dt, self.fft_data, self.gauge)         # :679-682
        project_velocity(self.u, self.v, self.w, self.p, nx, ny, nz, self.dx, self.dy,
                         self.dz_c, self.dz_f, dt)                              # :692-694
        apply_bc_all(self.u, self.v, self.w, self.top)                          # :697
        ub = compute_bulk_velocity(self.u, self.cell_vol, self.total_volume)    # :701
        self.forcing = self.forcing + (self.U_bulk - ub) / dt * 0.1             # :707-708
        return ub, self.forcing

    def cfl_dt(self, cfl_target, dt_min=1e-4, dt_max=0.01):
        rate = compute_cfl_fused(self.u, self.v, self.w, self.nx, self.ny, self.nz, self.dx, self.dy,
                                 self.dz_f, self.dz_c)
        return cfl_dt(rate, cfl_target, dt_min, dt_max)

    def run(self, n_steps, dt, dt_update_interval=0, cfl_target=0.25, dt_min=1e-4, dt_max=0.01,
            stats=None, n_stats=0, t_stats=0.0, t_max=math.inf, n_out=0):
        """The stepping part of run_simulation (solver.py:793-859), up to step ``n_steps`` (an
        absolute step count, as the reference's n_steps) while time < t_max: adaptive-dt
        cadence with the 5% rule, then step, then time advance, then the n_out diagnostics row
        (step, time, u_bulk, u_tau, forcing; solver.py:823-842) and statistics every n_stats
        steps once time >= t_stats.  Returns (dt, rows)."""
        rows = []
        while self.current_step < n_steps and self.time < t_max:
            self.current_step += 1
            step = self.current_step
            if dt_update_interval > 0 and step % dt_update_interval == 0:
                new = self.cfl_dt(cfl_target, dt_min, dt_max)
                if abs(new - dt) / dt > 0.05:
                    dt = new
            ub, f = self.step_imex(dt)
            self.time += dt
            if n_out > 0 and step % n_out == 0:
                rows.append((step, self.time, float(ub), float(compute_u_tau(self.u, self.z_c, self.nu, self.top)),
                             float(f)))
            if stats is not None and n_stats > 0 and self.time >= t_stats and step % n_stats == 0:
                stats.accumulate_statistics(self.u, self.v, self.w)
        return dt, rows


PINS = {
    "generate_grid": "closed form tanh + sum(dz_f)=Lz + golden trajectory grid",
    "generate_hybrid_grid": "uniform part = linspace; stretched part: endpoints, sum(dz_f)=Lz, growth, tanh at a node",
    "initialize_flow": "golden fixture from the reference (seeded) + parabola closed form",
    "apply_bc_all": "brute-force loop over the reference's order + golden trajectory",
    "rhs_fused_imex": "golden fixture (reference fused op) + sin-field O(h^2) convergence",
    "rhs_separate": "golden fixture (reference separate ops) + interior equality with fused",
    "rhs_fused_v2": "fused-IMEX terms (pinned) + z second derivative: O(h^2) convergence on analytic profiles",
    "pcr_solve": "scipy.linalg.solve_banded on random diagonally dominant systems",
    "thomas_solve": "scipy.linalg.solve_banded on random diagonally dominant systems",
    "solve_implicit_diffusion_uv": "dense numpy solve of the CN system + golden trajectory",
    "solve_implicit_diffusion_w": "golden fixture (reference, shifted explicit half)",
    "compute_divergence": "brute-force loop + closed form on linear fields",
    "compute_bulk_velocity": "bulk of 1 = 1; forcing hits target (test_forcing_indexing.py:49,78)",
    "compute_u_tau": "golden fixture (reference, both top walls, one-sided grid, distinct ghost/face-0/wall-plane means) "
                     "+ planted mistakes (interior mean, ghost plane, top-wall distance, no abs) fail it + parabola",
    "project_velocity": "divergence <= 1e-12 after projection + golden trajectory",
    "initialize_fft_solver": "dense 7-point operator (projection.py:3-84) restated as a matrix",
    "solve_poisson_fft": "dense lstsq on the 7-point matrix (test_fft_vs_direct.py:97) + numpy.fft",
    "build_poisson_matrix": "brute-force loop over the stencil; constant and linear-in-x/y null behaviour",
    "solve_poisson": "residual of the dense system; FFT solve modulo the mean (unpinned matrix)",
    "compute_cfl_fused": "golden fixture (reference, random signed fields, wall and ghost-row spikes, two grids) "
                         "+ swapped dz_f/dz_c fails it + brute-force loop",
    "TurbulenceStats": "single-mode fields with known spectra + golden fixture from the reference",
    "Channel.step_imex": "golden 3-step trajectory from the reference ChannelFlow (cpu branch)",
    "Channel.run": "golden run_simulation from the reference: adaptive dt, statistics, t_max between saves, restart "
                   "(time series, final fields and dt, statistics)",
    "Channel(restart=...)": "golden restart leg from the reference (fields_init, time series, final fields)",
}
