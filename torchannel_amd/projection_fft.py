"""FFT + tridiagonal pressure Poisson solver (projection_fft.py) on the HIP library.

The solve is hand-written: LDS-resident batched FFTs along x (real, two z-rows per complex
transform) and y, one wall-normal Thomas solve per (kx, ky) system (its two sweeps evaluated as
parallel scans of affine maps across 16 or 32 lanes, factors precomputed per plan), inverse FFTs,
and the ghost images of p written by the last kernel.  The singular (0,0) mode is gauge-fixed in
float64.
"""
import math

import torch

from . import _lib as L
from .tridiag import thomas_solve


def initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f, top_wall_bc_type="dirichlet", dtype=None, _grid=None):
    """projection_fft.py:3-99: same dictionary keys (kx, ky, tri_a, tri_b, tri_c, nx, ny, nz,
    workspace_p, top_wall_bc_type), plus the private '_grid' that holds the HIP plan.
    ``dtype`` (default dz_c.dtype) selects the working precision of the solve; ``_grid`` (a
    solver's grid of the same sizes) makes the plan that grid's and workspace_p an array in its
    layout."""
    device = dz_c.device
    wdt = dtype or dz_c.dtype
    kx = torch.fft.fftfreq(nx, d=dx / (2 * math.pi), device=device, dtype=dz_c.dtype)
    ky = torch.fft.rfftfreq(ny, d=dy / (2 * math.pi), device=device, dtype=dz_c.dtype)
    kxm = (2.0 / dx) * torch.sin(kx * dx / 2.0)
    kym = (2.0 / dy) * torch.sin(ky * dy / 2.0)
    gx, gy = torch.meshgrid(kxm, kym, indexing="ij")
    k2 = gx ** 2 + gy ** 2
    cl = 1.0 / (dz_c[:-1] * dz_f)
    cr = 1.0 / (dz_c[1:] * dz_f)
    nkx, nky = len(kx), len(ky)
    ta = torch.zeros(nkx, nky, nz, device=device, dtype=dz_c.dtype)
    ta[:, :, 1:] = cl[1:].view(1, 1, -1)
    tc = torch.zeros(nkx, nky, nz, device=device, dtype=dz_c.dtype)
    tc[:, :, :-1] = cr[:-1].view(1, 1, -1)
    tb = -((cl + cr).view(1, 1, -1) + k2.unsqueeze(-1).expand(nkx, nky, nz))
    tb[:, :, 0] += cl[0]
    tb[:, :, -1] += cr[-1]
    grid = _grid
    workspace_p = L.empty_field((nx + 2, ny + 2, nz + 2), grid.pitch if grid is not None else 0, wdt, device)
    if device.type == "cuda":
        if grid is None:
            grid = L.get_grid(nx, ny, nz, dx, dy, 0.0, L.top_code(top_wall_bc_type), wdt, device,
                              dz_f=dz_f, dz_c=dz_c)
        grid.poisson()
    return {"kx": kx, "ky": ky, "tri_a": ta, "tri_b": tb, "tri_c": tc, "nx": nx, "ny": ny, "nz": nz,
            "workspace_p": workspace_p, "top_wall_bc_type": top_wall_bc_type, "_grid": grid}


def solve_poisson_fft(div, fft_data):
    """projection_fft.py:101-165: returns fft_data['workspace_p'] (an alias, as the reference)
    holding p with every ghost entry set (periodic x/y images, Neumann z walls)."""
    p = fft_data["workspace_p"]
    L.require_device(div)
    _, pitch = L.require_fields(p)
    g = fft_data.get("_grid")
    if g is None:
        raise RuntimeError("torchannel_amd: fft_data was initialised off the HIP device")
    if pitch != g.pitch and not (pitch == 0 and p.shape[2] == g.pitch):
        raise RuntimeError("torchannel_amd: fft_data's workspace_p does not have its plan's layout")
    if div.dtype != p.dtype:
        div = div.to(p.dtype)
    div = div.contiguous()
    L.call("tc_poisson_solve", g.poisson(), L.ptr(div), L.ptr(p), L.stream_of(p))
    return p


def solve_tridiagonal(a, b, c, d):
    """projection_fft.py:167-215: batched Thomas sweep with per-row diagonals, on the device."""
    return thomas_solve(a, b, c, d)
