"""Initial fields (initflow.py:4-104).

Setup only: the fields are built on the host in float64 with the reference's formulas and noise
draws (torch's CPU generator, order u, v, w), then moved to the target device and precision, so
that a seeded start is the same on every machine.
"""
import torch

from .utils import load_flow_fields


def initialize_flow(nx, ny, nz, z_c, Ly, Lz, U_bulk=1.0, init_type="parabolic", perturbation_intensity=0.0,
                    n_vortices=4, device="cpu", top_wall_bc_type="dirichlet", seed=None, dtype=torch.float64):
    f64 = torch.float64
    zc = z_c.detach().to("cpu", f64)
    u = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=f64)
    v = torch.zeros(nx + 2, ny + 1, nz + 2, dtype=f64)
    w = torch.zeros(nx + 2, ny + 2, nz + 1, dtype=f64)
    p = torch.zeros(nx + 2, ny + 2, nz + 2, dtype=f64)
    if init_type in ("parabolic", "vortices"):                      # initflow.py:11-26
        peak = 1.5 * U_bulk
        if top_wall_bc_type == "neumann":
            zn = zc / Lz
            prof = peak * zn * (2 - zn)
        else:
            zn = 2 * zc / Lz - 1
            prof = peak * (1 - zn ** 2)
        u[:, :, :] = prof[None, None, :]
    if init_type == "uniform":                                      # initflow.py:28-29
        u[:, :, :] = U_bulk
    elif init_type == "vortices":                                   # initflow.py:31-81
        A = perturbation_intensity * U_bulk
        ky = n_vortices * torch.pi / Ly
        y_v = torch.linspace(0, Ly, ny + 1, dtype=f64)
        y_c = 0.5 * (y_v[:-1] + y_v[1:])
        zin = zc[1:nz + 1]
        yy, zz = torch.meshgrid(y_v, zin, indexing="ij")
        v_pert = -A * torch.sin(ky * yy) * 2 * torch.sin(torch.pi * zz / Lz) * torch.cos(torch.pi * zz / Lz) * (torch.pi / Lz)
        v[1:nx + 1, :, 1:nz + 1] += v_pert[None, :, :]
        zf_loc = torch.cat([torch.tensor([0.0], dtype=f64), 0.5 * (zin[:-1] + zin[1:]), torch.tensor([Lz], dtype=f64)])
        yw, zw = torch.meshgrid(y_c, zf_loc, indexing="ij")
        w_pert = A * torch.cos(ky * yw) * ky * torch.sin(torch.pi * zw / Lz) ** 2
        w[1:nx + 1, 1:ny + 1, :] += w_pert[None, :, :]
    if perturbation_intensity > 0.0 and init_type != "vortices":    # initflow.py:84-92
        if seed is not None:
            torch.manual_seed(seed)
        amp = perturbation_intensity * U_bulk
        for q in (u, v, w):
            box = q[1:-1, 1:-1, 2:-2]
            box += (torch.rand_like(box) - 0.5) * 2 * amp
    u[:, :, 0] = -u[:, :, 1]                                        # initflow.py:95-102
    u[:, :, -1] = -u[:, :, -2]
    v[:, :, 0] = -v[:, :, 1]
    v[:, :, -1] = -v[:, :, -2]
    w[:, :, 0] = 0.0
    w[:, :, -1] = 0.0
    return tuple(t.to(device=device, dtype=dtype) for t in (u, v, w, p))


def initialize_flow_from_file(field_file, device="cpu", reset_time=False, dtype=None):
    """initflow.py:106-167: fields from a saved npz, step/time kept or reset, wall ghosts
    re-applied (no-slip reflections for u, v; zero w at the walls)."""
    f = load_flow_fields(field_file, device="cpu")
    u, v, w, p = (f[k] for k in ("u", "v", "w", "p"))
    u[:, :, 0] = -u[:, :, 1]
    u[:, :, -1] = -u[:, :, -2]
    v[:, :, 0] = -v[:, :, 1]
    v[:, :, -1] = -v[:, :, -2]
    w[:, :, 0] = 0.0
    w[:, :, -1] = 0.0
    step, time = (0, 0.0) if reset_time else (int(f["step"]), float(f["time"]))
    dt = dtype or u.dtype
    u, v, w, p = (t.to(device=device, dtype=dt).contiguous() for t in (u, v, w, p))
    return u, v, w, p, step, time
