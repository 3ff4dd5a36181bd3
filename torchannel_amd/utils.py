"""Grid generation and the scalar reductions of utils.py, on the HIP library.

generate_grid is one-off setup and uses the reference's own tensor formulas on the target
device; the per-step functions (divergence, bulk velocity, wall shear) run HIP kernels.
"""
import os

import numpy as np
import torch

from . import _lib as L


def generate_grid(gamma, nz, Lz, device="cpu", stretching_type="symmetric", dtype=torch.float64):
    """tanh-stretched wall-normal grid (utils.py:5-45): z_f[nz+1], z_c[nz+2], dz_f[nz], dz_c[nz+1]."""
    k = torch.linspace(0, nz, nz + 1, device=device, dtype=dtype)
    g = torch.tensor(gamma, device=device, dtype=dtype)
    if stretching_type == "bottom":
        xi = k / nz
        z_f = Lz * (1.0 - torch.tanh(gamma * (1.0 - xi)) / torch.tanh(g))
    else:
        xi = (2 * k / nz) - 1
        z_f = 0.5 * Lz * (1 + torch.tanh(gamma * xi) / torch.tanh(g))
    mid = 0.5 * (z_f[:-1] + z_f[1:])
    z_c = torch.cat([torch.tensor([-mid[0]], device=device, dtype=dtype), mid,
                     torch.tensor([2 * z_f[-1] - mid[-1]], device=device, dtype=dtype)])
    return z_f, z_c, z_f[1:] - z_f[:-1], z_c[1:] - z_c[:-1]


def generate_hybrid_grid(nz_uniform, nz_stretched, z_transition, Lz, gamma, device="cpu", dtype=torch.float64,
                         verbose=False):
    """Uniform cells on [0, z_transition] under one-sided tanh cells on [z_transition, Lz]
    (utils.py:47-113): z_f[nz+1], z_c[nz+2], dz_f[nz], dz_c[nz+1] with nz = nz_uniform +
    nz_stretched.  The upper region is fine at the transition and coarse at the top.  Raises
    ValueError for a non-monotone grid, as the reference does; with verbose, prints the relative
    jump of dz at the transition (the reference's C1 check, 1% threshold)."""
    z_lo = torch.linspace(0, z_transition, nz_uniform + 1, device=device, dtype=dtype)
    H = Lz - z_transition
    xi = torch.linspace(0, nz_stretched, nz_stretched + 1, device=device, dtype=dtype) / nz_stretched
    g = torch.tensor(gamma, device=device, dtype=dtype)
    z_hi = H * (1.0 - torch.tanh(gamma * (1.0 - xi)) / torch.tanh(g))
    z_f = torch.cat([z_lo, z_transition + z_hi[1:]])
    mid = 0.5 * (z_f[:-1] + z_f[1:])
    z_c = torch.cat([torch.tensor([-mid[0]], device=device, dtype=dtype), mid,
                     torch.tensor([2 * z_f[-1] - mid[-1]], device=device, dtype=dtype)])
    dz_f = z_f[1:] - z_f[:-1]
    dz_c = z_c[1:] - z_c[:-1]
    if verbose:
        before, after = float(dz_f[nz_uniform - 1]), float(dz_f[nz_uniform])
        jump = abs(after - before) / before
        state = "OK" if jump <= 0.01 else "consider adjusting gamma"
        print(f"Hybrid grid C1 continuity at z={z_transition}: {jump * 100:.3f}% ({state})", flush=True)
    if bool(torch.any(dz_f <= 0)):
        raise ValueError("Non-positive cell spacing detected in hybrid grid!")
    if bool(torch.any(z_f[1:] <= z_f[:-1])):
        raise ValueError("Non-monotonic grid detected!")
    return z_f, z_c, dz_f, dz_c


def _dims_u(u):
    return u.shape[0] - 1, u.shape[1] - 2, u.shape[2] - 2


def compute_divergence(u, v, w, nx, ny, nz, dx, dy, dz_f):
    """utils.py:226-240 -> new (nx, ny, nz) tensor."""
    dt, pitch = L.require_fields(u, v, w)
    g = L.get_grid(nx, ny, nz, dx, dy, 0.0, L.TC_TOP_DIRICHLET, dt, u.device, dz_f=dz_f, pitch=pitch)
    div = torch.empty((nx, ny, nz), dtype=dt, device=u.device)
    L.call("tc_divergence", g.handle, L.ptr(u), L.ptr(v), L.ptr(w), L.ptr(div), L.stream_of(u))
    return div


def compute_bulk_velocity(u, cell_vol_ratio, total_volume):
    """utils.py:207-223 -> 0-D float64 tensor on the device."""
    dt, pitch = L.require_fields(u)
    nx, ny, nz = cell_vol_ratio.shape
    wgt = cell_vol_ratio
    if wgt.dtype != torch.float64 or wgt.device != u.device:
        wgt = wgt.to(device=u.device, dtype=torch.float64)
    g = L.get_grid(nx, ny, nz, 1.0, 1.0, 0.0, L.TC_TOP_DIRICHLET, dt, u.device, pitch=pitch)
    out = torch.empty((), dtype=torch.float64, device=u.device)
    s = wgt.stride()
    L.call("tc_bulk_velocity", g.handle, L.ptr(u), ctypes_ptr(wgt), s[0], s[1], s[2],
           float(total_volume), L.ptr(out), L.stream_of(u))
    return out if dt == torch.float64 else out.to(dt)


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def compute_u_tau(u, z_c, nu, top_wall_bc_type="dirichlet"):
    """utils.py:176-204 (means over the whole k=1 and k=nz planes) -> 0-D tensor."""
    dt, pitch = L.require_fields(u)
    nx, ny, nz = _dims_u(u)
    g = L.get_grid(nx, ny, nz, 1.0, 1.0, float(nu), L.top_code(top_wall_bc_type), dt, u.device, pitch=pitch)
    out = torch.empty((), dtype=torch.float64, device=u.device)
    L.call("tc_u_tau", g.handle, L.ptr(u), float(z_c[1]), L.ptr(out), L.stream_of(u))
    return out if dt == torch.float64 else out.to(dt)


def save_flow_fields(u, v, w, p, z_c, z_f, Lx, Ly, step, time, u_tau, forcing, results_folder,
                     filename="fields.npz"):
    """utils.py:369-400: same npz keys (u, v, w, p, z_c, z_f, Lx, Ly, step, time, u_tau, forcing)."""
    def host(x):
        return x.detach().cpu().numpy() if torch.is_tensor(x) else x

    def scalar(x):
        return x.item() if torch.is_tensor(x) else x

    os.makedirs(results_folder, exist_ok=True)
    np.savez(os.path.join(results_folder, filename), u=host(u), v=host(v), w=host(w), p=host(p),
             z_c=host(z_c), z_f=host(z_f), Lx=Lx, Ly=Ly, step=step, time=time, u_tau=scalar(u_tau),
             forcing=scalar(forcing))


def load_flow_fields(filepath, device="cpu"):
    """utils.py:402-444 (npz loaded without pickles)."""
    if not os.path.exists(filepath):
        raise FileNotFoundError(f"Field file not found: {filepath}")
    data = np.load(filepath, allow_pickle=False)
    out = {k: torch.tensor(data[k], device=device) for k in ("u", "v", "w", "p", "z_c", "z_f")}
    out.update(Lx=float(data["Lx"]), Ly=float(data["Ly"]), step=int(data["step"]), time=float(data["time"]),
               u_tau=float(data["u_tau"]), forcing=float(data["forcing"]))
    return out


def append_timeseries(path, step, time, u_bulk, u_tau, forcing):
    """The timeseries.npz append of solver.py:889-929 (keys step, time, u_bulk, u_tau, forcing)."""
    chunk = {"step": np.asarray(step, dtype=np.int32), "time": np.asarray(time, dtype=np.float64),
             "u_bulk": np.asarray(u_bulk, dtype=np.float64), "u_tau": np.asarray(u_tau, dtype=np.float64),
             "forcing": np.asarray(forcing, dtype=np.float64)}
    if os.path.exists(path):
        old = np.load(path, allow_pickle=False)
        chunk = {k: np.concatenate([old[k], chunk[k]]) for k in chunk}
    np.savez_compressed(path, **chunk)
