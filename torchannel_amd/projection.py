"""Velocity correction q -= dt*grad(p) (projection.py:143-165) on the HIP library."""
import torch

from . import _lib as L


def project_velocity(u, v, w, p, nx, ny, nz, dx, dy, dz_c, dz_f, dt):
    """In place on the interior; returns (u, v, w) like the reference.  dt may be a float or a
    0-D float64 device tensor (solver.py:659)."""
    dtp, pitch = L.require_fields(u, v, w, p)
    g = L.get_grid(nx, ny, nz, dx, dy, 0.0, L.TC_TOP_DIRICHLET, dtp, u.device, dz_f=dz_f, dz_c=dz_c,
                   pitch=pitch)
    if isinstance(dt, torch.Tensor):
        dtd = dt.to(device=u.device, dtype=torch.float64).reshape(())
        L.call("tc_project", g.handle, L.ptr(u), L.ptr(v), L.ptr(w), L.ptr(p), 0.0, L.ptr(dtd),
               L.stream_of(u))
    else:
        L.call("tc_project", g.handle, L.ptr(u), L.ptr(v), L.ptr(w), L.ptr(p), float(dt), None,
               L.stream_of(u))
    return u, v, w
