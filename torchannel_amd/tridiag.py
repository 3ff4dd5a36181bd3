"""Batched tridiagonal solve by parallel cyclic reduction (tridiag.py:33-58) on the HIP library.

One 64-lane wavefront solves one system; rows live in registers and neighbours are exchanged
with lane shuffles.  a, b, c are real, shape (n,) (shared) or (batch, n); d is (batch, n), real
or complex.  n <= 512.
"""
import torch

from . import _lib as L


def _prep(a, b, c, d):
    n = d.shape[-1]
    cplx = d.is_complex()
    rd = d.real.dtype if cplx else d.dtype
    a, b, c = (t.to(rd).contiguous() for t in (a, b, c))
    dd = d.contiguous()
    x = torch.empty_like(dd)
    dv = torch.view_as_real(dd) if cplx else dd
    xv = torch.view_as_real(x) if cplx else x
    L.require_device(a, b, c, dv, xv)
    batch = dd.numel() // n if n > 0 else 0
    if a.dim() == 1:
        stride = 0
    elif a.numel() == batch * n and b.numel() == batch * n and c.numel() == batch * n:
        stride = n
    else:
        raise ValueError("tridiagonal solve: a, b, c must be (n,) or match d's batch")
    return n, batch, cplx, rd, a, b, c, dv, x, xv, stride


def thomas_solve(a, b, c, d):
    """Batched Thomas sweep (projection_fft.py:167-215) on the device, one thread per system."""
    n, batch, cplx, rd, a, b, c, dv, x, xv, stride = _prep(a, b, c, d)
    work = torch.empty(max(batch * n, 1), dtype=rd, device=x.device)
    L.call("tc_thomas_batched", n, batch, L.dtype_code(rd), int(cplx), L.ptr(a), L.ptr(b), L.ptr(c),
           stride, L.ptr(dv), L.ptr(xv), L.ptr(work), L.stream_of(x))
    return x


def pcr_solve(a, b, c, d):
    """Solve a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i (a_0, c_{n-1} ignored); x shaped like d."""
    n, batch, cplx, rd, a, b, c, dv, x, xv, stride = _prep(a, b, c, d)
    L.call("tc_pcr_batched", n, batch, L.dtype_code(rd), int(cplx), L.ptr(a), L.ptr(b), L.ptr(c),
           stride, L.ptr(dv), L.ptr(xv), L.stream_of(x))
    return x
