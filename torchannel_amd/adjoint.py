"""Gradient of the time-averaged wall shear stress with respect to a body-force field.

No counterpart in the reference (torChannel has no adjoint).  A ChannelFlow (float64, GPU-branch
explicit terms) is stepped N times at fixed dt from its current state with a body force f (shaped
like u) added to rhs_u next to the scalar forcing F (solver.py:626).  The objective is

    J(f) = 1/N * sum_{n=1..N} tau_w(u_n),
    tau_w(u) = nu/z_c[1] * (mean u[:,:,1] + mean u[:,:,nz]) / 2      (no-slip top)
             = nu/z_c[1] * mean u[:,:,1]                            (free-slip top)

(means over whole planes, ghosts included: the linear core of compute_u_tau, utils.py:187-204).
dJ/df comes from the hand-written discrete adjoint of the step (tc_adjoint_*, HIP), run backward
over the trajectory with the forward recomputed from checkpoints taken every `checkpoint_every`
steps.  `wall_shear_objective(runner, f)` wraps both in a torch.autograd.Function.
"""
import ctypes

import torch

from . import _lib as L


class WallShearAdjoint:
    """Objective and gradient runner bound to one ChannelFlow and its state at construction."""

    def __init__(self, channel, n_steps, dt, checkpoint_every=10):
        ch = channel
        if ch.wdtype != torch.float64:
            raise ValueError("the adjoint runs the float64 path (compute.precision: fp64)")
        if getattr(ch, "rhs_branch", "gpu") != "gpu":
            raise ValueError("the adjoint differentiates the GPU-branch explicit terms (compute.rhs_branch: gpu)")
        if n_steps < 1 or checkpoint_every < 1:
            raise ValueError("n_steps and checkpoint_every must be positive")
        self.ch = ch
        self.n_steps = int(n_steps)
        self.dt = float(dt)
        self.k = int(checkpoint_every)
        h = ctypes.c_void_p()
        L.call("tc_adjoint_create", ch._solver, float(ch.z_c[1]), ctypes.byref(h))
        self._h = h
        self._init = self._snapshot()
        self._ckpts = None
        self._ckpt_f = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                L.load().tc_adjoint_destroy(self._h)
        except Exception:
            pass

    # -- state plumbing ---------------------------------------------------------------------
    def _snapshot(self):
        ch = self.ch
        nw = ch._newest
        return {"q": [q.clone() for q in (ch.u, ch.v, ch.w)], "newest": nw,
                "rhs": [r.clone() for r in ch._rhs[nw]] if nw >= 0 else None, "scal": ch._scal.clone()}

    def _restore(self, snap):
        ch = self.ch
        for dst, src in zip((ch.u, ch.v, ch.w), snap["q"]):
            dst.copy_(src)
        nw = snap["newest"]
        if nw >= 0:
            for dst, src in zip(ch._rhs[nw], snap["rhs"]):
                dst.copy_(src)
        ch._scal.copy_(snap["scal"])
        ch._newest = nw
        L.call("tc_solver_set_newest", ch._solver, nw)

    def _step(self, f):
        ch = self.ch
        newest = ctypes.c_int(-1)
        first = 1 if ch._newest < 0 else 0
        L.call("tc_solver_step", ch._solver, self.dt, first, L.ptr(f), ctypes.byref(newest), ch._stream())
        ch._newest = newest.value

    def _check_f(self, f):
        """f in the solver's layout (a copy when the solver's arrays are padded)."""
        ch = self.ch
        if f.shape != ch.u.shape or f.dtype != torch.float64 or f.device != ch.u.device:
            raise ValueError(f"body force must be a float64 tensor of shape {tuple(ch.u.shape)} "
                             f"on {ch.u.device}")
        if L.field_pitch(ch.u) == 0 and f.is_contiguous():
            return f
        key = (f.data_ptr(), f._version)
        if getattr(self, "_f_key", None) != key:
            self._f_key, self._f_lay = key, L.clone_field_as(f, ch.u)
        return self._f_lay

    # -- objective and gradient --------------------------------------------------------------
    def objective(self, f):
        """J(f) as a 0-d float64 device tensor; keeps the checkpoints for gradient(f)."""
        fl = self._check_f(f)
        ch = self.ch
        self._restore(self._init)
        J = torch.zeros(1, dtype=torch.float64, device=ch.u.device)
        ckpts = []
        w = 1.0 / self.n_steps
        for n in range(self.n_steps):
            if n % self.k == 0:
                ckpts.append(self._snapshot())
            self._step(fl)
            L.call("tc_adjoint_objective", self._h, L.ptr(ch.u), w, L.ptr(J), ch._stream())
        self._ckpts = ckpts
        self._ckpt_f = (f.data_ptr(), f._version)
        return J[0]

    def gradient(self, f):
        """dJ/df (shaped like u, dense; zero on ghost entries) by the reverse sweep."""
        fl = self._check_f(f)
        if self._ckpts is None or self._ckpt_f != (f.data_ptr(), f._version):
            self.objective(f)
        ch = self.ch
        st = ch._stream()
        grad = L.field_like(ch.u, zero=True)
        w = 1.0 / self.n_steps
        fresh = self._init["newest"] < 0
        L.call("tc_adjoint_begin", self._h, self.dt, w, L.ptr(grad), st)
        for seg in range(len(self._ckpts) - 1, -1, -1):
            n0 = seg * self.k
            n1 = min(n0 + self.k, self.n_steps)
            self._restore(self._ckpts[seg])
            states = []
            for _ in range(n0, n1):
                states.append([L.clone_field(q) for q in (ch.u, ch.v, ch.w)])
                self._step(fl)
            for n in range(n1 - 1, n0 - 1, -1):
                q = states[n - n0]
                first = 1 if (n == 0 and fresh) else 0
                L.call("tc_adjoint_step", self._h, L.ptr(q[0]), L.ptr(q[1]), L.ptr(q[2]), first,
                       w if n >= 1 else 0.0, L.ptr(grad), st)
        return grad.contiguous()


class _WallShear(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, runner):
        ctx.runner = runner
        ctx.save_for_backward(f)
        return runner.objective(f.detach())

    @staticmethod
    def backward(ctx, gJ):
        (f,) = ctx.saved_tensors
        return gJ * ctx.runner.gradient(f.detach()), None


def wall_shear_objective(runner, f):
    """Differentiable J(f): `wall_shear_objective(runner, f).backward()` fills f.grad."""
    return _WallShear.apply(f, runner)
