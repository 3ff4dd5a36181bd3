"""The padded array layout (columns a whole number of 128-byte lines apart, row k = 1 of every
column on a line) gives bit-identical results to the dense layout the reference allocates: whole
solver steps (forward-Euler start, AB2 steps, graph replay, body force, both rhs branches, both
top walls, both precisions), the stand-alone operators on padded views against the same
operators on dense copies, and the adjoint gradient."""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def cfg(nx, ny, nz, precision, branch, top, layout):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": 0.005, "n_steps": 10},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": 5},
        "compute": {"device": "cuda", "precision": precision, "rhs_branch": branch, "layout": layout},
        "output": {"save_init": False},
        "boundary_conditions": {"top_wall": {"type": top}},
    }


@pytest.fixture(scope="module")
def tc():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import torchannel_amd
    torchannel_amd._lib.load()
    return torchannel_amd


SHAPES = [(4, 4, 4), (16, 12, 16), (18, 10, 48), (32, 24, 96), (24, 18, 128), (24, 18, 130), (16, 12, 256)]


def _state(ch):
    return [ch.u, ch.v, ch.w, ch.p, ch._scal] + [r for s in ch._rhs for r in s]


def _equal(a, b, what):
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y), what


def test_padded_arrays_are_line_aligned(tc):
    ch = tc.ChannelFlow(config=cfg(24, 18, 130, "fp64", "gpu", "dirichlet", "padded"))
    for q in [ch.u, ch.v, ch.w, ch.p] + [r for s in ch._rhs for r in s]:
        esz = q.element_size()
        assert q.stride(1) == ch.pitch and (ch.pitch * esz) % 128 == 0
        assert (q.data_ptr() + esz) % 128 == 0          # row k = 1 of column (0, 0)
    assert ch.pitch == 144 and not ch.u.is_contiguous()
    d = tc.ChannelFlow(config=cfg(24, 18, 130, "fp64", "gpu", "dirichlet", "dense"))
    assert d.pitch == 0 and d.u.is_contiguous()
    assert torch.equal(ch.u, d.u) and torch.equal(ch.p, d.p)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("branch,top", [("gpu", "dirichlet"), ("cpu", "neumann")])
def test_steps_padded_equal_dense(tc, shape, precision, branch, top):
    print(f"[layout] steps {shape} {precision} {branch} {top}", flush=True)
    a = tc.ChannelFlow(config=cfg(*shape, precision, branch, top, "dense"))
    b = tc.ChannelFlow(config=cfg(*shape, precision, branch, top, "padded"))
    _equal(a, b, "initial state")
    for n in range(4):
        a.step_imex(0.004)
        b.step_imex(0.004)
        _equal(a, b, f"step {n}")
    torch.cuda.synchronize()


@pytest.mark.parametrize("shape", [(16, 12, 16), (32, 24, 96)])
def test_graph_and_body_force_padded_equal_dense(tc, shape):
    print(f"[layout] graph + body force {shape}", flush=True)
    a = tc.ChannelFlow(config=cfg(*shape, "fp64", "gpu", "dirichlet", "dense"))
    b = tc.ChannelFlow(config=cfg(*shape, "fp64", "gpu", "dirichlet", "padded"))
    gen = torch.Generator().manual_seed(1)
    f = 1e-3 * torch.randn(tuple(a.u.shape), generator=gen, dtype=torch.float64)
    a.body_force = f.cuda()
    b.body_force = f.cuda()
    assert not b.body_force.is_contiguous() and torch.equal(a.body_force, b.body_force)
    for _ in range(3):
        a.step_imex(0.005)
        b.step_imex(0.005)
    _equal(a, b, "body-force steps")
    a.body_force = None
    b.body_force = None
    a.step_imex()
    b.step_imex()
    a.run_graph(2, 4, dt_update_interval=2)
    b.run_graph(2, 4, dt_update_interval=2)
    torch.cuda.synchronize()
    _equal(a, b, "graph chunks")


def _dense(t):
    return t.contiguous()


@pytest.mark.parametrize("shape", [(16, 12, 16), (24, 18, 130)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_operators_on_padded_views(tc, shape, precision):
    """Every stand-alone operator on the padded arrays equals the same operator on dense copies."""
    from torchannel_amd import operators as O, projection as PR, utils as U
    from torchannel_amd.solver import apply_bc_all
    print(f"[layout] operators {shape} {precision}", flush=True)
    nx, ny, nz = shape
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz, precision, "gpu", "dirichlet", "padded"))
    for _ in range(2):
        ch.step_imex(0.005)
    u, v, w, p = ch.u, ch.v, ch.w, ch.p
    ud, vd, wd, pd = (_dense(q) for q in (u, v, w, p))
    args = (nx, ny, nz, ch.dx, ch.dy, ch.dz_c, ch.dz_f, ch.nu)
    for fn in (O.compute_momentum_rhs_fused_imex, O.compute_momentum_rhs_separate,
               O.compute_momentum_rhs_fused_v2):
        rp = fn(u, v, w, *args)
        rd = fn(ud, vd, wd, *args)
        for x, y in zip(rp, rd):
            assert x.stride() == u.stride() or x.stride() == v.stride() or x.stride() == w.stride()
            assert torch.equal(x, y), fn.__name__
    for fn, q, qd in ((O.solve_implicit_diffusion_u, u, ud), (O.solve_implicit_diffusion_v, v, vd)):
        assert torch.equal(fn(q, 0.005, nx, ny, nz, ch.dz_c, ch.dz_f, ch.nu),
                           fn(qd, 0.005, nx, ny, nz, ch.dz_c, ch.dz_f, ch.nu)), fn.__name__
    assert torch.equal(O.solve_implicit_diffusion_w(w, 0.005, nx, ny, nz, ch.dz_c, ch.dz_f, ch.nu),
                       O.solve_implicit_diffusion_w(wd, 0.005, nx, ny, nz, ch.dz_c, ch.dz_f, ch.nu))
    assert O.compute_cfl_fused(u, v, w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f, ch.dz_c) == \
        O.compute_cfl_fused(ud, vd, wd, nx, ny, nz, ch.dx, ch.dy, ch.dz_f, ch.dz_c)
    assert torch.equal(U.compute_divergence(u, v, w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f),
                       U.compute_divergence(ud, vd, wd, nx, ny, nz, ch.dx, ch.dy, ch.dz_f))
    assert torch.equal(U.compute_bulk_velocity(u, ch.cell_vol_ratio, ch.total_volume),
                       U.compute_bulk_velocity(ud, ch.cell_vol_ratio, ch.total_volume))
    assert torch.equal(U.compute_u_tau(u, ch.z_c, ch.nu), U.compute_u_tau(ud, ch.z_c, ch.nu))
    # in-place operators: on clones in each layout
    from torchannel_amd import _lib as L
    up, vp, wp = (L.clone_field(q) for q in (u, v, w))
    uq, vq, wq = (q.clone() for q in (ud, vd, wd))
    PR.project_velocity(up, vp, wp, p, nx, ny, nz, ch.dx, ch.dy, ch.dz_c, ch.dz_f, 0.005)
    PR.project_velocity(uq, vq, wq, pd, nx, ny, nz, ch.dx, ch.dy, ch.dz_c, ch.dz_f, 0.005)
    apply_bc_all(up, vp, wp, "neumann")
    apply_bc_all(uq, vq, wq, "neumann")
    for x, y in ((up, uq), (vp, vq), (wp, wq)):
        assert torch.equal(x, y), "project + bc"
    # the Poisson solve through the solver's fft_data (padded workspace) and a dense one
    from torchannel_amd.projection_fft import initialize_fft_solver, solve_poisson_fft
    div = U.compute_divergence(u, v, w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f) / 0.005
    fd = initialize_fft_solver(nx, ny, nz, ch.dx, ch.dy, ch.dz_c, ch.dz_f, dtype=ch.wdtype)
    p1 = solve_poisson_fft(div, ch.fft_data)
    assert p1 is ch.p and not p1.is_contiguous()
    p2 = solve_poisson_fft(div, fd)
    assert torch.equal(p1, p2), "poisson"
    # statistics on padded and dense fields
    from torchannel_amd.turbstats import TurbulenceStats
    sts = []
    for fields in ((u, v, w), (ud, vd, wd)):
        s = TurbulenceStats(nx, ny, nz, ch.Lx, ch.Ly, ch.Lz, ch.z_c, ch.z_f, ch.dz_c, ch.dz_f, ch.dx, ch.dy,
                            ch.nu, 180.0, device=ch.device)
        s.accumulate_statistics(*fields)
        sts.append(s)
    assert torch.equal(sts[0]._prof, sts[1]._prof) and torch.equal(sts[0]._spec, sts[1]._spec)


def test_adjoint_gradient_padded_equal_dense(tc):
    print("[layout] adjoint 16x12x16", flush=True)
    out = []
    for layout in ("dense", "padded"):
        ch = tc.ChannelFlow(config=cfg(16, 12, 16, "fp64", "gpu", "dirichlet", layout))
        run = tc.WallShearAdjoint(ch, 6, 0.005, checkpoint_every=3)
        gen = torch.Generator().manual_seed(2)
        f = (1e-3 * torch.randn(tuple(ch.u.shape), generator=gen, dtype=torch.float64)).cuda()
        J = run.objective(f)
        g = run.gradient(f)
        assert g.is_contiguous() and g.shape == ch.u.shape
        out.append((J.clone(), g))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ensemble_of_padded_channels(tc, precision):
    """Channels in the padded layout stepped as an ensemble (per-channel strides in the channel
    descriptors) equal the same channels in the dense layout, bit for bit."""
    import copy
    shapes = [(16, 12, 16), (24, 18, 48), (20, 16, 130)]
    print(f"[layout] ensemble {precision} {shapes}", flush=True)
    out = []
    for layout in ("dense", "padded"):
        cfgs = []
        for q, sh in enumerate(shapes):
            c = cfg(*sh, precision, "gpu", "dirichlet", layout)
            c["initialization"]["seed"] = q + 1
            c["compute"]["private_plan"] = True
            cfgs.append(c)
        ens = tc.make_ensemble([copy.deepcopy(c) for c in cfgs])
        ens.step(0.005)
        ens.step()
        ens.run_graph(1, 4)
        torch.cuda.synchronize()
        out.append(ens.channels)
    for a, b in zip(*out):
        assert b.pitch > 0 and a.pitch == 0
        _equal(a, b, "ensemble")
