"""Parity of each HIP operator with the CPU oracle (float64), on seeded inputs.

Bar: <= 1e-12 relative (max|a-b| / max|b|) for every field of the plain FP64 path; p is compared
modulo its mean (the reference's p is defined up to a constant).
"""
import math
import os

import numpy as np
import pytest
import torch

from oracle import channel_oracle as O

pytestmark = pytest.mark.gpu

F64 = torch.float64
TOL = 1e-12


def rel(a, b):
    a = a.detach().cpu().double()
    b = b.detach().cpu().double()
    den = float(b.abs().max())
    return float((a - b).abs().max()) / (den if den > 0 else 1.0)


def fields(nx, ny, nz, seed, scale=0.5):
    g = torch.Generator().manual_seed(seed)
    u = (torch.rand(nx + 1, ny + 2, nz + 2, generator=g, dtype=F64) - 0.5) * 2 * scale + 1.0
    v = (torch.rand(nx + 2, ny + 1, nz + 2, generator=g, dtype=F64) - 0.5) * 2 * scale
    w = (torch.rand(nx + 2, ny + 2, nz + 1, generator=g, dtype=F64) - 0.5) * 2 * scale
    return u, v, w


def cuda(*ts):
    return [t.cuda() for t in ts]


@pytest.fixture(scope="module")
def tc():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import torchannel_amd
    torchannel_amd._lib.load()
    return torchannel_amd


# nz in 65-128 runs the two-rows-per-lane instances of the column kernels
SHAPES = [(12, 10, 14), (16, 12, 16), (64, 48, 40), (32, 24, 96), (24, 18, 128), (24, 18, 130)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("branch", ["gpu", "cpu"])
def test_rhs(tc, shape, branch):
    nx, ny, nz = shape
    print(f"[rhs] shape={shape} branch={branch}", flush=True)
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
    u, v, w = fields(nx, ny, nz, 1)
    dx, dy, nu = 4 * math.pi / nx, 4 * math.pi / 3 / ny, 1 / 2870.0
    fo = O.rhs_fused_imex if branch == "gpu" else O.rhs_separate
    ref = fo(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu)
    fg = tc.compute_momentum_rhs_fused_imex if branch == "gpu" else tc.compute_momentum_rhs_separate
    got = fg(*cuda(u, v, w), nx, ny, nz, dx, dy, dz_c.cuda(), dz_f.cuda(), nu)
    for a, b in zip(got, ref):
        assert rel(a, b) < TOL


def test_rhs_matches_reference_fixture(tc, golden_dir):
    g = np.load(os.path.join(golden_dir, "ops.npz"))
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, float(g["Lz"]))
    u, v, w = (torch.from_numpy(g[k]).cuda() for k in ("u", "v", "w"))
    got = tc.compute_momentum_rhs_fused_imex(u, v, w, nx, ny, nz, float(g["Lx"]) / nx, float(g["Ly"]) / ny,
                                             dz_c.cuda(), dz_f.cuda(), float(g["nu"]))
    for a, key in zip(got, ("rhs_u", "rhs_v", "rhs_w")):
        assert rel(a, torch.from_numpy(g[key])) < TOL, key
    w_cn = tc.solve_implicit_diffusion_w(w, float(g["cn_dt"]), nx, ny, nz, dz_c.cuda(), dz_f.cuda(), float(g["nu"]))
    assert rel(w_cn, torch.from_numpy(g["w_cn"])) < TOL


def test_fused_v2_rhs(tc, golden_dir):
    """compute_momentum_rhs_fused_v2 against the reference's own output and, on the larger
    shapes, against the oracle."""
    g = np.load(os.path.join(golden_dir, "ops_v2.npz"))
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, float(g["Lz"]))
    u, v, w = (torch.from_numpy(g[k]).cuda() for k in ("u", "v", "w"))
    got = tc.compute_momentum_rhs_fused_v2(u, v, w, nx, ny, nz, float(g["Lx"]) / nx, float(g["Ly"]) / ny,
                                           dz_c.cuda(), dz_f.cuda(), float(g["nu"]))
    for a, key in zip(got, ("rhs_u", "rhs_v", "rhs_w")):
        assert rel(a, torch.from_numpy(g[key])) < TOL, key
    for shape in SHAPES:
        nx, ny, nz = shape
        print(f"[rhs_v2] shape={shape}", flush=True)
        _, _, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
        u, v, w = fields(nx, ny, nz, 3)
        want = O.rhs_fused_v2(u, v, w, nx, ny, nz, 0.1, 0.12, dz_c, dz_f, 1 / 2870.0)
        got = tc.compute_momentum_rhs_fused_v2(u.cuda(), v.cuda(), w.cuda(), nx, ny, nz, 0.1, 0.12, dz_c.cuda(),
                                               dz_f.cuda(), 1 / 2870.0)
        for a, b_, key in zip(got, want, ("u", "v", "w")):
            assert rel(a, b_) < TOL, (shape, key)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_cn(tc, shape, top):
    nx, ny, nz = shape
    print(f"[cn] shape={shape} top={top}", flush=True)
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
    u, v, w = fields(nx, ny, nz, 2)
    nu, dt = 1 / 2870.0, 0.01
    ru = O.solve_implicit_diffusion_uv(u, dt, nx, ny, nz, dz_c, dz_f, nu, 0.5, top)
    rv = O.solve_implicit_diffusion_uv(v, dt, nx, ny, nz, dz_c, dz_f, nu, 0.5, top)
    rw = O.solve_implicit_diffusion_w(w, dt, nx, ny, nz, dz_c, dz_f, nu)
    gu, gv, gw = cuda(u, v, w)
    dtt = torch.tensor(dt, dtype=F64, device="cuda")
    assert rel(tc.solve_implicit_diffusion_u(gu, dtt, nx, ny, nz, dz_c.cuda(), dz_f.cuda(), nu,
                                             top_wall_bc_type=top), ru) < TOL
    assert rel(tc.solve_implicit_diffusion_v(gv, dt, nx, ny, nz, dz_c.cuda(), dz_f.cuda(), nu,
                                             top_wall_bc_type=top), rv) < TOL
    assert rel(tc.solve_implicit_diffusion_w(gw, dt, nx, ny, nz, dz_c.cuda(), dz_f.cuda(), nu), rw) < TOL


@pytest.mark.parametrize("n", [1, 2, 7, 64, 65, 128, 200, 256, 300, 512])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shared", [False, True])
def test_pcr(tc, n, cplx, shared):
    print(f"[pcr] n={n} complex={cplx} shared={shared}", flush=True)
    rng = np.random.default_rng(n)
    batch = 37
    shp = (n,) if shared else (batch, n)
    a = torch.from_numpy(rng.uniform(-1, 1, shp))
    c = torch.from_numpy(rng.uniform(-1, 1, shp))
    b = torch.from_numpy(3.0 + rng.uniform(0, 1, shp))
    a[..., 0] = 0.0
    c[..., -1] = 0.0
    d = torch.from_numpy(rng.standard_normal((batch, n)))
    if cplx:
        d = torch.complex(d, torch.from_numpy(rng.standard_normal((batch, n))))
    ref = O.pcr_solve(a, b, c, d)
    got = tc.pcr_solve(*cuda(a, b, c, d))
    assert rel(torch.view_as_real(got) if cplx else got, torch.view_as_real(ref) if cplx else ref) < TOL


@pytest.mark.parametrize("n", [1, 2, 48, 128, 257])
@pytest.mark.parametrize("cplx", [False, True])
def test_thomas(tc, n, cplx):
    """solve_tridiagonal (projection_fft.py:167-215): the device Thomas sweep vs the oracle's."""
    print(f"[thomas] n={n} complex={cplx}", flush=True)
    rng = np.random.default_rng(100 + n)
    batch = 29
    a = torch.from_numpy(rng.uniform(-1, 1, (batch, n)))
    c = torch.from_numpy(rng.uniform(-1, 1, (batch, n)))
    b = torch.from_numpy(3.0 + rng.uniform(0, 1, (batch, n)))
    d = torch.from_numpy(rng.standard_normal((batch, n)))
    if cplx:
        d = torch.complex(d, torch.from_numpy(rng.standard_normal((batch, n))))
        ref = O.thomas_solve(a.to(d.dtype), b.to(d.dtype), c.to(d.dtype), d)
    else:
        ref = O.thomas_solve(a, b, c, d)
    got = tc.solve_tridiagonal(*cuda(a, b, c, d))
    assert rel(torch.view_as_real(got) if cplx else got, torch.view_as_real(ref) if cplx else ref) < TOL


@pytest.mark.parametrize("where", ["interior", "corner_lo", "corner_hi"])
def test_cfl_maximum_anywhere(tc, where):
    """The CFL rate is a maximum over every cell and direction: a spike placed in the interior or
    at either corner of the grid (first / last i, j, k) is found, at a size with many blocks."""
    nx, ny, nz = 66, 300, 20
    print(f"[cfl] where={where}", flush=True)
    Lx, Ly = 4 * math.pi, 4 * math.pi / 3
    dx, dy = Lx / nx, Ly / ny
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
    u, v, w = fields(nx, ny, nz, 3)
    i, j, k = {"interior": (31, 157, 9), "corner_lo": (1, 1, 1), "corner_hi": (nx, ny, nz)}[where]
    u[i, j, k] = 40.0
    v[i, j - 1, k] = -35.0
    w[i, j, k - 1] = 30.0
    ref = O.compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c)
    got = tc.compute_cfl_fused(u.cuda(), v.cuda(), w.cuda(), nx, ny, nz, dx, dy, dz_f.cuda(), dz_c.cuda())
    assert abs(got - ref) < 1e-13 * ref


@pytest.mark.parametrize("shape", SHAPES)
def test_bc_div_project_bulk_cfl_utau(tc, shape):
    nx, ny, nz = shape
    print(f"[small ops] shape={shape}", flush=True)
    z_f, z_c, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
    u, v, w = fields(nx, ny, nz, 3)
    dx, dy, nu = 0.3, 0.2, 1 / 2870.0
    gu, gv, gw = cuda(u, v, w)
    for top in ("dirichlet", "neumann"):
        a = [q.clone() for q in (u, v, w)]
        ga = [q.clone() for q in (gu, gv, gw)]
        O.apply_bc_all(*a, top)
        tc.apply_bc_all(*ga, top)
        for x, y in zip(ga, a):
            assert torch.equal(x.cpu(), y)
    O.apply_bc_all(u, v, w)
    tc.apply_bc_all(gu, gv, gw)
    assert rel(tc.compute_divergence(gu, gv, gw, nx, ny, nz, dx, dy, dz_f.cuda()),
               O.compute_divergence(u, v, w, nx, ny, nz, dx, dy, dz_f)) < TOL
    p = torch.from_numpy(np.random.default_rng(4).standard_normal((nx + 2, ny + 2, nz + 2)))
    pu, pv, pw = u.clone(), v.clone(), w.clone()
    O.project_velocity(pu, pv, pw, p, nx, ny, nz, dx, dy, dz_c, dz_f, 0.005)
    qu, qv, qw = (q.clone() for q in (gu, gv, gw))
    tc.project_velocity(qu, qv, qw, p.cuda(), nx, ny, nz, dx, dy, dz_c.cuda(), dz_f.cuda(),
                        torch.tensor(0.005, dtype=F64, device="cuda"))
    for x, y in ((qu, pu), (qv, pv), (qw, pw)):
        assert rel(x, y) < TOL
    vol = (dx * dy * dz_f.view(1, 1, -1)).expand(nx, ny, nz)
    ub = tc.compute_bulk_velocity(gu, vol.cuda(), nx * dx * ny * dy * 2.0)
    assert abs(float(ub) - float(O.compute_bulk_velocity(u, vol, nx * dx * ny * dy * 2.0))) < 1e-14
    cfl = tc.compute_cfl_fused(gu, gv, gw, nx, ny, nz, dx, dy, dz_f.cuda(), dz_c.cuda())
    ref = O.compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c)
    assert abs(cfl - ref) < 1e-13 * ref
    for top in ("dirichlet", "neumann"):
        ut = tc.compute_u_tau(gu, z_c.cuda(), nu, top)
        assert abs(float(ut) - float(O.compute_u_tau(u, z_c, nu, top))) < 1e-14


POISSON = [(16, 12, 16), (8, 6, 10), (64, 48, 40), (48, 96, 32), (192, 64, 20), (24, 18, 130), (10, 8, 129),
           (16, 12, 200), (8, 6, 256), (320, 6, 18), (16, 12, 128), (12, 10, 64), (8, 6, 32),
           (4, 4, 4), (1536, 4, 6), (4, 1536, 6)]


@pytest.mark.parametrize("shape", POISSON)
def test_poisson(tc, shape):
    nx, ny, nz = shape
    print(f"[poisson] shape={shape}", flush=True)
    Lx, Ly = 4 * math.pi, 4 * math.pi / 3
    dx, dy = Lx / nx, Ly / ny
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
    rng = np.random.default_rng(7)
    div = rng.standard_normal((nx, ny, nz))
    div -= (div * dz_f.numpy()).sum() / (dz_f.numpy().sum() * nx * ny)
    div = torch.from_numpy(div)
    fd = O.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f)
    ref = O.solve_poisson_fft(div, fd, gauge_zero_mode=True).clone()
    gfd = tc.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c.cuda(), dz_f.cuda())
    p = tc.solve_poisson_fft(div.cuda(), gfd)
    assert p.data_ptr() == gfd["workspace_p"].data_ptr()       # alias, projection_fft.py:142-165
    pc = p.cpu()
    assert rel(pc - pc.mean(), ref - ref.mean()) < 1e-12
    # ghosts are the periodic images / Neumann walls of the interior
    assert torch.equal(pc[0], pc[nx]) and torch.equal(pc[nx + 1], pc[1])
    assert torch.equal(pc[:, 0], pc[:, ny]) and torch.equal(pc[:, ny + 1], pc[:, 1])
    assert torch.equal(pc[:, :, 0], pc[:, :, 1]) and torch.equal(pc[:, :, nz + 1], pc[:, :, nz])
    # the discrete Laplacian of p reproduces the right-hand side
    lap = ((pc[2:, 1:-1, 1:-1] - 2 * pc[1:-1, 1:-1, 1:-1] + pc[:-2, 1:-1, 1:-1]) / dx ** 2
           + (pc[1:-1, 2:, 1:-1] - 2 * pc[1:-1, 1:-1, 1:-1] + pc[1:-1, :-2, 1:-1]) / dy ** 2
           + ((pc[1:-1, 1:-1, 2:] - pc[1:-1, 1:-1, 1:-1]) / dz_c[1:].view(1, 1, -1)
              - (pc[1:-1, 1:-1, 1:-1] - pc[1:-1, 1:-1, :-2]) / dz_c[:-1].view(1, 1, -1)) / dz_f.view(1, 1, -1))
    assert float((lap - div).abs().max()) < 1e-7 * float(div.abs().max())


@pytest.mark.parametrize("shape", [(16, 12, 40), (10, 8, 129), (16, 12, 200), (8, 6, 256), (320, 6, 18),
                                   (16, 12, 128), (12, 10, 64), (48, 96, 32), (4, 1536, 6), (384, 4, 40)])
def test_poisson_fp32(tc, shape):
    """The float32 solve (float64 gauge mode) against the float64 oracle: float32 round-off
    amplified by the low-wavenumber systems stays well below 1e-4."""
    nx, ny, nz = shape
    print(f"[poisson32] shape={shape}", flush=True)
    Lx, Ly = 4 * math.pi, 4 * math.pi / 3
    dx, dy = Lx / nx, Ly / ny
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
    rng = np.random.default_rng(8)
    div = rng.standard_normal((nx, ny, nz))
    div -= (div * dz_f.numpy()).sum() / (dz_f.numpy().sum() * nx * ny)
    div = torch.from_numpy(div)
    fd = O.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f)
    ref = O.solve_poisson_fft(div.float().double(), fd, gauge_zero_mode=True).clone()
    gfd = tc.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c.cuda(), dz_f.cuda(), dtype=torch.float32)
    p = tc.solve_poisson_fft(div.float().cuda(), gfd)
    assert p.dtype == torch.float32
    pc = p.cpu().double()
    assert rel(pc - pc.mean(), ref - ref.mean()) < 1e-4


@pytest.mark.parametrize("shape", [(16, 12, 16), (64, 48, 40)])
def test_stats(tc, shape):
    nx, ny, nz = shape
    print(f"[stats] shape={shape}", flush=True)
    Lx, Ly, Lz, nu = 4 * math.pi, 4 * math.pi / 3, 2.0, 1 / 2870.0
    z_f, z_c, dz_f, dz_c = O.generate_grid(2.6, nz, Lz)
    st = O.TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c, Lx / nx, Ly / ny, nu, 180.0, 15.0)
    gs = tc.TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c.cuda(), z_f.cuda(), dz_c.cuda(), dz_f.cuda(),
                            Lx / nx, Ly / ny, nu, 180.0, 15.0, device="cuda")
    assert (gs.k_bot, gs.k_top) == (st.k_bot, st.k_top)
    for seed in (5, 6):
        u, v, w = fields(nx, ny, nz, seed)
        st.accumulate_statistics(u, v, w)
        gs.accumulate_statistics(*cuda(u, v, w))
    a, b = gs.finalize_statistics(), st.finalize_statistics()
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean", "E_uu_2d", "E_vv_2d", "E_ww_2d", "E_uw_2d"):
        assert rel(torch.from_numpy(np.asarray(a[k])), torch.from_numpy(np.asarray(b[k]))) < TOL, k


def test_stats_matches_reference_fixture(tc, golden_dir):
    g = np.load(os.path.join(golden_dir, "ops.npz"))
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    Lx, Ly, Lz, nu = float(g["Lx"]), float(g["Ly"]), float(g["Lz"]), float(g["nu"])
    z_f, z_c, dz_f, dz_c = O.generate_grid(2.6, nz, Lz)
    if nx % 2 or ny % 2:
        pytest.skip("odd grid")
    gs = tc.TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c.cuda(), z_f.cuda(), dz_c.cuda(), dz_f.cuda(),
                            Lx / nx, Ly / ny, nu, 180.0, 15.0, device="cuda")
    gs.accumulate_statistics(*(torch.from_numpy(g[k]).cuda() for k in ("u", "v", "w")))
    a = gs.finalize_statistics()
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean", "E_uu_2d", "E_vv_2d", "E_ww_2d", "E_uw_2d"):
        assert rel(torch.from_numpy(np.asarray(a[k])), torch.from_numpy(g[k])) < TOL, k


def test_product_rejects_cpu_tensors(tc):
    u, v, w = fields(8, 6, 8, 0)
    with pytest.raises(RuntimeError):
        tc.apply_bc_all(u, v, w)
