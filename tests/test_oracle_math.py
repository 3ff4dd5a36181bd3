"""Pins of the CPU oracle by mathematics: closed forms, dense solves, brute-force loops."""
import math

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import channel_oracle as O


def test_grid_closed_form_and_volume():
    nz, Lz, g = 48, 2.0, 2.6
    z_f, z_c, dz_f, dz_c = O.generate_grid(g, nz, Lz)
    k = np.arange(nz + 1)
    ref = 0.5 * Lz * (1 + np.tanh(g * (2 * k / nz - 1)) / np.tanh(g))
    assert np.abs(z_f.numpy() - ref).max() < 1e-15
    assert abs(float(dz_f.sum()) - Lz) < 1e-12                    # test_volume_integration.py:47
    assert abs(float(z_c[0]) + float(z_c[1])) < 1e-15               # mirrored ghost centre
    assert len(dz_c) == nz + 1 and len(z_c) == nz + 2


@pytest.mark.parametrize("nzu,nzs,zt,g", [(75, 75, 0.25, 1.8), (10, 14, 0.4, 1.5), (3, 40, 0.05, 2.6)])
def test_hybrid_grid_structure(nzu, nzs, zt, g):
    """utils.py:47-113: uniform cells up to z_transition; above it, faces whose inverse tanh map
    recovers equally spaced xi (an independent inversion of the stretching law)."""
    Lz = 2.0
    z_f, z_c, dz_f, dz_c = O.generate_hybrid_grid(nzu, nzs, zt, Lz, g)
    assert len(z_f) == nzu + nzs + 1 and len(z_c) == nzu + nzs + 2 and len(dz_c) == nzu + nzs + 1
    assert np.abs(z_f[:nzu + 1].numpy() - np.linspace(0, zt, nzu + 1)).max() < 1e-15
    assert float(z_f[nzu]) == zt and abs(float(z_f[-1]) - Lz) < 1e-15
    assert abs(float(dz_f.sum()) - Lz) < 1e-13
    up = dz_f[nzu:].numpy()
    assert np.all(np.diff(up) > 0)                      # fine at the transition, coarse at the top
    H = Lz - zt
    zeta = (z_f[nzu:].numpy() - zt) / H
    xi = 1.0 - np.arctanh((1.0 - zeta) * np.tanh(g)) / g
    assert np.abs(xi - np.arange(nzs + 1) / nzs).max() < 1e-12
    assert abs(float(z_c[0]) + float(z_c[1])) < 1e-15
    assert abs(float(z_c[-1]) - (2 * Lz - float(z_c[-2]))) < 1e-14


def test_hybrid_grid_rejects_folded_grid():
    # a transition above the top wall makes the upper region run downwards
    with pytest.raises(ValueError):
        O.generate_hybrid_grid(8, 8, 2.5, 2.0, 1.5)


def test_package_hybrid_grid_matches_oracle():
    import torchannel_amd as tc
    for args in ((75, 75, 0.25, 2.0, 1.8), (10, 14, 0.4, 2.0, 1.5)):
        for a, b in zip(tc.generate_hybrid_grid(*args), O.generate_hybrid_grid(*args)):
            assert torch.equal(a, b)


def _v2_error(nz, gamma, comp):
    """max error of the oracle's FE-scheme RHS on a single-mode z profile against its analytic
    value: u = cos(pi z) (v = w = 0) gives nu u''; w = sin(pi z / 2)^2 on faces (u = v = 0)
    gives nu w'' - (w^2)'."""
    nx, ny, Lz, nu = 4, 4, 2.0, 0.1
    z_f, z_c, dz_f, dz_c = O.generate_grid(gamma, nz, Lz)
    u = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=torch.float64)
    v = torch.zeros(nx + 2, ny + 1, nz + 2, dtype=torch.float64)
    w = torch.zeros(nx + 2, ny + 2, nz + 1, dtype=torch.float64)
    if comp == "u":
        u[:] = torch.cos(math.pi * z_c).view(1, 1, -1)
        ru, _, _ = O.rhs_fused_v2(u, v, w, nx, ny, nz, 1.0, 1.0, dz_c, dz_f, nu)
        exact = -nu * math.pi ** 2 * torch.cos(math.pi * z_c[1:nz + 1])
        return float((ru[2, 2, 1:nz + 1] - exact).abs().max())
    f = lambda z: torch.sin(0.5 * math.pi * z) ** 2
    w[:] = f(z_f).view(1, 1, -1)
    _, _, rw = O.rhs_fused_v2(u, v, w, nx, ny, nz, 1.0, 1.0, dz_c, dz_f, nu)
    z = z_f[1:nz]
    w2 = 0.5 * math.pi ** 2 * torch.cos(math.pi * z)                 # d2/dz2 sin^2(pi z/2)
    ww = 2 * f(z) * 0.5 * math.pi * torch.sin(math.pi * z)          # d/dz (w^2) = 2 w w'
    return float((rw[2, 2, 1:nz] - (nu * w2 - ww)).abs().max())


@pytest.mark.parametrize("comp,gamma,order", [("u", 1e-6, 2), ("w", 1e-6, 2), ("u", 2.6, 1), ("w", 2.6, 1)])
def test_fused_v2_converges_on_smooth_profiles(comp, gamma, order):
    """operators.py:904-909 / 1021-1026: the explicit z terms converge at second order on a
    uniform grid and at first order on the tanh-stretched one (three-point differences of
    centre values over uneven cells)."""
    e = [_v2_error(nz, gamma, comp) for nz in (32, 64, 128)]
    lo = 3.5 if order == 2 else 1.8
    assert e[0] / e[1] > lo and e[1] / e[2] > lo, e


def _banded(a, b, c, d):
    n = len(b)
    ab = np.zeros((3, n))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    return scipy.linalg.solve_banded((1, 1), ab, d)


@pytest.mark.parametrize("n", [1, 2, 5, 64, 127, 128, 200])
def test_pcr_and_thomas_match_banded_solver(n):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 3.0 + rng.uniform(0, 1, n)
    a[0] = 0.0
    c[-1] = 0.0
    d = rng.standard_normal((4, n))
    ref = np.stack([_banded(a, b, c, row) for row in d])
    t = lambda x: torch.from_numpy(x)
    x = O.pcr_solve(t(a), t(b), t(c), t(d)).numpy()
    assert np.abs(x - ref).max() < 1e-13
    x2 = O.thomas_solve(t(np.tile(a, (4, 1))), t(np.tile(b, (4, 1))), t(np.tile(c, (4, 1))), t(d)).numpy()
    assert np.abs(x2 - ref).max() < 1e-13


def test_cn_uv_solves_the_theta_scheme():
    """The CN output x satisfies (I - a D2) x = q + b D2 q with wall closures (dense check)."""
    nx, ny, nz, nu, dt = 3, 2, 24, 1e-2, 0.05
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, 2.0)
    rng = np.random.default_rng(0)
    for top, sgn in (("dirichlet", -1.0), ("neumann", 1.0)):
        q = torch.from_numpy(rng.standard_normal((nx + 1, ny + 2, nz + 2)))
        x = O.solve_implicit_diffusion_uv(q, dt, nx, ny, nz, dz_c, dz_f, nu, 0.5, top)
        D = np.zeros((nz, nz))
        dzc, dzf = dz_c.numpy(), dz_f.numpy()
        for m in range(nz):                     # second difference with ghost elimination
            l, r = 1.0 / (dzc[m] * dzf[m]), 1.0 / (dzc[m + 1] * dzf[m])
            D[m, m] -= l + r
            if m > 0:
                D[m, m - 1] += l
            else:
                D[m, m] -= l                    # ghost = -q[1]
            if m < nz - 1:
                D[m, m + 1] += r
            else:
                D[m, m] += sgn * r              # ghost = -q[nz] (no-slip) / +q[nz] (free-slip)
        A = np.eye(nz) - 0.5 * dt * nu * D
        B = np.eye(nz) + 0.5 * dt * nu * D
        for i in range(1, nx + 1):
            for j in range(1, ny + 1):
                col = q[i, j, 1:nz + 1].numpy()
                assert np.abs(A @ x[i, j, 1:nz + 1].numpy() - B @ col).max() < 1e-12


def test_divergence_of_linear_field_and_loop():
    nx, ny, nz = 5, 4, 6
    _, _, dz_f, _ = O.generate_grid(2.0, nz, 2.0)
    dx, dy = 0.3, 0.7
    rng = np.random.default_rng(3)
    u = torch.from_numpy(rng.standard_normal((nx + 1, ny + 2, nz + 2)))
    v = torch.from_numpy(rng.standard_normal((nx + 2, ny + 1, nz + 2)))
    w = torch.from_numpy(rng.standard_normal((nx + 2, ny + 2, nz + 1)))
    div = O.compute_divergence(u, v, w, nx, ny, nz, dx, dy, dz_f).numpy()
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            for k in range(1, nz + 1):
                ref = ((u[i, j, k] - u[i - 1, j, k]) / dx + (v[i, j, k] - v[i, j - 1, k]) / dy
                       + (w[i, j, k] - w[i, j, k - 1]) / dz_f[k - 1])
                assert abs(div[i - 1, j - 1, k - 1] - float(ref)) < 1e-12
    # linear u = 2x -> du/dx = 2 exactly
    ul = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=torch.float64)
    ul[:] = 2.0 * dx * torch.arange(nx + 1, dtype=torch.float64).view(-1, 1, 1)
    d2 = O.compute_divergence(ul, torch.zeros_like(v), torch.zeros_like(w), nx, ny, nz, dx, dy, dz_f)
    assert torch.allclose(d2, torch.full_like(d2, 2.0), atol=1e-13)


def test_bulk_of_one_and_forcing_target():
    """test_forcing_indexing.py:40-78."""
    nx, ny, nz, Lx, Ly, Lz = 16, 12, 20, 2.67, 0.8, 2.0
    _, _, dz_f, _ = O.generate_grid(2.5, nz, Lz)
    vol = (Lx / nx * Ly / ny * dz_f.view(1, 1, -1)).expand(nx, ny, nz)
    u = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=torch.float64)
    u[1:nx + 1, 1:ny + 1, 1:nz + 1] = 1.0
    u[0] = 999.0
    assert abs(float(O.compute_bulk_velocity(u, vol, Lx * Ly * Lz)) - 1.0) < 1e-12
    u2 = torch.from_numpy(np.random.default_rng(1).standard_normal((nx + 1, ny + 2, nz + 2)))
    ub = O.compute_bulk_velocity(u2, vol, Lx * Ly * Lz)
    F = (1.5 - ub) / 0.01
    u2[1:nx + 1, 1:ny + 1, 1:nz + 1] += 0.01 * F
    assert abs(float(O.compute_bulk_velocity(u2, vol, Lx * Ly * Lz)) - 1.5) < 1e-12


def test_u_tau_of_parabola():
    nx, ny, nz, Lz, nu = 4, 4, 64, 2.0, 1 / 2870.0
    _, z_c, _, _ = O.generate_grid(2.6, nz, Lz)
    u, _, _, _ = O.initialize_flow(nx, ny, nz, z_c, Lz, 1.0)
    z1 = float(z_c[1])
    expect = math.sqrt(nu * 1.5 * (1 - (2 * z1 / Lz - 1) ** 2) / z1)
    assert abs(float(O.compute_u_tau(u, z_c, nu)) - expect) < 1e-14


def _poisson_matrix(nx, ny, nz, dx, dy, dz_c, dz_f):
    """Dense 7-point operator, periodic x/y and Neumann z (projection.py:18-77), i fastest."""
    N = nx * ny * nz
    A = np.zeros((N, N))
    idx = lambda i, j, k: (i % nx) + (j % ny) * nx + k * nx * ny
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                for di, dj, cf in ((1, 0, 1 / dx ** 2), (-1, 0, 1 / dx ** 2), (0, 1, 1 / dy ** 2), (0, -1, 1 / dy ** 2)):
                    A[r, idx(i + di, j + dj, k)] += cf
                    A[r, r] -= cf
                if k > 0:
                    cf = 1.0 / (dz_f[k] * dz_c[k])
                    A[r, idx(i, j, k - 1)] += cf
                    A[r, r] -= cf
                if k < nz - 1:
                    cf = 1.0 / (dz_f[k] * dz_c[k + 1])
                    A[r, idx(i, j, k + 1)] += cf
                    A[r, r] -= cf
    return A


def test_dense_poisson_matrix_matches_stencil_loop():
    """projection.py:3-84 restated by vectorised indexing: against a brute-force stencil loop,
    and A applied to a constant vanishes (periodic x/y, Neumann z) before the pin."""
    nx, ny, nz = 5, 4, 6
    _, _, dz_f, dz_c = O.generate_grid(2.0, nz, 2.0)
    A = O.build_poisson_matrix(nx, ny, nz, 0.3, 0.4, dz_c, dz_f, pin_pressure=False).numpy()
    assert np.array_equal(A, _poisson_matrix(nx, ny, nz, 0.3, 0.4, dz_c.numpy(), dz_f.numpy()))
    assert np.abs(A @ np.ones(nx * ny * nz)).max() < 1e-10
    Ap = O.build_poisson_matrix(nx, ny, nz, 0.3, 0.4, dz_c, dz_f).numpy()
    assert Ap[0, 0] == 1.0 and np.count_nonzero(Ap[0]) == 1 and np.array_equal(Ap[1:], A[1:])


@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_direct_poisson_solve(top):
    """projection.py:87-140: the pinned system is solved exactly (row 0 reads p_0 = b_0), then
    the ghost fill; without the pin the solve matches the FFT solve modulo the mean."""
    nx, ny, nz = 6, 4, 8
    _, _, dz_f, dz_c = O.generate_grid(1.5, nz, 2.0)
    rng = np.random.default_rng(9)
    div = rng.standard_normal((nx, ny, nz))
    div -= (div * dz_f.numpy()).sum() / (dz_f.numpy().sum() * nx * ny)
    A = O.build_poisson_matrix(nx, ny, nz, 0.2, 0.3, dz_c, dz_f)
    p = O.solve_poisson(A, torch.from_numpy(div), nx, ny, nz, top).numpy()
    x = p[1:nx + 1, 1:ny + 1, 1:nz + 1].transpose(2, 1, 0).reshape(-1)
    b = div.transpose(2, 1, 0).reshape(-1)
    assert np.abs(A.numpy() @ x - b).max() < 1e-9 * np.abs(b).max()
    assert np.array_equal(p[0], p[nx]) and np.array_equal(p[:, ny + 1], p[:, 1])
    assert np.array_equal(p[:, :, 0], p[:, :, 1])
    assert np.all(p[:, :, nz + 1] == 0.0) if top == "neumann" else np.array_equal(p[:, :, nz + 1], p[:, :, nz])
    A0 = O.build_poisson_matrix(nx, ny, nz, 0.2, 0.3, dz_c, dz_f, pin_pressure=False)
    q = O.solve_poisson(A0, torch.from_numpy(div), nx, ny, nz)[1:nx + 1, 1:ny + 1, 1:nz + 1].numpy()
    fd = O.initialize_fft_solver(nx, ny, nz, 0.2, 0.3, dz_c, dz_f)
    f = O.solve_poisson_fft(torch.from_numpy(div), fd).numpy()[1:nx + 1, 1:ny + 1, 1:nz + 1]
    assert np.abs((q - q.mean()) - (f - f.mean())).max() < 1e-9


@pytest.mark.parametrize("gauge", [False, True])
def test_fft_poisson_matches_dense_operator(gauge):
    """test_fft_vs_direct.py:97: FFT solve equals the dense solve modulo the mean (<=1e-9)."""
    nx, ny, nz = 8, 6, 10
    Lx, Ly, Lz = 0.9, 0.7, 2.0
    dx, dy = Lx / nx, Ly / ny
    _, z_c, dz_f, dz_c = O.generate_grid(1.5, nz, Lz)
    rng = np.random.default_rng(5)
    div = rng.standard_normal((nx, ny, nz))
    div -= (div * dz_f.numpy()).sum() / (dz_f.numpy().sum() * nx * ny)   # compatible RHS
    fd = O.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f)
    p = O.solve_poisson_fft(torch.from_numpy(div), fd, gauge_zero_mode=gauge).numpy()
    A = _poisson_matrix(nx, ny, nz, dx, dy, dz_c.numpy(), dz_f.numpy())
    b = div.transpose(2, 1, 0).reshape(-1)
    sol = np.linalg.lstsq(A, b, rcond=None)[0].reshape(nz, ny, nx).transpose(2, 1, 0)
    pin = p[1:nx + 1, 1:ny + 1, 1:nz + 1]
    assert np.abs((pin - pin.mean()) - (sol - sol.mean())).max() < 1e-9
    # ghosts: periodic copies then Neumann walls (projection_fft.py:150-163)
    assert np.array_equal(p[0], p[nx]) and np.array_equal(p[nx + 1], p[1])
    assert np.array_equal(p[:, 0], p[:, ny]) and np.array_equal(p[:, :, 0], p[:, :, 1])


def test_fft_convention_is_numpy_rfft2():
    x = np.random.default_rng(2).standard_normal((12, 10, 3))
    got = torch.fft.rfft2(torch.from_numpy(x), dim=(0, 1)).numpy()
    assert np.abs(got - np.fft.rfft2(x, axes=(0, 1))).max() < 1e-12


def test_projection_makes_field_divergence_free():
    nx, ny, nz = 12, 10, 16
    ch = O.Channel(nx, ny, nz, 2.0, 1.0, seed=4, perturbation=0.2)
    div = O.compute_divergence(ch.u, ch.v, ch.w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f)
    assert float(div.abs().max()) < 1e-12                         # test_divergence_by_cell.py:96
    ch.step_imex(0.005)
    div = O.compute_divergence(ch.u, ch.v, ch.w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f)
    assert float(div.abs().max()) < 1e-10


def test_cfl_rate_matches_loop():
    nx, ny, nz = 4, 3, 5
    _, _, dz_f, dz_c = O.generate_grid(2.0, nz, 2.0)
    rng = np.random.default_rng(9)
    u = torch.from_numpy(rng.standard_normal((nx + 1, ny + 2, nz + 2)))
    v = torch.from_numpy(rng.standard_normal((nx + 2, ny + 1, nz + 2)))
    w = torch.from_numpy(rng.standard_normal((nx + 2, ny + 2, nz + 1)))
    dx, dy = 0.5, 0.25
    best = 0.0
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            for k in range(1, nz + 1):
                ex = (abs(u[i, j, k]) / dx + 0.25 * abs(v[i, j, k] + v[i, j - 1, k] + v[i + 1, j, k] + v[i + 1, j - 1, k]) / dy
                      + 0.25 * abs(w[i, j, k] + w[i, j, k - 1] + w[i + 1, j, k] + w[i + 1, j, k - 1]) / dz_f[k - 1])
                ey = (0.25 * abs(u[i, j, k] + u[i, j + 1, k] + u[i - 1, j + 1, k] + u[i - 1, j, k]) / dx + abs(v[i, j, k]) / dy
                      + 0.25 * abs(w[i, j, k] + w[i, j + 1, k] + w[i, j + 1, k - 1] + w[i, j, k - 1]) / dz_f[k - 1])
                best = max(best, float(ex), float(ey))
                if k < nz:
                    ez = (0.25 * abs(u[i, j, k] + u[i - 1, j, k] + u[i - 1, j, k + 1] + u[i, j, k + 1]) / dx
                          + 0.25 * abs(v[i, j, k] + v[i, j - 1, k] + v[i, j - 1, k + 1] + v[i, j, k + 1]) / dy
                          + abs(w[i, j, k]) / dz_c[k])
                    best = max(best, float(ez))
    got = O.compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c)
    assert abs(got - best) < 1e-12 * best


def test_spectra_of_single_mode():
    """w' = cos(3x - 2y) on every face: the folded spectrum equals numpy.fft's |W|^2/(nx ny)^2
    at (kx,ky)=(3,2) (plus its -kx image), zero elsewhere, and u'w' vanishes with u=0."""
    nx, ny, nz = 16, 12, 24
    Lx, Ly, Lz, nu = 2 * math.pi, 2 * math.pi, 2.0, 1 / 2870.0
    _, z_c, _, _ = O.generate_grid(2.6, nz, Lz)
    st = O.TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c, Lx / nx, Ly / ny, nu, 180.0, 15.0)
    x = (np.arange(nx) + 0.5) * Lx / nx
    y = (np.arange(ny) + 0.5) * Ly / ny
    mode = np.cos(3 * x[:, None] - 2 * y[None, :])
    u = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=torch.float64)
    v = torch.zeros(nx + 2, ny + 1, nz + 2, dtype=torch.float64)
    w = torch.zeros(nx + 2, ny + 2, nz + 1, dtype=torch.float64)
    w[1:nx + 1, 1:ny + 1, :] = torch.from_numpy(mode)[:, :, None]
    st.accumulate_statistics(u, v, w)
    E = np.abs(np.fft.rfft2(mode)) ** 2 / (nx * ny) ** 2
    fold = E[1:nx // 2 + 1, 1:ny // 2 + 1] + E[nx - nx // 2:nx, 1:ny // 2 + 1][::-1]
    got = st.spec["ww"].numpy()
    assert np.abs(got - fold).max() < 1e-15
    assert abs(got[2, 1] - 0.25) < 1e-15 and abs(got.sum() - 0.25) < 1e-14
    assert st.spec["uw"].abs().max() == 0.0 and st.spec["uu"].abs().max() == 0.0
    assert abs(float(st.sums["ww"][5]) - 0.5) < 1e-14
