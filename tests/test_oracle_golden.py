"""The CPU oracle against fixtures produced by the reference itself (tests/golden/make_golden.py)."""
import math
import os

import numpy as np
import pytest
import torch

from oracle import channel_oracle as O


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def ops(golden_dir):
    return np.load(os.path.join(golden_dir, "ops.npz"))


def test_fused_rhs_matches_reference(ops):
    nx, ny, nz = int(ops["nx"]), int(ops["ny"]), int(ops["nz"])
    Lx, Ly, Lz, nu = float(ops["Lx"]), float(ops["Ly"]), float(ops["Lz"]), float(ops["nu"])
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, Lz)
    u, v, w = (torch.from_numpy(ops[k]) for k in ("u", "v", "w"))
    ru, rv, rw = O.rhs_fused_imex(u, v, w, nx, ny, nz, Lx / nx, Ly / ny, dz_c, dz_f, nu)
    for got, key in ((ru, "rhs_u"), (rv, "rhs_v"), (rw, "rhs_w")):
        assert _rel(got.numpy(), ops[key]) < 1e-13, key
    # Q1/Q2: the fused op leaves u-face nx and v-face ny at zero
    assert np.all(ops["rhs_u"][nx] == 0.0) and np.all(ops["rhs_v"][:, ny] == 0.0)


def test_fused_v2_rhs_matches_reference(golden_dir):
    """compute_momentum_rhs_fused_v2 on a stretched grid (its w control volume dz_c[k-1])."""
    g = np.load(os.path.join(golden_dir, "ops_v2.npz"))
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, float(g["Lz"]))
    u, v, w = (torch.from_numpy(g[k]) for k in ("u", "v", "w"))
    got = O.rhs_fused_v2(u, v, w, nx, ny, nz, float(g["Lx"]) / nx, float(g["Ly"]) / ny, dz_c, dz_f, float(g["nu"]))
    for a, key in zip(got, ("rhs_u", "rhs_v", "rhs_w")):
        assert _rel(a.numpy(), g[key]) < 1e-13, key


def test_separate_rhs_equals_fused_on_interior(ops):
    """test_fused_kernels.py:81-83/176-178 idea: the two branches agree off the faces nx/ny."""
    nx, ny, nz = int(ops["nx"]), int(ops["ny"]), int(ops["nz"])
    Lx, Ly, Lz, nu = float(ops["Lx"]), float(ops["Ly"]), float(ops["Lz"]), float(ops["nu"])
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, Lz)
    u, v, w = (torch.from_numpy(ops[k]) for k in ("u", "v", "w"))
    a = O.rhs_separate(u, v, w, nx, ny, nz, Lx / nx, Ly / ny, dz_c, dz_f, nu)
    b = O.rhs_fused_imex(u, v, w, nx, ny, nz, Lx / nx, Ly / ny, dz_c, dz_f, nu)
    assert torch.allclose(a[0][1:nx], b[0][1:nx], rtol=0, atol=1e-12)
    assert torch.allclose(a[1][:, 1:ny], b[1][:, 1:ny], rtol=0, atol=1e-12)
    assert torch.allclose(a[2], b[2], rtol=0, atol=1e-12)
    assert a[0][nx].abs().max() > 0 and a[1][:, ny].abs().max() > 0   # Q2


def test_cn_w_matches_reference(ops):
    nx, ny, nz = int(ops["nx"]), int(ops["ny"]), int(ops["nz"])
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, float(ops["Lz"]))
    w = torch.from_numpy(ops["w"])
    got = O.solve_implicit_diffusion_w(w, float(ops["cn_dt"]), nx, ny, nz, dz_c, dz_f, float(ops["nu"]))
    assert _rel(got.numpy(), ops["w_cn"]) < 1e-14


def test_turbstats_matches_reference(ops):
    nx, ny, nz = int(ops["nx"]), int(ops["ny"]), int(ops["nz"])
    Lx, Ly, Lz, nu = float(ops["Lx"]), float(ops["Ly"]), float(ops["Lz"]), float(ops["nu"])
    _, z_c, _, _ = O.generate_grid(2.6, nz, Lz)
    st = O.TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c, Lx / nx, Ly / ny, nu, 180.0, 15.0)
    assert (st.k_bot, st.k_top) == (int(ops["k_bot"]), int(ops["k_top"]))
    st.accumulate_statistics(*(torch.from_numpy(ops[k]) for k in ("u", "v", "w")))
    fin = st.finalize_statistics()
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean",
              "E_uu_2d", "E_vv_2d", "E_ww_2d", "E_uw_2d"):
        assert _rel(fin[k], ops[k]) < 1e-12, k


def test_three_step_trajectory_matches_reference(golden_dir):
    """ChannelFlow composition (CPU branch), 16x12x16, seed 0: init + 3 IMEX steps."""
    g = np.load(os.path.join(golden_dir, "step_trajectory.npz"))
    ch = O.Channel(16, 12, 16, 4 * math.pi, 4 * math.pi / 3, seed=0, branch="cpu")
    for s in range(0, 4):
        if s > 0:
            ub, f = ch.step_imex(0.005)
            assert abs(float(ub) - float(g[f"ubulk{s}"])) < 1e-13
            assert abs(float(f) - float(g[f"forcing{s}"])) < 1e-10
            assert _rel(ch.rhs_prev[0].numpy(), g[f"rhs_u_prev{s}"]) < 1e-11
        for q, key in ((ch.u, "u"), (ch.v, "v"), (ch.w, "w")):
            assert _rel(q.numpy(), g[f"{key}{s}"]) < 1e-12, (key, s)
        pr = g[f"p{s}"]
        po = ch.p.numpy()
        assert _rel(po - po.mean(), pr - pr.mean()) < 1e-9, ("p", s)


@pytest.mark.slow_cpu
def test_c1_200_steps_matches_reference(golden_dir):
    """64x64x48 plain method, dt=0.005, 200 steps on the CPU branch (the CPU-runnable case)."""
    g = np.load(os.path.join(golden_dir, "c1_run.npz"))
    nx, ny, nz = 64, 64, 48
    ch = O.Channel(nx, ny, nz, 4 * math.pi, 4 * math.pi / 3, seed=0, branch="cpu")
    ub = []
    for _ in range(200):
        b, _ = ch.step_imex(0.005)
        ub.append(float(b))
    assert np.abs(np.array(ub) - g["u_bulk"]).max() < 1e-12
    u = ch.u[1:nx + 1, 1:ny + 1, 1:nz + 1]
    assert _rel(u.mean(dim=(0, 1)).numpy(), g["U_mean"]) < 1e-9
    assert _rel(u.std(dim=(0, 1), unbiased=False).numpy(), g["u_rms"]) < 1e-9
    w = ch.w[1:nx + 1, 1:ny + 1, 1:nz]
    assert _rel(w.std(dim=(0, 1), unbiased=False).numpy(), g["w_rms"]) < 1e-9


# ---- compute_u_tau and compute_cfl_fused against the reference (utau_cfl.npz) -----------------

@pytest.fixture(scope="module")
def uc(golden_dir):
    return np.load(os.path.join(golden_dir, "utau_cfl.npz"))


def _utau_variant(u, z_c, nu, top, interior=False, top_dist=None, top_plane=-2, no_abs=False):
    """compute_u_tau with one plausible mistake planted (for the pin checks below)."""
    nx, ny = u.shape[0] - 1, u.shape[1] - 2
    plane = (lambda k: u[1:nx + 1, 1:ny + 1, k]) if interior else (lambda k: u[:, :, k])
    sq = (lambda x: torch.sqrt(x)) if no_abs else (lambda x: torch.sqrt(torch.abs(x)))
    ub = sq(nu * torch.mean(plane(1)) / z_c[1])
    if top == "neumann":
        return float(ub)
    d_top = z_c[1] if top_dist is None else top_dist
    return float(0.5 * (ub + sq(nu * torch.mean(plane(top_plane)) / d_top)))


@pytest.mark.parametrize("grid_kind", ["symmetric", "bottom"])
@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_u_tau_matches_reference(uc, grid_kind, top):
    """Whole-plane means (ghost rows and face 0 included), z_c[1] for both walls, abs inside
    the square root (utils.py:176-204).  Each planted mistake moves the result off the
    reference value, so the fixture pins every one of them."""
    nu, nz = float(uc["nu"]), int(uc["nz"])
    u = torch.from_numpy(uc[f"{grid_kind}_u"])
    z_c = torch.from_numpy(uc[f"{grid_kind}_z_c"])
    want = float(uc[f"{grid_kind}_utau_{top}"])
    got = float(O.compute_u_tau(u, z_c, nu, top))
    assert abs(got - want) <= 1e-14 * abs(want)
    assert _utau_variant(u, z_c, nu, top) == pytest.approx(want, rel=1e-14)
    wrong = [_utau_variant(u, z_c, nu, top, interior=True)]
    if top == "dirichlet":
        wrong.append(_utau_variant(u, z_c, nu, top, top_plane=-1))              # ghost row, not k = nz
        wrong.append(_utau_variant(u, z_c, nu, top, no_abs=True))               # NaN: negative top mean
        if grid_kind == "bottom":
            wrong.append(_utau_variant(u, z_c, nu, top, top_dist=2.0 - z_c[nz]))  # distance to the top wall
    for x in wrong:
        assert not (abs(x - want) <= 1e-6 * abs(want)), (x, want)


@pytest.mark.parametrize("grid_kind", ["symmetric", "bottom"])
@pytest.mark.parametrize("case", [0, 1, 2])
def test_cfl_rate_matches_reference(uc, grid_kind, case):
    """compute_cfl_fused (operators.py:1232-1333) on random signed fields, with spikes next to
    a wall and in a ghost row; swapping the cell and face spacings (a plausible mistake) moves
    the rate off the reference value on the spike-free case."""
    nx, ny, nz = int(uc["nx"]), int(uc["ny"]), int(uc["nz"])
    Lx, Ly, Lz = float(uc["Lx"]), float(uc["Ly"]), float(uc["Lz"])
    _, _, dz_f, dz_c = O.generate_grid(2.6, nz, Lz, grid_kind)
    key = f"{grid_kind}_cfl{case}"
    u, v, w = (torch.from_numpy(uc[key + s]) for s in ("_u", "_v", "_w"))
    want = float(uc[key])
    got = O.compute_cfl_fused(u, v, w, nx, ny, nz, Lx / nx, Ly / ny, dz_f, dz_c)
    assert abs(got - want) <= 1e-14 * want
    if case == 0:
        bad = O.compute_cfl_fused(u, v, w, nx, ny, nz, Lx / nx, Ly / ny, dz_c[1:], torch.cat([dz_f[:1], dz_f]))
        assert abs(bad - want) > 1e-6 * want


# ---- run_simulation cadence and restart against the reference (runloop.npz) -------------------

@pytest.fixture(scope="module")
def rl(golden_dir):
    return np.load(os.path.join(golden_dir, "runloop.npz"))


def _run_leg(rl, leg):
    """The oracle's Channel.run over one leg of the reference's run_simulation fixture."""
    nx, ny, nz = int(rl["cfg_nx"]), int(rl["cfg_ny"]), int(rl["cfg_nz"])
    Lx, Ly = 4 * math.pi, 4 * math.pi / 3
    restart = None
    if leg == "b":
        restart = (rl["a_fields__u"], rl["a_fields__v"], rl["a_fields__w"], int(rl["a_fields__step"]),
                   float(rl["a_fields__time"]))
    ch = O.Channel(nx, ny, nz, Lx, Ly, seed=int(rl["cfg_seed"]), branch="cpu", restart=restart)
    st = O.TurbulenceStats(nx, ny, nz, Lx, Ly, 2.0, ch.z_c, ch.dx, ch.dy, ch.nu, 180.0, 15.0)
    if leg == "b":                                                  # turbstats.py:366-402
        st.n_samples = int(rl["a_turbulence_stats_state__n_samples"])
        for k in ("U", "uu", "vv", "ww", "uw"):
            st.sums[k] = torch.from_numpy(rl[f"a_turbulence_stats_state__{k}_sum"]).clone()
        for k in ("uu", "vv", "ww", "uw"):
            st.spec[k] = torch.from_numpy(rl[f"a_turbulence_stats_state__E_{k}_2d_sum"]).clone()
    n_steps = int(rl["cfg_n_steps"]) if leg == "a" else int(rl["cfg_restart_n_steps"])
    t_max = float(rl["cfg_t_max"]) if leg == "a" else 1000.0
    dt, rows = ch.run(n_steps, float(rl["cfg_dt"]), int(rl["cfg_dt_update_interval"]), float(rl["cfg_CFL_target"]),
                      stats=st, n_stats=int(rl["cfg_n_stats"]), t_stats=float(rl["cfg_t_stats"]), t_max=t_max,
                      n_out=int(rl["cfg_n_out"]))
    return ch, st, dt, rows


@pytest.mark.parametrize("leg", ["a", "b"])
def test_run_loop_matches_reference(rl, leg):
    """Leg a: fresh start; the adaptive-dt rule changes dt, statistics start past t_stats and
    t_max stops the loop between two saves.  Leg b: restart from leg a's last checkpoint and
    statistics state (step and time kept, forcing and AB2 history reset)."""
    ch, st, dt, rows = _run_leg(rl, leg)
    p = f"{leg}_"
    # fields_init.npz holds the state before the initial projection (solver.py:337-340)
    for q, k in zip(ch.init_state, ("u", "v", "w")):
        assert _rel(q.numpy(), rl[p + "fields_init__" + k]) < 1e-13, k
    ts = {k: rl[p + "timeseries__" + k] for k in ("step", "time", "u_bulk", "u_tau", "forcing")}
    n = len(ts["step"])
    rows = np.array(rows[:n])
    assert np.array_equal(rows[:, 0].astype(int), ts["step"])
    assert np.abs(rows[:, 1] - ts["time"]).max() < 1e-15
    assert np.abs(rows[:, 2] - ts["u_bulk"]).max() < 1e-13
    assert np.abs(rows[:, 3] - ts["u_tau"]).max() < 1e-12
    assert np.abs(rows[:, 4] - ts["forcing"]).max() < 1e-10
    assert ch.current_step == int(rl[p + "fields_final__step"])
    assert abs(ch.time - float(rl[p + "final_time"])) < 1e-15
    assert dt == float(rl[p + "final_dt"])
    for q, k in ((ch.u, "u"), (ch.v, "v"), (ch.w, "w")):
        assert _rel(q.numpy(), rl[p + "fields_final__" + k]) < 1e-12, k
    pr, po = rl[p + "fields_final__p"], ch.p.numpy()
    assert _rel(po - po.mean(), pr - pr.mean()) < 1e-9
    fin = st.finalize_statistics()
    assert fin["n_samples"] == int(rl[p + "turbulence_stats__n_samples"])
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean", "E_uu_2d", "E_vv_2d", "E_ww_2d", "E_uw_2d"):
        assert _rel(fin[k], rl[p + "turbulence_stats__" + k]) < 1e-9, k
