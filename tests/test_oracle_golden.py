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
