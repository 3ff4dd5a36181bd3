"""Write the reference-generated fixtures under tests/golden/ (run in the build container only).

The reference (pure Python/PyTorch) is imported from /root/reference to produce input/output
vectors for the parts of the IMEX path that closed forms and invariants leave open:

* ``step_trajectory.npz`` -- ChannelFlow.step_imex composition (BC order, AB2 buffers, forcing
  controller, CN, FFT-Poisson, projection) on a 16x12x16 seeded channel, CPU branch, 3 steps.
* ``c1_run.npz`` -- the 64x64x48 plain-method case (dt=0.005, CPU branch) for 200 steps:
  per-step u_bulk and forcing, plus final mean/RMS profiles.
* ``ops.npz`` -- compute_momentum_rhs_fused_imex (GPU-branch RHS, including the u-face nx /
  v-face ny behaviour), solve_implicit_diffusion_w (shifted explicit half), and
  TurbulenceStats.accumulate_statistics on one random field.
* ``ops_v2.npz`` -- compute_momentum_rhs_fused_v2 (FE-scheme RHS) on random fields.
* ``utau_cfl.npz`` -- compute_u_tau (both top walls) on fields whose ghost rows, face 0 and two
  wall planes all have different means, over the one-sided ('bottom') grid where z_c[1] differs
  from Lz - z_c[nz]; compute_cfl_fused on random signed fields with and without a corner spike.
* ``runloop.npz`` -- run_simulation (solver.py:738-970) on a 16x12x16 channel (CPU branch) with
  the adaptive-dt rule changing dt, statistics past t_stats and t_max reached between two saves:
  every file it writes (fields_init/fields/fields_final, timeseries, statistics and their state);
  then a restart from that fields.npz and statistics state, run to a fixed step count.

Only data is written: inputs and the reference's outputs.  Usage:
    PYTORCH_JIT=0 MPLBACKEND=Agg python tests/golden/make_golden.py [ops] [trajectory] [c1] [v2] [utau_cfl] [runloop]
"""
import os
import sys
import tempfile

os.environ.setdefault("PYTORCH_JIT", "0")      # the reference assumes eager mode (operators.py:1339)
os.environ.setdefault("MPLBACKEND", "Agg")

import numpy as np
import torch
import yaml

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REF)

import solver as ref_solver                     # noqa: E402
import operators as ref_ops                     # noqa: E402
import turbstats as ref_turb                    # noqa: E402
from utils import generate_grid                 # noqa: E402

torch.set_default_dtype(torch.float64)


def channel(tmp, nx, ny, nz, Lx, Ly, seed, n_stats=0, dt=0.005):
    cfg = {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": Lx, "Ly": Ly, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": dt, "n_steps": 1, "CFL_target": 0.25},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1},
        "output": {"results_folder": os.path.join(tmp, "res"), "n_out": 10, "n_save": 1000},
        "compute": {"device": "cpu"},
        "statistics": {"n_stats": n_stats},
    }
    path = os.path.join(tmp, "cfg.yaml")
    with open(path, "w") as f:
        yaml.safe_dump(cfg, f)
    torch.manual_seed(seed)
    return ref_solver.ChannelFlow(path)


def np64(t):
    return t.detach().cpu().numpy().astype(np.float64)


def make_trajectory(tmp):
    cf = channel(tmp, 16, 12, 16, 4 * np.pi, 4 * np.pi / 3, seed=0)
    out = {"u0": np64(cf.u), "v0": np64(cf.v), "w0": np64(cf.w), "p0": np64(cf.p)}
    for s in range(1, 4):
        ub, f = cf.step_imex(0.005)
        out[f"u{s}"], out[f"v{s}"], out[f"w{s}"], out[f"p{s}"] = (np64(cf.u), np64(cf.v),
                                                                    np64(cf.w), np64(cf.p))
        out[f"ubulk{s}"] = float(ub)
        out[f"forcing{s}"] = float(f)
        out[f"rhs_u_prev{s}"] = np64(cf.rhs_u_prev)
    np.savez_compressed(os.path.join(HERE, "step_trajectory.npz"), **out)


def make_c1(tmp):
    nx, ny, nz = 64, 64, 48
    cf = channel(tmp, nx, ny, nz, 4 * np.pi, 4 * np.pi / 3, seed=0)
    ub, fo = [], []
    for _ in range(200):
        b, f = cf.step_imex(0.005)
        ub.append(float(b))
        fo.append(float(f))
    u = cf.u[1:nx + 1, 1:ny + 1, 1:nz + 1]
    v = cf.v[1:nx + 1, 1:ny + 1, 1:nz + 1]
    w = cf.w[1:nx + 1, 1:ny + 1, 1:nz]
    div = ref_solver.compute_divergence(cf.u, cf.v, cf.w, nx, ny, nz, cf.dx, cf.dy, cf.dz_f)
    np.savez_compressed(
        os.path.join(HERE, "c1_run.npz"),
        u_bulk=np.array(ub), forcing=np.array(fo),
        U_mean=np64(u.mean(dim=(0, 1))), u_rms=np64(u.std(dim=(0, 1), unbiased=False)),
        v_rms=np64(v.std(dim=(0, 1), unbiased=False)), w_rms=np64(w.std(dim=(0, 1), unbiased=False)),
        max_div=float(div.abs().max()))


def make_ops():
    g = torch.Generator().manual_seed(1234)
    nx, ny, nz = 12, 10, 14
    Lx, Ly, Lz, nu = 2.0, 1.5, 2.0, 1.0 / 2870.0
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c, dz_f, dz_c = generate_grid(2.6, nz, Lz)
    u = torch.rand(nx + 1, ny + 2, nz + 2, generator=g) - 0.5
    v = torch.rand(nx + 2, ny + 1, nz + 2, generator=g) - 0.5
    w = torch.rand(nx + 2, ny + 2, nz + 1, generator=g) - 0.5
    ru, rv, rw = ref_ops.compute_momentum_rhs_fused_imex(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu)
    w_cn = ref_ops.solve_implicit_diffusion_w(w, torch.tensor(0.01), nx, ny, nz, dz_c, dz_f, nu)
    stats = ref_turb.TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c, z_f, dz_c, dz_f, dx, dy, nu, 180.0,
                                     z_plus_target=15.0, device="cpu")
    stats.accumulate_statistics(u, v, w, torch.tensor(0.05))
    fin = stats.finalize_statistics()
    np.savez_compressed(
        os.path.join(HERE, "ops.npz"),
        nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Lz=Lz, nu=nu,
        u=np64(u), v=np64(v), w=np64(w),
        rhs_u=np64(ru), rhs_v=np64(rv), rhs_w=np64(rw), w_cn=np64(w_cn), cn_dt=0.01,
        k_bot=stats.k_bot, k_top=stats.k_top,
        U_mean=fin["U_mean"], uu_mean=fin["uu_mean"], vv_mean=fin["vv_mean"],
        ww_mean=fin["ww_mean"], uw_mean=fin["uw_mean"], E_uu_2d=fin["E_uu_2d"],
        E_vv_2d=fin["E_vv_2d"], E_ww_2d=fin["E_ww_2d"], E_uw_2d=fin["E_uw_2d"],
        u_tau=fin["u_tau"])


def make_v2():
    """ops_v2.npz: compute_momentum_rhs_fused_v2 (FE-scheme RHS, full explicit 3-D diffusion) on
    random fields over a stretched grid, where its w control volume dz_c[k-1] differs from
    dz_c[k]."""
    g = torch.Generator().manual_seed(4321)
    nx, ny, nz = 12, 10, 14
    Lx, Ly, Lz, nu = 2.0, 1.5, 2.0, 1.0 / 2870.0
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c, dz_f, dz_c = generate_grid(2.6, nz, Lz)
    u = torch.rand(nx + 1, ny + 2, nz + 2, generator=g) - 0.5
    v = torch.rand(nx + 2, ny + 1, nz + 2, generator=g) - 0.5
    w = torch.rand(nx + 2, ny + 2, nz + 1, generator=g) - 0.5
    ru, rv, rw = ref_ops.compute_momentum_rhs_fused_v2(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu)
    np.savez_compressed(os.path.join(HERE, "ops_v2.npz"), nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Lz=Lz, nu=nu,
                        u=np64(u), v=np64(v), w=np64(w), rhs_u=np64(ru), rhs_v=np64(rv), rhs_w=np64(rw))


def make_utau_cfl():
    """utau_cfl.npz: compute_u_tau and compute_cfl_fused on random fields (reference functions)."""
    from utils import compute_u_tau
    g = torch.Generator().manual_seed(2468)
    nx, ny, nz = 12, 10, 14
    Lx, Ly, Lz, nu = 2.0, 1.5, 2.0, 1.0 / 2870.0
    out = {"nx": nx, "ny": ny, "nz": nz, "Lx": Lx, "Ly": Ly, "Lz": Lz, "nu": nu}
    for grid_kind in ("symmetric", "bottom"):
        z_f, z_c, dz_f, dz_c = generate_grid(2.6, nz, Lz, stretching_type=grid_kind)
        u = torch.rand(nx + 1, ny + 2, nz + 2, generator=g) + 0.05 * torch.arange(nz + 2).view(1, 1, -1)
        u[0] += 1.0                 # the duplicate periodic face
        u[:, 0] += 2.0              # y ghost rows
        u[:, ny + 1] -= 1.5
        u[:, :, nz] -= 3.0          # top-wall plane: negative mean (abs inside the square root)
        out[f"{grid_kind}_u"] = np64(u)
        out[f"{grid_kind}_z_c"] = np64(z_c)
        for top in ("dirichlet", "neumann"):
            out[f"{grid_kind}_utau_{top}"] = float(compute_u_tau(u, z_c, nu, top_wall_bc_type=top))
        for case in range(3):
            u = (torch.rand(nx + 1, ny + 2, nz + 2, generator=g) - 0.3) * 2.0
            v = (torch.rand(nx + 2, ny + 1, nz + 2, generator=g) - 0.5) * 0.7
            w = (torch.rand(nx + 2, ny + 2, nz + 1, generator=g) - 0.5) * 0.4
            if case == 1:
                w[nx, ny, nz - 1] = 25.0           # spike next to the top wall in the last column
            if case == 2:
                v[1, 0, 1] = -40.0                 # spike in a y ghost row next to the bottom wall
            dx, dy = Lx / nx, Ly / ny
            rate = ref_ops.compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c)
            key = f"{grid_kind}_cfl{case}"
            out[key + "_u"], out[key + "_v"], out[key + "_w"] = np64(u), np64(v), np64(w)
            out[key] = float(rate)
    np.savez_compressed(os.path.join(HERE, "utau_cfl.npz"), **out)


RUNLOOP = {
    "nx": 16, "ny": 12, "nz": 16, "seed": 0, "dt": 0.005, "dt_update_interval": 5, "CFL_target": 0.01,
    "n_out": 4, "n_save": 8, "n_stats": 4, "t_stats": 0.05, "t_max": 0.15, "n_steps": 1000,
    "restart_n_steps": 44,
}


def runloop_config(res, n_steps, t_max, field_file=None, stats_file=None):
    """The reference YAML of the run-loop fixture (also rebuilt by the tests for this package)."""
    r = RUNLOOP
    cfg = {
        "grid": {"nx": r["nx"], "ny": r["ny"], "nz": r["nz"]},
        "domain": {"Lx": 4 * np.pi, "Ly": 4 * np.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": r["dt"], "n_steps": n_steps, "t_max": t_max, "CFL_target": r["CFL_target"],
                 "dt_update_interval": r["dt_update_interval"]},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1},
        "output": {"results_folder": res, "n_out": r["n_out"], "n_save": r["n_save"]},
        "compute": {"device": "cpu"},
        "statistics": {"n_stats": r["n_stats"], "t_stats": r["t_stats"], "z_plus_target": 15.0},
    }
    if field_file is not None:
        cfg["initialization"]["field_file"] = field_file
    if stats_file is not None:
        cfg["statistics"]["restart_state_file"] = stats_file
    return cfg


def _npz(path):
    d = np.load(path, allow_pickle=False)
    return {k: d[k] for k in d.files}


def make_runloop(tmp):
    """runloop.npz: every file run_simulation writes, for a fresh run and for a restart."""
    r = RUNLOOP
    out = {}
    for leg in ("a", "b"):
        res = os.path.join(tmp, f"run_{leg}")
        if leg == "a":
            cfg = runloop_config(res, r["n_steps"], r["t_max"])
        else:
            cfg = runloop_config(res, r["restart_n_steps"], 1000.0,
                                 field_file=os.path.join(tmp, "a_fields.npz"),
                                 stats_file=os.path.join(tmp, "a_stats_state.npz"))
        path = os.path.join(tmp, f"cfg_{leg}.yaml")
        with open(path, "w") as f:
            yaml.safe_dump(cfg, f)
        torch.manual_seed(r["seed"])
        cf = ref_solver.ChannelFlow(path)
        cf.run_simulation()
        out[f"{leg}_final_dt"] = float(cf.dt)
        out[f"{leg}_final_time"] = float(cf.time)
        out[f"{leg}_files"] = np.array(sorted(f for f in os.listdir(res) if f.endswith(".npz")))
        for name in ("fields_init", "fields", "fields_final", "timeseries", "turbulence_stats",
                     "turbulence_stats_state"):
            fp = os.path.join(res, name + ".npz")
            if os.path.exists(fp):
                for k, val in _npz(fp).items():
                    out[f"{leg}_{name}__{k}"] = val
        if leg == "a":
            # the restart inputs: this run's last checkpoint and statistics state
            import shutil
            shutil.copy(os.path.join(res, "fields.npz"), os.path.join(tmp, "a_fields.npz"))
            shutil.copy(os.path.join(res, "turbulence_stats_state.npz"), os.path.join(tmp, "a_stats_state.npz"))
    for k, val in r.items():
        out[f"cfg_{k}"] = val
    np.savez_compressed(os.path.join(HERE, "runloop.npz"), **out)


if __name__ == "__main__":
    only = sys.argv[1:]
    with tempfile.TemporaryDirectory() as tmp:
        if not only or "ops" in only:
            make_ops()
        if not only or "trajectory" in only:
            make_trajectory(tmp)
        if not only or "c1" in only:
            make_c1(tmp)
        if not only or "v2" in only:
            make_v2()
        if not only or "utau_cfl" in only:
            make_utau_cfl()
        if not only or "runloop" in only:
            make_runloop(tmp)
    print("fixtures written to", HERE)
