"""run_simulation's loop (solver.py:738-970) on the device: graph chunks between the n_out
events, the n_out diagnostics as device time-series rows, host reads once per n_save."""
import math
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def cfg(folder, n_steps=40, n_out=10, n_save=20, dt_interval=10, nx=16, ny=12, nz=16):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": 0.005, "n_steps": n_steps, "CFL_target": 0.25, "dt_update_interval": dt_interval},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": 0},
        "compute": {"device": "cuda", "precision": "fp64"},
        "output": {"results_folder": folder, "n_out": n_out, "n_save": n_save},
    }


@pytest.fixture(scope="module")
def tc():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import torchannel_amd
    torchannel_amd._lib.load()
    return torchannel_amd


def test_diagnostics_row_matches_host_functions(tc, tmp_path):
    from torchannel_amd import _lib as L
    ch = tc.ChannelFlow(config=cfg(str(tmp_path)))
    for s in range(3):
        ch.step_imex(0.005)
    ch.update_dt_device()
    rows = torch.zeros(8 * 3, dtype=torch.float64, device="cuda")
    ch.record_diagnostics(rows, 2)
    ch.step_imex(None)
    ch.record_diagnostics(rows, 2)
    ch.record_diagnostics(rows, 2)          # past capacity: counted, not written
    h = rows.cpu().numpy()
    assert int(h[0]) == 3
    r = h[16:24]                            # second row: the state after the fourth step
    scal = ch._scal.cpu().numpy()
    assert r[0] == scal[L.S_STEP] and r[1] == scal[L.S_TIME] and r[2] == scal[L.S_UBULK]
    assert r[4] == scal[L.S_F] and r[5] == scal[L.S_DT] and r[7] == scal[L.S_CFL]
    ut = float(tc.compute_u_tau(ch.u, ch.z_c, ch.nu, "dirichlet"))
    assert abs(r[3] - ut) <= 1e-14 * abs(ut)
    div = tc.compute_divergence(ch.u, ch.v, ch.w, ch.nx, ch.ny, ch.nz, ch.dx, ch.dy, ch.dz_f)
    assert r[6] == float(div.abs().max())


@pytest.mark.parametrize("n_steps,n_out,n_save,dt_interval", [(40, 10, 20, 10), (30, 5, 15, 5)])
def test_run_simulation_matches_eager_loop(tc, tmp_path, n_steps, n_out, n_save, dt_interval):
    """n_out = 10 runs graph chunks between the events, n_out = 5 (odd) eager steps."""
    folder = str(tmp_path / "run")
    print(f"[runloop] run_simulation {n_steps} steps, n_out {n_out}, n_save {n_save}, dt every {dt_interval}",
          flush=True)
    kw = dict(n_steps=n_steps, n_out=n_out, n_save=n_save, dt_interval=dt_interval)
    ch = tc.ChannelFlow(config=cfg(folder, **kw))
    assert ch.run_simulation() == n_steps
    # the same cadence, eagerly, recording what the reference loop records
    ref = tc.ChannelFlow(config=cfg(str(tmp_path / "eager"), **kw))
    want = []
    for step in range(1, n_steps + 1):
        if step % dt_interval == 0:
            ref.update_dt_device()
        ub, f = ref.step_imex(None)
        if step % n_out == 0:
            ut = tc.compute_u_tau(ref.u, ref.z_c, ref.nu, "dirichlet")
            want.append((step, float(ref._scal[3]), float(ub), float(ut), float(f)))
    for a, b in ((ch.u, ref.u), (ch.v, ref.v), (ch.w, ref.w), (ch.p, ref.p)):
        assert torch.equal(a, b)
    ts = np.load(os.path.join(folder, "timeseries.npz"))
    assert list(ts["step"]) == [w[0] for w in want]
    for key, col in (("time", 1), ("u_bulk", 2), ("forcing", 4)):
        assert np.array_equal(ts[key], np.array([w[col] for w in want])), key
    assert np.allclose(ts["u_tau"], [w[3] for w in want], rtol=1e-14, atol=0)
    fields = np.load(os.path.join(folder, "fields.npz"))
    assert int(fields["step"]) == n_steps
    assert np.array_equal(fields["u"], ch.u.cpu().numpy())


def test_statistics_share_the_solver_plan(tc, tmp_path):
    """The statistics sample on the solver's own grid and plan ('cuda' and 'cuda:0' are one cache
    key), so no second plan is built at the first sample inside the time loop."""
    from torchannel_amd import _lib as L
    c = cfg(str(tmp_path))
    c["statistics"] = {"n_stats": 10, "t_stats": 0.0, "z_plus_target": 15.0}
    ch = tc.ChannelFlow(config=c)
    assert L.device_key("cuda") == L.device_key(ch.u.device)
    assert ch.turbulence_stats._plan(ch.u.dtype, ch.u.device, ch.pitch).value == ch._grid.poisson().value
    n = len(L._grid_cache)
    ch.step_imex(0.005)
    ch.turbulence_stats.accumulate_statistics(ch.u, ch.v, ch.w)
    assert len(L._grid_cache) == n and ch.turbulence_stats.n_samples == 1


# ---- the reference's own run_simulation output files (tests/golden/runloop.npz) ---------------

def _rl_config(g, folder, n_steps, t_max, field_file=None, stats_file=None):
    """The configuration the fixture's reference runs used (tests/golden/make_golden.py), on the
    CPU-branch right-hand side the reference takes on a CPU device."""
    c = {
        "grid": {"nx": int(g["cfg_nx"]), "ny": int(g["cfg_ny"]), "nz": int(g["cfg_nz"])},
        "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": float(g["cfg_dt"]), "n_steps": int(n_steps), "t_max": float(t_max),
                 "CFL_target": float(g["cfg_CFL_target"]), "dt_update_interval": int(g["cfg_dt_update_interval"])},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": int(g["cfg_seed"])},
        "output": {"results_folder": folder, "n_out": int(g["cfg_n_out"]), "n_save": int(g["cfg_n_save"])},
        "compute": {"device": "cuda", "precision": "fp64", "rhs_branch": "cpu"},
        "statistics": {"n_stats": int(g["cfg_n_stats"]), "t_stats": float(g["cfg_t_stats"]), "z_plus_target": 15.0},
    }
    if field_file is not None:
        c["initialization"]["field_file"] = field_file
    if stats_file is not None:
        c["statistics"]["restart_state_file"] = stats_file
    return c


def _relerr(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _check_files(g, leg, folder):
    """Every npz the reference wrote in this leg, against the files written here."""
    p = f"{leg}_"
    files = sorted(f for f in os.listdir(folder) if f.endswith(".npz"))
    assert files == [str(x) for x in g[p + "files"]]
    for name in ("fields_init", "fields", "fields_final"):
        f = np.load(os.path.join(folder, name + ".npz"))
        want = lambda k: g[f"{p}{name}__{k}"]
        for k in ("u", "v", "w"):
            assert _relerr(f[k], want(k)) < 1e-12, (name, k)
        pa, pb = f["p"], want("p")
        assert _relerr(pa - pa.mean(), pb - pb.mean()) < 1e-9, (name, "p")
        for k in ("z_c", "z_f"):
            assert _relerr(f[k], want(k)) < 1e-15, (name, k)
        assert int(f["step"]) == int(want("step")), name
        assert abs(float(f["time"]) - float(want("time"))) < 1e-15, name
        assert float(f["Lx"]) == float(want("Lx")) and float(f["Ly"]) == float(want("Ly"))
        assert abs(float(f["u_tau"]) - float(want("u_tau"))) <= 1e-12 * abs(float(want("u_tau"))), name
        assert abs(float(f["forcing"]) - float(want("forcing"))) < 1e-10, name
    ts = np.load(os.path.join(folder, "timeseries.npz"))
    assert np.array_equal(ts["step"], g[p + "timeseries__step"])
    assert np.abs(ts["time"] - g[p + "timeseries__time"]).max() < 1e-15
    assert np.abs(ts["u_bulk"] - g[p + "timeseries__u_bulk"]).max() < 1e-13
    assert _relerr(ts["u_tau"], g[p + "timeseries__u_tau"]) < 1e-12
    assert np.abs(ts["forcing"] - g[p + "timeseries__forcing"]).max() < 1e-10
    st = np.load(os.path.join(folder, "turbulence_stats.npz"))
    assert int(st["n_samples"]) == int(g[p + "turbulence_stats__n_samples"])
    assert int(st["z_plus_target"]) == int(g[p + "turbulence_stats__z_plus_target"])
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean", "E_uu_2d", "E_vv_2d", "E_ww_2d", "E_uw_2d",
              "z_c", "kx", "ky"):
        assert _relerr(st[k], g[p + "turbulence_stats__" + k]) < 1e-9, k
    for k in ("Re_tau_target", "nu", "Re", "u_tau"):
        assert abs(float(st[k]) - float(g[p + "turbulence_stats__" + k])) <= 1e-9 * abs(float(st[k])), k
    ss = np.load(os.path.join(folder, "turbulence_stats_state.npz"))
    for k in ("n_samples", "nx", "ny", "nz"):
        assert int(ss[k]) == int(g[p + "turbulence_stats_state__" + k]), k
    for k in ("U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum", "E_uu_2d_sum", "E_vv_2d_sum", "E_ww_2d_sum",
              "E_uw_2d_sum"):
        assert _relerr(ss[k], g[p + "turbulence_stats_state__" + k]) < 1e-9, k


def test_run_simulation_writes_the_reference_files(tc, tmp_path, golden_dir):
    """Fresh start: the adaptive-dt rule changes dt, statistics begin past t_stats, t_max stops
    the loop between two saves (so the last rows stay unwritten), and every file the reference
    wrote is written here with the same keys and values."""
    g = np.load(os.path.join(golden_dir, "runloop.npz"))
    folder = str(tmp_path / "a")
    print("[runloop] reference run_simulation fixture, fresh start", flush=True)
    ch = tc.ChannelFlow(config=_rl_config(g, folder, int(g["cfg_n_steps"]), float(g["cfg_t_max"])))
    assert ch.run_simulation() == int(g["a_fields_final__step"])
    assert abs(ch.time - float(g["a_final_time"])) < 1e-15
    assert abs(ch.dt - float(g["a_final_dt"])) <= 1e-14 * float(g["a_final_dt"])
    _check_files(g, "a", folder)


def test_restart_from_reference_checkpoint(tc, tmp_path, golden_dir):
    """Restart from the fields.npz and statistics state the reference wrote: step and time are
    kept, forcing and AB2 history start from zero, and the continued run writes the reference's
    files."""
    g = np.load(os.path.join(golden_dir, "runloop.npz"))
    keys = ("u", "v", "w", "p", "z_c", "z_f", "Lx", "Ly", "step", "time", "u_tau", "forcing")
    ff = str(tmp_path / "a_fields.npz")
    np.savez(ff, **{k: g["a_fields__" + k] for k in keys})
    sf = str(tmp_path / "a_state.npz")
    np.savez_compressed(sf, **{k[len("a_turbulence_stats_state__"):]: g[k] for k in g.files
                               if k.startswith("a_turbulence_stats_state__")})
    folder = str(tmp_path / "b")
    print("[runloop] reference run_simulation fixture, restart", flush=True)
    ch = tc.ChannelFlow(config=_rl_config(g, folder, int(g["cfg_restart_n_steps"]), 1000.0, ff, sf))
    assert ch.current_step == int(g["a_fields__step"]) and ch.time == float(g["a_fields__time"])
    assert float(ch.forcing) == 0.0 and ch.rhs_u_prev is None
    assert ch.run_simulation() == int(g["cfg_restart_n_steps"])
    _check_files(g, "b", folder)


def test_non_finite_state_stops_at_next_save(tc, tmp_path):
    """A NaN in the state (a plain value) is seen at the next n_save step (solver.py:862-881):
    fields_error.npz is written, the loop stops, fields.npz keeps the last good checkpoint and
    the time-series rows after it are not written; fields_final.npz holds the state at the stop."""
    folder = str(tmp_path / "nan")
    ch = tc.ChannelFlow(config=cfg(folder, n_steps=10, n_out=5, n_save=10, dt_interval=0))
    assert ch.run_simulation() == 10
    ch.u[4, 4, 4] = float("nan")
    ch.n_steps = 40
    print("[runloop] NaN planted at step 10", flush=True)
    assert ch.run_simulation() == 20
    err = np.load(os.path.join(folder, "fields_error.npz"))
    assert int(err["step"]) == 20 and np.isnan(err["u"]).any()
    good = np.load(os.path.join(folder, "fields.npz"))
    assert int(good["step"]) == 10 and np.isfinite(good["u"]).all()
    fin = np.load(os.path.join(folder, "fields_final.npz"))
    assert int(fin["step"]) == 20
    ts = np.load(os.path.join(folder, "timeseries.npz"))
    assert list(ts["step"]) == [5, 10]


def test_graph_after_odd_eager_steps(tc, tmp_path):
    """Graph chunks, then an odd number of eager steps (the AB2 history's parity flips), then
    graph chunks again: the second replay recaptures for the new parity, and the states equal an
    all-eager run.  The library refuses a replay of a graph recorded for the other parity."""
    from torchannel_amd import _lib as L
    nx, ny, nz = 16, 12, 16
    a = tc.ChannelFlow(config=cfg(str(tmp_path), nx=nx, ny=ny, nz=nz))
    b = tc.ChannelFlow(config=cfg(str(tmp_path), nx=nx, ny=ny, nz=nz))
    for ch in (a, b):
        ch.step_imex(0.005)
    a.run_graph(1, 4)
    for _ in range(3):
        a.step_imex(None)
    with pytest.raises(RuntimeError, match="captured graph"):
        L.call("tc_solver_replay", a._solver, a._stream())
    a.run_graph(2, 4)
    for _ in range(4 + 3 + 8):
        b.step_imex(None)
    for q, r in ((a.u, b.u), (a.v, b.v), (a.w, b.w), (a.p, b.p), (a._scal, b._scal)):
        assert torch.equal(q, r)
    a.body_force = torch.zeros_like(a.u)
    with pytest.raises(RuntimeError, match="body force"):
        a.run_graph(1, 4)
