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
    assert ch.turbulence_stats._plan(ch.u.dtype, ch.u.device).value == ch._grid.poisson().value
    n = len(L._grid_cache)
    ch.step_imex(0.005)
    ch.turbulence_stats.accumulate_statistics(ch.u, ch.v, ch.w)
    assert len(L._grid_cache) == n and ch.turbulence_stats.n_samples == 1
