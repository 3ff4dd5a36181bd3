"""Whole IMEX steps on the HIP path against the oracle and against the reference's own
trajectories (tests/golden): per step, per field, <= 1e-12 relative in float64."""
import math
import os

import numpy as np
import pytest
import torch

from oracle import channel_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a = torch.as_tensor(a).detach().cpu().double()
    b = torch.as_tensor(b).detach().cpu().double()
    den = float(b.abs().max())
    return float((a - b).abs().max()) / (den if den > 0 else 1.0)


def cfg(nx, ny, nz, Lx=4 * math.pi, Ly=4 * math.pi / 3, dt=0.005, seed=0, branch="gpu", precision="fp64",
        amp=0.1, top="dirichlet"):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": Lx, "Ly": Ly, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": dt, "n_steps": 10, "CFL_target": 0.25},
        "initialization": {"type": "parabolic", "perturbation_intensity": amp, "seed": seed},
        "compute": {"device": "cuda", "precision": precision, "rhs_branch": branch},
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


@pytest.mark.parametrize("shape", [(14, 12, 16), (16, 15, 16), (16, 12, 300), (16, 12, 3), (2, 12, 16),
                                   (1600, 4, 8)])
def test_unsupported_grid_is_rejected(tc, shape):
    """x/y sizes must be even, in [4, 1536], with prime factors 2, 3, 5 only, and 4 <= nz <= 256:
    anything else raises (with the library's message) before any step runs."""
    print(f"[reject] shape={shape}", flush=True)
    with pytest.raises(RuntimeError, match="unsupported grid size"):
        tc.ChannelFlow(config=cfg(*shape))


def _pmod(p):
    p = torch.as_tensor(p).double().cpu()
    return p - p.mean()


@pytest.mark.parametrize("shape", [(4, 4, 4), (16, 12, 16), (32, 24, 48), (32, 24, 96), (24, 18, 128),
                                   (24, 18, 130), (16, 12, 256), (288, 6, 12)])
@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_steps_match_oracle(tc, shape, top):
    nx, ny, nz = shape
    print(f"[step] shape={shape} top={top}", flush=True)
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz, top=top))
    ref = O.Channel(nx, ny, nz, 4 * math.pi, 4 * math.pi / 3, seed=0, branch="gpu", gauge_zero_mode=True,
                    top_wall_bc_type=top)
    for q, r in ((ch.u, ref.u), (ch.v, ref.v), (ch.w, ref.w)):
        assert rel(q, r) < TOL
    f_used = 0.0
    for s in range(6):
        # the explicit terms this step evaluates, from the HIP path's own (bit-identical) state
        st = [q.detach().cpu().double().clone() for q in (ch.u, ch.v, ch.w)]
        O.apply_bc_all(*st, ref.top)
        want = O.rhs_fused_imex(*st, nx, ny, nz, ref.dx, ref.dy, ref.dz_c, ref.dz_f, ref.nu)
        ub, f = ch.step_imex(0.005)
        rub, rf = ref.step_imex(0.005)
        for q, r, name in ((ch.u, ref.u, "u"), (ch.v, ref.v, "v"), (ch.w, ref.w, "w")):
            assert rel(q, r) < TOL, (name, s)
        # p comes from per-mode wall-normal solves whose worst-conditioned (small |k|) systems
        # amplify round-off: the oracle's float64 Thomas sweep itself sits ~5e-13 from an
        # extended-precision solve, so p is held to 1e-10 while u, v, w (which only see grad p
        # times dt) stay at 1e-12
        assert rel(_pmod(ch.p), _pmod(ref.p)) < 1e-10, ("p", s)
        assert abs(float(ub) - float(rub)) < 1e-13
        # F += 0.1*(U_b - u_b)/dt cancels: a last-bit difference in u_b (summation order) moves F
        # by ~1e-14 absolute, so F is checked absolutely
        assert abs(float(f) - float(rf)) < 1e-12
        # the stored rhs against the oracle's explicit terms on the same input state: the
        # derivative operators amplify the (<= 1e-12) state differences to the oracle's own
        # trajectory by up to 1/dz_min, so the kernel itself is pinned on identical inputs
        assert rel(ch.rhs_u_prev - f_used, want[0]) < TOL, ("rhs_u", s)
        assert rel(ch.rhs_v_prev, want[1]) < TOL and rel(ch.rhs_w_prev, want[2]) < TOL, ("rhs_vw", s)
        f_used = float(f)
        div = tc.compute_divergence(ch.u, ch.v, ch.w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f)
        assert float(div.abs().max()) < 1e-10


def test_hybrid_grid_steps_match_oracle(tc):
    """stretching_type 'hybrid' (solver.py:192-211): nz comes from the two regions."""
    nx, ny, nzu, nzs = 16, 12, 10, 14
    c = cfg(nx, ny, 8)
    c["grid"].update({"nz_uniform": nzu, "nz_stretched": nzs})
    c["domain"].update({"stretching_type": "hybrid", "z_transition": 0.4, "gamma_stretched": 1.5})
    print("[step] hybrid grid 16x12x(10+14)", flush=True)
    ch = tc.ChannelFlow(config=c)
    assert ch.nz == nzu + nzs
    grid = O.generate_hybrid_grid(nzu, nzs, 0.4, 2.0, 1.5)
    ref = O.Channel(nx, ny, nzu + nzs, 4 * math.pi, 4 * math.pi / 3, seed=0, gauge_zero_mode=True, grid=grid)
    assert torch.equal(ch.dz_f.cpu(), ref.dz_f)
    for s in range(4):
        ch.step_imex(0.005)
        ref.step_imex(0.005)
        for q, r, name in ((ch.u, ref.u, "u"), (ch.v, ref.v, "v"), (ch.w, ref.w, "w")):
            assert rel(q, r) < TOL, (name, s)
        assert rel(_pmod(ch.p), _pmod(ref.p)) < 1e-10


def test_three_steps_match_reference_trajectory(tc, golden_dir):
    """The reference's own ChannelFlow (CPU branch) on a seeded 16x12x16 channel."""
    g = np.load(os.path.join(golden_dir, "step_trajectory.npz"))
    ch = tc.ChannelFlow(config=cfg(16, 12, 16, branch="cpu"))
    for s in range(4):
        if s > 0:
            ub, f = ch.step_imex(0.005)
            assert abs(float(ub) - float(g[f"ubulk{s}"])) < 1e-13
            assert abs(float(f) - float(g[f"forcing{s}"])) < 1e-10
            assert rel(ch.rhs_u_prev, g[f"rhs_u_prev{s}"]) < 1e-11
        for q, key in ((ch.u, "u"), (ch.v, "v"), (ch.w, "w")):
            assert rel(q, g[f"{key}{s}"]) < TOL, (key, s)
        assert rel(_pmod(ch.p), _pmod(g[f"p{s}"])) < 1e-9


def test_c1_200_steps_match_reference(tc, golden_dir):
    """The CPU-runnable case (64x64x48, dt = 0.005, 200 steps, CPU branch) against the reference's
    own run: bulk velocity every step, mean and RMS profiles at the end (<= 1e-9)."""
    g = np.load(os.path.join(golden_dir, "c1_run.npz"))
    nx, ny, nz = 64, 64, 48
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz, branch="cpu"))
    ubs = []
    for _ in range(200):
        ub, _ = ch.step_imex(0.005)
        ubs.append(ub)
    ubs = torch.stack(ubs).cpu().numpy()
    assert np.abs(ubs - g["u_bulk"]).max() < 1e-12
    u = ch.u[1:nx + 1, 1:ny + 1, 1:nz + 1].cpu()
    w = ch.w[1:nx + 1, 1:ny + 1, 1:nz].cpu()
    v = ch.v[1:nx + 1, 1:ny + 1, 1:nz + 1].cpu()
    assert rel(u.mean(dim=(0, 1)), g["U_mean"]) < 1e-9
    assert rel(u.std(dim=(0, 1), unbiased=False), g["u_rms"]) < 1e-9
    assert rel(v.std(dim=(0, 1), unbiased=False), g["v_rms"]) < 1e-9
    assert rel(w.std(dim=(0, 1), unbiased=False), g["w_rms"]) < 1e-9
    div = tc.compute_divergence(ch.u, ch.v, ch.w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f)
    assert float(div.abs().max()) < 1e-10


def test_graph_replay_equals_eager(tc):
    nx, ny, nz = 32, 24, 48
    a = tc.ChannelFlow(config=cfg(nx, ny, nz))
    b = tc.ChannelFlow(config=cfg(nx, ny, nz))
    for ch in (a, b):
        ch.step_imex(0.005)
    for _ in range(2):
        a.run_graph(1, 10, dt_update_interval=5)
        torch.cuda.synchronize()
        for i in range(10):
            if (i + 1) % 5 == 0:
                b.update_dt_device()
            b.step_imex(None)
        for q, r in ((a.u, b.u), (a.v, b.v), (a.w, b.w), (a.p, b.p), (a._scal, b._scal)):
            assert torch.equal(q, r)


def test_adaptive_dt_rule_matches_oracle(tc):
    nx, ny, nz = 32, 24, 48
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz, amp=0.5))
    ref = O.Channel(nx, ny, nz, 4 * math.pi, 4 * math.pi / 3, seed=0, perturbation=0.5, gauge_zero_mode=True)
    host = ch.compute_cfl_dt()
    expect = ref.cfl_dt(0.25)
    assert abs(host - expect) < 1e-13 * expect
    ch.update_dt_device()
    dev = ch.device_dt()
    want = expect if abs(expect - 0.005) / 0.005 > 0.05 else 0.005
    assert abs(dev - want) < 1e-13 * want


def test_fp32_bulk_and_wall_shear_track_fp64(tc):
    """FP32 stencils/FFT/PCR with float64 gauge mode, bulk reduction and forcing: u_b and u_tau
    within 1e-5 of the float64 oracle after 50 steps (64x64x48)."""
    nx, ny, nz = 64, 64, 48
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz, precision="fp32"))
    ref = O.Channel(nx, ny, nz, 4 * math.pi, 4 * math.pi / 3, seed=0, gauge_zero_mode=True)
    for _ in range(50):
        ub, _ = ch.step_imex(0.005)
        rub, _ = ref.step_imex(0.005)
    assert abs(float(ub) - float(rub)) < 1e-5
    ut = tc.compute_u_tau(ch.u, ch.z_c, ch.nu)
    assert abs(float(ut) - float(O.compute_u_tau(ref.u, ref.z_c, ref.nu))) < 1e-5
    assert rel(ch.u.double(), ref.u) < 1e-4
