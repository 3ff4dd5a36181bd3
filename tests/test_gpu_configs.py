"""Parity at the production sizes: the 192x192x128 channel (float64 against the oracle, float32
against the float64 oracle and, over a longer run, against the oracle-pinned float64 path), the
384x384x256 channel, and the 48-channel ensemble.

These are the launch shapes the benchmark times (two rows per lane at nz = 128, four at
nz = 256), so every kernel instance of the production runs is compared with the oracle here.
The float32 errors printed by these tests are the ones quoted for the float32 form.
"""
import copy
import math

import pytest
import torch

from oracle import channel_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = torch.as_tensor(a).detach().cpu().double()
    b = torch.as_tensor(b).detach().cpu().double()
    den = float(b.abs().max())
    return float((a - b).abs().max()) / (den if den > 0 else 1.0)


def pmod(p):
    p = torch.as_tensor(p).detach().cpu().double()
    return p - p.mean()


def cfg(nx, ny, nz, Lx, Ly, re, precision="fp64", dt=0.005, seed=0, n_stats=0):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": Lx, "Ly": Ly, "Lz": 2.0},
        "flow": {"Re": re, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": dt, "n_steps": 10, "CFL_target": 0.25},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": seed},
        "compute": {"device": "cuda", "precision": precision},
        "output": {"save_init": False},
        "statistics": {"n_stats": n_stats, "t_stats": 0.0, "z_plus_target": 15.0},
    }


@pytest.fixture(scope="module")
def tc():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import torchannel_amd
    torchannel_amd._lib.load()
    return torchannel_amd


def oracle_run(nx, ny, nz, Lx, Ly, re, dt, steps):
    """The oracle's states after 0..steps steps (float64, CPU), with u_b and F per step."""
    ref = O.Channel(nx, ny, nz, Lx, Ly, Re=re, seed=0, gauge_zero_mode=True)
    out = [dict(u=ref.u.clone(), v=ref.v.clone(), w=ref.w.clone(), p=ref.p.clone(), ub=None, f=None,
                utau=float(O.compute_u_tau(ref.u, ref.z_c, ref.nu)))]
    for s in range(steps):
        ub, f = ref.step_imex(dt)
        out.append(dict(u=ref.u.clone(), v=ref.v.clone(), w=ref.w.clone(), p=ref.p.clone(), ub=float(ub),
                        f=float(f), utau=float(O.compute_u_tau(ref.u, ref.z_c, ref.nu))))
        print(f"[oracle] {nx}x{ny}x{nz} step {s + 1}", flush=True)
    return ref, out


C2 = (192, 192, 128, 4 * math.pi, 4 * math.pi / 3, 2870.0)
C3 = (384, 384, 256, 2 * math.pi, math.pi, 10000.0)


@pytest.fixture(scope="module")
def c2_oracle():
    return oracle_run(*C2, dt=0.005, steps=3)


@pytest.mark.timeout(900)
def test_c2_fp64_steps_match_oracle(tc, c2_oracle):
    """192x192x128 float64, 3 steps: u, v, w <= 1e-12, p <= 1e-10 modulo its mean, u_b <= 1e-13,
    F <= 1e-12, div <= 1e-10; the stored rhs of the first step against the oracle's explicit
    terms on the same state."""
    nx, ny, nz = C2[:3]
    ref, traj = c2_oracle
    print("[c2] fp64", flush=True)
    ch = tc.ChannelFlow(config=cfg(*C2))
    for q, k in ((ch.u, "u"), (ch.v, "v"), (ch.w, "w")):
        assert rel(q, traj[0][k]) < 1e-12, k
    st = [q.detach().cpu().double().clone() for q in (ch.u, ch.v, ch.w)]
    O.apply_bc_all(*st)
    want = O.rhs_fused_imex(*st, nx, ny, nz, ref.dx, ref.dy, ref.dz_c, ref.dz_f, ref.nu)
    for s in range(1, 4):
        ub, f = ch.step_imex(0.005)
        for q, k in ((ch.u, "u"), (ch.v, "v"), (ch.w, "w")):
            assert rel(q, traj[s][k]) < 1e-12, (k, s)
        assert rel(pmod(ch.p), pmod(traj[s]["p"])) < 1e-10, ("p", s)
        assert abs(float(ub) - traj[s]["ub"]) < 1e-13
        assert abs(float(f) - traj[s]["f"]) < 1e-12
        if s == 1:
            for a, b, k in zip((ch.rhs_u_prev, ch.rhs_v_prev, ch.rhs_w_prev), want, "uvw"):
                assert rel(a, b) < 1e-12, ("rhs", k)
        div = tc.compute_divergence(ch.u, ch.v, ch.w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f)
        assert float(div.abs().max()) < 1e-10


@pytest.mark.timeout(900)
def test_c2_fp32_tracks_fp64_oracle(tc, c2_oracle):
    """192x192x128 float32 against the float64 oracle over 3 steps: u_b and u_tau within 1e-5,
    every velocity field within 1e-5 of its maximum, p within 1e-4 modulo its mean."""
    ref, traj = c2_oracle
    print("[c2] fp32", flush=True)
    ch = tc.ChannelFlow(config=cfg(*C2, precision="fp32"))
    for s in range(1, 4):
        ub, _ = ch.step_imex(0.005)
        ut = float(tc.compute_u_tau(ch.u, ch.z_c, ch.nu))
        e = {k: rel(getattr(ch, k).double(), traj[s][k]) for k in ("u", "v", "w")}
        e["p"] = rel(pmod(ch.p), pmod(traj[s]["p"]))
        e["u_b"] = abs(float(ub) - traj[s]["ub"])
        e["u_tau"] = abs(ut - traj[s]["utau"])
        print(f"[c2 fp32] step {s}: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()), flush=True)
        assert e["u_b"] < 1e-5 and e["u_tau"] < 1e-5
        assert max(e["u"], e["v"], e["w"]) < 1e-5
        assert e["p"] < 1e-4


@pytest.mark.timeout(900)
def test_c2_fp32_long_run_against_fp64(tc):
    """200 steps at 192x192x128, statistics every 20 steps: float32 against the float64 path
    (itself pinned to the oracle above).  u_b and u_tau stay within 1e-5 at every sample; the
    mean profile within 1e-5 and the Reynolds-stress profiles within 1e-4 of their maxima."""
    print("[c2] fp32 vs fp64, 200 steps", flush=True)
    a = tc.ChannelFlow(config=cfg(*C2, n_stats=20))
    b = tc.ChannelFlow(config=cfg(*C2, precision="fp32", n_stats=20))
    worst = {"u_b": 0.0, "u_tau": 0.0}
    for s in range(1, 201):
        ua, _ = a.step_imex(0.005)
        ubb, _ = b.step_imex(0.005)
        if s % 20 == 0:
            for ch in (a, b):
                ch.turbulence_stats.accumulate_statistics(ch.u, ch.v, ch.w)
            ta = float(tc.compute_u_tau(a.u, a.z_c, a.nu))
            tb = float(tc.compute_u_tau(b.u, b.z_c, b.nu))
            worst["u_b"] = max(worst["u_b"], abs(float(ua) - float(ubb)))
            worst["u_tau"] = max(worst["u_tau"], abs(ta - tb))
    fa = a.turbulence_stats.finalize_statistics()
    fb = b.turbulence_stats.finalize_statistics()
    e = dict(worst)
    for k in ("u", "v", "w"):
        e[k] = rel(getattr(b, k).double(), getattr(a, k))
    e["p"] = rel(pmod(b.p), pmod(a.p))
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean", "E_uu_2d", "E_ww_2d"):
        e[k] = rel(torch.as_tensor(fb[k]), torch.as_tensor(fa[k]))
    print("[c2 fp32 200 steps] " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()), flush=True)
    assert e["u_b"] < 1e-5 and e["u_tau"] < 1e-5
    assert e["U_mean"] < 1e-5
    assert max(e["uu_mean"], e["vv_mean"], e["ww_mean"], e["uw_mean"]) < 1e-4


@pytest.mark.timeout(1200)
def test_c3_steps_match_oracle(tc):
    """384x384x256 (Lx = 2 pi, Ly = pi, Re_b = 10000), 2 steps at dt = 0.0025: float64 against the
    oracle at the same bars as the small grids; float32 against the float64 oracle with u_b and
    u_tau within 1e-5 and the velocity fields within 1e-5 of their maxima."""
    nx, ny, nz = C3[:3]
    ref, traj = oracle_run(*C3, dt=0.0025, steps=2)
    print("[c3] fp64", flush=True)
    ch = tc.ChannelFlow(config=cfg(*C3, dt=0.0025))
    for s in range(1, 3):
        ub, f = ch.step_imex(0.0025)
        for q, k in ((ch.u, "u"), (ch.v, "v"), (ch.w, "w")):
            assert rel(q, traj[s][k]) < 1e-12, (k, s)
        assert rel(pmod(ch.p), pmod(traj[s]["p"])) < 1e-10, ("p", s)
        assert abs(float(ub) - traj[s]["ub"]) < 1e-13
        div = tc.compute_divergence(ch.u, ch.v, ch.w, nx, ny, nz, ch.dx, ch.dy, ch.dz_f)
        assert float(div.abs().max()) < 1e-10
    del ch
    torch.cuda.empty_cache()
    print("[c3] fp32", flush=True)
    ch = tc.ChannelFlow(config=cfg(*C3, precision="fp32", dt=0.0025))
    for s in range(1, 3):
        ub, _ = ch.step_imex(0.0025)
        ut = float(tc.compute_u_tau(ch.u, ch.z_c, ch.nu))
        e = {k: rel(getattr(ch, k).double(), traj[s][k]) for k in ("u", "v", "w")}
        e["p"] = rel(pmod(ch.p), pmod(traj[s]["p"]))
        e["u_b"] = abs(float(ub) - traj[s]["ub"])
        e["u_tau"] = abs(ut - traj[s]["utau"])
        print(f"[c3 fp32] step {s}: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()), flush=True)
        assert e["u_b"] < 1e-5 and e["u_tau"] < 1e-5
        assert max(e["u"], e["v"], e["w"]) < 1e-5


def _state(ch):
    return [ch.u, ch.v, ch.w, ch.p, ch._scal] + [r for s in ch._rhs for r in s]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("form", ["phased", pytest.param("persistent", marks=pytest.mark.skip(
    reason="the persistent ensemble form is switched off: its first GPU run made no progress for minutes and "
           "the cause is not established"))])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_ensemble_48_channels(tc, precision, form):
    """The 48-channel workload (sizes, Re_b and gamma drawn with default_rng(2026)), in both
    ensemble forms: each channel after one eager, three batched and eight graph-replayed steps is
    bit-identical to stepping it alone; in float64 the first two channels also match the oracle
    on their own grids."""
    cfgs = tc.ensemble_configs(48, precision=precision)
    print(f"[ensemble48] {precision} {form}", flush=True)
    ens = tc.make_ensemble(cfgs, form=form)
    singles = [tc.ChannelFlow(config=copy.deepcopy(c)) for c in cfgs]
    refs = []
    if precision == "fp64":
        for c in cfgs[:2]:
            g = c["grid"]
            refs.append(O.Channel(g["nx"], g["ny"], g["nz"], c["domain"]["Lx"], c["domain"]["Ly"],
                                  gamma=c["flow"]["gamma"], Re=c["flow"]["Re"],
                                  seed=c["initialization"]["seed"], gauge_zero_mode=True))
    ens.step(0.005)
    for s in singles:
        s.step_imex(0.005)
    for _ in range(3):
        ens.step()
        for s in singles:
            s.step_imex(None)
    ens.run_graph(2, 4)
    ens.check()
    for _ in range(8):
        for s in singles:
            s.step_imex(None)
    torch.cuda.synchronize()
    for q, (a, b) in enumerate(zip(ens.channels, singles)):
        for x, y in zip(_state(a), _state(b)):
            assert torch.equal(x, y), q
    for q, r in enumerate(refs):
        for _ in range(12):
            r.step_imex(0.005)
        ch = ens.channels[q]
        for x, y, k in ((ch.u, r.u, "u"), (ch.v, r.v, "v"), (ch.w, r.w, "w")):
            assert rel(x, y) < 1e-12, (q, k)
        assert rel(pmod(ch.p), pmod(r.p)) < 1e-10, (q, "p")
