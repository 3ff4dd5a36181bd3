"""Discrete adjoint of the fixed-dt step (float64): dJ/df of the time-averaged wall shear stress
with respect to a body force, against autograd through the oracle and central differences."""
import math

import pytest
import torch

from oracle import channel_oracle as O

pytestmark = pytest.mark.gpu


def cfg(nx, ny, nz, top="dirichlet"):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": 0.005, "n_steps": 10},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": 0},
        "compute": {"device": "cuda", "precision": "fp64", "rhs_branch": "gpu"},
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


def body_force(shape, seed, sigma=1e-3):
    g = torch.Generator().manual_seed(seed)
    return (sigma * torch.randn(shape, generator=g, dtype=torch.float64))


def interior_direction(shape, seed):
    d = body_force(shape, seed, 1.0)
    mask = torch.zeros(shape, dtype=torch.float64)
    mask[1:, 1:-1, 1:-1] = 1.0          # u faces 1..nx, rows 1..ny, cells 1..nz
    return d * mask


def oracle_objective_and_grad(nx, ny, nz, f, n_steps, dt, top):
    ref = O.Channel(nx, ny, nz, 4 * math.pi, 4 * math.pi / 3, seed=0, gauge_zero_mode=True, top_wall_bc_type=top)
    fr = f.clone().requires_grad_(True)
    J = 0.0
    coef = (0.5 if top == "dirichlet" else 1.0) * ref.nu / ref.z_c[1]
    for _ in range(n_steps):
        ref.step_imex(dt, body_force=fr)
        tw = ref.u[:, :, 1].mean() + (ref.u[:, :, -2].mean() if top == "dirichlet" else 0.0)
        J = J + coef * tw / n_steps
    J.backward()
    return float(J), fr.grad


@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_adjoint_matches_autograd_through_oracle(tc, top):
    nx, ny, nz, n, dt = 16, 12, 16, 6, 0.005
    print(f"[adjoint] oracle autograd {nx}x{ny}x{nz} top={top}", flush=True)
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz, top))
    f = body_force(tuple(ch.u.shape), 0)
    run = tc.WallShearAdjoint(ch, n, dt, checkpoint_every=4)
    J = float(run.objective(f.cuda()))
    g = run.gradient(f.cuda()).cpu()
    Jr, gr = oracle_objective_and_grad(nx, ny, nz, f, n, dt, top)
    assert abs(J - Jr) <= 1e-12 * abs(Jr)
    err = float((g - gr).abs().max()) / float(gr.abs().max())
    print(f"[adjoint] J rel diff {abs(J - Jr) / abs(Jr):.2e}, grad rel diff {err:.2e}", flush=True)
    assert err < 1e-9, err
    assert float(g[0].abs().max()) == 0.0 and float(g[:, 0].abs().max()) == 0.0


def test_adjoint_central_differences_c5(tc):
    """96x96x64, 100 fixed-dt steps, checkpoints every 10, 16 seeded directions, step 1e-6."""
    nx, ny, nz, n, dt, eps = 96, 96, 64, 100, 0.005, 1e-6
    print(f"[adjoint] central differences {nx}x{ny}x{nz}, {n} steps", flush=True)
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz))
    f = body_force(tuple(ch.u.shape), 0).cuda()
    run = tc.WallShearAdjoint(ch, n, dt, checkpoint_every=10)
    g = run.gradient(f)
    worst = 0.0
    for seed in range(1, 17):
        d = interior_direction(tuple(ch.u.shape), seed).cuda()
        jp = float(run.objective(f + eps * d))
        jm = float(run.objective(f - eps * d))
        fd = (jp - jm) / (2 * eps)
        ad = float((g * d).sum())
        worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"[adjoint] worst relative FD mismatch {worst:.3e}", flush=True)
    assert worst < 1e-5


def test_autograd_function_fills_grad(tc):
    nx, ny, nz = 16, 12, 16
    ch = tc.ChannelFlow(config=cfg(nx, ny, nz))
    run = tc.WallShearAdjoint(ch, 4, 0.005, checkpoint_every=3)
    f = body_force(tuple(ch.u.shape), 3).cuda().requires_grad_(True)
    J = tc.wall_shear_objective(run, f)
    (2.0 * J).backward()
    g = run.gradient(f.detach())
    assert torch.equal(f.grad, 2.0 * g)


def test_adjoint_rejects_unsupported(tc):
    c = cfg(16, 12, 16)
    c["compute"]["precision"] = "fp32"
    ch = tc.ChannelFlow(config=c)
    with pytest.raises(ValueError):
        tc.WallShearAdjoint(ch, 4, 0.005)
