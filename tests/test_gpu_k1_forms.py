"""The forms of the explicit + AB2 + CN kernel give bit-identical results: the tiled form (three
components of a band per block, neighbour columns staged through LDS) and the three-component
pencils (one wavefront marching u, v and w of one i) against the per-component pencils, over
whole steps (forward-Euler start, AB2 steps, body force, both rhs branches, both top walls, both
precisions) and for the explicit terms alone."""
import ctypes
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def cfg(nx, ny, nz, precision, branch, top):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": 0.005, "n_steps": 10},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": 3},
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


SHAPES = [(4, 4, 4), (16, 12, 16), (18, 10, 48), (32, 24, 96), (24, 18, 128), (24, 18, 130), (16, 12, 256),
          (288, 6, 12), (10, 30, 64)]


def _state(ch):
    return [ch.u, ch.v, ch.w, ch.p, ch._scal] + [r for s in ch._rhs for r in s]


# (form, pencil length or 0 for the default): the three-component pencils march only interior
# bands (1 <= j0, j0 + jb - 1 <= ny - 1), so a short pencil gives them more of each grid
@pytest.mark.parametrize("form", [(1, 0), (2, 0), (2, 5)])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("branch,top", [("gpu", "dirichlet"), ("cpu", "neumann"), ("gpu", "neumann")])
def test_forms_steps_equal_pencils(tc, shape, precision, branch, top, form):
    from torchannel_amd import _lib as L
    print(f"[k1 forms] form {form} {shape} {precision} {branch} {top}", flush=True)
    form, jb = form
    a = tc.ChannelFlow(config=cfg(*shape, precision, branch, top))
    b = tc.ChannelFlow(config=cfg(*shape, precision, branch, top))
    a._grid.set_tuning(L.TUNE_K1_TILE, 0)
    b._grid.set_tuning(L.TUNE_K1_TILE, form)
    if jb:
        b._grid.set_tuning(L.TUNE_K1_JB, jb)
    assert b._grid.get_tuning(L.TUNE_K1_TILE) == form and a._grid.get_tuning(L.TUNE_K1_TILE) == 0
    g = torch.Generator().manual_seed(5)
    f = (1e-3 * torch.randn(tuple(a.u.shape), generator=g, dtype=torch.float64)).to(a.u.dtype).cuda()
    for s in range(5):
        body = f if s == 2 else None
        a.body_force = body
        b.body_force = body
        a.step_imex(0.005)
        b.step_imex(0.005)
        for x, y in zip(_state(a), _state(b)):
            assert torch.equal(x, y), s


@pytest.mark.parametrize("shape", [(16, 12, 16), (24, 18, 128), (16, 12, 256)])
@pytest.mark.parametrize("precision", [torch.float64, torch.float32])
@pytest.mark.parametrize("branch", [0, 1])
def test_forms_rhs_equal_pencils(tc, shape, precision, branch):
    """The rhs-only launch (compute_momentum_rhs_fused_imex and the CPU-branch terms)."""
    from torchannel_amd import _lib as L
    from torchannel_amd.utils import generate_grid
    nx, ny, nz = shape
    _, _, dz_f, dz_c = generate_grid(2.6, nz, 2.0)
    gen = torch.Generator().manual_seed(7)
    u = (torch.rand(nx + 1, ny + 2, nz + 2, generator=gen, dtype=torch.float64) - 0.5).to(precision).cuda()
    v = (torch.rand(nx + 2, ny + 1, nz + 2, generator=gen, dtype=torch.float64) - 0.5).to(precision).cuda()
    w = (torch.rand(nx + 2, ny + 2, nz + 1, generator=gen, dtype=torch.float64) - 0.5).to(precision).cuda()
    outs = []
    for tile in (0, 1, 2):
        grid = L.Grid(nx, ny, nz, 0.3, 0.2, 1 / 2870.0, 0, precision, dz_f.tolist(), dz_c.tolist(), u.device)
        grid.set_tuning(L.TUNE_K1_TILE, tile)
        r = [torch.full_like(q, float("nan")) for q in (u, v, w)]
        L.call("tc_rhs_imex", grid.handle, branch, L.ptr(u), L.ptr(v), L.ptr(w), L.ptr(r[0]), L.ptr(r[1]),
               L.ptr(r[2]), L.stream_of(u))
        outs.append(r)
    torch.cuda.synchronize()
    for other in outs[1:]:
        for x, y in zip(outs[0], other):
            assert torch.equal(x, y)
