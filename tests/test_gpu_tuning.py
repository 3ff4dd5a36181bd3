"""Launch-shape knobs and the autotuner: every setting gives bit-identical steps."""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def cfg(nx, ny, nz, precision="fp64", autotune=False):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": 0.005, "n_steps": 10},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": 0},
        "compute": {"device": "cuda", "precision": precision, "autotune": autotune},
        "output": {"save_init": False},
    }


@pytest.fixture(scope="module")
def tc():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import torchannel_amd
    torchannel_amd._lib.load()
    return torchannel_amd


def run(tc, knobs, precision, steps=4, shape=(32, 24, 48)):
    ch = tc.ChannelFlow(config=cfg(*shape, precision=precision))
    for k, v in knobs.items():
        ch._grid.set_tuning(k, v)
    for _ in range(steps):
        ch.step_imex(0.005)
    return [q.clone() for q in (ch.u, ch.v, ch.w, ch.p, ch._scal)]


# nz = 48: rows per lane padded (scalar row accesses); 64: 16 lanes x 4 rows and 256: 32 lanes x
# 8 rows, both as 16-byte row vectors (where the LDS-staged Poisson form applies); 17 x 24 = 408
# systems leave a partial last block at every systems-per-block setting
@pytest.mark.parametrize("shape", [(32, 24, 48), (32, 24, 64), (32, 24, 256)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_knobs_do_not_change_results(tc, precision, shape):
    from torchannel_amd import _lib as L
    ref = run(tc, {}, precision, shape=shape)
    for knobs in ({L.TUNE_K1_JB: 1}, {L.TUNE_K1_JB: 3}, {L.TUNE_K1_JB: 26},
                  {L.TUNE_THOMAS_SB: 4}, {L.TUNE_THOMAS_SB: 8}, {L.TUNE_K1_JB: 5, L.TUNE_THOMAS_SB: 8},
                  {L.TUNE_K1_TILE: 0}, {L.TUNE_K1_TILE: 1}, {L.TUNE_K1_TILE: 1, L.TUNE_K1_JB: 3},
                  {L.TUNE_K1_TILE: 2}, {L.TUNE_K1_TILE: 2, L.TUNE_K1_JB: 3},
                  {L.TUNE_THOMAS_LDS: 1}, {L.TUNE_THOMAS_LDS: 1, L.TUNE_THOMAS_SB: 4},
                  {L.TUNE_THOMAS_LDS: 1, L.TUNE_THOMAS_SB: 8}):
        print(f"[tuning] {precision} {shape} {knobs}", flush=True)
        got = run(tc, knobs, precision, shape=shape)
        for a, b in zip(got, ref):
            assert torch.equal(a, b), knobs


def test_autotune_keeps_state_and_results(tc, tmp_path):
    from torchannel_amd import _lib as L
    ref = run(tc, {}, "fp64", steps=3)
    ch = tc.ChannelFlow(config=cfg(32, 24, 48, autotune=True))
    for _ in range(3):
        ch.step_imex(0.005)
    assert ch.tuning is not None and L.TUNE_K1_JB in ch.tuning
    for a, b in zip([ch.u, ch.v, ch.w, ch.p, ch._scal], ref):
        assert torch.equal(a, b)
    path = str(tmp_path / "tune.json")
    chosen = tc.tune(ch, path=path, force=True)
    again = tc.tune(ch, path=path)
    assert chosen == again


def test_bad_knob_rejected(tc):
    from torchannel_amd import _lib as L
    ch = tc.ChannelFlow(config=cfg(16, 12, 16))
    with pytest.raises(RuntimeError):
        ch._grid.set_tuning(L.TUNE_K1_JB, 10 ** 6)
    with pytest.raises(RuntimeError):
        ch._grid.set_tuning(L.TUNE_THOMAS_SB, 17)
    with pytest.raises(RuntimeError):
        ch._grid.set_tuning(L.TUNE_THOMAS_LDS, 2)
    with pytest.raises(RuntimeError):
        ch._grid.set_tuning(L.TUNE_K1_TILE, 3)
    with pytest.raises(RuntimeError):
        ch._grid.set_tuning(99, 1)


def test_default_pencil_length_follows_ny(tc):
    """Untuned K1 pencils: about ny/24 j-columns in steps of 4, within [8, 16]."""
    from torchannel_amd import _lib as L
    dzf = torch.full((8,), 0.25, dtype=torch.float64, device="cuda")
    dzc = torch.full((9,), 0.25, dtype=torch.float64, device="cuda")
    for ny, jb in ((12, 8), (192, 8), (288, 12), (384, 16), (768, 16)):
        g = L.get_grid(4, ny, 8, 0.1, 0.1, 1e-3, L.TC_TOP_DIRICHLET, torch.float64, "cuda", dz_f=dzf, dz_c=dzc)
        assert g.get_tuning(L.TUNE_K1_JB) == jb


def test_default_k1_form(tc):
    """Untuned K1 form: the three-component pencils where a lane holds two or more rows and the
    form is built for the row count (float64 up to 128 rows, float32 up to 256), with their own
    default pencil length; elsewhere the per-component pencils."""
    from torchannel_amd import _lib as L
    for dt, nz, ny, form, jb in ((torch.float64, 48, 192, 0, 8), (torch.float64, 128, 192, 2, 8),
                                 (torch.float64, 256, 384, 0, 16), (torch.float32, 128, 192, 2, 4),
                                 (torch.float32, 256, 384, 2, 6), (torch.float32, 64, 192, 0, 8)):
        dzf = torch.full((nz,), 2.0 / nz, dtype=torch.float64, device="cuda")
        dzc = torch.full((nz + 1,), 2.0 / nz, dtype=torch.float64, device="cuda")
        g = L.get_grid(4, ny, nz, 0.1, 0.1, 1e-3, L.TC_TOP_DIRICHLET, dt, "cuda", dz_f=dzf, dz_c=dzc)
        assert g.get_tuning(L.TUNE_K1_TILE) == form, (dt, nz)
        assert g.get_tuning(L.TUNE_K1_JB) == jb, (dt, nz)
