"""TurbulenceStats' checkpoint (turbstats.py:334-402) on the host: save_state / load_state round
trip and the grid check.  No kernel runs, so this is a CPU test."""
import math

import numpy as np
import pytest
import torch

from torchannel_amd.turbstats import TurbulenceStats
from torchannel_amd.utils import generate_grid


def _stats(nx, ny, nz):
    z_f, z_c, dz_f, dz_c = generate_grid(2.6, nz, 2.0)
    return TurbulenceStats(nx, ny, nz, 4 * math.pi, 4 * math.pi / 3, 2.0, z_c, z_f, dz_c, dz_f,
                           4 * math.pi / nx, 4 * math.pi / 3 / ny, 1 / 2870.0, 180.0, device="cpu")


def test_state_round_trip(tmp_path):
    a = _stats(16, 12, 16)
    g = torch.Generator().manual_seed(3)
    a._prof.copy_(torch.rand(a._prof.shape, generator=g, dtype=torch.float64))
    a._spec.copy_(torch.rand(a._spec.shape, generator=g, dtype=torch.float64))
    a.n_samples = 7
    path = str(tmp_path / "state.npz")
    a.save_state(path)
    d = np.load(path)
    assert sorted(d.files) == sorted(["n_samples", "U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum", "E_uu_2d_sum",
                                      "E_vv_2d_sum", "E_ww_2d_sum", "E_uw_2d_sum", "nx", "ny", "nz"])
    b = _stats(16, 12, 16)
    b.load_state(path)
    assert b.n_samples == 7
    assert torch.equal(b._prof, a._prof) and torch.equal(b._spec, a._spec)
    fin = b.finalize_statistics()
    assert np.array_equal(fin["U_mean"], (a._prof[0] / 7).numpy())


def test_state_grid_mismatch_raises(tmp_path):
    a = _stats(16, 12, 16)
    a.n_samples = 1
    path = str(tmp_path / "state.npz")
    a.save_state(path)
    with pytest.raises(ValueError, match="Grid dimension mismatch"):
        _stats(16, 12, 18).load_state(path)
    with pytest.raises(FileNotFoundError):
        _stats(16, 12, 16).load_state(str(tmp_path / "missing.npz"))
