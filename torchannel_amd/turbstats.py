"""On-the-fly turbulence statistics (turbstats.py:20-303) with the reductions and plane FFTs on
the HIP library; running sums are float64 on the device."""
import os

import numpy as np
import torch

from . import _lib as L


class TurbulenceStats:
    def __init__(self, nx, ny, nz, Lx, Ly, Lz, z_c, z_f, dz_c, dz_f, dx, dy, nu, Re_tau_target,
                 z_plus_target=15.0, device="cpu"):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.z_c, self.z_f, self.dz_c, self.dz_f = z_c, z_f, dz_c, dz_f
        self.dx, self.dy, self.nu = dx, dy, nu
        self.device = torch.device(device)
        self.u_tau_target = Re_tau_target * nu / (Lz / 2.0)                  # turbstats.py:52-53
        zt = z_plus_target * nu / self.u_tau_target
        zin = z_c[1:nz + 1].detach().to("cpu", torch.float64)
        self.k_bot = int(torch.argmin(torch.abs(zin - zt)).item()) + 1        # turbstats.py:63-69
        self.k_top = int(torch.argmin(torch.abs(zin - (Lz - zt))).item()) + 1
        self.n_samples = 0
        f64 = torch.float64
        self._prof = torch.zeros(5, nz, dtype=f64, device=self.device)
        self._spec = torch.zeros(4, nx // 2, ny // 2, dtype=f64, device=self.device)
        self.U_sum, self.uu_sum, self.vv_sum, self.ww_sum, self.uw_sum = self._prof.unbind(0)
        self.E_uu_2d_sum, self.E_vv_2d_sum, self.E_ww_2d_sum, self.E_uw_2d_sum = self._spec.unbind(0)
        self.kx = 2 * np.pi * np.fft.rfftfreq(nx, d=dx)[1:]
        self.ky = 2 * np.pi * np.fft.rfftfreq(ny, d=dy)[1:]
        self._grids = {}

    def use_grid(self, grid, dtype, device):
        """Take the plane transforms of an existing grid of the same nx, ny, dx, dy (the solver's),
        so that the first sample builds no plan of its own inside the time loop."""
        self._grids[(dtype, L.device_key(device), grid.pitch)] = grid

    def _plan(self, dt, device, pitch=0):
        key = (dt, L.device_key(device), pitch)
        if key not in self._grids:
            g = L.get_grid(self.nx, self.ny, self.nz, self.dx, self.dy, self.nu, L.TC_TOP_DIRICHLET, dt,
                           device, dz_f=self.dz_f, dz_c=self.dz_c, pitch=pitch)
            self._grids[key] = g
        return self._grids[key].poisson()

    def accumulate_statistics(self, u, v, w, u_tau_current=None):
        """turbstats.py:101-228: one sample into the running sums."""
        dt, pitch = L.require_fields(u, v, w)
        L.call("tc_stats_accumulate", self._plan(dt, u.device, pitch), L.ptr(u), L.ptr(v), L.ptr(w),
               self.k_bot, self.k_top, L.ptr(self._prof), L.ptr(self._spec), L.stream_of(u))
        self.n_samples += 1

    def finalize_statistics(self):
        """turbstats.py:230-303."""
        if self.n_samples == 0:
            raise RuntimeError("No samples accumulated - cannot finalize statistics")
        n = self.n_samples
        prof = (self._prof / n).cpu().numpy()
        spec = (self._spec / n).cpu().numpy()
        dz = float(self.z_c[1])
        dudz = float(0.5 * (prof[0][0] + prof[0][-1]) / dz)
        return {
            "n_samples": n,
            "z_c": self.z_c[1:self.nz + 1].detach().cpu().numpy(),
            "U_mean": prof[0], "uu_mean": prof[1], "vv_mean": prof[2], "ww_mean": prof[3], "uw_mean": prof[4],
            "kx": self.kx[: self.nx // 2], "ky": self.ky[: self.ny // 2],
            "E_uu_2d": spec[0], "E_vv_2d": spec[1], "E_ww_2d": spec[2], "E_uw_2d": spec[3],
            "z_plus_target": self.k_bot,
            "Re_tau_target": self.u_tau_target * (self.Lz / 2) / self.nu,
            "nu": self.nu, "Re": 1.0 / self.nu,
            "u_tau": float(np.sqrt(self.nu * dudz)) if self.nu * dudz >= 0 else float("nan"),
        }

    def save_statistics(self, filepath):
        """turbstats.py:305-314."""
        np.savez_compressed(filepath, **self.finalize_statistics())

    @staticmethod
    def load_statistics(filepath):
        """turbstats.py:316-332."""
        if not os.path.exists(filepath):
            raise FileNotFoundError(f"Statistics file not found: {filepath}")
        data = np.load(filepath, allow_pickle=False)
        return {k: data[k] for k in data.files}

    def save_state(self, filepath):
        """turbstats.py:334-364: running sums, n_samples and the grid for validation."""
        host = lambda t: np.asarray(t.detach().cpu().numpy())
        np.savez_compressed(filepath, n_samples=self.n_samples,
                            U_sum=host(self.U_sum), uu_sum=host(self.uu_sum), vv_sum=host(self.vv_sum),
                            ww_sum=host(self.ww_sum), uw_sum=host(self.uw_sum),
                            E_uu_2d_sum=host(self.E_uu_2d_sum), E_vv_2d_sum=host(self.E_vv_2d_sum),
                            E_ww_2d_sum=host(self.E_ww_2d_sum), E_uw_2d_sum=host(self.E_uw_2d_sum),
                            nx=self.nx, ny=self.ny, nz=self.nz)

    def load_state(self, filepath):
        """turbstats.py:366-402: restore the sums (grid must match)."""
        if not os.path.exists(filepath):
            raise FileNotFoundError(f"Statistics state file not found: {filepath}")
        data = np.load(filepath, allow_pickle=False)
        if data["nx"] != self.nx or data["ny"] != self.ny or data["nz"] != self.nz:
            raise ValueError(
                f"Grid dimension mismatch! State file has (nx={data['nx']}, ny={data['ny']}, nz={data['nz']}) "
                f"but current simulation has (nx={self.nx}, ny={self.ny}, nz={self.nz})")
        self.n_samples = int(data["n_samples"])
        for dst, key in ((self.U_sum, "U_sum"), (self.uu_sum, "uu_sum"), (self.vv_sum, "vv_sum"),
                         (self.ww_sum, "ww_sum"), (self.uw_sum, "uw_sum"), (self.E_uu_2d_sum, "E_uu_2d_sum"),
                         (self.E_vv_2d_sum, "E_vv_2d_sum"), (self.E_ww_2d_sum, "E_ww_2d_sum"),
                         (self.E_uw_2d_sum, "E_uw_2d_sum")):
            dst.copy_(torch.as_tensor(data[key], dtype=torch.float64))
