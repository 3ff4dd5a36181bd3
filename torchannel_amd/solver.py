"""ChannelFlow and apply_bc_all (solver.py) on the HIP library.

ChannelFlow keeps the reference's configuration keys, attributes and step semantics; every step
runs the fused HIP path:

  k_cn_tables  -> CN/PCR coefficient tables for the current dt (device scalar)
  k_explicit_cn-> explicit RHS + forcing, AB2, Crank-Nicolson z-diffusion      (solver.py:588-668)
  k_fwd_x, k_fft_y, k_poisson_thomas, k_fft_y, k_inv_x -> pressure Poisson solve (solver.py:672-682)
  k_project_bc -> projection, boundary images, bulk velocity, forcing update     (solver.py:692-708)

dt, the forcing and the bulk velocity live in a float64 device scalar block, so steps need no
host synchronisation and can be captured in a HIP graph (run_graph).
"""
import ctypes
import math
import os
import time as _time

import numpy as np
import torch
import yaml

from . import _lib as L
from .initflow import initialize_flow, initialize_flow_from_file
from .projection_fft import initialize_fft_solver
from .turbstats import TurbulenceStats
from .utils import (append_timeseries, compute_bulk_velocity, compute_divergence, compute_u_tau, generate_grid,
                    generate_hybrid_grid, save_flow_fields)


def apply_bc_all(u, v, w, top_wall_bc_type="dirichlet"):
    """solver.py:23-79, in place."""
    dt, pitch = L.require_fields(u, v, w)
    nx, ny, nz = u.shape[0] - 1, u.shape[1] - 2, u.shape[2] - 2
    g = L.get_grid(nx, ny, nz, 1.0, 1.0, 0.0, L.top_code(top_wall_bc_type), dt, u.device, pitch=pitch)
    L.call("tc_apply_bc", g.handle, L.ptr(u), L.ptr(v), L.ptr(w), L.stream_of(u))


class ChannelFlow:
    """Drop-in for solver.py:81-970's IMEX path.

    Configuration: the reference's YAML keys (grid, domain, flow, initialization, time, compute,
    output, statistics, boundary_conditions) plus ``compute.precision`` ('fp64' | 'fp32') and
    ``initialization.seed`` (torch.manual_seed on the CPU generator before the noise draws) and
    ``compute.rhs_branch`` ('gpu' = the fused RHS the reference runs on any HIP/CUDA device,
    'cpu' = the separate kernels of its CPU branch, solver.py:593-622), ``compute.autotune``
    (pick the kernels' launch shapes by timing after the first step; results do not change) and
    ``compute.private_plan`` (a Poisson workspace of its own, for channels stepped together) and
    ``output.save_init`` (default true: write fields_init.npz at construction, as the reference
    does; callers that only step, such as benchmarks, turn it off) and ``compute.layout``
    ('dense', the default: contiguous tensors as the reference allocates them; 'padded': u, v,
    w, p, the AB2 history and the body force are strided views, indexed exactly like the dense
    arrays, whose wall-normal columns sit a whole number of 128-byte lines apart with row k = 1
    on a line; results are bit-identical, the explicit + CN kernel's loads stream faster, the
    arrays take 11-23 % more memory).
    ``config`` may be given as a dict instead of a file path.
    """

    def __init__(self, config_file="config.yaml", config=None, verbose=False):
        if config is None:
            with open(config_file, "r") as f:
                config = yaml.safe_load(f)
        self.config = config
        self.verbose = verbose
        dev = config.get("compute", {}).get("device", "auto")
        if dev == "cpu" or not torch.cuda.is_available():
            raise RuntimeError("torchannel_amd.ChannelFlow runs on a HIP device only")
        self.device = torch.device("cuda")
        prec = config.get("compute", {}).get("precision", "fp64")
        if prec not in ("fp64", "fp32"):
            raise ValueError(f"Invalid precision: {prec}")
        self.precision = prec
        self.autotune = bool(config.get("compute", {}).get("autotune", False))
        branch = config.get("compute", {}).get("rhs_branch", "gpu")
        if branch not in ("gpu", "cpu"):
            raise ValueError(f"Invalid rhs_branch: {branch}")
        self.rhs_branch = branch
        self.wdtype = torch.float64 if prec == "fp64" else torch.float32
        layout = config.get("compute", {}).get("layout", "dense")
        if layout not in ("padded", "dense"):
            raise ValueError(f"Invalid layout: {layout}. Must be 'padded' or 'dense'")
        self.layout = layout

        g = config["grid"]
        self.nx, self.ny, self.nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
        d = config["domain"]
        self.Lx, self.Ly, self.Lz = float(d["Lx"]), float(d["Ly"]), float(d["Lz"])
        self.stretching_type = d.get("stretching_type", "symmetric")
        if self.stretching_type not in ("symmetric", "bottom", "hybrid"):
            raise ValueError(f"Invalid stretching type: {self.stretching_type}. "
                             "Must be 'symmetric', 'bottom', or 'hybrid'")
        fl = config["flow"]
        self.nu = 1.0 / fl["Re"]
        self.Re_tau = fl.get("Re_tau", 180.0)
        self.U_bulk = float(fl["U_bulk"])
        self.gamma = float(fl["gamma"])
        self.top_wall_bc_type = config.get("boundary_conditions", {}).get("top_wall", {}).get("type", "dirichlet")
        L.top_code(self.top_wall_bc_type)
        t = config["time"]
        self.dt = float(t["dt"])
        self.n_steps = int(t["n_steps"])
        self.t_max = t.get("t_max", 1000.0)
        self.cfl_target = t.get("CFL_target", 0.25)
        self.dt_update_interval = t.get("dt_update_interval", 0)
        self.dt_max = t.get("dt_max", 0.01)
        self.dt_min = t.get("dt_min", 0.0001)
        self.time_scheme = t.get("scheme", "IMEX")
        # the step's Poisson solve is the FFT one (solver.py:281-294); the reference's dense
        # direct solver (projection.py:3-140) is small-grid test tooling here, not a step option
        self.solver_type = config.get("solver", {}).get("type", "fft")
        if self.solver_type != "fft":
            raise ValueError(f"Unsupported solver type: {self.solver_type}. torchannel_amd steps with the FFT "
                             "Poisson solver ('fft')")
        out = config.get("output", {})
        self.results_folder = out.get("results_folder", "results")
        self.n_out = out.get("n_out", out.get("print_interval", 10))
        self.n_save = out.get("n_save", out.get("slice_interval", 100))
        self.save_init = bool(out.get("save_init", True))

        # grid (utils.py:5-113), built on the host in float64 and moved to the device; the hybrid
        # grid sets nz from its two regions (solver.py:192-211)
        if self.stretching_type == "hybrid":
            nz_u, nz_s = int(g.get("nz_uniform", 75)), int(g.get("nz_stretched", 75))
            z_f, z_c, dz_f, dz_c = generate_hybrid_grid(nz_u, nz_s, float(d.get("z_transition", 0.25)), self.Lz,
                                                        float(d.get("gamma_stretched", 1.8)), device="cpu",
                                                        verbose=verbose)
            self.nz = nz_u + nz_s
        else:
            z_f, z_c, dz_f, dz_c = generate_grid(self.gamma, self.nz, self.Lz, device="cpu",
                                                 stretching_type=self.stretching_type)
        self.z_f, self.z_c, self.dz_f, self.dz_c = (x.to(self.device) for x in (z_f, z_c, dz_f, dz_c))
        self._zc1 = float(z_c[1])      # wall distance of compute_u_tau (utils.py:188)
        self.dx, self.dy = self.Lx / self.nx, self.Ly / self.ny
        self.cell_vol = (self.dx * self.dy * self.dz_f.view(1, 1, -1)).expand(self.nx, self.ny, self.nz)
        self.cell_vol_ratio = self.cell_vol
        self.total_volume = self.Lx * self.Ly * self.Lz

        ini = config["initialization"]
        field_file = ini.get("field_file")
        self.initial_step, self.initial_time = 0, 0.0
        if field_file is not None:                                  # restart (solver.py:246-249)
            u0, v0, w0, p0, self.initial_step, self.initial_time = initialize_flow_from_file(
                field_file, device=self.device, reset_time=ini.get("reset_time", False), dtype=self.wdtype)
        else:
            u0, v0, w0, p0 = initialize_flow(
                self.nx, self.ny, self.nz, z_c, self.Ly, self.Lz, U_bulk=self.U_bulk,
                init_type=ini.get("type", "parabolic"),
                perturbation_intensity=ini.get("perturbation_intensity", 0.0),
                n_vortices=ini.get("n_vortices", 4), device=self.device,
                top_wall_bc_type=self.top_wall_bc_type, seed=ini.get("seed"), dtype=self.wdtype)

        # the state in the solver's layout, and the grid whose calls take that layout
        self.pitch = L.padded_pitch(self.nz, self.wdtype) if layout == "padded" else 0
        self.u, self.v, self.w = (self._field(q.shape, q) for q in (u0, v0, w0))
        self._grid = L.get_grid(self.nx, self.ny, self.nz, self.dx, self.dy, self.nu,
                                L.top_code(self.top_wall_bc_type), self.wdtype, self.device,
                                dz_f=self.dz_f, dz_c=self.dz_c, pitch=self.pitch)
        self.fft_data = initialize_fft_solver(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz_c,
                                              self.dz_f, top_wall_bc_type=self.top_wall_bc_type,
                                              dtype=self.wdtype, _grid=self._grid)
        self.p = self.fft_data["workspace_p"]
        self.p.copy_(p0)
        # a plan of its own when the channel is stepped together with others (ensembles)
        self._own_plan = bool(config.get("compute", {}).get("private_plan", False))
        self._plan = self._grid.new_poisson() if self._own_plan else self._grid.poisson()

        st = config.get("statistics", {})
        self.n_stats = st.get("n_stats", 0)
        self.t_stats = st.get("t_stats", 10.0)
        self.turbulence_stats = None
        self.stats_output_path = os.path.join(self.results_folder, st.get("output_file", "turbulence_stats.npz"))
        self.stats_state_path = os.path.join(self.results_folder, st.get("state_file", "turbulence_stats_state.npz"))
        if self.n_stats > 0:
            self.turbulence_stats = TurbulenceStats(
                self.nx, self.ny, self.nz, self.Lx, self.Ly, self.Lz, self.z_c, self.z_f, self.dz_c,
                self.dz_f, self.dx, self.dy, self.nu, self.Re_tau,
                z_plus_target=st.get("z_plus_target", 15.0), device=self.device)
            self.turbulence_stats.use_grid(self._grid, self.wdtype, self.device)
            self.turbulence_stats._plan(self.wdtype, self.device)      # built here, not in the time loop
            if st.get("restart_state_file") is not None:
                self.turbulence_stats.load_state(st["restart_state_file"])

        # AB2 history (two rhs sets swapped each step) and the device scalar block
        self._rhs = [[self._field(q.shape) for q in (self.u, self.v, self.w)] for _ in range(2)]
        self._scal = torch.zeros(L.TC_NSCAL, dtype=torch.float64, device=self.device)
        self._scal[L.S_DT] = self.dt
        self._scal[L.S_UTARGET] = self.U_bulk
        self._newest = -1
        self.body_force = None
        self.tuning = None
        ra = (ctypes.c_void_p * 3)(*[q.data_ptr() for q in self._rhs[0]])
        rb = (ctypes.c_void_p * 3)(*[q.data_ptr() for q in self._rhs[1]])
        # the solver's scratch (post-CN velocities, CN tables, reduction partials) in a
        # torch-allocated workspace, kept alive with the solver
        nbytes = ctypes.c_size_t()
        L.call("tc_solver_workspace_size", self._grid.handle, ctypes.byref(nbytes))
        self._workspace = torch.empty(nbytes.value + 256, dtype=torch.uint8, device=self.device)
        ws_ptr = (self._workspace.data_ptr() + 255) & ~255
        h = ctypes.c_void_p()
        L.call("tc_solver_create", self._grid.handle, self._plan, L.ptr(self.u), L.ptr(self.v),
               L.ptr(self.w), L.ptr(self.p), ra, rb, L.ptr(self._scal), float(self.total_volume),
               L.TC_BRANCH_CPU if branch == "cpu" else L.TC_BRANCH_GPU, ctypes.c_void_p(ws_ptr),
               ctypes.c_size_t(nbytes.value), ctypes.byref(h))
        self._solver = h
        self._graph_key = None

        # rescale to U_bulk unless restarting (solver.py:268-276), fields_init.npz of the state
        # before the projection (solver.py:337-340: step 0, time 0, forcing 0 also on restart),
        # and the initial projection (solver.py:342-369); forcing and AB2 history start from zero
        # (solver.py:249, 377-383)
        if field_file is None:
            L.call("tc_solver_rescale", self._solver, self._stream())
        if self.save_init:
            ut0 = compute_u_tau(self.u, self.z_c, self.nu, self.top_wall_bc_type)
            save_flow_fields(self.u, self.v, self.w, self.p, self.z_c, self.z_f, self.Lx, self.Ly, 0, 0.0, ut0,
                             0.0, self.results_folder, "fields_init.npz")
        L.call("tc_solver_project_initial", self._solver, 0, self._stream())
        self.time = self.initial_time
        self._scal[L.S_TIME] = self.time
        self.current_step = self.initial_step
        self._scal[L.S_STEP] = self.current_step      # stamps the diagnostics rows (kept on restart)

    # -- plumbing --------------------------------------------------------------------------
    def _stream(self):
        return L.stream_of(self.u)

    def _field(self, shape, src=None):
        """A zeroed array of `shape` in the solver's layout, holding a copy of `src` if given."""
        out = L.empty_field(shape, self.pitch, self.wdtype, self.device)
        if src is not None:
            out.copy_(src)
        return out

    @property
    def body_force(self):
        """Body force f(x, y, z) added to rhs_u (shaped like u), or None.  Assigning copies it
        into the solver's layout."""
        return self._body

    @body_force.setter
    def body_force(self, f):
        if f is None:
            self._body = None
            return
        if tuple(f.shape) != tuple(self.u.shape):
            raise ValueError(f"body force must have the shape of u, {tuple(self.u.shape)}")
        self._body = self._field(self.u.shape, f.to(device=self.device, dtype=self.wdtype))

    def __del__(self):
        try:
            if getattr(self, "_solver", None) is not None:
                L.load().tc_solver_destroy(self._solver)
            if getattr(self, "_own_plan", False) and getattr(self, "_plan", None) is not None:
                L.Grid.destroy_poisson(self._plan)
        except Exception:
            pass

    @property
    def forcing(self):
        return self._scal[L.S_F]

    @property
    def rhs_u_prev(self):
        return self._rhs[self._newest][0] if self._newest >= 0 else None

    @property
    def rhs_v_prev(self):
        return self._rhs[self._newest][1] if self._newest >= 0 else None

    @property
    def rhs_w_prev(self):
        return self._rhs[self._newest][2] if self._newest >= 0 else None

    @property
    def rhs_u_curr(self):
        return self._rhs[1 - self._newest][0] if self._newest >= 0 else None

    @property
    def rhs_v_curr(self):
        return self._rhs[1 - self._newest][1] if self._newest >= 0 else None

    @property
    def rhs_w_curr(self):
        return self._rhs[1 - self._newest][2] if self._newest >= 0 else None

    def device_dt(self):
        return float(self._scal[L.S_DT].item())

    # -- the step ------------------------------------------------------------------------------
    def apply_bc_uvw(self):
        apply_bc_all(self.u, self.v, self.w, self.top_wall_bc_type)

    def step_imex(self, dt=None):
        """One IMEX step (solver.py:580-710).  dt (host float) sets the device dt first; None keeps
        the device value.  Returns (u_bulk, forcing) as 0-D float64 device tensors."""
        newest = ctypes.c_int(-1)
        first = 1 if self._newest < 0 else 0
        body = L.ptr(self.body_force) if self.body_force is not None else None
        L.call("tc_solver_step", self._solver, float(dt) if dt is not None else -1.0, first, body,
               ctypes.byref(newest), self._stream())
        self._newest = newest.value
        if dt is not None:
            self.dt = float(dt)
        out = self._scal[L.S_UBULK].clone(), self._scal[L.S_F].clone()
        if self.autotune and self.tuning is None:
            from .autotune import tune
            self.tuning = tune(self)
        return out

    def compute_cfl_dt(self):
        """solver.py:465-496 (host value; one device sync)."""
        from .operators import compute_cfl_fused
        dti = compute_cfl_fused(self.u, self.v, self.w, self.nx, self.ny, self.nz, self.dx, self.dy,
                                self.dz_f, self.dz_c)
        if dti < 1e-10:
            dti = 1.0
        return min(max(self.cfl_target * (1.0 / dti), self.dt_min), self.dt_max)

    def update_dt_device(self):
        """The adaptive-dt rule (solver.py:808-812) evaluated on the device, no host sync."""
        L.call("tc_solver_update_dt", self._solver, float(self.cfl_target), float(self.dt_min),
               float(self.dt_max), self._stream())

    def run_graph(self, n_chunks, chunk_steps, dt_update_interval=0):
        """Replay `n_chunks` HIP-graph chunks of `chunk_steps` IMEX steps each (even), with the
        device-side adaptive dt before every dt_update_interval-th step of a chunk.  The AB2
        history must exist (take at least one eager step first)."""
        if self._newest < 0:
            raise RuntimeError("run_graph needs one eager step first (forward-Euler start)")
        if self.body_force is not None:
            raise RuntimeError("run_graph steps without a body force; use step_imex while body_force is set")
        # the captured steps read R_{n-1} from the rhs set that is newest at capture: the parity
        # of the AB2 history is part of what a graph was recorded for
        key = (chunk_steps, dt_update_interval, self.cfl_target, self.dt_min, self.dt_max, self._newest)
        if self._graph_key != key:
            L.call("tc_solver_capture", self._solver, int(chunk_steps), int(dt_update_interval),
                   float(self.cfl_target), float(self.dt_min), float(self.dt_max), self._stream())
            self._graph_key = key
        for _ in range(n_chunks):
            L.call("tc_solver_replay", self._solver, self._stream())
        self.current_step += n_chunks * chunk_steps

    def record_diagnostics(self, rows, capacity):
        """Append the n_out diagnostics row [step, time, u_bulk, u_tau, F, dt, max|div|, CFL rate]
        to the device buffer `rows` (rows[0] counts rows; see tc_solver_diagnostics), no sync."""
        L.call("tc_solver_diagnostics", self._solver, self._zc1, L.ptr(rows), int(capacity), self._stream())

    def _advance(self, k):
        """k IMEX steps with the adaptive-dt cadence: HIP-graph chunks when the chunk boundaries
        line up with the dt cadence and no body force is set, eager steps otherwise."""
        iv = self.dt_update_interval
        if self._newest < 0 and k > 0:            # the forward-Euler start is an eager step
            self.current_step += 1
            if iv > 0 and self.current_step % iv == 0:
                self.update_dt_device()
            self.step_imex(None)
            k -= 1
        chunk = self.n_out if self.n_out >= 2 and self.n_out % 2 == 0 else 0
        if (chunk and k > 0 and self.body_force is None and k % chunk == 0
                and (iv <= 0 or (chunk % iv == 0 and self.current_step % iv == 0))):
            self.run_graph(k // chunk, chunk, dt_update_interval=max(iv, 0))
            return
        for _ in range(k):
            self.current_step += 1
            if iv > 0 and self.current_step % iv == 0:
                self.update_dt_device()
            self.step_imex(None)

    def _drain(self, rows, ts):
        """One host sync: move the device time-series rows into `ts` and reset the buffer."""
        host = rows.cpu().numpy()
        n = min(int(host[0]), (len(host) // 8) - 1)
        for r in range(n):
            step, tm, ub, ut, f, dt, div, _ = host[8 * (r + 1): 8 * (r + 2)]
            for key, val in (("step", int(step)), ("time", tm), ("u_bulk", ub), ("u_tau", ut), ("forcing", f)):
                ts[key].append(val)
            if self.verbose:
                print(f"{int(step):6d} {tm:10.6f} {dt:10.6f} {div:12.3e} {ub:10.6f} {ut:10.6f} {f:12.3e}",
                      flush=True)
        rows[0] = 0.0

    def _time_refresh(self):
        """One host sync: the device time and dt, and the largest dt any later step can take
        (the adaptive rule clamps to dt_max, or keeps the current dt)."""
        self.time = float(self._scal[L.S_TIME].item())
        self.dt = float(self._scal[L.S_DT].item())
        self._t_step = self.current_step
        self._dt_bound = max(self.dt, self.dt_max) if self.dt_update_interval > 0 else self.dt

    def _steps_below_tmax(self):
        """How many further steps certainly start below t_max (the reference takes a step while
        time < t_max, solver.py:793), from the last device time read and the dt bound."""
        if not math.isfinite(self.t_max):
            return 1 << 62
        x = (self.t_max - self.time) / self._dt_bound - (self.current_step - self._t_step)
        return int(math.floor(x * (1.0 - 1e-12))) if x > 0 else 0

    def run_simulation(self):
        """The loop of solver.py:738-970: steps while step < n_steps and time < t_max, adaptive
        dt every dt_update_interval steps (evaluated on the device), diagnostics every n_out
        steps, statistics every n_stats steps once time >= t_stats, and every n_save steps the
        non-finite check (a device flag set by the projection kernel), timeseries.npz append,
        fields.npz and the statistics state; fields_final.npz at the end.

        Output files follow the reference: the time-series rows recorded after the last n_save
        step are not written (the reference writes them only at n_save steps), and a non-finite
        state writes fields_error.npz, stops the loop and leaves fields.npz at the last good
        checkpoint.

        The steps between two events run as HIP-graph chunks of n_out steps where the cadences
        allow it.  The n_out diagnostics (u_bulk, u_tau, forcing, max|div|) are rows of a device
        time series written without a host sync; the host reads them once per n_save steps and
        reads the device time at the statistics cadence and near t_max: segments are cut so that
        no step starts at or past t_max."""
        if self.time_scheme != "IMEX":
            raise ValueError(f"Unknown time stepping scheme: {self.time_scheme}. Use 'IMEX'.")
        os.makedirs(self.results_folder, exist_ok=True)
        ts = {k: [] for k in ("step", "time", "u_bulk", "u_tau", "forcing")}
        n_out, n_save = max(1, int(self.n_out)), max(1, int(self.n_save))
        stats_on = self.turbulence_stats is not None and self.n_stats > 0
        cap = n_save // n_out + 4
        rows = torch.zeros(8 * (cap + 1), dtype=torch.float64, device=self.device)
        t0 = _time.time()
        step = self.current_step
        self._time_refresh()

        def ceil_to(x, m):
            return ((x + m - 1) // m) * m

        while step < self.n_steps and self.time < self.t_max:
            nxt = min(ceil_to(step + 1, n_out), ceil_to(step + 1, n_save), self.n_steps)
            if stats_on:
                nxt = min(nxt, ceil_to(step + 1, self.n_stats))
            k = nxt - step
            room = self._steps_below_tmax()
            if room < k:
                if room == 0:
                    if self._t_step != step:
                        self._time_refresh()
                        if self.time >= self.t_max:
                            break
                        room = self._steps_below_tmax()
                    room = max(room, 1)          # time < t_max was just read: one step is allowed
                k = min(k, room)
            self._advance(k)
            step = self.current_step
            if step % n_out == 0:
                self.record_diagnostics(rows, cap)
            if stats_on and step % self.n_stats == 0:
                self._time_refresh()
                if self.time >= self.t_stats:
                    self.turbulence_stats.accumulate_statistics(self.u, self.v, self.w)
            if step % n_save == 0:
                self._drain(rows, ts)
                self._time_refresh()
                ut = compute_u_tau(self.u, self.z_c, self.nu, self.top_wall_bc_type)
                if float(self._scal[L.S_NONFINITE].item()) != 0.0:          # solver.py:862-881
                    if self.verbose:
                        print(f"ERROR: NaN or Inf detected in solution at step {step}, time = {self.time:.6f}",
                              flush=True)
                    ube = compute_bulk_velocity(self.u, self.cell_vol_ratio, self.total_volume)
                    save_flow_fields(self.u, self.v, self.w, self.p, self.z_c, self.z_f, self.Lx, self.Ly, step,
                                     self.time, ut, (self.U_bulk - float(ube)) / self.dt, self.results_folder,
                                     "fields_error.npz")
                    ts = {k: [] for k in ts}
                    break
                if ts["step"]:
                    append_timeseries(os.path.join(self.results_folder, "timeseries.npz"), **ts)
                    ts = {k: [] for k in ts}
                save_flow_fields(self.u, self.v, self.w, self.p, self.z_c, self.z_f, self.Lx, self.Ly, step,
                                 self.time, ut, float(self._scal[L.S_F].item()), self.results_folder, "fields.npz")
                if self.turbulence_stats is not None and self.turbulence_stats.n_samples > 0:
                    self.turbulence_stats.save_state(self.stats_state_path)
        self._drain(rows, ts)                   # rows after the last n_save step are not written
        self._time_refresh()
        if self.verbose:
            print(f"Simulation complete: {step} steps, total time = {self.time:.6f}, final dt = {self.dt:.6f}, "
                  f"wall {(_time.time() - t0):.2f}s", flush=True)
        if self.turbulence_stats is not None and self.turbulence_stats.n_samples > 0:
            self.turbulence_stats.save_state(self.stats_state_path)
            self.turbulence_stats.save_statistics(self.stats_output_path)
        ut = compute_u_tau(self.u, self.z_c, self.nu, self.top_wall_bc_type)
        ubf = compute_bulk_velocity(self.u, self.cell_vol_ratio, self.total_volume)
        save_flow_fields(self.u, self.v, self.w, self.p, self.z_c, self.z_f, self.Lx, self.Ly, step, self.time, ut,
                         (self.U_bulk - float(ubf)) / self.dt, self.results_folder, "fields_final.npz")
        return step
