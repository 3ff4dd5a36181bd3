"""Benchmark: Mcell-steps/s of the IMEX step at 192x192x128 on one MI355X.

Workload (the Re_tau~180 production case): 192x192x128, Lx=4pi, Ly=4pi/3, Lz=2, tanh gamma=2.6,
Re_bulk=2870, constant-flow-rate forcing, adaptive dt at CFL 0.25 (dt update every 10 steps, on
the device), turbulence statistics every 200 steps; seeded perturbed parabola (synthetic).
`value` is the plain float64 path; the float32 path (float64 gauge mode, bulk/forcing and
statistic sums) is reported beside it.  A "step" is one full IMEX step of the whole grid.  After
W eager warm-up steps, exactly K steps are timed: eager steps up to the next multiple of 10, then
HIP-graph chunks of 10 (the dt update sits inside the chunk, at every 10th step counted from the
start), then eager steps for the rest; statistics run every 200th step, all inside the timed
region.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--no-fp32] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX, NY, NZ = 192, 192, 128
HBM_PEAK_GBS = 8000.0
PHASES = ["cn_tables", "explicit_ab2_cn", "div_fft_x", "fft_y_fwd", "poisson_thomas", "fft_y_inv",
          "ifft_x_p", "project_bc_bulk", "cfl_max"]
KERNEL_OF = {"explicit_ab2_cn": "k_explicit_cn_uvw", "div_fft_x": "k_fwd_x", "fft_y_fwd": "k_fft_y",
             "poisson_thomas": "k_poisson_thomas", "fft_y_inv": "k_fft_y", "ifft_x_p": "k_inv_x",
             "project_bc_bulk": "k_project_bc", "cfl_max": "k_cfl"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "imex_step_192x192x128_hbm_traffic.json")
LAYOUT = "dense"      # the solver's array layout (compute.layout), --layout
TUNE = True           # pick the launch-shape knobs with the autotuner after the first warm-up step
TUNE_CACHE = None     # --tune-cache: a JSON file of earlier choices (read if it has this workload)
KNOB_NAMES = {1: "k1_pencil_length", 2: "thomas_systems_per_block", 3: "k1_form", 4: "thomas_lds"}


def _tune(ch):
    """The autotuner on a channel that has taken a step (results do not depend on the knobs);
    returns the chosen knobs by name, or None when tuning is off or no step was taken yet."""
    if not TUNE or ch._newest < 0:
        return None
    from torchannel_amd.autotune import tune
    return {KNOB_NAMES.get(k, str(k)): v for k, v in tune(ch, path=TUNE_CACHE).items()}


def config(precision, nx=NX, ny=NY, nz=NZ, steps=1000):
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
        "flow": {"Re": 2870.0, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": 2.6},
        "time": {"dt": 0.005, "n_steps": steps, "CFL_target": 0.25, "dt_update_interval": 10,
                 "dt_max": 0.01, "dt_min": 1e-4},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.1, "seed": 0},
        "compute": {"device": "cuda", "precision": precision, "layout": LAYOUT},
        "output": {"save_init": False},
        "statistics": {"n_stats": 200, "t_stats": 0.0, "z_plus_target": 15.0},
    }


def algorithmic_words(nx, ny, nz):
    """Compulsory HBM words per interior cell for each kernel (ghost cells excluded)."""
    spec = 2.0 * (nx // 2 + 1) / nx          # half-spectrum complex words per cell
    return {
        "cn_tables": 0.0,
        "explicit_ab2_cn": 12.0,              # read u, v, w, rhs_prev x3; write rhs x3, q* x3
        "div_fft_x": 3.0 + spec,              # read q* x3; write x half spectra
        "fft_y_fwd": 2 * spec,
        "poisson_thomas": 2 * spec,
        "fft_y_inv": 2 * spec,
        "ifft_x_p": spec + 1.0,               # read spectra; write p
        "project_bc_bulk": 7.0,               # read q* x3, p; write u, v, w
        "cfl_max": 3.0,
    }


def phase_times(ch, reps):
    """Average duration (ms) of each kernel as it runs inside the step sequence: the CFL pass, the
    CN tables, K1, the five Poisson passes and K5, in step order, `reps` times, with HIP events
    recorded between the launches on the solver's stream (torch's current stream), so every
    kernel meets the cache state the preceding kernels of a step leave behind."""
    import torch
    from torchannel_amd import _lib as L

    order = [8, 0, 1, 2, 3, 4, 5, 6, 7]
    st = ch._stream()
    for i in order:
        L.call("tc_solver_phase", ch._solver, i, 1, st)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(order) + 1)] for _ in range(reps)]
    for r in range(reps):
        ev[r][0].record()
        for q, i in enumerate(order):
            L.call("tc_solver_phase", ch._solver, i, 1, st)
            ev[r][q + 1].record()
    torch.cuda.synchronize()
    return {PHASES[i]: sum(ev[r][q].elapsed_time(ev[r][q + 1]) for r in range(reps)) / reps
            for q, i in enumerate(order)}


def _eager(ch, n, interval):
    """n eager steps with the reference cadence: the dt update before every interval-th step and
    statistics after every n_stats-th step, both counted from the start of the run."""
    for _ in range(n):
        ch.current_step += 1
        if ch.current_step % interval == 0:
            ch.update_dt_device()
        ch.step_imex(None)
        if ch.current_step % ch.n_stats == 0:
            ch.turbulence_stats.accumulate_statistics(ch.u, ch.v, ch.w)


def _precapture(ch, chunk, interval, newest):
    """Record the chunk graph for the AB2 parity the chunks will start from (capture launches
    nothing, so the state is untouched)."""
    from torchannel_amd import _lib as L
    if newest < 0:
        return
    saved = ch._newest
    L.call("tc_solver_set_newest", ch._solver, newest)
    ch._newest = newest
    ch.run_graph(0, chunk, dt_update_interval=interval)
    L.call("tc_solver_set_newest", ch._solver, saved)
    ch._newest = saved


def run_gpu(precision, steps, warmup, phase_reps=20):
    import torch
    import torchannel_amd as tc
    from torchannel_amd import _lib as L

    ch = tc.ChannelFlow(config=config(precision, steps=steps))
    interval = ch.dt_update_interval
    chunk = 10
    # warm-up: forward-Euler start then AB2, eager, with the reference cadence; the autotuner
    # after the first warm-up step
    _eager(ch, min(warmup, 1), interval)
    tuning = _tune(ch)
    _eager(ch, warmup - min(warmup, 1), interval)
    # the timed K steps: eager up to a multiple of the chunk, chunks, eager rest
    # (with no warm-up the first chunk's worth is eager: a graph needs the AB2 history)
    lead = min(steps, (-warmup) % chunk if warmup > 0 else chunk)
    n_chunks = (steps - lead) // chunk
    tail = steps - lead - n_chunks * chunk
    if n_chunks:
        newest = ch._newest if lead % 2 == 0 or ch._newest < 0 else 1 - ch._newest
        if ch._newest < 0 and lead > 0:
            newest = (lead - 1) % 2        # the forward-Euler start fills set 0, then they alternate
        _precapture(ch, chunk, interval, newest)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record()
    _eager(ch, lead, interval)
    for _ in range(n_chunks):
        ch.run_graph(1, chunk, dt_update_interval=interval)
        if ch.current_step % ch.n_stats == 0:
            ch.turbulence_stats.accumulate_statistics(ch.u, ch.v, ch.w)
    _eager(ch, tail, interval)
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1)
    done = steps
    n_cells = NX * NY * NZ
    # fixed-dt kernel throughput: graph chunks with no CFL evaluation (at least one chunk)
    fixed_chunks = max(1, steps // chunk)
    ch.run_graph(0, chunk, dt_update_interval=0)
    torch.cuda.synchronize()
    t0.record()
    ch.run_graph(fixed_chunks, chunk, dt_update_interval=0)
    t1.record()
    torch.cuda.synchronize()
    ms_fixed = t0.elapsed_time(t1)
    out = {
        "mcell_steps_per_s": n_cells * done / (ms * 1e-3) / 1e6,
        "ms_per_step": ms / done,
        "wall_s": wall,
        "steps": done,
        "fixed_dt_ms_per_step": ms_fixed / (fixed_chunks * chunk),
        "dt_final": ch.device_dt(),
        "u_bulk": float(ch._scal[L.S_UBULK]),
        "stats_samples": ch.turbulence_stats.n_samples,
    }
    # per-kernel average duration inside the step sequence, HIP events on the launch stream
    per = {}
    words = algorithmic_words(NX, NY, NZ)
    wsize = 8 if precision == "fp64" else 4
    for name, avg_ms in phase_times(ch, phase_reps).items():
        byts = words[name] * wsize * n_cells
        per[name] = {"avg_us": avg_ms * 1e3, "alg_bytes": byts,
                     "gbs": byts / (avg_ms * 1e-3) / 1e9 if byts else 0.0}
    out["kernels"] = per
    out["pitch"] = ch.pitch
    out["tuning"] = tuning
    out["k1_form"] = ch._grid.get_tuning(L.TUNE_K1_TILE)
    out["k1_jb"] = ch._grid.get_tuning(L.TUNE_K1_JB)
    step_kernels = [k for k in PHASES if k not in ("cfl_max",)]
    dom = max(step_kernels, key=lambda k: per[k]["avg_us"])
    out["dominant"] = dom
    tot_bytes = sum(words[k] * wsize * n_cells for k in step_kernels)
    out["step_alg_bytes"] = tot_bytes
    out["step_gbs"] = tot_bytes / (out["ms_per_step"] * 1e-3) / 1e9
    del ch
    torch.cuda.empty_cache()
    return out


def column_ceiling(precision, nx, ny, nz, jb, pitch, reps=20):
    """Bandwidth of the explicit + CN kernel's column access patterns without its arithmetic
    (tc_calib_columns): the three-component pencils' pattern (the form the kernel runs at these
    sizes), the per-component pencils' pattern with their neighbour-column loads, and the
    compulsory streams alone; GB/s of the kernel's 12 algorithmic words per cell, from HIP events
    on the launch stream.  jb: the pencil length of the three-component form; pitch: the column
    stride of the solver's arrays (0 dense), which the twelve arrays here share."""
    import ctypes
    import torch
    from torchannel_amd import _lib as L

    dt = torch.float64 if precision == "fp64" else torch.float32
    arrs = [L.empty_field((nx + 2, ny + 2, nz + 2), pitch, dt, "cuda") for _ in range(12)]
    for a in arrs:
        a.uniform_()
    ptrs = (ctypes.c_void_p * 12)(*[a.data_ptr() for a in arrs])
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {"pitch": pitch}
    for pat, name, pjb in ((2, "k1_pattern", jb), (0, "pencil_pattern", 8), (1, "compulsory", 8)):
        for _ in range(3):
            L.call("tc_calib_columns", L.dtype_code(dt), pat, nx, ny, nz, pjb, pitch, ptrs, st)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            L.call("tc_calib_columns", L.dtype_code(dt), pat, nx, ny, nz, pjb, pitch, ptrs, st)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        byts = 12 * (8 if precision == "fp64" else 4) * nx * ny * nz
        out[name] = {"avg_us": round(us, 2), "gbs": round(byts / (us * 1e-6) / 1e9, 1)}
    del arrs
    torch.cuda.empty_cache()
    return out


def run_adjoint(n_steps=100, ckpt=10, reps=3):
    """Adjoint/forward cost at 96x96x64: 100 fixed-dt float64 steps, checkpoints every 10, a
    seeded Gaussian body force (sigma 1e-3).  forward = objective(f) (the steps, the per-step
    objective and the checkpoints); adjoint = gradient(f) (recompute from the checkpoints + the
    reverse sweep).  Medians over `reps` runs."""
    import torch
    import torchannel_amd as tc

    c = config("fp64", 96, 96, 64, steps=n_steps)
    c["compute"]["rhs_branch"] = "gpu"
    ch = tc.ChannelFlow(config=c)
    gen = torch.Generator().manual_seed(0)
    f = (1e-3 * torch.randn(tuple(ch.u.shape), generator=gen, dtype=torch.float64)).cuda()
    run = tc.WallShearAdjoint(ch, n_steps, 0.005, checkpoint_every=ckpt)
    run.gradient(f)
    torch.cuda.synchronize()
    tf, tg = [], []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e2 = torch.cuda.Event(enable_timing=True)
        e0.record()
        run.objective(f)
        e1.record()
        run.gradient(f)
        e2.record()
        torch.cuda.synchronize()
        tf.append(e0.elapsed_time(e1))
        tg.append(e1.elapsed_time(e2))
    tf.sort()
    tg.sort()
    fw, ad = tf[len(tf) // 2], tg[len(tg) // 2]
    return {"workload": "96x96x64, 100 fixed-dt (0.005) float64 steps, checkpoints every 10",
            "forward_ms": round(fw, 3), "adjoint_ms": round(ad, 3),
            "forward_step_us": round(fw * 1e3 / n_steps, 2), "adjoint_step_us": round(ad * 1e3 / n_steps, 2),
            "adjoint_over_forward": round(ad / fw, 3)}


def run_c3(steps=500, warmup=20, phase_reps=10):
    """Re_tau~550 single channel: 384x384x256, Lx=2pi, Ly=pi, gamma=2.6, Re_bulk=10000, adaptive
    dt at CFL 0.25 (every 10 steps), float32 (float64 gauge mode), seeded perturbed parabola;
    500 steps as graph chunks of 10, then per-kernel average durations and algorithmic GB/s
    (the wall-normal Poisson solve is one Thomas sweep per (kx, ky) system: 193 x 384 systems of
    length 256 per launch)."""
    import torch
    import torchannel_amd as tc
    from torchannel_amd import _lib as L

    nx, ny, nz = 384, 384, 256
    c = config("fp32", nx, ny, nz, steps=steps)
    c["domain"] = {"Lx": 2 * math.pi, "Ly": math.pi, "Lz": 2.0}
    c["flow"]["Re"] = 10000.0
    c["flow"]["Re_tau"] = 550.0
    c["statistics"] = {"n_stats": 0}
    ch = tc.ChannelFlow(config=c)
    chunk = 10
    # the first dt from the CFL rule (a white-noise start at this resolution is far above CFL 0.25
    # at the nominal dt = 0.005), then the regular cadence
    ch.update_dt_device()
    tuning = None
    for s_ in range(warmup):
        if (s_ + 1) % 10 == 0:
            ch.update_dt_device()
        ch.step_imex(None)
        if s_ == 0:
            tuning = _tune(ch)
    ch.run_graph(0, chunk, dt_update_interval=10)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    ch.run_graph(steps // chunk, chunk, dt_update_interval=10)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / ((steps // chunk) * chunk)
    # state after the timed steps (the per-kernel timing below re-runs phases out of order)
    finite = float(ch._scal[L.S_NONFINITE]) == 0.0
    dt_final, u_bulk = ch.device_dt(), float(ch._scal[L.S_UBULK])
    n_cells = nx * ny * nz
    words = algorithmic_words(nx, ny, nz)
    per = {}
    for name, avg in phase_times(ch, phase_reps).items():
        byts = words[name] * 4 * n_cells
        per[name] = {"avg_us": round(avg * 1e3, 2), "gbs": round(byts / (avg * 1e-3) / 1e9, 1) if byts else 0.0}
    jb, pitch = ch._grid.get_tuning(L.TUNE_K1_JB), ch.pitch
    del ch
    torch.cuda.empty_cache()
    return {"workload": "384x384x256, Lx=2pi, Ly=pi, Re_bulk=10000, float32, adaptive dt CFL 0.25 (first dt "
                        f"from the CFL rule), {(steps // chunk) * chunk} steps", "finite": finite,
            "value": round(n_cells / (ms * 1e-3) / 1e6, 2), "unit": "Mcell-steps/s", "ms_per_step": round(ms, 4),
            "dt_final": dt_final, "u_bulk": u_bulk, "kernels": per, "tuning": tuning,
            "column_pattern_ceiling": column_ceiling("fp32", nx, ny, nz, jb, pitch)}


def run_ensemble(n=48, steps=300, warmup=10, chunk=10):
    """Batched ensemble: 48 channels, nz from {32, 48, 64, 96}, nx = ny from {48, 64, 96},
    Re_bulk in [2000, 4000], gamma in [2.0, 2.8] (numpy default_rng(2026)), instance i seeded
    with i, float32, fixed dt 0.005; 300 steps as graph chunks of 10 ensemble steps.  Baselines:
    the same channels stepped one after another, eagerly (tc_solver_step per channel) and by
    replaying each channel's own 10-step graph."""
    import torch
    import torchannel_amd as tc
    from torchannel_amd import _lib as L

    ens = tc.make_ensemble(tc.ensemble_configs(n))
    ens.step(0.005)
    for _ in range(warmup - 1):
        ens.step()
    ens.run_graph(0, chunk)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    ens.run_graph(steps // chunk, chunk)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    cells = ens.n_cells
    done = (steps // chunk) * chunk
    # results after the timed steps: every channel finite (the device non-finite flag its
    # projection kernel sets) and its bulk velocity held at the target by the forcing controller
    scal = torch.stack([ch._scal for ch in ens.channels]).cpu()
    check = {"all_finite": bool((scal[:, L.S_NONFINITE] == 0).all()),
             "max_abs_u_bulk_error": float((scal[:, L.S_UBULK] - scal[:, L.S_UTARGET]).abs().max()),
             "steps_per_channel": int(scal[:, L.S_STEP].min())}
    # looping the single-channel path
    chans = ens.channels
    loop_steps = 20
    e0.record()
    for _ in range(loop_steps):
        for ch in chans:
            ch.step_imex(None)
    e1.record()
    torch.cuda.synchronize()
    ms_loop = e0.elapsed_time(e1) / loop_steps
    for ch in chans:
        ch.run_graph(0, chunk)
    torch.cuda.synchronize()
    loop_chunks = 3
    e0.record()
    for _ in range(loop_chunks):
        for ch in chans:
            ch.run_graph(1, chunk)
    e1.record()
    torch.cuda.synchronize()
    ms_loop_graph = e0.elapsed_time(e1) / (loop_chunks * chunk)
    ms_step = ms / done
    return {"workload": f"{n} channels ({cells} cells), float32, fixed dt 0.005, {done} steps in graph chunks of {chunk}",
            "cells": cells, "value": round(cells / (ms_step * 1e-3) / 1e6, 2), "unit": "Mcell-steps/s",
            "ms_per_step": round(ms_step, 4),
            "loop_eager_ms_per_step": round(ms_loop, 4),
            "loop_graph_ms_per_step": round(ms_loop_graph, 4),
            "speedup_vs_loop_eager": round(ms_loop / ms_step, 2),
            "speedup_vs_loop_graph": round(ms_loop_graph / ms_step, 2),
            "check": check}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=15.0, c1_steps=200):
    """The float64 CPU oracle (a port of the reference op sequence) on the host cores, following
    the reference's CPU-device branch (separate RHS kernels, solver.py:601-622): a bounded sample
    of the same 192x192x128 workload at fixed dt (after 2 warm-up steps), and the CPU-runnable
    64x64x48 case (200 steps at dt = 0.005) in full.  Threads: the box's CPU share
    (OMP_NUM_THREADS), not every core the host reports."""
    import torch
    from oracle import channel_oracle as O

    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    torch.set_num_threads(threads)
    ch = O.Channel(NX, NY, NZ, 4 * math.pi, 4 * math.pi / 3, seed=0, branch="cpu")
    for _ in range(2):                        # warm-up (forward-Euler start, first AB2 step)
        ch.step_imex(0.005)
    n = 0
    t0 = time.perf_counter()
    while n < 20 and (time.perf_counter() - t0) < budget_s:
        ch.step_imex(0.005)
        n += 1
    dt = time.perf_counter() - t0
    del ch
    c1 = O.Channel(64, 64, 48, 4 * math.pi, 4 * math.pi / 3, seed=0, branch="cpu")
    t1 = time.perf_counter()
    for _ in range(c1_steps):
        c1.step_imex(0.005)
    dt1 = time.perf_counter() - t1
    return {"value": NX * NY * NZ * n / dt / 1e6, "unit": "Mcell-steps/s", "cores": torch.get_num_threads(),
            "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"{n} float64 IMEX steps of the 192x192x128 channel (after 2 warm-up steps), CPU-branch "
                      f"right-hand side, fixed dt=0.005, torch CPU, {dt:.1f} s",
            "c1_64x64x48_200_steps": {"seconds": round(dt1, 2), "ms_per_step": round(dt1 * 1e3 / c1_steps, 2),
                                      "value": round(64 * 64 * 48 * c1_steps / dt1 / 1e6, 3)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--c3-only", action="store_true",
                    help="run only the 384x384x256 leg (for profiling it on its own)")
    ap.add_argument("--c3-steps", type=int, default=500)
    ap.add_argument("--layout", default="dense", choices=["dense", "padded"],
                    help="the solver's array layout (compute.layout); dense is the default")
    ap.add_argument("--no-tune", action="store_true",
                    help="run the default launch shapes instead of the autotuner's choice")
    ap.add_argument("--tune-cache", default=None,
                    help="JSON file of the autotuner's choices: read when it holds this workload (no "
                         "timing launches), else written after tuning")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (5 times), the reference "
                         "kernel for scaling HBM byte counters")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: measures one GPU only (the IMEX step of one channel on one MI355X); "
              "runs on several GPUs are not supported.", flush=True)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("bench.py: --steps must be >= 1 and --warmup >= 0", flush=True)
        sys.exit(2)
    steps, warmup = args.steps, args.warmup
    global LAYOUT, TUNE, TUNE_CACHE
    LAYOUT = args.layout
    TUNE = not args.no_tune
    TUNE_CACHE = args.tune_cache
    print(f"[bench] layout {LAYOUT}", file=sys.stderr, flush=True)
    print(f"[bench] {args.precision} {NX}x{NY}x{NZ} steps={steps} warmup={warmup}", file=sys.stderr, flush=True)
    if args.copy_calib_mb > 0:
        import torch
        n = args.copy_calib_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: 5 x {args.copy_calib_mb} MiB read + write", file=sys.stderr, flush=True)
        del src, dst
    if args.c3_only:
        c3 = run_c3(steps=max(10, (args.c3_steps // 10) * 10))
        print(json.dumps({"re_tau_550": c3}), flush=True)
        return
    main_run = run_gpu(args.precision, steps, warmup)
    print(f"[bench] {args.precision}: {main_run['mcell_steps_per_s']:.1f} Mcell-steps/s", file=sys.stderr, flush=True)
    other = None
    if not args.no_fp32 and args.precision == "fp64":
        other = run_gpu("fp32", steps, warmup)
        print(f"[bench] fp32: {other['mcell_steps_per_s']:.1f} Mcell-steps/s", file=sys.stderr, flush=True)
    ceiling = column_ceiling(args.precision, NX, NY, NZ, main_run["k1_jb"], main_run["pitch"])
    print(f"[bench] column pattern: {ceiling['k1_pattern']['gbs']} GB/s, compulsory streams "
          f"{ceiling['compulsory']['gbs']} GB/s", file=sys.stderr, flush=True)
    c3 = None
    if not args.no_c3:
        c3 = run_c3(steps=max(10, (args.c3_steps // 10) * 10))
        print(f"[bench] 384x384x256 fp32: {c3['value']:.1f} Mcell-steps/s", file=sys.stderr, flush=True)
    ens = None
    if not args.no_ensemble:
        ens = run_ensemble()
        print(f"[bench] ensemble: {ens['value']:.1f} Mcell-steps/s, x{ens['speedup_vs_loop_graph']} vs looped graphs",
              file=sys.stderr, flush=True)
    adj = None
    if not args.no_adjoint and args.precision == "fp64":
        adj = run_adjoint()
        print(f"[bench] adjoint/forward: {adj['adjoint_over_forward']:.2f}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()
        print(f"[bench] cpu oracle: {cpu['value']:.2f} Mcell-steps/s", file=sys.stderr, flush=True)
    dom = main_run["dominant"]
    k = main_run["kernels"][dom]
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh).get(args.precision, {}).get(KERNEL_OF[dom])
        if t is not None:
            traffic = round(t["read_bytes"] + t["write_bytes"])
    line = {
        "metric": "Mcell-steps/s at 192x192x128 (fp32 vs fp64 plain); adjoint/forward step cost",
        "value": round(main_run["mcell_steps_per_s"], 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded perturbed laminar parabola)",
        "config": {"workload": "Re_tau~180 production step: 192x192x128, Lx=4pi, Ly=4pi/3, gamma=2.6, "
                               "Re_bulk=2870, adaptive dt CFL 0.25 every 10 steps, constant flow rate, "
                               "turbstats every 200 steps",
                   "precision": args.precision, "graph_chunk_steps": 10, "layout": LAYOUT,
                   "column_pitch": main_run["pitch"], "tuning": main_run["tuning"]},
        "roofline": {"bound": "hbm", "kernel": dom, "k1_form": main_run["k1_form"], "k1_jb": main_run["k1_jb"],
                     "achieved": round(k["gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(k["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes": k["alg_bytes"], "avg_us": round(k["avg_us"], 2),
                     "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (copy-calibrated), "
                                       "profiles/imex_step_192x192x128_hbm_traffic.md",
                     "column_pattern_ceiling": ceiling},
        "fixed_dt": {"value": round(NX * NY * NZ / (main_run["fixed_dt_ms_per_step"] * 1e-3) / 1e6, 2),
                     "ms_per_step": round(main_run["fixed_dt_ms_per_step"], 5)},
        "step_roofline": {"alg_bytes_per_step": main_run["step_alg_bytes"],
                          "achieved_gbs": round(main_run["step_gbs"], 1),
                          "frac": round(main_run["step_gbs"] / HBM_PEAK_GBS, 4)},
        "kernels_us": {n: round(v["avg_us"], 2) for n, v in main_run["kernels"].items()},
        "cpu_baseline": cpu,
        "adjoint": adj,
        "ensemble": ens,
        "re_tau_550": c3,
    }
    if other is not None:
        line["fp32"] = {"value": round(other["mcell_steps_per_s"], 2), "ms_per_step": round(other["ms_per_step"], 5),
                        "speedup_vs_fp64": round(other["mcell_steps_per_s"] / main_run["mcell_steps_per_s"], 3),
                        "kernels_us": {n: round(v["avg_us"], 2) for n, v in other["kernels"].items()},
                        "step_gbs": round(other["step_gbs"], 1)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
