"""Launch-shape autotuner for the step kernels (no counterpart in the reference, whose kernels
are TorchScript-fused).

Each knob changes only how work is dealt to workgroups, never the arithmetic, so every setting
gives bit-identical results; the tuner times the kernel a knob drives with HIP events on the
solver's stream and keeps the fastest setting on the grid.  Knobs:

* ``TUNE_K1_TILE``   form of the explicit/AB2/CN kernel: per-component pencils, the tiled form
  with the neighbour columns staged through LDS, or three-component pencils (one wavefront marching
  u, v and w of one i; searched jointly with the next knob)
* ``TUNE_K1_JB``     j-columns per explicit/AB2/CN pencil or tile
* ``TUNE_THOMAS_SB`` (kx, ky) systems per block of the wall-normal Poisson solve
* ``TUNE_THOMAS_LDS`` that solve's rows copied through LDS with whole-block 16-byte accesses, or
  loaded by each lane directly (searched jointly with the previous knob)

Chosen settings are cached per (nx, ny, nz, dtype, top wall, device architecture and CU count) in
memory and, when ``path`` (or the ``TORCHANNEL_TUNE_CACHE`` environment variable) names a JSON
file, on disk.
"""
import json
import os

import torch

from . import _lib as L

# phase indices of tc_solver_phase
_PHASE_K1 = 1
_PHASE_THOMAS = 4

_memory = {}


def _key(ch):
    """Cache key: grid, dtype, top wall and the device's architecture and CU count (the marketing
    name is not used: it reads empty under some tools, e.g. a profiler)."""
    props = torch.cuda.get_device_properties(ch.u.device)
    arch = (getattr(props, "gcnArchName", "") or props.name or "device").split(":")[0]
    return (f"{ch.nx}x{ch.ny}x{ch.nz}:{str(ch.wdtype).split('.')[-1]}:{ch.top_wall_bc_type}:"
            f"{arch}:{props.multi_processor_count}cu")


def _candidates(ch):
    """Groups of settings searched together: (list of {knob: value}, phase the group drives).
    The pencil forms are searched over every length from 2 to 24 (and 32): the block count of a
    length against the device's resident blocks makes the time a sawtooth in the length (at
    192x192x128 in float64, 13 beats 8 by 2 % and 9-12 lose 8-15 %), so sparse candidates miss
    the good lengths.  The tiled form, which measured slower everywhere, gets three lengths."""
    jb = sorted({j for j in list(range(2, 25)) + [32] if j <= ch.ny + 2})
    k1 = [{L.TUNE_K1_TILE: t, L.TUNE_K1_JB: j} for t in (2, 0) for j in jb]
    k1 += [{L.TUNE_K1_TILE: 1, L.TUNE_K1_JB: j} for j in (4, 8, 16) if j <= ch.ny + 2]
    sb = [{L.TUNE_THOMAS_LDS: s, L.TUNE_THOMAS_SB: v} for s in (1, 0) for v in (4, 8, 16)]
    return [(k1, _PHASE_K1), (sb, _PHASE_THOMAS)]


def _time_phase(ch, phase, reps):
    st = ch._stream()
    L.call("tc_solver_phase", ch._solver, phase, 2, st)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    L.call("tc_solver_phase", ch._solver, phase, reps, st)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _load(path):
    if path and os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def tune(channel, reps=10, path=None, force=False):
    """Pick the knobs for `channel`'s grid; returns {knob: value}.  The solver must have taken
    one step (the phases time the step's own intermediates); the state is not advanced by the
    timing launches except for scratch buffers the next step rewrites."""
    ch = channel
    if ch._newest < 0:
        raise RuntimeError("autotune needs one step taken first")
    path = path or os.environ.get("TORCHANNEL_TUNE_CACHE")
    key = _key(ch)
    if not force:
        hit = _memory.get(key) or _load(path).get(key)
        if hit:
            chosen = {int(k): int(v) for k, v in hit.items()}
            for k, v in chosen.items():
                ch._grid.set_tuning(k, v)
            _memory[key] = chosen
            return chosen
    # the timing launches overwrite the step's intermediates and the newest rhs set: keep the state
    saved = [q.clone() for q in (ch.u, ch.v, ch.w, ch.p, ch._scal)] + [r.clone() for s in ch._rhs for r in s]
    chosen = {}
    for group, phase in _candidates(ch):
        # two passes over the group, each setting keeping its faster pass (drifts of a few percent
        # over the search would otherwise decide between close settings)
        times = {}
        for _ in range(2):
            for q, setting in enumerate(group):
                try:
                    for k, v in setting.items():
                        ch._grid.set_tuning(k, v)
                    t = _time_phase(ch, phase, reps)
                except RuntimeError:
                    continue                 # setting outside the kernel's limits on this grid
                times[q] = min(times.get(q, float("inf")), t)
        best, best_t = None, float("inf")
        for q, t in sorted(times.items()):
            if t < best_t:
                best, best_t = group[q], t
        best = best if best is not None else {k: 0 for k in group[0]}
        for k, v in best.items():
            ch._grid.set_tuning(k, v)
        chosen.update(best)
    dst = [ch.u, ch.v, ch.w, ch.p, ch._scal] + [r for s in ch._rhs for r in s]
    for d, s in zip(dst, saved):
        d.copy_(s)
    _memory[key] = chosen
    if path:
        data = _load(path)
        data[key] = {str(k): v for k, v in chosen.items()}
        with open(path, "w") as fh:
            json.dump(data, fh, indent=1)
    return chosen
