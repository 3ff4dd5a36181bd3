"""Ensembles of independent channels advanced together (no counterpart in the reference, which
runs one channel per process).

Every phase of the IMEX step is one launch over all channels ('phased', tc_ensemble_set_form;
the 'persistent' form, one launch per step whose blocks take (channel, phase, block) work items
from a ticket counter, is built but switched off in the library): small channels share the GPU
instead of each launching its own under-filled kernels, and a chunk of steps is one HIP-graph
launch.  Each
channel is an ordinary ChannelFlow (with ``compute.private_plan`` so the Poisson workspaces are
distinct); results are bit-identical to stepping it alone, in both forms.
"""
import copy
import ctypes
import math

import numpy as np

from . import _lib as L
from .solver import ChannelFlow


FORMS = {"phased": 0, "persistent": 1}


class ChannelEnsemble:
    def __init__(self, channels, form="phased"):
        if not channels:
            raise ValueError("empty ensemble")
        c0 = channels[0]
        for ch in channels:
            if ch.wdtype != c0.wdtype or ch.rhs_branch != c0.rhs_branch or ch.device != c0.device:
                raise ValueError("ensemble channels must share dtype, rhs branch and device")
            if ch._newest != c0._newest:
                raise ValueError("ensemble channels must have AB2 histories in step")
        self.channels = list(channels)
        arr = (ctypes.c_void_p * len(channels))(*[ch._solver.value for ch in channels])
        h = ctypes.c_void_p()
        L.call("tc_ensemble_create", arr, len(channels), ctypes.byref(h))
        self._h = h
        self._graph_steps = None
        self.form = "phased"
        self.set_form(form)

    def set_form(self, form):
        """'phased' (one launch per phase).  'persistent' (one launch per step or chunk) is built
        but switched off in the library and raises."""
        if form not in FORMS:
            raise ValueError(f"Invalid ensemble form: {form}. Must be one of {sorted(FORMS)}")
        L.call("tc_ensemble_set_form", self._h, FORMS[form])
        if form != self.form:
            self._graph_steps = None
        self.form = form

    def check(self):
        """Wait for the ensemble's stream; raise if a persistent launch since the last check had a
        dependency wait overrun its bound (its results are then invalid)."""
        L.call("tc_ensemble_check", self._h, self._stream())

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                L.load().tc_ensemble_destroy(self._h)
        except Exception:
            pass

    @property
    def n_cells(self):
        return sum(ch.nx * ch.ny * ch.nz for ch in self.channels)

    def _stream(self):
        return self.channels[0]._stream()

    def step(self, dt=None):
        """One IMEX step of every channel; dt (host float) sets every channel's dt first."""
        c0 = self.channels[0]
        first = 1 if c0._newest < 0 else 0
        L.call("tc_ensemble_step", self._h, float(dt) if dt is not None else -1.0, first, self._stream())
        cur = 0 if first else 1 - c0._newest
        for ch in self.channels:
            ch._newest = cur
            ch.current_step += 1
            if dt is not None:
                ch.dt = float(dt)

    def run_graph(self, n_chunks, chunk_steps):
        """Replay `n_chunks` graph chunks of `chunk_steps` (even) steps of every channel."""
        if self.channels[0]._newest < 0:
            raise RuntimeError("run_graph needs one eager step first (forward-Euler start)")
        # a graph is recorded for one parity of the AB2 history (which rhs set is newest)
        key = (chunk_steps, self.channels[0]._newest)
        if self._graph_steps != key:
            L.call("tc_ensemble_capture", self._h, int(chunk_steps), self._stream())
            self._graph_steps = key
        for _ in range(n_chunks):
            L.call("tc_ensemble_replay", self._h, self._stream())
        for ch in self.channels:
            ch.current_step += n_chunks * chunk_steps


def ensemble_configs(n=48, seed=2026, precision="fp32", nz_choices=(32, 48, 64, 96), nxy_choices=(48, 64, 96),
                     re_range=(2000.0, 4000.0), gamma_range=(2.0, 2.8), dt=0.005, perturbation=0.1):
    """Configurations of the batched-ensemble workload: per instance nz, nx = ny, Re_bulk and the
    stretching gamma drawn in that order from numpy.random.default_rng(seed); instance i is a
    perturbed parabola seeded with i."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nz = int(rng.choice(nz_choices))
        nxy = int(rng.choice(nxy_choices))
        re = float(rng.uniform(*re_range))
        gamma = float(rng.uniform(*gamma_range))
        out.append({
            "grid": {"nx": nxy, "ny": nxy, "nz": nz},
            "domain": {"Lx": 4 * math.pi, "Ly": 4 * math.pi / 3, "Lz": 2.0},
            "flow": {"Re": re, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": gamma},
            "time": {"dt": dt, "n_steps": 300},
            "initialization": {"type": "parabolic", "perturbation_intensity": perturbation, "seed": i},
            "compute": {"device": "cuda", "precision": precision, "private_plan": True},
            "output": {"save_init": False},
        })
    return out


def make_ensemble(configs, form="phased"):
    chans = [ChannelFlow(config=copy.deepcopy(c)) for c in configs]
    return ChannelEnsemble(chans, form=form)
