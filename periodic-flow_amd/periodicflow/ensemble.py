"""Batched ensembles: many independent members, of uneven grid sizes, on one GPU.

Members with the same shape and parameters share one plan (one launch per
pass for all of them).  Groups of different shapes advance together through
pf_step_multi: every stage issues each pass of all groups at once, and two
groups whose transform lengths form a compiled pair (e.g. 128^3 and 96^3
under the 3/2 rule: 192 and 144) share one launch of that pass, unless
autotune() finds the groups faster stepped one after another.  A step can
also be captured into one HIP graph and replayed.  Every member keeps its
own seed and RNG step counter, so its result does not depend on the other
members.  The reference runs one solve per MPI job; this layer has no
counterpart there.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .plan import Plan, step_multi


@dataclass(frozen=True)
class MemberSpec:
    """Parameters of one member (the Solver keywords that shape a plan)."""

    n: tuple
    dt: float
    integrator: str = "rk3"
    viscosity: float = 1e-3
    nonlinear: bool = True
    dealias: str = "3/2"
    noise_type: str = "thermal"
    noise_mag: float = 0.0
    forcing: str | None = None
    forcing_power: float = 0.1
    forcing_band: tuple = (1.0, 2.5)
    length: tuple | None = None

    def key(self, precision):
        return (tuple(self.n), self.dt, self.integrator, self.viscosity, self.nonlinear, self.dealias,
                self.noise_type, self.noise_mag, self.forcing, self.forcing_power, tuple(self.forcing_band),
                None if self.length is None else tuple(self.length), precision)


@dataclass
class Group:
    spec: MemberSpec
    plan: Plan
    state: torch.Tensor
    members: list = field(default_factory=list)  # global member ids, slot order
    seeds: list = field(default_factory=list)    # per slot


class Ensemble:
    """A set of members advanced together.

    ``members``: list of (MemberSpec, seed).  States are device tensors, one
    (batch, d, *Nf) tensor per group; ``state(i)`` is member i's view.
    """

    def __init__(self, members, precision="fp32", correlation=None, autotune=False):
        self.precision = precision
        self._tune = autotune
        self.groups: dict = {}
        self.where = []  # member id -> (group key, slot)
        order = {}
        for i, (spec, seed) in enumerate(members):
            k = spec.key(precision)
            order.setdefault(k, []).append((i, spec, seed))
        self.where = [None] * len(members)
        for k, lst in order.items():
            spec = lst[0][1]
            plan = Plan(spec.n, spec.dt, length=spec.length, precision=precision, dealias=spec.dealias,
                        integrator=spec.integrator, nonlinear=spec.nonlinear, viscosity=spec.viscosity,
                        noise_type=spec.noise_type, noise_mag=spec.noise_mag, forcing=spec.forcing,
                        forcing_power=spec.forcing_power, forcing_band=spec.forcing_band, batch=len(lst),
                        seeds=[s for _, _, s in lst])
            # without a correlation function the plan keeps C = 0 (zero
            # noise), as the reference's zero correlation_func does
            if spec.noise_type == "correlated" and spec.noise_mag != 0 and correlation is not None:
                C = torch.as_tensor(np.asarray(correlation(spec)), device="cuda", dtype=plan.complex_dtype)
                plan.set_correlation(C.contiguous())
            g = Group(spec, plan, plan.new_state(), [i for i, _, _ in lst], [s for _, _, s in lst])
            self.groups[k] = g
            for slot, (i, _, _) in enumerate(lst):
                self.where[i] = (k, slot)
        self._graph = None
        # groups of different grids step pass by pass together, with paired
        # launches (pf_step_multi); autotune() keeps this only where it is
        # not slower than stepping the groups one after another
        self.pair = True

    def __len__(self):
        return len(self.where)

    def state(self, i):
        k, slot = self.where[i]
        return self.groups[k].state[slot]

    def plan_of(self, i):
        return self.groups[self.where[i][0]].plan

    def set_fourier(self, i, u_hat):
        self.state(i).copy_(torch.as_tensor(u_hat, device="cuda", dtype=self.plan_of(i).complex_dtype))

    def points_per_step(self):
        return sum(int(np.prod(g.spec.n)) * g.plan.batch for g in self.groups.values())

    def autotune(self, reps=5):
        """Choose every group's launch plans by timing (Plan.autotune), then,
        with several groups, paired launches against the groups stepped one
        after another (``self.pair``: kept unless stepping apart is at least
        1% faster), timed on the states themselves, which are then put back
        with every member's RNG step counter.  Returns {group key: {pass
        name: chosen ms}}.  A member then equals the same member run alone
        bit for bit only under the same plan (the radix factorisation changes
        rounding)."""
        self._tune = False
        self._graph = None
        out = {k: g.plan.autotune(g.state, reps) for k, g in self.groups.items()}
        gs = list(self.groups.values())
        if len(gs) > 1 and all(g.spec.nonlinear for g in gs):
            saved = [(g.state.clone(), g.plan.step_counter, g.plan.member_steps()) for g in gs]
            self.pair = True
            self._step(1)
            t = {}
            for mode in (True, False, True, False):
                self.pair = mode
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self._step(2)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                t[mode] = min(t.get(mode, dt), dt)
            self.pair = not t[False] < 0.99 * t[True]
            for g, (st, host, msteps) in zip(gs, saved):
                g.state.copy_(st)
                g.plan.step_counter = host
                for slot, (seed, m) in enumerate(zip(g.seeds, msteps)):
                    g.plan.set_member(slot, seed, m)
        return out

    def _step(self, n_steps):
        gs = list(self.groups.values())
        if len(gs) == 1 or not self.pair:
            for g in gs:
                g.plan.step(g.state, n_steps)
            return 0
        return step_multi([g.plan for g in gs], [g.state for g in gs], n_steps)

    def step(self, n_steps=1):
        """Advance every member n_steps; returns the number of paired launches."""
        if self._tune:
            self.autotune()
        return self._step(n_steps)

    # ---------------------------------------------------------------- graphs
    def capture(self):
        """Capture one step of every group into a HIP graph (replayed by run)."""
        if self._tune:
            self.autotune()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for grp in self.groups.values():
                grp.plan.step(grp.state, 1)
        # capture records the step without running it: its host step count
        # is taken back (the device counters did not move)
        for grp in self.groups.values():
            grp.plan.add_step_counter(-1)
        self._graph = g
        return g

    def run(self, n_steps, graph=False):
        if graph:
            if self._graph is None:
                self.capture()
            for _ in range(n_steps):
                self._graph.replay()
            # each replay advanced every member's device counter by one step
            for grp in self.groups.values():
                grp.plan.add_step_counter(n_steps)
        else:
            self.step(n_steps)

    def diagnostics(self):
        """[(energy, enstrophy)] per member, FP64."""
        out = [None] * len(self.where)
        for g in self.groups.values():
            d = g.plan.diagnostics(g.state).cpu().numpy()
            for slot, i in enumerate(g.members):
                out[i] = (float(d[slot, 0]), float(d[slot, 1]))
        return out


def random_solenoidal(plan, seed, energy=0.5, kp=4.0):
    """Seeded solenoidal field with E(k) ~ k^4 exp(-2 k^2 / kp^2), total energy ``energy``.

    Built on the host in FP64 (a set-up step, not the time step) and
    returned as one member's (d, *Nf) spectrum.
    """
    rs = np.random.RandomState(seed)
    d = plan.dim
    n = plan.n
    ks = []
    for i, ni in enumerate(n):
        k = np.fft.fftfreq(ni, 1.0 / ni) if i < d - 1 else np.fft.rfftfreq(ni, 1.0 / ni)
        shp = [1] * d
        shp[i] = k.size
        ks.append(k.reshape(shp))
    kv = np.array(np.broadcast_arrays(*ks))
    k2 = np.sum(kv**2, axis=0)
    kk = np.sqrt(k2)
    u = rs.standard_normal((d,) + plan.shape_fourier) + 1j * rs.standard_normal((d,) + plan.shape_fourier)
    amp = np.where(kk > 0, kk**2 * np.exp(-kk**2 / kp**2), 0.0) / np.maximum(kk, 1e-30) ** ((d - 1) / 2)
    u *= amp
    # project
    p = np.sum(kv * u, axis=0) / np.where(k2 == 0, 1, k2)
    u -= kv * p
    w = np.full(plan.shape_fourier, 2.0)
    w[..., 0] = 1.0
    if n[-1] % 2 == 0:
        w[..., -1] = 1.0
    e = 0.5 * np.sum(w * np.sum(np.abs(u) ** 2, axis=0))
    u *= math.sqrt(energy / e) if e > 0 else 0.0
    return u
