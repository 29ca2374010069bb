"""Request loop with continuous batching on one GPU.

Solve requests (grid, case parameters, seed, step count) arrive over time.
Requests whose plans would match (same shape and parameters) share a pool:
one batched plan with a fixed number of slots, its members advanced
together.  The engine repeatedly

1. admits waiting requests in arrival order, into a free slot of their pool
   or a new pool, while the pools' device footprint stays within the HBM
   budget (idle pools are released to make room);
2. advances pools by chunks of steps, never past the smallest remaining
   step count among a pool's members.  Under the default policy ('srpt')
   the pool whose shortest member has the least work left goes first, with a
   chunk of about ``round_s`` seconds (from the backend's per-step cost
   estimate); further pools, in the same order, join the round while the
   estimated work of the round stays within ``overlap`` times the first
   pool's.  Under 'all' every pool advances by about ``round_s`` seconds
   each round.  The GPU backend gives each pool one of a few HIP streams, so
   the launches of different pools in a round run side by side on the
   device;
3. retires members that finished (or went non-finite), moving the pool's
   last active member into the freed slot so active members stay in
   slots [0, n).  Energies are read back only for pools with a finishing
   member and, to catch non-finite members, every ``check_every`` rounds.

A member's arithmetic and noise stream (keyed by its seed and its own step
counter) do not depend on its neighbours or on when it joined, so each
result is the same alone and in the mix.  Latency is measured from arrival
to completion, queueing included.

The reference has no counterpart (it runs one solve per MPI job); the
engine is written against a small backend interface (GpuBackend here; the
tests drive the scheduling logic with a host mock).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import logger
from .ensemble import MemberSpec


@dataclass
class Request:
    rid: int
    spec: MemberSpec
    seed: int
    steps: int
    init: object = None          # (d, *Nf) complex array, or callable(plan_like) -> array
    arrival: float = 0.0         # seconds after the run starts
    t_admit: float | None = None
    t_done: float | None = None
    status: str = "waiting"      # waiting | running | done | failed | cancelled
    result: object = None        # host copy of the final u_hat
    energy: float | None = None
    steps_done: int = 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    def points(self):
        return int(np.prod(self.spec.n)) * self.steps


@dataclass
class Pool:
    key: tuple
    spec: MemberSpec
    handle: object
    capacity: int
    footprint: int
    slots: list = field(default_factory=list)  # request per active slot, slots [0, n)
    last_used: float = 0.0


class GpuBackend:
    """Pools as batched HIP plans; states are device tensors.

    Pools are dealt round-robin to ``n_streams`` HIP streams (default
    priority; HIP gives a process 4 hardware queues by default), so the steps
    of different pools overlap on the device.  Loads, moves and read-backs
    run on the default stream; each round waits for them before stepping and
    the engine synchronises the device at the end of every round."""

    def __init__(self, precision="fp32", correlation=None, n_streams=4):
        import torch

        from .plan import Plan

        self.torch = torch
        self.Plan = Plan
        self.precision = precision
        self.correlation = correlation
        self.streams = [torch.cuda.Stream() for _ in range(max(1, int(n_streams)))]
        self._next = 0

    def step_cost(self, spec, n):
        """Estimated seconds of one step of n members: algorithmic bytes at
        4 TB/s plus launch overhead for Navier-Stokes; register-resident
        Stokes flow at 300 G point-steps/s."""
        from .perfmodel import step_bytes

        stages = {"rk3": 3, "rk4": 4}.get(spec.integrator, 2)
        if spec.nonlinear:
            b = step_bytes(spec.n, spec.dealias, self.precision, stages, batch=n)
            return b / 4.0e12 + stages * 5 * 4e-6
        return n * int(np.prod(spec.n)) / 3.0e11 + 5e-6

    def footprint(self, spec, capacity):
        """Device bytes of a pool: plan workspace + state (estimated before creation)."""
        d = len(spec.n)
        c = 16 if self.precision == "fp64" else 8
        nh = spec.n[-1] // 2 + 1
        nk = int(np.prod(spec.n[:-1])) * nh
        M = [n if spec.dealias in ("2/3", "truncate") else 3 * n // 2 for n in spec.n]
        if spec.nonlinear:
            if d == 3:
                w = (5 * M[0] * spec.n[1] * nh + 6 * M[0] * M[1] * nh) * c
            else:
                w = (3 * M[0] * nh + 2 * M[0] * nh) * c
        else:
            w = d * nk * c
        return int(capacity * (w + 3 * d * nk * c))

    def create(self, spec, capacity):
        p = self.Plan(spec.n, spec.dt, length=spec.length, precision=self.precision, dealias=spec.dealias,
                      integrator=spec.integrator, nonlinear=spec.nonlinear, viscosity=spec.viscosity,
                      noise_type=spec.noise_type, noise_mag=spec.noise_mag, forcing=spec.forcing,
                      forcing_power=spec.forcing_power, forcing_band=spec.forcing_band, batch=capacity)
        if spec.noise_type == "correlated" and spec.noise_mag != 0 and self.correlation is not None:
            C = self.torch.as_tensor(np.asarray(self.correlation(spec)), device="cuda", dtype=p.complex_dtype)
            p.set_correlation(C.contiguous())
        p.active = 0
        st = self.streams[self._next % len(self.streams)]
        self._next += 1
        return {"plan": p, "state": p.new_state(), "stream": st}

    def load(self, h, slot, req):
        """The request's initial spectrum (a host array, or a device tensor
        its init callable built on the GPU) into slot ``slot``."""
        p = h["plan"]
        init = req.init(p) if callable(req.init) else req.init
        t = init if isinstance(init, self.torch.Tensor) else self.torch.as_tensor(np.asarray(init))
        h["state"][slot].copy_(t.to(device="cuda", dtype=p.complex_dtype))
        p.set_member(slot, req.seed, 0)

    def move(self, h, src, dst, req):
        """Slot src -> dst (the request keeps its seed and step counter)."""
        h["state"][dst].copy_(h["state"][src])
        h["plan"].set_member(dst, req.seed, req.steps_done)

    def set_active(self, h, n):
        h["plan"].active = n

    def step(self, h, k):
        s = h["stream"]
        s.wait_stream(self.torch.cuda.default_stream())
        with self.torch.cuda.stream(s):
            h["plan"].step(h["state"], k)

    def energies(self, h, n):
        if n == 0:
            return np.zeros(0)
        return h["plan"].diagnostics(h["state"]).cpu().numpy()[:n, 0]

    def fetch(self, h, slot):
        return h["state"][slot].detach().cpu().numpy().copy()

    def sync(self):
        self.torch.cuda.synchronize()

    def release(self, h):
        self.torch.cuda.synchronize()
        h.clear()


class Engine:
    """``chunk``: most steps a pool advances per round; ``round_s``: target
    device time of a pool's share of a round (with a backend step_cost);
    ``overlap``: under 'srpt', how much estimated work the round may hold
    relative to its first pool's; ``check_every``: rounds between energy
    read-backs of pools with no finishing member; ``policy``: 'srpt' or
    'all' (see the module docstring)."""

    def __init__(self, backend, hbm_budget=64 << 30, chunk=200, max_slots=64, keep_results=True,
                 clock=time.perf_counter, policy="srpt", round_s=0.05, check_every=8, overlap=1.25,
                 log_every=0):
        self.be = backend
        self.budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.max_slots = int(max_slots)
        self.keep_results = keep_results
        self.clock = clock
        self.pools: dict = {}
        self.waiting: deque = deque()
        self.finished: list = []
        self.t0 = None
        self.in_use = 0
        self.rounds = 0
        self.round_s = float(round_s)
        self.check_every = max(1, int(check_every))
        self.overlap = float(overlap)
        self.log_every = int(log_every)  # report every n-th finished request (0: never)
        if policy not in ("srpt", "all"):
            raise ValueError("policy must be 'srpt' or 'all'")
        self.policy = policy

    # ------------------------------------------------------------ admission
    def _capacity(self, spec):
        one = self.be.footprint(spec, 1)
        return int(max(1, min(self.max_slots, self.budget // 4 // max(one, 1))))

    def _key(self, spec):
        return spec.key(getattr(self.be, "precision", "fp32"))

    def _free_idle(self, need):
        idle = sorted((p for p in self.pools.values() if not p.slots), key=lambda p: p.last_used)
        for p in idle:
            if self.budget - self.in_use >= need:
                break
            self.be.release(p.handle)
            self.in_use -= p.footprint
            del self.pools[p.key]

    def _admit(self, req, now):
        key = self._key(req.spec)
        pool = self.pools.get(key)
        if pool is None:
            cap = self._capacity(req.spec)
            fp = self.be.footprint(req.spec, cap)
            while cap > 1 and fp > self.budget:
                cap //= 2
                fp = self.be.footprint(req.spec, cap)
            if fp > self.budget:
                req.status = "failed"
                req.t_done = now
                return True  # can never fit: reject
            if self.budget - self.in_use < fp:
                self._free_idle(fp)
            if self.budget - self.in_use < fp:
                return False
            pool = Pool(key, req.spec, self.be.create(req.spec, cap), cap, fp)
            self.pools[key] = pool
            self.in_use += fp
        if len(pool.slots) >= pool.capacity:
            return False
        slot = len(pool.slots)
        self.be.load(pool.handle, slot, req)
        pool.slots.append(req)
        self.be.set_active(pool.handle, len(pool.slots))
        req.status = "running"
        req.t_admit = now
        req.steps_done = 0
        return True

    def submit(self, req):
        self.waiting.append(req)

    def cancel(self, rid):
        for r in list(self.waiting):
            if r.rid == rid:
                self.waiting.remove(r)
                r.status = "cancelled"
                r.t_done = self._now()
                self.finished.append(r)
                return True
        for pool in self.pools.values():
            for slot, r in enumerate(pool.slots):
                if r.rid == rid:
                    r.status = "cancelled"
                    self._retire(pool, slot, keep=False)
                    return True
        return False

    def _now(self):
        return self.clock() - self.t0 if self.t0 is not None else 0.0

    # ------------------------------------------------------------ stepping
    def _retire(self, pool, slot, keep=True):
        r = pool.slots[slot]
        if keep and self.keep_results:
            r.result = self.be.fetch(pool.handle, slot)
        last = len(pool.slots) - 1
        if slot != last:
            moved = pool.slots[last]
            self.be.move(pool.handle, last, slot, moved)
            pool.slots[slot] = moved
        pool.slots.pop()
        self.be.set_active(pool.handle, len(pool.slots))
        r.t_done = self._now()
        self.finished.append(r)
        pool.last_used = r.t_done
        if self.log_every and len(self.finished) % self.log_every == 0:
            logger.info(f"serving: {len(self.finished)} requests finished at {r.t_done:.1f} s")

    def _remaining(self, pool):
        """Work left before the pool's shortest member finishes: estimated
        seconds with a backend step_cost, else point-steps."""
        steps = min((r.steps - r.steps_done) for r in pool.slots)
        cost = getattr(self.be, "step_cost", None)
        if cost is not None:
            return steps * cost(pool.spec, len(pool.slots))
        return steps * int(np.prod(pool.spec.n))

    def _steps_this_round(self, pool):
        k = self.chunk
        cost = getattr(self.be, "step_cost", None)
        if cost is not None:
            k = min(k, max(1, int(self.round_s / max(cost(pool.spec, len(pool.slots)), 1e-9))))
        return min(k, min(r.steps - r.steps_done for r in pool.slots))

    def _round(self):
        """[(pool, steps)] to advance this round."""
        act = [p for p in self.pools.values() if p.slots]
        if self.policy == "all":
            return [(p, self._steps_this_round(p)) for p in act]
        act.sort(key=self._remaining)
        if not act:
            return []
        head = act[0]
        out = [(head, self._steps_this_round(head))]
        cost = getattr(self.be, "step_cost", None)
        if cost is None:
            return out
        used = out[0][1] * cost(head.spec, len(head.slots))
        limit = self.overlap * used
        for p in act[1:]:
            k = self._steps_this_round(p)
            c = k * cost(p.spec, len(p.slots))
            if used + c <= limit:
                out.append((p, k))
                used += c
        return out

    def _advance(self):
        rnd = self._round()
        if not rnd:
            return False
        act = [p for p, _ in rnd]
        for pool, k in rnd:
            if k > 0:
                self.be.step(pool.handle, k)
                for r in pool.slots:
                    r.steps_done += k
        self.be.sync()
        self.rounds += 1
        check = self.rounds % self.check_every == 0
        for pool in act:
            n = len(pool.slots)
            if not (check or any(r.steps_done >= r.steps for r in pool.slots)):
                continue
            e = self.be.energies(pool.handle, n)
            # retire from the back so slot indices stay valid
            for slot in range(n - 1, -1, -1):
                r = pool.slots[slot]
                if not np.isfinite(e[slot]):
                    r.status = "failed"
                    r.energy = float(e[slot])
                    self._retire(pool, slot, keep=False)
                elif r.steps_done >= r.steps:
                    r.status = "done"
                    r.energy = float(e[slot])
                    self._retire(pool, slot)
        return True

    def run(self, requests, realtime=True, timeout=None, serial=False):
        """Serve ``requests`` (arrival times in seconds from now); returns stats.
        serial=True admits a request only when nothing is running: the
        one-at-a-time baseline through the same code."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.t0 = self.clock()
        while pending or self.waiting or any(p.slots for p in self.pools.values()):
            now = self._now()
            if timeout is not None and now > timeout:
                break
            while pending and (pending[0].arrival <= now or not realtime):
                self.waiting.append(pending.popleft())
            for r in list(self.waiting):
                if serial and any(p.slots for p in self.pools.values()):
                    break
                if self._admit(r, self._now()):
                    self.waiting.remove(r)
                    if r.status == "failed":
                        self.finished.append(r)
            if not self._advance() and pending and realtime:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - self._now())))
        return self.stats()

    def stats(self):
        done = [r for r in self.finished if r.status == "done"]
        lat = np.array([r.latency for r in done]) if done else np.zeros(0)
        span = max((r.t_done for r in self.finished), default=0.0)
        pts = sum(r.points() for r in done)
        return {
            "requests": len(self.finished),
            "done": len(done),
            "failed": sum(r.status == "failed" for r in self.finished),
            "cancelled": sum(r.status == "cancelled" for r in self.finished),
            "makespan_s": span,
            "Mpoint_steps_per_s": pts / span / 1e6 if span > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
        }


def serve_one_at_a_time(backend, requests, hbm_budget=64 << 30, chunk=1 << 20, keep_results=True, log_every=0):
    """The plain baseline: each request alone (batch 1) in arrival order,
    its whole step count in one call, through the same engine; latency
    includes queueing."""
    eng = Engine(backend, hbm_budget=hbm_budget, chunk=chunk, max_slots=1, round_s=1e9, keep_results=keep_results,
                 log_every=log_every)
    return eng.run(requests, serial=True), eng.finished
