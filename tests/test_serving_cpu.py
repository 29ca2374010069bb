"""Scheduling logic of the request loop, driven from the host with a mock backend.

The mock advances a member deterministically from (seed, step), like the
device kernels do, so the engine's slot moves, admission, retirement and
failure isolation can be checked exactly without a GPU.
"""
import numpy as np
import pytest

from periodicflow.ensemble import MemberSpec
from periodicflow.serving import Engine, Request, serve_one_at_a_time


class MockBackend:
    precision = "fp32"

    def __init__(self, bytes_per_member=1000):
        self.bpm = bytes_per_member
        self.created = 0
        self.log = []

    def footprint(self, spec, capacity):
        return capacity * self.bpm * int(np.prod(spec.n))

    def create(self, spec, capacity):
        self.created += 1
        return {"x": np.zeros(capacity), "seed": np.zeros(capacity, dtype=np.int64),
                "step": np.zeros(capacity, dtype=np.int64), "n": 0}

    def load(self, h, slot, req):
        init = req.init if req.init is not None else 1.0
        h["x"][slot] = init
        h["seed"][slot] = req.seed
        h["step"][slot] = 0

    def move(self, h, src, dst, req):
        h["x"][dst] = h["x"][src]
        h["seed"][dst] = req.seed
        h["step"][dst] = req.steps_done
        assert h["step"][src] == req.steps_done

    def set_active(self, h, n):
        h["n"] = n

    def step(self, h, k):
        for _ in range(k):
            for s in range(h["n"]):
                # depends only on the member's own (seed, step)
                h["x"][s] = 0.9 * h["x"][s] + 1e-3 * ((h["seed"][s] * 7919 + h["step"][s] * 104729) % 1000)
                h["step"][s] += 1
        self.log.append(("step", h["n"], k))

    def energies(self, h, n):
        return h["x"][:n].copy()

    def fetch(self, h, slot):
        return float(h["x"][slot])

    def sync(self):
        pass

    def release(self, h):
        h.clear()


def _alone(seed, steps, init=1.0):
    x, st = init, 0
    for _ in range(steps):
        x = 0.9 * x + 1e-3 * ((seed * 7919 + st * 104729) % 1000)
        st += 1
    return x


def _reqs(n=12, rs=None):
    rs = rs or np.random.RandomState(0)
    sizes = [(8, 8), (16, 16), (8, 8, 8)]
    out = []
    for i in range(n):
        spec = MemberSpec(n=sizes[i % 3], dt=0.01)
        out.append(Request(i, spec, seed=100 + i, steps=int(rs.randint(5, 60)), arrival=0.0))
    return out


def test_results_equal_alone_and_in_the_mix():
    be = MockBackend()
    eng = Engine(be, hbm_budget=10**9, chunk=7, max_slots=4)
    reqs = _reqs(15)
    st = eng.run(reqs, realtime=False)
    assert st["done"] == 15 and st["failed"] == 0
    for r in reqs:
        assert r.result == pytest.approx(_alone(r.seed, r.steps), rel=0, abs=0)
    # several members really shared a launch
    assert max(n for _, n, _ in be.log) > 1


def test_admission_is_bounded_by_footprint():
    be = MockBackend(bytes_per_member=1000)
    # one 16x16 member = 256 kB; budget admits one pool of capacity 1 at a time
    spec = MemberSpec(n=(16, 16), dt=0.01)
    budget = 300_000
    eng = Engine(be, hbm_budget=budget, chunk=5, max_slots=8)
    reqs = [Request(i, spec, seed=i, steps=10) for i in range(4)]
    st = eng.run(reqs, realtime=False)
    assert st["done"] == 4
    assert all(n <= 1 for _, n, _ in be.log)
    assert eng.in_use <= budget
    # a request that can never fit is rejected, not queued forever
    big = Request(99, MemberSpec(n=(64, 64), dt=0.01), seed=1, steps=3)
    eng2 = Engine(MockBackend(), hbm_budget=budget)
    st2 = eng2.run([big], realtime=False)
    assert st2["failed"] == 1 and big.status == "failed"


def test_nan_member_fails_alone():
    be = MockBackend()
    eng = Engine(be, hbm_budget=10**9, chunk=4, max_slots=8)
    spec = MemberSpec(n=(8, 8), dt=0.01)
    reqs = [Request(i, spec, seed=i, steps=12, init=(np.nan if i == 2 else 1.0)) for i in range(5)]
    st = eng.run(reqs, realtime=False)
    assert st["failed"] == 1 and reqs[2].status == "failed"
    for r in reqs:
        if r.rid != 2:
            assert r.status == "done"
            assert r.result == _alone(r.seed, r.steps)


def test_cancel_waiting_and_running():
    be = MockBackend()
    eng = Engine(be, hbm_budget=10**9, chunk=3, max_slots=2)
    spec = MemberSpec(n=(8, 8), dt=0.01)
    reqs = [Request(i, spec, seed=i, steps=30) for i in range(4)]
    for r in reqs:
        eng.submit(r)
    eng.t0 = 0.0
    for r in list(eng.waiting):
        if eng._admit(r, 0.0):
            eng.waiting.remove(r)
    assert reqs[0].status == reqs[1].status == "running" and reqs[2].status == "waiting"
    eng._advance()
    assert eng.cancel(reqs[0].rid) and reqs[0].status == "cancelled"
    assert eng.cancel(reqs[3].rid) and reqs[3].status == "cancelled"
    assert not eng.cancel(12345)
    st = eng.run([], realtime=False)
    assert reqs[1].status == "done" and reqs[1].result == _alone(1, 30)
    assert reqs[2].status == "done" and reqs[2].result == _alone(2, 30)


def test_serial_baseline_runs_one_at_a_time():
    be = MockBackend()
    reqs = _reqs(6)
    st, fin = serve_one_at_a_time(be, reqs)
    assert st["done"] == 6
    assert all(n == 1 for _, n, _ in be.log)
    for r in reqs:
        assert r.result == _alone(r.seed, r.steps)


class CostedMock(MockBackend):
    """The mock with a per-step cost estimate: 1 us per grid point and member."""

    def step_cost(self, spec, n):
        return 1e-6 * n * int(np.prod(spec.n))


def test_srpt_round_takes_the_shortest_pool_first_and_fills_with_others():
    be = CostedMock()
    eng = Engine(be, hbm_budget=10**9, chunk=50, max_slots=4, round_s=0.01, overlap=1.25)
    big = MemberSpec(n=(64, 64), dt=0.01)      # 4096 us per step
    small = MemberSpec(n=(8, 8), dt=0.01)      # 64 us per step
    reqs = [Request(0, big, seed=1, steps=3), Request(1, small, seed=2, steps=40),
            Request(2, small, seed=3, steps=40)]
    for r in reqs:
        assert eng._admit(r, 0.0)
    rnd = eng._round()
    # estimated time left: big 3 x 4.1 ms, small 40 x 128 us, so the small pool
    # goes first, with round_s / cost = 0.01 / 128e-6 = 78 steps, capped at 40
    assert rnd[0][0].spec == small and rnd[0][1] == 40
    # the big pool's chunk (2 steps = 8.2 ms) would exceed 1.25 x 5.1 ms: not in this round
    assert len(rnd) == 1
    st = eng.run([], realtime=False)
    assert st["done"] == 3
    for r in reqs:
        assert r.result == _alone(r.seed, r.steps)
    eng2 = Engine(CostedMock(), hbm_budget=10**9, policy="all", round_s=0.01)
    for r in [Request(0, big, seed=1, steps=3), Request(1, small, seed=2, steps=40)]:
        eng2._admit(r, 0.0)
    assert len(eng2._round()) == 2
