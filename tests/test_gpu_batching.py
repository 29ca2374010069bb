"""Batched ensembles, graph replay and the request loop on the GPU.

* A member's result is bitwise the same alone (batch 1) and in a batch of
  uneven-size members, with on-device noise keyed by (seed, member step).
* A captured HIP graph of a whole step, replayed several times with eager
  work between replays, matches the eager path exactly.
* The request loop returns every result bitwise equal to the same request
  run alone.
* Fluctuating Stokes flow (thermal noise, implicit PC, FP32): the static
  structure factor is flat and at the closed-form level.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-batch]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from periodicflow import ensemble, plan, serving

    return ensemble, plan, serving


def _specs(ensemble):
    S = ensemble.MemberSpec
    return [
        (S(n=(16, 16, 16), dt=0.005, integrator="rk3", viscosity=0.01, noise_type="thermal", noise_mag=0.01,
           forcing="constant_power"), 11),
        (S(n=(12, 12, 12), dt=0.005, integrator="rk3", viscosity=0.01, noise_type="thermal", noise_mag=0.01,
           forcing="constant_power"), 12),
        (S(n=(16, 16, 16), dt=0.005, integrator="rk3", viscosity=0.01, noise_type="thermal", noise_mag=0.01,
           forcing="constant_power"), 13),
        (S(n=(32, 32), dt=0.01, integrator="implicit_pc", viscosity=0.5, nonlinear=False, noise_mag=0.2), 14),
        (S(n=(12, 12, 12), dt=0.005, integrator="rk3", viscosity=0.01, noise_type="thermal", noise_mag=0.01,
           forcing="constant_power"), 15),
    ]


def test_batched_equals_alone(mods):
    ensemble, plan, _ = mods
    _say("batched vs alone")
    members = _specs(ensemble)
    ens = ensemble.Ensemble(members, precision="fp32")
    assert len(ens.groups) == 3
    inits = []
    for i, (spec, seed) in enumerate(members):
        u = ensemble.random_solenoidal(ens.plan_of(i), 500 + i, energy=0.3)
        inits.append(u)
        ens.set_fourier(i, u)
    ens.step(6)
    for i, (spec, seed) in enumerate(members):
        p = plan.Plan(spec.n, spec.dt, precision="fp32", integrator=spec.integrator, viscosity=spec.viscosity,
                      nonlinear=spec.nonlinear, noise_type=spec.noise_type, noise_mag=spec.noise_mag,
                      forcing=spec.forcing, seeds=[seed])
        st = torch.as_tensor(inits[i][None], device="cuda", dtype=p.complex_dtype).contiguous()
        p.step(st, 6)
        assert torch.equal(st[0], ens.state(i)), i


def test_graph_replay_matches_eager(mods):
    ensemble, _, _ = mods
    _say("graph replay")
    members = _specs(ensemble)
    a = ensemble.Ensemble(members, precision="fp32")
    b = ensemble.Ensemble(members, precision="fp32")
    for i in range(len(members)):
        u = ensemble.random_solenoidal(a.plan_of(i), 700 + i, energy=0.3)
        a.set_fourier(i, u)
        b.set_fourier(i, u)
    b.capture()
    for rep in range(4):
        a.run(2)
        b.run(2, graph=True)
        torch.cuda.synchronize()
        # eager work between replays
        _ = [g.plan.diagnostics(g.state) for g in b.groups.values()]
        torch.cuda.synchronize()
        for i in range(len(members)):
            assert torch.equal(a.state(i), b.state(i)), (rep, i)
        # host and device step counters agree with the eager run's
        for ka, kb in zip(a.groups, b.groups):
            pa, pb = a.groups[ka].plan, b.groups[kb].plan
            assert pa.step_counter == pb.step_counter == 2 * (rep + 1)
            assert pa.member_steps() == pb.member_steps()


def test_serving_results_equal_alone(mods):
    ensemble, _, serving = mods
    from periodicflow import workloads
    _say("serving")
    S = ensemble.MemberSpec
    kinds = [
        S(n=(16, 16, 16), dt=0.01, integrator="rk3", viscosity=0.02),
        S(n=(32, 32), dt=0.05, integrator="implicit_pc", viscosity=1.0, nonlinear=False, noise_mag=0.3),
        S(n=(24, 24), dt=0.01, integrator="rk3", viscosity=0.01, noise_mag=0.01, forcing="constant_power"),
    ]
    rs = np.random.RandomState(1)
    reqs = []
    for i in range(14):
        spec = kinds[i % 3]
        # host-built and device-built initial fields alternate
        init = ((lambda p, s=i: ensemble.random_solenoidal(p, 900 + s, energy=0.2)) if i % 2 else
                (lambda p, s=i: workloads.solenoidal_device(p, 900 + s, energy=0.2)))
        reqs.append(serving.Request(i, spec, seed=3000 + i, steps=int(rs.randint(5, 40)), init=init,
                                    arrival=0.002 * i))
    be = serving.GpuBackend("fp32")
    eng = serving.Engine(be, hbm_budget=8 << 30, chunk=8, max_slots=4)
    st = eng.run(reqs)
    _say(st)
    assert st["done"] == 14 and st["p99_latency_s"] is not None
    alone = [serving.Request(r.rid, r.spec, r.seed, r.steps, r.init, 0.0) for r in reqs]
    st1, fin = serving.serve_one_at_a_time(serving.GpuBackend("fp32"), alone)
    by = {r.rid: r for r in fin}
    for r in reqs:
        assert np.array_equal(r.result, by[r.rid].result), r.rid


def test_structure_factor_flat(mods):
    """Thermal-noise Stokes, implicit PC: S(k) = nm^2 d / (2 nu prod N) for every k."""
    _, plan, _ = mods
    _say("structure factor")
    n, B, nu, dt, kT = 64, 64, 1.0, 0.05, 1e-3
    dx = 2 * np.pi / n
    nm = np.sqrt(2 * nu * kT / dx**2)
    p = plan.Plan((n, n), dt, precision="fp32", integrator="implicit_pc", nonlinear=False, viscosity=nu,
                  noise_mag=nm, batch=B, seeds=[1000 + r for r in range(B)])
    u = p.new_state()
    p.step(u, 200)
    acc = torch.zeros(p.shape_fourier, dtype=torch.float64, device="cuda")
    nsamp = 400
    p.step_accumulate_sk(u, nsamp, acc)
    S = acc.cpu().numpy() / (nsamp * B)
    theory = nm**2 * 2 / (2 * nu * n * n)
    k0 = np.fft.fftfreq(n, 1.0 / n)[:, None]
    k1 = np.arange(n // 2 + 1)[None, :]
    kk = np.sqrt(k0**2 + k1**2)
    sel = kk > 0
    ratio = S[sel] / theory
    _say("S(k)/theory mean", ratio.mean(), "std", ratio.std())
    assert abs(ratio.mean() - 1) < 0.02
    shells = np.rint(kk[sel]).astype(int)
    for s in range(2, n // 2):
        m = shells == s
        if m.sum() >= 50:
            assert abs(ratio[m].mean() - 1) < 0.05, s


# (256, 256) with 140 members runs 8 members per thread with a partial last group
@pytest.mark.parametrize("N,ntype,prec,B", [((256, 256), "thermal", "fp32", 140), ((64, 64), "thermal", "fp64", 3),
                                           ((16, 16, 16), "correlated", "fp32", 5),
                                           ((12, 16, 8), "thermal", "fp64", 2)])
def test_register_resident_stokes_equals_per_step(mods, N, ntype, prec, B):
    """Implicit-PC Stokes steps with the state in registers (default) give the
    state of one launch per step (stage variant 1) bit for bit, and the FP64
    structure factor summed in registers matches per-step accumulation."""
    _, plan, _ = mods
    _say("resident stokes", N, ntype, prec, B)
    kw = dict(precision=prec, integrator="implicit_pc", nonlinear=False, viscosity=0.7, noise_type=ntype,
              noise_mag=0.4, batch=B, seeds=[50 + b for b in range(B)])
    pr = plan.Plan(N, 0.03, **kw)
    ps = plan.Plan(N, 0.03, **kw)
    ps.set_tiles({"stage": 1})
    if ntype == "correlated":
        C = torch.rand(pr.shape_fourier, dtype=pr.real_dtype, device="cuda").to(pr.complex_dtype)
        pr.set_correlation(C)
        ps.set_correlation(C)
    rs = np.random.RandomState(3)
    u0 = rs.standard_normal(pr.state_shape()) + 1j * rs.standard_normal(pr.state_shape())
    ur = torch.from_numpy(u0).to("cuda", pr.complex_dtype)
    us = ur.clone()
    pr.step(ur, 7)
    ps.step(us, 7)
    assert torch.equal(ur, us)
    acc_r = torch.zeros(pr.shape_fourier, dtype=torch.float64, device="cuda")
    acc_s = torch.zeros_like(acc_r)
    pr.step_accumulate_sk(ur, 230, acc_r)  # crosses the 200-step launch chunk
    for _ in range(230):
        ps.step(us, 1)
        ps.accumulate_sk(us, acc_s)
    assert torch.equal(ur, us)
    assert pr.step_counter == ps.step_counter == 237
    assert pr.member_steps() == ps.member_steps() == [237] * B
    rel = float(torch.max(torch.abs(acc_r - acc_s) / acc_s.clamp_min(1e-300)))
    _say("   S(k) accumulation max rel diff", rel)
    assert rel < 1e-12


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_paired_launches_equal_separate_plans(mods, prec):
    """An ensemble of a 128^3 and a 96^3 group (3/2 rule, RK3, thermal noise,
    forcing): every nonlinear pass of both groups in one paired launch (FP32;
    FP64 has no paired kernels and steps the groups in turn), results bitwise
    those of each plan stepped alone."""
    ensemble, plan, _ = mods
    _say("paired launches", prec)
    S = ensemble.MemberSpec
    members = [(S(n=(128,) * 3, dt=5e-3, integrator="rk3", viscosity=0.005, noise_mag=0.01,
                  forcing="constant_power"), 21),
               (S(n=(96,) * 3, dt=5e-3, integrator="rk3", viscosity=0.005, noise_mag=0.01,
                  forcing="constant_power"), 22),
               (S(n=(96,) * 3, dt=5e-3, integrator="rk3", viscosity=0.005, noise_mag=0.01,
                  forcing="constant_power"), 23)]
    ens = ensemble.Ensemble(members, precision=prec)
    inits = []
    for i, (spec, seed) in enumerate(members):
        u = ensemble.random_solenoidal(ens.plan_of(i), 40 + i, energy=0.5)
        inits.append(u)
        ens.set_fourier(i, u)
    paired = ens.step(3)
    # 3 steps x 3 stages x 5 passes, all paired in FP32
    assert paired == (45 if prec == "fp32" else 0), paired
    for i, (spec, seed) in enumerate(members):
        p = plan.Plan(spec.n, spec.dt, precision=prec, integrator="rk3", viscosity=spec.viscosity,
                      noise_mag=spec.noise_mag, forcing=spec.forcing, seeds=[seed])
        st = torch.as_tensor(inits[i][None], device="cuda", dtype=p.complex_dtype).contiguous()
        p.step(st, 3)
        assert torch.equal(st[0], ens.state(i)), i
        assert ens.plan_of(i).step_counter == 3


def test_ensemble_autotune_leaves_members_and_counters(mods):
    """Ensemble.autotune times paired against separate stepping on the states
    themselves: afterwards every member and RNG step counter is as before,
    and a step equals each member's plan stepped alone (same plan)."""
    ensemble, plan, _ = mods
    _say("ensemble autotune restores states")
    S = ensemble.MemberSpec
    members = [(S(n=(32,) * 3, dt=5e-3, integrator="rk3", viscosity=0.01, noise_mag=0.01), 41),
               (S(n=(24,) * 3, dt=5e-3, integrator="rk3", viscosity=0.01, noise_mag=0.01), 42)]
    ens = ensemble.Ensemble(members, precision="fp32")
    for i in range(2):
        ens.set_fourier(i, ensemble.random_solenoidal(ens.plan_of(i), 60 + i, energy=0.5))
    ens.step(1)
    before = [ens.state(i).clone() for i in range(2)]
    counters = [(ens.plan_of(i).step_counter, ens.plan_of(i).member_steps()) for i in range(2)]
    ens.autotune(reps=2)
    assert ens.pair in (True, False)
    for i in range(2):
        assert torch.equal(ens.state(i), before[i]), i
        assert (ens.plan_of(i).step_counter, ens.plan_of(i).member_steps()) == counters[i]
    ens.step(2)
    for i, (spec, seed) in enumerate(members):
        tiles, fused = ens.plan_of(i).tiles
        p = plan.Plan(spec.n, spec.dt, precision="fp32", integrator="rk3", viscosity=spec.viscosity,
                      noise_mag=spec.noise_mag, seeds=[seed])
        p.set_tiles(tiles)
        p.set_fusion(fused)
        p.step_counter = 1
        st = before[i][None].clone().contiguous()
        p.step(st, 2)
        assert torch.equal(st[0], ens.state(i)), i


@pytest.mark.parametrize("fused", [True, False])
def test_paired_launches_cache_policy(mods, fused):
    """Paired launches at the nontemporal launch variants (both groups at the
    same variant of every pass): still one paired launch per pass, and the
    members equal each plan stepped alone at the default variants.  Unfused:
    the axis-0 forward passes pair and each group's stage kernel follows."""
    ensemble, plan, _ = mods
    _say("paired launches, nontemporal variants, fused", fused)
    S = ensemble.MemberSpec
    members = [(S(n=(128,) * 3, dt=5e-3, integrator="rk3", viscosity=0.005, noise_mag=0.01), 31),
               (S(n=(96,) * 3, dt=5e-3, integrator="rk3", viscosity=0.005, noise_mag=0.01), 32)]
    ens = ensemble.Ensemble(members, precision="fp32")
    inits = []
    for i, (spec, seed) in enumerate(members):
        u = ensemble.random_solenoidal(ens.plan_of(i), 50 + i, energy=0.5)
        inits.append(u)
        ens.set_fourier(i, u)
    nt = {"inv_axis0_curl": 6, "inv_axis1": 7, "row_c2r_cross_r2c": 4, "fwd_axis1": 6, "fwd_axis0_stage": 6,
          "fwd_axis0": 7, "stage": 7}
    for g in ens.groups.values():
        g.plan.set_tiles(nt)
        g.plan.set_fusion(fused)
    paired = ens.step(2)
    assert paired == 30, paired  # 2 steps x 3 stages x 5 passes
    for i, (spec, seed) in enumerate(members):
        p = plan.Plan(spec.n, spec.dt, precision="fp32", integrator="rk3", viscosity=spec.viscosity,
                      noise_mag=spec.noise_mag, seeds=[seed])
        p.set_fusion(fused)
        st = torch.as_tensor(inits[i][None], device="cuda", dtype=p.complex_dtype).contiguous()
        p.step(st, 2)
        assert torch.equal(st[0], ens.state(i)), i
