#!/usr/bin/env python
"""Mpoint-steps/s of the Fourier-Galerkin time step on one MI355X.

Workload (one GPU, the plain single-instance configuration): 3D Taylor-Green
vortex on a 256^3 periodic grid, nu = 1/1600, dt = 2.5e-3, low-storage RK4,
2/3 truncation, FP32 arithmetic, synthetic initial field u = sin x cos y cos z,
v = -cos x sin y cos z, w = 0.  A step is one full time step (four RK stages,
each: nonlinear term through five FFT passes, then the fused stage update).

Prints one JSON line.  ``value`` = 256^3 * steps / seconds / 1e6 with the
state resident in HBM.  ``roofline`` prices the most expensive pass against
the 8 TB/s HBM peak using the algorithmic bytes of periodicflow.perfmodel
and the pass's mean duration from HIP events that each of its launches
records at the kernel's start and end inside the timed region.  ``cpu_baseline`` times the FP64 NumPy
oracle (scipy.fft over all host cores) on a bounded sample of the same
workload, and ``accuracy_vs_oracle`` compares the GPU's FP64 and FP32 states
with the oracle's after those steps from the same start.  The FP32 result is
also compared with the FP64 GPU path after all the timed steps
(``fp32_rel_l2_vs_fp64``).

Beside the headline: ``batched_vs_plain`` (thermal-noise Stokes 512^2 x 256
realisations; forced turbulence 32 x 64^3; forced turbulence 16 x 128^3 +
16 x 96^3 at its own sizes; each against the same members one at a time)
and ``serving`` (a bounded request stream: continuous batching against one
at a time, throughput and p50/p99 latency).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "periodic-flow_amd"))

N = 256
NU = 1.0 / 1600.0
DT = 2.5e-3


def _tgv_state(plan, torch):
    """TGV on the device: physical field -> forward transform -> projection."""
    import math

    n = plan.n[0]
    x = torch.arange(n, dtype=plan.real_dtype, device="cuda") * (2 * math.pi / n)
    X = x.view(n, 1, 1)
    Y = x.view(1, n, 1)
    Z = x.view(1, 1, n)
    u = torch.zeros(plan.physical_shape(), dtype=plan.real_dtype, device="cuda")
    u[0, 0] = torch.sin(X) * torch.cos(Y) * torch.cos(Z)
    u[0, 1] = -torch.cos(X) * torch.sin(Y) * torch.cos(Z)
    uh = plan.new_state()
    plan.forward(u, uh)
    plan.project(uh)
    del u
    return uh


def cpu_baseline(steps, torch=None):
    """The oracle timed on the host's cores; with ``torch`` also the GPU's
    FP64 and FP32 states after the same steps from the same start, against
    the oracle's."""
    from oracle import pf_oracle as po

    cores = len(os.sched_getaffinity(0))
    po.set_fft_workers(cores)
    s = po.OracleSolver([N] * 3, DT, time_integrator="rk4", viscosity=NU, dealias="2/3")
    s.initialize(po.taylor_green_3d(s.g))
    start = s.u_hat.copy()
    t0 = time.perf_counter()
    s.solve(steps)
    dt = time.perf_counter() - t0
    acc = None
    if torch is not None:
        from periodicflow.plan import Plan

        acc = {"steps": steps}
        for prec in ("fp64", "fp32"):
            p = Plan([N] * 3, DT, precision=prec, integrator="rk4", viscosity=NU, dealias="2/3")
            uh = torch.from_numpy(start[None].copy()).to("cuda", p.complex_dtype)
            p.step(uh, steps)
            acc[f"{prec}_rel_l2_vs_oracle"] = po.rel_l2(uh.cpu().numpy()[0], s.u_hat)
            del p, uh
    return acc, {
        "value": N**3 * steps / dt / 1e6,
        "unit": "Mpoint-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 NumPy/scipy.fft restatement of the reference path (CPU restatement standing in "
                  f"for the reference mpirun path), TGV 256^3 RK4 2/3, {steps} step(s) incl. the "
                  f"reference's per-step random_fields draws, {dt:.1f} s",
    }


def _rate(torch, fn, points):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return points / (time.perf_counter() - t0) / 1e6


def batched_vs_plain(torch, members=256, steps=100, plain_members=32):
    """Batched FP32 ensemble stepping vs the plain path (the same kernels at
    batch 1, one solve after another), on the 2D thermal-noise Stokes
    ensemble (512^2, implicit PC, dt = 0.05, nu = 1, kT = 1e-3, seeds
    1000 + r), on a small 3D forced-turbulence ensemble (64^3, RK3), on the
    forced 16 x 128^3 + 16 x 96^3 ensemble, and on members too small to fill
    the GPU alone (small_batched_vs_plain)."""
    import math

    from periodicflow.ensemble import random_solenoidal
    from periodicflow.plan import Plan

    out = {}
    n, nu, dt, kT = 512, 1.0, 0.05, 1e-3
    nm = math.sqrt(2 * nu * kT / (2 * math.pi / n) ** 2)
    kw = dict(precision="fp32", integrator="implicit_pc", nonlinear=False, viscosity=nu, noise_mag=nm)
    pb = Plan((n, n), dt, batch=members, seeds=[1000 + r for r in range(members)], **kw)
    ub = pb.new_state()
    pb.step(ub, 2)
    ub.zero_()
    pb.step_counter = 0
    rb = _rate(torch, lambda: pb.step(ub, steps), n * n * members * steps)
    pp = Plan((n, n), dt, batch=1, **kw)
    up = pp.new_state()
    pp.step(up, 2)

    def plain():
        for m in range(plain_members):
            pp.set_member(0, 1000 + m, 0)
            up.zero_()
            pp.step(up, steps)

    rp = _rate(torch, plain, n * n * plain_members * steps)
    p64 = Plan((n, n), dt, batch=1, seeds=[1000], **dict(kw, precision="fp64"))
    u64 = p64.new_state()
    p64.step(u64, steps)
    e = float(torch.linalg.vector_norm(ub[0].to(torch.complex128) - u64[0]) / torch.linalg.vector_norm(u64[0]))
    out["stokes2d_512_thermal_implicit_pc"] = {
        "members": members, "steps": steps, "batched": round(rb, 1), "plain": round(rp, 1),
        "plain_members_timed": plain_members, "speedup": round(rb / rp, 2), "fp32_rel_l2_vs_fp64": e}
    del pb, ub, pp, up, p64, u64

    m3, n3, st3 = 32, 64, 10
    kw3 = dict(precision="fp32", integrator="rk3", viscosity=0.005, forcing="constant_power", forcing_power=0.1)
    pb = Plan((n3,) * 3, 5e-3, batch=m3, seeds=list(range(m3)), **kw3)
    ub = pb.new_state()
    inits = [random_solenoidal(pb, 2000 + m, energy=0.5) for m in range(m3)]
    for m in range(m3):
        ub[m].copy_(torch.as_tensor(inits[m], device="cuda", dtype=pb.complex_dtype))
    u_start = ub.clone()
    pb.autotune(ub)
    pb.step(ub, 1)
    ub.copy_(u_start)
    rb = _rate(torch, lambda: pb.step(ub, st3), n3**3 * m3 * st3)
    pp = Plan((n3,) * 3, 5e-3, batch=1, **kw3)
    up = pp.new_state()

    def plain3():
        for m in range(8):
            up.copy_(u_start[m:m + 1])
            pp.step(up, st3)

    up.copy_(u_start[0:1])
    pp.autotune(up)
    pp.step(up, 1)
    rp = _rate(torch, plain3, n3**3 * 8 * st3)
    out["forced3d_64_rk3"] = {"members": m3, "steps": st3, "batched": round(rb, 1), "plain": round(rp, 1),
                              "plain_members_timed": 8, "speedup": round(rb / rp, 2)}
    del pb, ub, pp, up
    out["forced3d_16x128_16x96_rk3"] = forced_batched_vs_plain(torch)
    out.update(small_batched_vs_plain(torch))
    return out


def small_batched_vs_plain(torch):
    """Members too small to fill the GPU alone, the regime batching is for
    (the request stream's 128^2 and 256^2 grids): thermal-noise Stokes 128^2
    (implicit PC, dt = 0.05, nu = 1, kT = 1e-3) x 2048 realisations, and
    nonlinear 2D flow 256^2 (3/2 rule, RK3, nu = 1e-3, dt = 1e-3, seeded
    solenoidal starts) x 128 members; each against the same members one at a
    time, with member 0's FP32 state against FP64 after the timed steps."""
    import math

    from periodicflow.ensemble import random_solenoidal
    from periodicflow.plan import Plan

    out = {}

    def rel(a, b):
        return float(torch.linalg.vector_norm(a.to(torch.complex128) - b) / torch.linalg.vector_norm(b))

    # thermal-noise Stokes 128^2
    n, nu, dt, kT, B, steps, pm = 128, 1.0, 0.05, 1e-3, 2048, 200, 16
    nm = math.sqrt(2 * nu * kT / (2 * math.pi / n) ** 2)
    kw = dict(integrator="implicit_pc", nonlinear=False, viscosity=nu, noise_mag=nm)
    pb = Plan((n, n), dt, precision="fp32", batch=B, seeds=[1000 + r for r in range(B)], **kw)
    ub = pb.new_state()
    pb.step(ub, 2)
    ub.zero_()
    pb.step_counter = 0
    rb = _rate(torch, lambda: pb.step(ub, steps), n * n * B * steps)
    pp = Plan((n, n), dt, precision="fp32", batch=1, **kw)
    up = pp.new_state()
    pp.step(up, 2)

    def plain():
        for m in range(pm):
            pp.set_member(0, 1000 + m, 0)
            up.zero_()
            pp.step(up, steps)

    rp = _rate(torch, plain, n * n * pm * steps)
    p64 = Plan((n, n), dt, precision="fp64", batch=1, seeds=[1000], **kw)
    u64 = p64.new_state()
    p64.step(u64, steps)
    out["stokes2d_128_thermal_implicit_pc"] = {
        "members": B, "steps": steps, "batched": round(rb, 1), "plain": round(rp, 1), "plain_members_timed": pm,
        "speedup": round(rb / rp, 2), "fp32_rel_l2_vs_fp64": rel(ub[0], u64[0])}
    del pb, ub, pp, up, p64, u64

    # nonlinear 2D 256^2
    n, B, steps, pm = 256, 128, 50, 8
    kw = dict(integrator="rk3", viscosity=1e-3, dealias="3/2")
    pb = Plan((n, n), 1e-3, precision="fp32", batch=B, **kw)
    ub = pb.new_state()
    inits = [random_solenoidal(pb, 500 + m, energy=0.5, kp=8.0) for m in range(B)]
    for m in range(B):
        ub[m].copy_(torch.as_tensor(inits[m], device="cuda", dtype=pb.complex_dtype))
    u_start = ub.clone()
    pb.autotune(ub)
    pb.step(ub, 1)
    ub.copy_(u_start)
    rb = _rate(torch, lambda: pb.step(ub, steps), n * n * B * steps)
    pp = Plan((n, n), 1e-3, precision="fp32", batch=1, **kw)
    up = pp.new_state()
    up.copy_(u_start[0:1])
    pp.autotune(up)
    pp.step(up, 1)

    def plain2():
        for m in range(pm):
            up.copy_(u_start[m:m + 1])
            pp.step(up, steps)

    rp = _rate(torch, plain2, n * n * pm * steps)
    p64 = Plan((n, n), 1e-3, precision="fp64", batch=1, **kw)
    u64 = torch.as_tensor(inits[0][None], device="cuda", dtype=p64.complex_dtype).contiguous()
    p64.step(u64, steps)
    out["nonlinear2d_256_rk3"] = {
        "members": B, "steps": steps, "batched": round(rb, 1), "plain": round(rp, 1), "plain_members_timed": pm,
        "speedup": round(rb / rp, 2), "fp32_rel_l2_vs_fp64": rel(ub[0], u64[0])}
    return out


def forced_batched_vs_plain(torch, steps=5):
    """The forced-turbulence ensemble at its sizes: 16 members at 128^3 and
    16 at 96^3 (3/2 rule, RK3, correlated noise, constant-power forcing,
    FP32), launch plans autotuned, against one member of each size alone
    (batch 1, also autotuned) timed for the same steps and counted 16 times."""
    from periodicflow import workloads as W
    from periodicflow.ensemble import Ensemble, random_solenoidal

    members = W.forced_members()

    def build(ms):
        e = Ensemble(ms, precision="fp32", correlation=W.forced_correlation)
        for i, (spec, seed) in enumerate(ms):
            e.set_fourier(i, random_solenoidal(e.plan_of(i), seed, energy=0.5, kp=4.0))
        e.autotune()
        e.step(1)
        return e

    ens = build(members)
    pts = ens.points_per_step() * steps
    rb = _rate(torch, lambda: ens.step(steps), pts)
    del ens
    t_plain = 0.0
    for m in (members[0], members[16]):
        e1 = build([m])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e1.step(steps)
        torch.cuda.synchronize()
        t_plain += 16 * (time.perf_counter() - t0)
        del e1
    rp = pts / t_plain / 1e6
    return {"members": len(members), "steps": steps, "batched": round(rb, 1), "plain": round(rp, 1),
            "speedup": round(rb / rp, 2)}


def serving_bench(torch, count=40, max_steps=300):
    """A bounded run of the request stream (Poisson arrivals at 20/s, seed
    42; grids 128^2..1024^2 and 64^3..256^3; Taylor-Green, thermal Stokes or
    forced turbulence; steps log-uniform in [100, max_steps]; FP32):
    continuous batching against the same requests one at a time."""
    from periodicflow import workloads as W
    from periodicflow.serving import Engine, GpuBackend, serve_one_at_a_time

    reqs = W.serving_requests(count=count, max_steps=max_steps)
    batched = Engine(GpuBackend("fp32"), hbm_budget=96 << 30, keep_results=False).run(reqs)
    reqs = W.serving_requests(count=count, max_steps=max_steps)
    serial = serve_one_at_a_time(GpuBackend("fp32"), reqs, hbm_budget=96 << 30, keep_results=False)[0]
    keep = ("done", "makespan_s", "Mpoint_steps_per_s", "p50_latency_s", "p99_latency_s")
    return {"requests": count, "max_steps": max_steps,
            "batched": {k: batched[k] for k in keep}, "one_at_a_time": {k: serial[k] for k in keep},
            "throughput_speedup": round(batched["Mpoint_steps_per_s"] / max(serial["Mpoint_steps_per_s"], 1e-9), 2)}


def _note(*a):
    print("[suite]", *a, file=sys.stderr, flush=True)


def tgv64_vs_oracle(torch):
    """Taylor-Green 64^3 (3/2 rule, RK3, FP64, 200 steps): energy and
    dissipation every step on the GPU and in the oracle (scipy.fft over the
    host's cores), and the final fields."""
    import numpy as np

    from oracle import pf_oracle as po
    from periodicflow import workloads as W

    cores = len(os.sched_getaffinity(0))
    po.set_fft_workers(cores)
    r = W.tgv64_rk3(200)
    # the GPU's 200 FP64 steps timed on their own (the run above reads the
    # diagnostics back every step)
    from periodicflow.plan import Plan

    p = Plan([64] * 3, 0.01, precision="fp64", integrator="rk3", viscosity=1 / 400, dealias="3/2")
    uh = W.tgv_state(p)
    p.step(uh, 5)
    uh = W.tgv_state(p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.step(uh, 200)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    o = po.OracleSolver([64] * 3, 0.01, time_integrator="rk3", viscosity=1 / 400, dealias="3/2")
    o.initialize(po.taylor_green_3d(o.g))
    E, eps = [po.energy(o.g, o.u_hat)], [po.dissipation(o.g, o.u_hat, 1 / 400)]
    t_cpu = 0.0
    for _ in range(200):
        t0 = time.perf_counter()
        o.solve(1, noise=[(0.0, 0.0)])
        t_cpu += time.perf_counter() - t0
        E.append(po.energy(o.g, o.u_hat))
        eps.append(po.dissipation(o.g, o.u_hat, 1 / 400))
    E, eps = np.array(E), np.array(eps)
    pts = 64**3 * 200
    return {"gpu_fp64_Mpoint_steps_per_s": pts / t_gpu / 1e6, "gpu_seconds": t_gpu,
            "cpu_oracle_Mpoint_steps_per_s": pts / t_cpu / 1e6, "cpu_oracle_seconds": t_cpu, "cpu_cores": cores,
            "E0": float(r["energy"][0]), "E200": float(r["energy"][-1]),
            "eps0": float(r["dissipation"][0]), "eps200": float(r["dissipation"][-1]),
            "E_max_rel_err_vs_oracle": float(np.max(np.abs(r["energy"] - E) / E)),
            "eps_max_rel_err_vs_oracle": float(np.max(np.abs(r["dissipation"] - eps) / eps)),
            "u_hat_rel_l2_vs_oracle": po.rel_l2(r["u_hat"].cpu().numpy()[0], o.u_hat),
            "history_every_20_steps": {"E": [float(x) for x in r["energy"][::20]],
                                       "eps": [float(x) for x in r["dissipation"][::20]]}}


def forced_fp32_vs_fp64(torch, steps=1000, every=50):
    """The forced-turbulence configuration's FP32 error against FP64 over its
    1000 steps: one 128^3 and one 96^3 member (seeds 2000 and 2100, on-device
    noise identical in both precisions), relative L2 of u_hat every ``every``
    steps.  Chaos amplifies rounding, so the error grows with the steps."""
    from periodicflow import workloads as W
    from periodicflow.ensemble import random_solenoidal
    from periodicflow.plan import Plan

    out = {}
    for spec, seed in (W.forced_members(1, 1)):
        plans, states = {}, {}
        for prec in ("fp32", "fp64"):
            p = Plan(spec.n, spec.dt, precision=prec, integrator="rk3", viscosity=spec.viscosity,
                     noise_type="correlated", noise_mag=spec.noise_mag, forcing="constant_power",
                     forcing_power=spec.forcing_power, forcing_band=spec.forcing_band, seeds=[seed])
            p.set_correlation(torch.as_tensor(W.forced_correlation(spec), device="cuda", dtype=p.complex_dtype))
            u = torch.as_tensor(random_solenoidal(p, seed, energy=0.5, kp=4.0)[None], device="cuda",
                                dtype=p.complex_dtype).contiguous()
            plans[prec], states[prec] = p, u
        curve = []
        for done in range(0, steps + 1, every):
            if done:
                for prec in ("fp32", "fp64"):
                    plans[prec].step(states[prec], every)
            a, b = states["fp32"].to(torch.complex128), states["fp64"]
            curve.append([done, float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))])
        out[str(spec.n[0])] = curve
        _note("forced fp32 vs fp64", spec.n, curve[-1])
    return out


def tgv256_fp64_rate(torch, steps=50):
    """The plain Taylor-Green 256^3 workload in FP64 (validation mode)."""
    from periodicflow.plan import Plan

    p = Plan([N] * 3, DT, precision="fp64", integrator="rk4", viscosity=NU, dealias="2/3")
    u = _tgv_state(p, torch)
    p.step(u, 2)
    r = _rate(torch, lambda: p.step(u, steps), N**3 * steps)
    return {"steps": steps, "Mpoint_steps_per_s": round(r, 1)}


def suite(torch, path):
    """The long configurations, written to ``path`` (one JSON object)."""
    from periodicflow import workloads as W
    from periodicflow.serving import Engine, GpuBackend, serve_one_at_a_time

    import logging

    logging.basicConfig(level=logging.INFO, stream=sys.stderr, format="[%(name)s] %(message)s")
    res = {}
    _note("tgv64 vs oracle")
    res["tgv64_rk3_fp64"] = tgv64_vs_oracle(torch)
    _note("tgv256 fp64")
    res["tgv256_rk4_fp64"] = tgv256_fp64_rate(torch)
    _note("stokes512")
    res["stokes512_thermal"] = W.stokes512()
    _note("forced3d")
    f = W.forced3d(steps=100)
    f.pop("ensemble")
    res["forced3d_16x128_16x96"] = f
    res["forced3d_fp32_vs_fp64"] = forced_fp32_vs_fp64(torch)
    for count, cap in ((400, 2000),):
        _note("serving", count, cap)
        reqs = W.serving_requests(count=count, max_steps=cap)
        b = Engine(GpuBackend("fp32"), hbm_budget=96 << 30, keep_results=False, log_every=20).run(reqs)
        _note("serving one at a time")
        reqs = W.serving_requests(count=count, max_steps=cap)
        one = serve_one_at_a_time(GpuBackend("fp32"), reqs, hbm_budget=96 << 30, keep_results=False,
                                  log_every=20)[0]
        res[f"serving_{count}x{cap}"] = {"batched": b, "one_at_a_time": one}
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1, default=float)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch plans")
    ap.add_argument("--plan", default=None,
                    help="launch plan instead of autotuning, as printed in launch_plan: "
                         "'pass=variant,...' and optionally 'fused=0|1' (e.g. to profile a tuned plan)")
    ap.add_argument("--suite", default=None, help="also run the long configurations, results to this JSON file")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB device-to-device copy (the counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(members shard across GPUs with no exchange; see the design notes).", file=sys.stderr)
        sys.exit(2)

    import torch

    from periodicflow import perfmodel
    from periodicflow.plan import Plan

    torch.cuda.set_device(0)
    plan = Plan([N] * 3, DT, precision="fp32", integrator="rk4", viscosity=NU, dealias="2/3")
    u_init = _tgv_state(plan, torch)
    uh = u_init.clone()
    tune_s = None
    tune_ms = {}
    if args.plan:
        kv = dict(item.split("=") for item in args.plan.split(",") if item)
        fused_arg = kv.pop("fused", None)
        plan.set_tiles({k: int(v) for k, v in kv.items()})
        if fused_arg is not None:
            plan.set_fusion(bool(int(fused_arg)))
    elif not args.no_autotune:
        # launch plans chosen by timing (a one-off, like FFTW_MEASURE planning)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        tune_ms = plan.autotune(u_init)
        tune_s = time.perf_counter() - ta
    tiles, fused_plan = plan.tiles
    plan.step(uh, args.warmup)

    # Per-pass breakdown: every pass records events (two event records per
    # launch), in a short run of its own before the timed region.
    bd_steps = max(1, min(args.steps, 50))
    torch.cuda.synchronize()
    plan.profile(True)
    tb0 = time.perf_counter()
    plan.step(uh, bd_steps)
    torch.cuda.synchronize()
    bd_sec = time.perf_counter() - tb0
    prof_all = plan.profile_read()
    plan.profile(False)
    fused = "fwd_axis0_stage" in {k for k, (ms, c) in prof_all.items() if c}
    pb = perfmodel.pass_bytes([N] * 3, "2/3", "fp32", fused=fused)
    # per right-hand-side evaluation (4 per RK4 step): a pass may be several
    # launches (velocity and vorticity)
    per_pass = {}
    for name, (ms, cnt) in prof_all.items():
        if cnt:
            avg = ms / (4 * bd_steps)
            per_pass[name] = {
                "avg_us": round(avg * 1e3, 2),
                "launches": cnt,
                "share": round(ms / (bd_sec * 1e3), 4),
                "GBps": round(pb.get(name, 0) / (avg * 1e-3) / 1e9, 1) if name in pb else None,
            }
    dom = max((k for k in prof_all if k in pb), key=lambda k: prof_all[k][0])

    # Timed region: only the dominant pass records events (its live duration
    # gives roofline.achieved); the trajectory restarts from the initial field.
    uh.copy_(u_init)
    plan.step_counter = 0
    torch.cuda.synchronize()
    plan.profile(True, passes=[dom])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step(uh, args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = plan.profile_read()
    plan.profile(False)
    sec = t1 - t0

    value = N**3 * args.steps / sec / 1e6
    # mean over the launches the dominant pass recorded in the timed region
    dom_avg_s = prof[dom][0] / max(1, prof[dom][1]) * 1e-3
    achieved = pb[dom] / dom_avg_s / 1e9
    # HBM bytes per launch of the dominant pass, from the committed profile
    # of this workload (rocprofv3 counters need a run of their own); None
    # when that profile's plan ran another variant of the pass
    traffic, traffic_plan = None, None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        traffic_plan = t.get("plan", "")
        prof_v = dict(kv.split("=") for kv in traffic_plan.split(",") if kv)
        if int(prof_v.get(dom, 0)) == tiles.get(dom, 0):
            traffic = t.get("bytes_per_launch", {}).get(dom)
    # the step's roofline: the fewest algorithmic bytes a step moves (the
    # fused sequence), whichever sequence the launch plan runs
    step_b = perfmodel.step_bytes([N] * 3, "2/3", "fp32", 4, fused=True)

    if args.calibrate_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    energy32 = float(plan.diagnostics(uh)[0, 0].item())
    err = None
    fp64_rate = None
    if not args.no_fp64_check:
        # the FP64 validation mode over the same steps, timed (autotuned:
        # FP64 has the default tile geometry only, and chooses the stage
        # kernel's cache policy and fused against unfused)
        p64 = Plan([N] * 3, DT, precision="fp64", integrator="rk4", viscosity=NU, dealias="2/3")
        u64 = _tgv_state(p64, torch)
        if not args.no_autotune:
            p64.autotune(u64)
        p64.step(u64, 1)
        u64 = _tgv_state(p64, torch)
        p64.step_counter = 0
        fp64_rate = _rate(torch, lambda: p64.step(u64, args.steps), N**3 * args.steps)
        diff = uh.to(torch.complex128) - u64
        err = float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(u64))
        del u64, p64

    out = {
        "metric": "Mpoint-steps/s",
        "value": round(value, 2),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (Taylor-Green initial field)",
        "config": {
            "workload": "tgv3d_256_rk4_trunc23_fp32",
            "grid": [N, N, N],
            "nu": NU,
            "dt": DT,
            "integrator": "rk4",
            "dealias": "2/3",
            "members": 1,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": perfmodel.HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / perfmodel.HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": "profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload "
                              f"with the launch plan {traffic_plan}, scaled by a 1 GiB copy's counters (an earlier "
                              "profiling run, not this one)",
            "algorithmic_bytes": pb[dom],
        },
        "step_roofline_frac": round(step_b / (sec / args.steps) / 1e9 / perfmodel.HBM_PEAK_GBS, 4),
        "step_bytes": step_b,
        "passes": per_pass,
        "launch_plan": {"tiles": {k: v for k, v in tiles.items() if v}, "fused_stage": fused_plan,
                        "autotune_s": None if tune_s is None else round(tune_s, 3),
                        "autotune_us": {k: round(v * 1e3, 1) for k, v in tune_ms.items() if v}},
        "fp32_rel_l2_vs_fp64": err,
        "fp64_Mpoint_steps_per_s": None if fp64_rate is None else round(fp64_rate, 2),
        "energy_end": energy32,
    }
    if not args.no_batched:
        out["batched_vs_plain"] = batched_vs_plain(torch)
    if not args.no_serving:
        out["serving"] = serving_bench(torch)
    if args.suite:
        suite(torch, args.suite)
        out["suite"] = args.suite
    if not args.no_cpu_baseline:
        acc, out["cpu_baseline"] = cpu_baseline(args.cpu_steps, torch)
        out["accuracy_vs_oracle"] = acc
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
