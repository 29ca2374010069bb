"""The benchmark configurations as runnable workloads.

tgv64_rk3     3D Taylor-Green 64^3, nu = 1/400, dt = 0.01, RK3, 3/2 rule, FP64,
              200 steps, energy and dissipation history every step.
tgv256_rk4    3D Taylor-Green 256^3, nu = 1/1600, dt = 2.5e-3, RK4, 2/3 rule,
              500 steps, FP32 and FP64 (the plain single-instance case).
stokes512     2D fluctuating Stokes 512^2, thermal noise with amplitude
              sqrt(2 nu kT / (dt dx^2)) (noise_mag = sqrt(2 nu kT / dx^2) since
              the integrator scales by sqrt(dt)), kT = 1e-3, nu = 1,
              implicit PC, dt = 0.05, 256 seeded realisations (seeds 1000+r)
              in one launch, FP32 state, FP64 structure factor after a
              200-step burn-in, 2000 steps.
forced3d      16 members at 128^3 (seeds 2000+m) and 16 at 96^3 (seeds
              2100+m), constant-power forcing P = 0.1 on 1 <= |k| <= 2.5,
              correlated noise C(k) = |k|^-1 for 0 < |k| <= k_max/3 (variance
              ~ |k|^-2), nu = 0.005, RK3, 3/2 rule, FP32, dt = 5e-3,
              noise_mag = 1e-2, seeded solenoidal start with
              E(k) ~ k^4 exp(-2 k^2 / 16), E0 = 0.5.
serving       Poisson arrivals (rate 20/s, seed 42) of requests with grids
              from {128^2, 256^2, 512^2, 1024^2, 64^3, 128^3, 256^3}, case
              from {Taylor-Green, thermal Stokes, forced turbulence}, steps
              log-uniform in [100, 2000], request i seeded 3000 + i, FP32.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from .ensemble import Ensemble, MemberSpec, random_solenoidal
from .plan import Plan


def tgv_state(plan):
    """Taylor-Green start on the device: physical field, forward transform, projection."""
    n = plan.n
    x = [torch.arange(m, dtype=plan.real_dtype, device="cuda") * (2 * math.pi / m) for m in n]
    u = torch.zeros(plan.physical_shape(), dtype=plan.real_dtype, device="cuda")
    if plan.dim == 3:
        X, Y, Z = x[0].view(-1, 1, 1), x[1].view(1, -1, 1), x[2].view(1, 1, -1)
        u[:, 0] = torch.sin(X) * torch.cos(Y) * torch.cos(Z)
        u[:, 1] = -torch.cos(X) * torch.sin(Y) * torch.cos(Z)
    else:
        X, Y = x[0].view(-1, 1), x[1].view(1, -1)
        u[:, 0] = torch.sin(X) * torch.cos(Y)
        u[:, 1] = -torch.cos(X) * torch.sin(Y)
    uh = plan.new_state()
    plan.forward(u, uh)
    plan.project(uh)
    return uh


def tgv64_rk3(steps=200):
    p = Plan([64] * 3, 0.01, precision="fp64", integrator="rk3", viscosity=1 / 400, dealias="3/2")
    uh = tgv_state(p)
    hist = [p.diagnostics(uh)[0].cpu().numpy()]
    for _ in range(steps):
        p.step(uh, 1)
        hist.append(p.diagnostics(uh)[0].cpu().numpy())
    h = np.array(hist)
    return {"energy": h[:, 0], "dissipation": 2 * (1 / 400) * h[:, 1], "u_hat": uh}


def _timed(fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def stokes512(members=256, steps=2000, burn=200, n=512, nu=1.0, kT=1e-3, dt=0.05, sk=True):
    dx = 2 * np.pi / n
    nm = math.sqrt(2 * nu * kT / dx**2)
    p = Plan((n, n), dt, precision="fp32", integrator="implicit_pc", nonlinear=False, viscosity=nu, noise_mag=nm,
             batch=members, seeds=[1000 + r for r in range(members)])
    u = equilibrium_state(p, nm**2 * 2 / (2 * nu * n * n))
    acc = torch.zeros(p.shape_fourier, dtype=torch.float64, device="cuda")
    t_burn = _timed(lambda: p.step(u, burn))

    def run():
        if sk:
            p.step_accumulate_sk(u, steps - burn, acc)
        else:
            p.step(u, steps - burn)

    t = _timed(run)
    S = acc.cpu().numpy() / ((steps - burn) * members)
    theory = nm**2 * 2 / (2 * nu * n * n)
    k0 = np.fft.fftfreq(n, 1.0 / n)[:, None]
    k1 = np.arange(n // 2 + 1)[None, :]
    kk = np.sqrt(k0**2 + k1**2)
    r = S[kk > 0] / theory
    shells = np.rint(kk[kk > 0]).astype(int)
    sh = np.array([r[shells == s].mean() for s in range(1, n // 2)])
    return {
        "members": members, "steps": steps, "seconds": t + t_burn,
        "Mpoint_steps_per_s": n * n * members * steps / (t + t_burn) / 1e6,
        "S_over_theory_mean": float(r.mean()), "S_shell_max_dev": float(np.max(np.abs(sh - 1))),
        "flat_within_2pct": bool(np.max(np.abs(sh - 1)) < 0.02),
    }


def equilibrium_state(plan, s_total):
    """Members drawn from the stationary distribution of fluctuating Stokes
    flow: divergence-free complex Gaussian modes with E sum_a |u_a(k)|^2 =
    s_total for k != 0 (as the reference's thermal-noise example starts,
    examples/fluctuating_stokes_thermal_noise.py:5-12).  Crank-Nicolson damps
    its high modes by |R| close to 1 at large nu k^2 dt, so a start from rest
    would take thousands of steps to equilibrate there."""
    d = plan.dim
    raw = torch.empty(plan.noise_shape(), dtype=plan.complex_dtype, device="cuda")
    saved = plan.step_counter
    plan.step_counter = 1 << 40  # a stream of its own
    plan.draw_noise(raw)
    plan.step_counter = saved
    u = plan.new_state()
    u[: raw.shape[0]] = raw[:, 0, :d] * math.sqrt(s_total / (2 * (d - 1)))
    plan.project(u)
    plan.mask_zero_mode(u)
    return u


def forced_members(n_big=16, n_small=16, big=128, small=96):
    def spec(n):
        return MemberSpec(n=(n, n, n), dt=5e-3, integrator="rk3", viscosity=0.005, nonlinear=True, dealias="3/2",
                          noise_type="correlated", noise_mag=1e-2, forcing="constant_power", forcing_power=0.1,
                          forcing_band=(1.0, 2.5))
    return ([(spec(big), 2000 + m) for m in range(n_big)] + [(spec(small), 2100 + m) for m in range(n_small)])


def forced_correlation(spec):
    """C(k) = |k|^-1 for 0 < |k| <= k_max / 3 with k_max = N/2."""
    n = spec.n
    d = len(n)
    ks = []
    for i, ni in enumerate(n):
        k = np.fft.fftfreq(ni, 1.0 / ni) if i < d - 1 else np.fft.rfftfreq(ni, 1.0 / ni)
        shp = [1] * d
        shp[i] = k.size
        ks.append(k.reshape(shp))
    kk = np.sqrt(sum(k**2 for k in np.broadcast_arrays(*ks)))
    kmax = n[0] / 2
    return np.where((kk > 0) & (kk <= kmax / 3), 1.0 / np.maximum(kk, 1e-30), 0.0) + 0j


def forced3d(steps=1000, members=None, graph=False):
    members = members or forced_members()
    ens = Ensemble(members, precision="fp32", correlation=forced_correlation)
    for i, (spec, seed) in enumerate(members):
        ens.set_fourier(i, random_solenoidal(ens.plan_of(i), seed, energy=0.5, kp=4.0))
    tune = ens.autotune()  # tile plans chosen by timing, before the timed run
    t = _timed(lambda: ens.run(steps, graph=graph))
    return {"members": len(members), "steps": steps, "seconds": t,
            "Mpoint_steps_per_s": ens.points_per_step() * steps / t / 1e6, "ensemble": ens,
            "tiles": {str(k): g.plan.tiles for k, g in ens.groups.items()},
            "tuning_ms": {str(k): v for k, v in tune.items()}}


def serving_requests(count=400, rate=20.0, seed=42, max_steps=2000, sizes=None):
    """The request stream: arrival times, grids, cases, step counts, seeds."""
    from .serving import Request

    rs = np.random.RandomState(seed)
    sizes = sizes or [(128, 128), (256, 256), (512, 512), (1024, 1024), (64, 64, 64), (128, 128, 128),
                      (256, 256, 256)]
    t = 0.0
    out = []
    for i in range(count):
        t += rs.exponential(1.0 / rate)
        n = sizes[rs.randint(len(sizes))]
        case = ("tgv", "stokes", "forced")[rs.randint(3)]
        steps = int(round(math.exp(rs.uniform(math.log(100), math.log(max_steps)))))
        if case == "tgv":
            spec = MemberSpec(n=n, dt=2.5e-3, integrator="rk4", viscosity=1 / 1600, dealias="2/3")
            init = tgv_device
        elif case == "stokes":
            dx = 2 * np.pi / n[0]
            spec = MemberSpec(n=n, dt=0.05, integrator="implicit_pc", viscosity=1.0, nonlinear=False,
                              noise_mag=math.sqrt(2 * 1e-3 / dx**2))
            init = zero_device
        else:
            # CFL ~ 0.15 at |u| ~ 2
            dt = min(5e-3, 0.15 * (2 * np.pi / n[0]) / 2.0)
            spec = MemberSpec(n=n, dt=float(f"{dt:.3g}"), integrator="rk3", viscosity=0.005,
                              forcing="constant_power", forcing_power=0.1)
            init = (lambda p, s=3000 + i: solenoidal_device(p, s, energy=0.5))
        out.append(Request(i, spec, 3000 + i, steps, init, t))
    return out


def tgv_host(plan):
    """Taylor-Green spectrum of one member (host, FP64): the 4 (2D) or 8 (3D) modes."""
    u = np.zeros((plan.dim,) + plan.shape_fourier, dtype=complex)
    n = plan.n
    if plan.dim == 3:
        # u = sin x cos y cos z: coefficients +-1/8 (i) on k = (+-1, +-1, 1) in the half spectrum
        for sy in (1, -1):
            j = 1 if sy > 0 else n[1] - 1
            u[0, 1, j, 1] = -0.125j
            u[0, n[0] - 1, j, 1] = 0.125j
            u[1, 1, j, 1] = (0.125j if sy > 0 else -0.125j)
            u[1, n[0] - 1, j, 1] = (0.125j if sy > 0 else -0.125j)
    else:
        for sx in (1, -1):
            i = 1 if sx > 0 else n[0] - 1
            u[0, i, 1] = -0.25j * sx
            u[1, i, 1] = 0.25j
    return u


def zero_host(plan):
    return np.zeros((plan.dim,) + plan.shape_fourier, dtype=complex)


def tgv_device(plan):
    """tgv_host's spectrum built on the device (no host array to transfer)."""
    u = torch.zeros((plan.dim,) + plan.shape_fourier, dtype=plan.complex_dtype, device="cuda")
    n = plan.n
    if plan.dim == 3:
        for sy in (1, -1):
            j = 1 if sy > 0 else n[1] - 1
            u[0, 1, j, 1] = -0.125j
            u[0, n[0] - 1, j, 1] = 0.125j
            u[1, 1, j, 1] = (0.125j if sy > 0 else -0.125j)
            u[1, n[0] - 1, j, 1] = (0.125j if sy > 0 else -0.125j)
    else:
        for sx in (1, -1):
            i = 1 if sx > 0 else n[0] - 1
            u[0, i, 1] = -0.25j * sx
            u[1, i, 1] = 0.25j
    return u


def zero_device(plan):
    return torch.zeros((plan.dim,) + plan.shape_fourier, dtype=plan.complex_dtype, device="cuda")


def solenoidal_device(plan, seed, energy=0.5, kp=4.0):
    """A seeded solenoidal field with E(k) ~ k^4 exp(-2 k^2 / kp^2) and total
    energy ``energy``, generated on the device (torch's generator seeded with
    ``seed``; the same shaping, projection and normalisation as
    random_solenoidal, whose host draws it does not reproduce)."""
    d = plan.dim
    dev = "cuda"
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    ks = []
    for i, ni in enumerate(plan.n):
        k = torch.fft.fftfreq(ni, 1.0 / ni, device=dev, dtype=torch.float64) if i < d - 1 else \
            torch.fft.rfftfreq(ni, 1.0 / ni, device=dev, dtype=torch.float64)
        shp = [1] * d
        shp[i] = k.numel()
        ks.append(k.view(shp))
    shape = plan.shape_fourier
    k2 = sum(k**2 for k in ks).expand(shape)
    kk = torch.sqrt(k2)
    amp = torch.where(kk > 0, kk**2 * torch.exp(-kk**2 / kp**2), torch.zeros_like(kk)) / \
        torch.clamp(kk, min=1e-30) ** ((d - 1) / 2)
    re = torch.randn((d,) + shape, generator=gen, device=dev, dtype=torch.float64)
    im = torch.randn((d,) + shape, generator=gen, device=dev, dtype=torch.float64)
    u = torch.complex(re, im) * amp
    p = sum(ks[i] * u[i] for i in range(d)) / torch.where(k2 == 0, torch.ones_like(k2), k2)
    u = torch.stack([u[i] - ks[i] * p for i in range(d)])
    w = torch.full(shape, 2.0, device=dev, dtype=torch.float64)
    w[..., 0] = 1.0
    if plan.n[-1] % 2 == 0:
        w[..., -1] = 1.0
    e = 0.5 * float(torch.sum(w * torch.sum(u.real**2 + u.imag**2, dim=0)))
    if e > 0:
        u = u * math.sqrt(energy / e)
    return u.to(plan.complex_dtype)


def _calibration_copy():
    """One 1 GiB device-to-device copy: the kernel whose counters scale the
    HBM traffic counters of a profile (known bytes read and written)."""
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def main(argv=None):
    """Run one configuration for a few steps (the target of kernel profiles):

    python -m periodicflow.workloads stokes512 --steps 600
        thermal-noise Stokes 512^2 x 256 realisations, register-resident
        steps with the FP64 structure factor accumulated
    python -m periodicflow.workloads forced --steps 5 [--plan PLAN] [--separate]
        the forced-turbulence ensemble 16 x 128^3 + 16 x 96^3, default plans
        or the given ones (no autotuning, so the trace holds only the steps);
        PLAN: 'N:pass=variant,...,fused=0|1;N:...' per group of grid size N;
        --separate steps the two groups one after the other instead of with
        paired launches
    python -m periodicflow.workloads tgv256 --steps 5
        Taylor-Green 256^3, RK4, 2/3 rule, FP32

    --calibrate-copy adds one 1 GiB device copy (__amd_rocclr_copyBuffer).
    """
    import argparse

    ap = argparse.ArgumentParser(prog="python -m periodicflow.workloads")
    ap.add_argument("workload", choices=["stokes512", "forced", "tgv256"])
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--plan", default=None, help="forced: launch plans per group, 'N:pass=v,...,fused=0|1;N:...'")
    ap.add_argument("--separate", action="store_true", help="forced: step the groups one after the other")
    args = ap.parse_args(argv)
    if args.workload == "stokes512":
        n, nu, dt, kT, B = 512, 1.0, 0.05, 1e-3, 256
        nm = math.sqrt(2 * nu * kT / (2 * math.pi / n) ** 2)
        p = Plan((n, n), dt, precision="fp32", integrator="implicit_pc", nonlinear=False, viscosity=nu,
                 noise_mag=nm, batch=B, seeds=[1000 + r for r in range(B)])
        u = p.new_state()
        acc = torch.zeros(p.shape_fourier, dtype=torch.float64, device="cuda")
        t = _timed(lambda: p.step_accumulate_sk(u, args.steps, acc))
        rate = n * n * B * args.steps / t / 1e6
    elif args.workload == "forced":
        members = forced_members()
        ens = Ensemble(members, precision="fp32", correlation=forced_correlation)
        for i, (spec, seed) in enumerate(members):
            ens.set_fourier(i, random_solenoidal(ens.plan_of(i), seed, energy=0.5, kp=4.0))
        for part in (args.plan or "").split(";"):
            if not part:
                continue
            size, kvs = part.split(":")
            kv = dict(item.split("=") for item in kvs.split(",") if item)
            fused = kv.pop("fused", None)
            for g in ens.groups.values():
                if g.spec.n[0] == int(size):
                    g.plan.set_tiles({k: int(v) for k, v in kv.items()})
                    if fused is not None:
                        g.plan.set_fusion(bool(int(fused)))
        ens.pair = not args.separate
        t = _timed(lambda: ens.step(args.steps))
        rate = ens.points_per_step() * args.steps / t / 1e6
    else:
        p = Plan([256] * 3, 2.5e-3, precision="fp32", integrator="rk4", viscosity=1 / 1600, dealias="2/3")
        u = tgv_state(p)
        t = _timed(lambda: p.step(u, args.steps))
        rate = 256**3 * args.steps / t / 1e6
    if args.calibrate_copy:
        _calibration_copy()
    print(f"{args.workload}: {args.steps} steps in {t:.3f} s, {rate:.1f} Mpoint-steps/s", flush=True)


if __name__ == "__main__":
    main()
