"""GPU parity against the FP64 oracle at the benchmark configurations' own sizes.

Every transform length the configurations use (64, 96, 128, 144, 192, 256,
384, 512, 768, 1024, 1536) runs here through the C ABI and is compared with
the oracle on identical inputs:

* Taylor-Green 64^3, 3/2 rule, RK3, FP64, all 200 steps: energy and
  dissipation every step, the final spectrum and the final physical field.
* Taylor-Green 256^3, 2/3 rule, RK4: two steps in FP64 and in FP32.
* Forced fluctuating turbulence, one 128^3 and one 96^3 member in one
  ensemble (3/2 rule, RK3, correlated noise |k|^-1 on 0 < |k| <= N/6,
  constant-power band forcing), two steps in FP64 and FP32, on-device noise
  restated by the oracle.
* Thermal-noise Stokes flow 512^2, implicit PC, two realisations, the
  reference's NumPy noise stream injected, FP64 and FP32.
* The nonlinear term of the served sizes 128^2 .. 1024^2 and 256^3 under the
  3/2 rule, FP64 and FP32.
* An anisotropic grid in a box that is not 2 pi per side.

Bars: relative L2 <= 1e-12 in FP64 and <= 1e-5 in FP32 (FP32 against the
FP64 oracle; the FP64 error is stated beside it in the output).  The oracle
runs on the host's cores through scipy.fft.
"""
import os
import sys

import numpy as np
import pytest

from oracle import pf_oracle as po

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _say(*a):
    print("[gpu-config]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from periodicflow import ensemble, plan, workloads
    po.set_fft_workers(max(1, min(16, len(os.sched_getaffinity(0)))))
    yield plan, ensemble, workloads
    po.set_fft_workers(None)


def _dev(x, cdt):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=cdt)


def _host(t):
    return t.detach().cpu().numpy()


def _silent(n):
    """Noise of n steps with noise_mag = 0: the draws are never used."""
    return [(0.0, 0.0)] * n


def _physical_err(plan, uh_dev, g, uh_ref):
    """Relative L2 of the physical velocity (GPU backward vs the oracle's)."""
    u = torch.empty(plan.physical_shape(), dtype=plan.real_dtype, device="cuda")
    plan.backward(uh_dev, u)
    ref = np.stack([g.backward(uh_ref[b]) for b in range(uh_ref.shape[0])])
    return po.rel_l2(_host(u), ref)


# --------------------------------------------------------- Taylor-Green 64^3
@pytest.mark.timeout(900)
def test_tgv64_rk3_fp64_200_steps(mods):
    plan_m, _, _ = mods
    steps, nu, dt = 200, 1 / 400, 0.01
    _say("TGV 64^3 3/2 RK3 FP64, 200 steps")
    o = po.OracleSolver([64] * 3, dt, time_integrator="rk3", viscosity=nu, dealias="3/2")
    o.initialize(po.taylor_green_3d(o.g))
    p = plan_m.Plan([64] * 3, dt, precision="fp64", integrator="rk3", viscosity=nu, dealias="3/2")
    uh = _dev(o.u_hat[None], p.complex_dtype)
    e_gpu, e_ref = [], []
    for s in range(steps + 1):
        if s:
            p.step(uh, 1)
            o.solve(1, noise=_silent(1))
        d = _host(p.diagnostics(uh))[0]
        e_gpu.append((d[0], 2 * nu * d[1]))
        e_ref.append((po.energy(o.g, o.u_hat), po.dissipation(o.g, o.u_hat, nu)))
        if s % 50 == 0:
            _say(f"  step {s}: E {d[0]:.15f} (oracle {e_ref[-1][0]:.15f})")
    e_gpu, e_ref = np.array(e_gpu), np.array(e_ref)
    rel = np.abs(e_gpu - e_ref) / np.abs(e_ref)
    _say("  max rel err energy", rel[:, 0].max(), "dissipation", rel[:, 1].max())
    assert abs(e_ref[0, 0] - 0.125) < 1e-15 and abs(e_ref[0, 1] - 1.875e-3) < 1e-15
    assert rel.max() < FP64_TOL
    err_k = po.rel_l2(_host(uh)[0], o.u_hat)
    err_x = _physical_err(p, uh, o.g, o.u_hat[None])
    _say("  final u_hat rel L2", err_k, "physical u rel L2", err_x)
    assert err_k < FP64_TOL and err_x < FP64_TOL


# -------------------------------------------------------- Taylor-Green 256^3
@pytest.fixture(scope="module")
def tgv256(mods):
    _say("oracle: TGV 256^3 2/3 RK4, 2 steps")
    o = po.OracleSolver([256] * 3, 2.5e-3, time_integrator="rk4", viscosity=1 / 1600, dealias="2/3")
    o.initialize(po.taylor_green_3d(o.g))
    start = o.u_hat.copy()
    o.solve(2, noise=_silent(2))
    return start, o


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tgv256_rk4_two_steps(mods, tgv256, prec):
    plan_m, _, workloads = mods
    start, o = tgv256
    p = plan_m.Plan([256] * 3, 2.5e-3, precision=prec, integrator="rk4", viscosity=1 / 1600, dealias="2/3")
    # the device's own Taylor-Green start (forward transform + projection)
    u0 = workloads.tgv_state(p)
    e0 = po.rel_l2(_host(u0)[0], start)
    uh = _dev(start[None], p.complex_dtype)
    p.step(uh, 2)
    err_k = po.rel_l2(_host(uh)[0], o.u_hat)
    err_x = _physical_err(p, uh, o.g, o.u_hat[None])
    _say(f"TGV 256^3 {prec}: start rel L2 {e0:.3e}, after 2 steps u_hat {err_k:.3e}, physical u {err_x:.3e}")
    tol = FP64_TOL if prec == "fp64" else FP32_TOL
    assert e0 < tol and err_k < tol and err_x < tol


# --------------------------------------- forced turbulence, 128^3 + 96^3
FORCED_STEPS = 2


@pytest.fixture(scope="module")
def forced_ref(mods):
    """Oracle runs of one 128^3 and one 96^3 member of the forced ensemble,
    with the on-device noise stream restated (Philox4x32-10 + Box-Muller)."""
    plan_m, ens_m, workloads = mods
    members = workloads.forced_members(n_big=1, n_small=1)
    out = []
    for spec, seed in members:
        _say(f"oracle: forced {spec.n} RK3 3/2, correlated noise + forcing, {FORCED_STEPS} steps")
        o = po.OracleSolver(spec.n, spec.dt, time_integrator="rk3", viscosity=spec.viscosity, dealias="3/2",
                            noise_type="correlated", noise_mag=spec.noise_mag, forcing="constant_power",
                            forcing_power=spec.forcing_power, forcing_band=spec.forcing_band)
        o.set_correlation_function(workloads.forced_correlation(spec))
        shape = o.g.shape_fourier
        nk, nc = int(np.prod(shape)), 3
        start = ens_m.random_solenoidal(_PlanShape(spec.n), seed, energy=0.5, kp=4.0)
        o.initialize(start, space="fourier", project=False, mask_zero_mode=False)
        noise = []
        for st in range(FORCED_STEPS):
            raw = po.device_noise_draws([seed], [st], nk, nc)[0]
            noise.append((raw[0].reshape((nc,) + shape), raw[1].reshape((nc,) + shape)))
        o.solve(FORCED_STEPS, noise=noise)
        out.append((start, o))
    return members, out


class _PlanShape:
    """What random_solenoidal reads of a plan: dim, n and the spectrum's shape."""

    def __init__(self, n):
        self.n = list(n)
        self.dim = len(n)
        self.shape_fourier = tuple(n[:-1]) + (n[-1] // 2 + 1,)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_forced_128_96_ensemble(mods, forced_ref, prec):
    plan_m, ens_m, workloads = mods
    members, refs = forced_ref
    ens = ens_m.Ensemble(members, precision=prec, correlation=workloads.forced_correlation)
    for i, (start, _) in enumerate(refs):
        ens.set_fourier(i, start)
    ens.step(FORCED_STEPS)
    errs = []
    for i, (spec, seed) in enumerate(members):
        p = ens.plan_of(i)
        o = refs[i][1]
        st = ens.state(i)
        assert p.batch == 1  # one member per size: its group's state is (1, 3, ...)
        err_k = po.rel_l2(_host(st), o.u_hat)
        err_x = _physical_err(p, st.unsqueeze(0).contiguous(), o.g, o.u_hat[None])
        _say(f"forced {spec.n} {prec}: u_hat rel L2 {err_k:.3e}, physical u {err_x:.3e}")
        errs += [err_k, err_x]
    tol = FP64_TOL if prec == "fp64" else FP32_TOL
    assert max(errs) < tol


# ---------------------------------------------- thermal-noise Stokes 512^2
@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stokes512_implicit_pc_injected_noise(mods, prec):
    plan_m, _, _ = mods
    n, nu, dt, kT, steps, B = 512, 1.0, 0.05, 1e-3, 5, 2
    nm = np.sqrt(2 * nu * kT / (2 * np.pi / n) ** 2)
    p = plan_m.Plan((n, n), dt, precision=prec, integrator="implicit_pc", nonlinear=False, viscosity=nu,
                    noise_mag=nm, batch=B)
    rs = np.random.RandomState(31)
    starts, refs = [], []
    raw = np.zeros((steps, B) + p.noise_shape()[1:], dtype=complex)
    for b in range(B):
        o = po.OracleSolver((n, n), dt, is_nonlinear=False, time_integrator="implicit_pc", viscosity=nu,
                            noise_mag=nm, seed=1000 + b)
        u0 = 1e-3 * (rs.standard_normal((2,) + o.g.shape_fourier) + 1j * rs.standard_normal((2,) + o.g.shape_fourier))
        o.initialize(u0, space="fourier", project=True)
        starts.append(o.u_hat.copy())
        noise = [o.noise_stream.next() for _ in range(steps)]
        for s in range(steps):
            raw[s, b, 0], raw[s, b, 1] = noise[s]
        o.solve(steps, noise=noise)
        refs.append(o.u_hat)
    uh = _dev(np.stack(starts), p.complex_dtype)
    p.step(uh, steps, noise=_dev(raw, p.complex_dtype))
    err = po.rel_l2(_host(uh), np.stack(refs))
    _say(f"Stokes 512^2 implicit PC {prec}, {steps} steps, NumPy noise injected: rel L2 {err:.3e}")
    assert err < (FP64_TOL if prec == "fp64" else FP32_TOL)


# ------------------------------------------------- served sizes: rhs only
SERVED = [((128, 128), "3/2"), ((256, 256), "3/2"), ((512, 512), "3/2"), ((1024, 1024), "3/2"),
          ((1024, 1024), "2/3"), ((128, 128, 128), "2/3"), ((256, 256, 256), "3/2")]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("N,dealias", SERVED, ids=[f"{'x'.join(map(str, n))}-{d}" for n, d in SERVED])
def test_rhs_nonlinear_served_sizes(mods, N, dealias):
    plan_m, _, _ = mods
    o = po.OracleSolver(N, 0.01, dealias=dealias)
    rs = np.random.RandomState(41)
    shp = (len(N),) + o.g.shape_fourier
    uh = rs.standard_normal(shp) + 1j * rs.standard_normal(shp)
    ref = o.rhs_nonlinear(uh)
    out = {}
    for prec in ("fp64", "fp32"):
        p = plan_m.Plan(N, 0.01, precision=prec, dealias=dealias)
        r = p.new_state()
        p.rhs_nonlinear(_dev(uh[None], p.complex_dtype), r)
        out[prec] = po.rel_l2(_host(r)[0], ref)
        del p, r
    _say(f"rhs_nonlinear {N} {dealias}: FP64 {out['fp64']:.3e}, FP32 {out['fp32']:.3e}")
    assert out["fp64"] < FP64_TOL and out["fp32"] < FP32_TOL


# -------------------------------------------------- box other than 2 pi
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_anisotropic_box_rk3_thermal(mods, prec):
    plan_m, _, _ = mods
    N, L = (16, 24, 32), (2 * np.pi, 4 * np.pi, 3 * np.pi)
    dt, nu, nm, steps, B = 0.01, 0.02, 0.05, 3, 2
    p = plan_m.Plan(N, dt, length=L, precision=prec, integrator="rk3", viscosity=nu, noise_mag=nm, batch=B)
    rs = np.random.RandomState(51)
    starts, refs = [], []
    raw = np.zeros((steps, B) + p.noise_shape()[1:], dtype=complex)
    for b in range(B):
        o = po.OracleSolver(N, dt, time_integrator="rk3", viscosity=nu, noise_mag=nm, seed=200 + b,
                            domain=[(0.0, v) for v in L])
        # amplitude small enough for the explicit step to stay stable (max
        # |u| about 12, CFL about 0.4 on the finest axis)
        u0 = 0.02 * (rs.standard_normal((3,) + o.g.shape_fourier) + 1j * rs.standard_normal((3,) + o.g.shape_fourier))
        o.initialize(u0, space="fourier", project=True)
        starts.append(o.u_hat.copy())
        noise = [o.noise_stream.next() for _ in range(steps)]
        for s in range(steps):
            raw[s, b, 0], raw[s, b, 1] = noise[s]
        o.solve(steps, noise=noise)
        assert np.abs(o.u_hat).max() < 2 * np.abs(starts[-1]).max()  # no blow-up
        refs.append(o.u_hat)
    uh = _dev(np.stack(starts), p.complex_dtype)
    p.step(uh, steps, noise=_dev(raw, p.complex_dtype))
    err = po.rel_l2(_host(uh), np.stack(refs))
    _say(f"box {L} grid {N} RK3 thermal {prec}: rel L2 {err:.3e}")
    assert err < (FP64_TOL if prec == "fp64" else FP32_TOL)


# ------------------------------------ on-device noise stream vs the oracle
def _single_draw_raw(seed, step, nk, shape, d, ntype, k):
    """The oracle's raw draws (w1, w2) that reproduce one step of the device's
    single-draw stream (implicit PC Stokes, whose only live noise stage has
    a = 1): entry e of the stream (po.device_single_draws: word pair step % 2
    of counter group e at step pair step // 2), scaled by sqrt(1 + a^2) = sqrt 2.

    Thermal: the device draws z (3 entries in 3D, one in 2D, along the third
    axis) and forms the projected noise as i k x z.  The oracle forms
    P (i W_sym k) from w1 with W_sym = symmetrize(w1) = sqrt 2 w1 for a
    symmetric w1 (space_solver.py:15-64), so w1 = (v k^T + k v^T) / (sqrt 2 k^2)
    with v = sqrt 2 (k x z) gives W_sym k = v, which P leaves as it is
    (v is orthogonal to k); w1 = 0 at k = 0.  Correlated: w1 = sqrt 2 g."""
    ne = (3 if d == 3 else 1) if ntype == "thermal" else d
    dev = po.device_single_draws(seed, step, nk, ne)  # (ne, nk)
    g = [dev[e].reshape(shape) for e in range(ne)]
    if ntype == "thermal":
        kk = [np.broadcast_to(k[i], shape) for i in range(d)]
        if d == 3:
            v = [kk[1] * g[2] - kk[2] * g[1], kk[2] * g[0] - kk[0] * g[2], kk[0] * g[1] - kk[1] * g[0]]
        else:
            v = [kk[1] * g[0], -kk[0] * g[0]]
        v = [np.sqrt(2.0) * x for x in v]
        k2 = sum(x * x for x in kk)
        inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0) / np.sqrt(2.0)
        w1 = np.zeros((d * d,) + shape, dtype=complex)
        for i in range(d):
            for j in range(d):
                w1[i * d + j] = (v[i] * kk[j] + kk[i] * v[j]) * inv
    else:
        w1 = np.sqrt(2.0) * np.array(g)
    return w1, np.zeros_like(w1)


@pytest.mark.parametrize("N,ntype", [((512, 512), "thermal"), ((24, 16, 32), "thermal"), ((64, 64), "correlated")])
def test_resident_stokes_device_noise_vs_oracle(mods, N, ntype):
    """Register-resident Stokes steps with the on-device noise stream, FP64,
    against the oracle fed the same draws: 6 steps, 2 members."""
    plan_m, _, _ = mods
    steps, B, nu, dt, nm = 6, 2, 1.0, 0.05, 0.3
    seeds = [77, 78]
    p = plan_m.Plan(N, dt, precision="fp64", integrator="implicit_pc", nonlinear=False, viscosity=nu,
                    noise_type=ntype, noise_mag=nm, batch=B, seeds=seeds)
    rs = np.random.RandomState(71)
    starts, refs = [], []
    C = None
    for b in range(B):
        o = po.OracleSolver(N, dt, is_nonlinear=False, time_integrator="implicit_pc", viscosity=nu,
                            noise_type=ntype, noise_mag=nm)
        if ntype == "correlated":
            C = np.where(o.g.k2 > 0, 1.0 / np.sqrt(np.where(o.g.k2 > 0, o.g.k2, 1.0)), 0.0) + 0j
            o.set_correlation_function(C)
        d = len(N)
        u0 = 0.1 * (rs.standard_normal((d,) + o.g.shape_fourier) + 1j * rs.standard_normal((d,) + o.g.shape_fourier))
        o.initialize(u0, space="fourier", project=True)
        starts.append(o.u_hat.copy())
        nk = int(np.prod(o.g.shape_fourier))
        noise = [_single_draw_raw(seeds[b], s, nk, o.g.shape_fourier, d, ntype, o.g.k) for s in range(steps)]
        o.solve(steps, noise=noise)
        refs.append(o.u_hat)
    if C is not None:
        p.set_correlation(_dev(C, p.complex_dtype))
    uh = _dev(np.stack(starts), p.complex_dtype)
    p.step(uh, steps)
    err = po.rel_l2(_host(uh), np.stack(refs))
    _say(f"resident Stokes {N} {ntype}, device noise restated: rel L2 {err:.3e}")
    assert err < FP64_TOL


# ------------------------------------------- served requests vs the oracle
def test_served_requests_vs_oracle(mods):
    """One request of each case, served in FP64 among others, against the
    oracle: Taylor-Green (RK4, 2/3 rule), forced turbulence (RK3, 3/2 rule,
    constant-power forcing) and thermal-noise Stokes flow (implicit PC,
    on-device noise restated for the oracle)."""
    plan_m, ens_m, workloads = mods
    from periodicflow import serving
    S = ens_m.MemberSpec
    tgv = S(n=(32, 32, 32), dt=2.5e-3, integrator="rk4", viscosity=1 / 1600, dealias="2/3")
    forced = S(n=(64, 64), dt=5e-3, integrator="rk3", viscosity=0.005, forcing="constant_power", forcing_power=0.1)
    stokes = S(n=(64, 64), dt=0.05, integrator="implicit_pc", viscosity=1.0, nonlinear=False, noise_mag=0.3)
    reqs = [
        serving.Request(0, tgv, 3000, 40, workloads.tgv_host, 0.0),
        serving.Request(1, forced, 3001, 30, (lambda p: ens_m.random_solenoidal(p, 11, energy=0.5)), 0.0),
        serving.Request(2, stokes, 3002, 25, workloads.zero_host, 0.0),
        # neighbours sharing the pools
        serving.Request(3, forced, 3003, 17, (lambda p: ens_m.random_solenoidal(p, 12, energy=0.5)), 0.001),
        serving.Request(4, stokes, 3004, 9, workloads.zero_host, 0.001),
        serving.Request(5, tgv, 3005, 23, workloads.tgv_host, 0.002),
    ]
    eng = serving.Engine(serving.GpuBackend("fp64"), hbm_budget=8 << 30, chunk=7, max_slots=4)
    st = eng.run(reqs)
    assert st["done"] == len(reqs)
    shape_of = {r.rid: r for r in reqs}
    for rid in (0, 1, 2):
        r = shape_of[rid]
        sp = r.spec
        o = po.OracleSolver(sp.n, sp.dt, is_nonlinear=sp.nonlinear, time_integrator=sp.integrator,
                            viscosity=sp.viscosity, dealias=sp.dealias, noise_mag=sp.noise_mag,
                            forcing=sp.forcing, forcing_power=sp.forcing_power, forcing_band=sp.forcing_band)
        like = _PlanShape(sp.n)
        u0 = r.init(like)
        o.initialize(u0, space="fourier", project=False, mask_zero_mode=False)
        if sp.noise_mag:
            nk = int(np.prod(o.g.shape_fourier))
            noise = [_single_draw_raw(r.seed, s, nk, o.g.shape_fourier, len(sp.n), sp.noise_type, o.g.k)
                     for s in range(r.steps)]
        else:
            noise = _silent(r.steps)
        o.solve(r.steps, noise=noise)
        err = po.rel_l2(r.result, o.u_hat)
        _say(f"served request {rid} {sp.n} {sp.integrator}: rel L2 vs oracle {err:.3e}")
        assert err < FP64_TOL
