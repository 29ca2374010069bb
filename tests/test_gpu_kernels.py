"""GPU parity of the HIP kernels against the FP64 oracle, through the C ABI.

Bars: FP64 relative L2 <= 1e-12; FP32 relative L2 <= 1e-5 (stated per test).
Inputs are seeded random spectra that fill every stored mode, including the
Nyquist planes and the non-Hermitian kz = 0 plane that the reference's noise
creates, so the padding / truncation / c2r conventions are all exercised.
"""
import sys

import numpy as np
import pytest

from oracle import pf_oracle as po

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pf():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from periodicflow.plan import Plan  # noqa: F401  (loads libpfhip.so; raises if missing)
    import periodicflow.plan as m
    return m


def _rand_spec(rs, shape):
    return rs.standard_normal(shape) + 1j * rs.standard_normal(shape)


def _dev(x, cdt):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=cdt)


def _host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("N", [(16, 24), (8, 12, 16), (12, 12, 8)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_forward_backward(pf, N, prec):
    _say("forward/backward", N, prec)
    g = po.Grid(N)
    plan = pf.Plan(N, 0.01, precision=prec, nonlinear=False, batch=2)
    rs = np.random.RandomState(0)
    u = rs.standard_normal((2, len(N)) + g.shape_physical)
    ud = torch.from_numpy(u).to("cuda", plan.real_dtype)
    uh = plan.new_state()
    plan.forward(ud, uh)
    ref = g.forward(u)
    tol = FP64_TOL if prec == "fp64" else FP32_TOL
    assert po.rel_l2(_host(uh), ref) < tol
    # backward of a full random spectrum (non-Hermitian planes included)
    sh = _rand_spec(rs, (2, len(N)) + g.shape_fourier)
    back = torch.empty_like(ud)
    plan.backward(_dev(sh, plan.complex_dtype), back)
    refb = np.stack([g.backward(sh[b]) for b in range(2)])
    assert po.rel_l2(_host(back), refb) < tol


@pytest.mark.parametrize("N,dealias", [((16, 16), "3/2"), ((16, 24), "2/3"), ((24, 16), "3/2"),
                                       ((8, 12, 16), "3/2"), ((12, 8, 8), "2/3"), ((16, 16, 16), "3/2")])
@pytest.mark.parametrize("mask_nyquist", [False, True])
def test_rhs_nonlinear_fp64(pf, N, dealias, mask_nyquist):
    _say("rhs_nonlinear", N, dealias, mask_nyquist)
    o = po.OracleSolver(N, 0.01, dealias=dealias, mask_nyquist=mask_nyquist)
    plan = pf.Plan(N, 0.01, precision="fp64", dealias=dealias, mask_nyquist=mask_nyquist, batch=2)
    rs = np.random.RandomState(1)
    uh = _rand_spec(rs, (2, len(N)) + o.g.shape_fourier)
    out = plan.new_state()
    plan.rhs_nonlinear(_dev(uh, plan.complex_dtype), out)
    ref = np.stack([o.rhs_nonlinear(uh[b]) for b in range(2)])
    err = po.rel_l2(_host(out), ref)
    _say("  rel err", err)
    assert err < FP64_TOL


CASES = [
    # (N, integrator, nonlinear, noise_type, noise_mag, dealias)
    ((16, 16), "explicit_pc", True, "thermal", 0.3, "3/2"),
    ((16, 16), "implicit_pc", False, "thermal", 0.5, "3/2"),
    ((16, 24), "rk2", True, "thermal", 0.0, "2/3"),
    ((16, 16), "rk3", True, "correlated", 0.2, "3/2"),
    ((16, 16), "rk4", True, "thermal", 0.0, "3/2"),
    ((8, 8, 12), "rk3", True, "thermal", 0.05, "3/2"),
    ((8, 12, 8), "implicit_pc", True, "correlated", 0.1, "2/3"),
    ((12, 8, 8), "explicit_pc", False, "thermal", 0.2, "3/2"),
    ((8, 8, 8), "rk4", True, "thermal", 0.0, "2/3"),
    ((8, 8, 8), "rk2", True, "thermal", 0.0, "3/2"),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[1]}-{len(c[0])}d-{c[3]}-{c[5]}" for c in CASES])
def test_step_parity_fp64(pf, case):
    N, integ, nonlin, ntype, nmag, dealias = case
    _say("step", case)
    nsteps, B, nu, dt = 3, 2, 0.05, 0.02
    plan = pf.Plan(N, dt, precision="fp64", integrator=integ, nonlinear=nonlin, viscosity=nu,
                   noise_type=ntype, noise_mag=nmag, dealias=dealias, batch=B)
    rs = np.random.RandomState(2)
    refs, noises = [], []
    u0 = None
    states = []
    for b in range(B):
        o = po.OracleSolver(N, dt, is_nonlinear=nonlin, time_integrator=integ, viscosity=nu,
                            noise_type=ntype, noise_mag=nmag, dealias=dealias, seed=100 + b)
        if ntype == "correlated":
            C = np.exp(-0.1 * o.g.k2) * (o.g.k2 > 0) + 0j
            o.set_correlation_function(C)
        uh = _rand_spec(rs, (len(N),) + o.g.shape_fourier) * 0.3
        o.initialize(uh, space="fourier", project=True)
        states.append(o.u_hat.copy())
        steps = [o.noise_stream.next() for _ in range(nsteps)]
        noises.append(steps)
        o.solve(nsteps, noise=steps)
        refs.append(o.u_hat)
    if ntype == "correlated":
        plan.set_correlation(_dev(C, plan.complex_dtype))
    raw = np.zeros((nsteps, B, 2, plan.noise_components) + plan.shape_fourier, dtype=complex)
    for b in range(B):
        for s in range(nsteps):
            raw[s, b, 0], raw[s, b, 1] = noises[b][s]
    ud = _dev(np.stack(states), plan.complex_dtype)
    plan.step(ud, nsteps, noise=_dev(raw, plan.complex_dtype))
    err = po.rel_l2(_host(ud), np.stack(refs))
    _say("  rel err", err)
    assert err < FP64_TOL
    assert plan.step_counter == nsteps


def test_tgv3d_fp32_vs_fp64_oracle(pf):
    N = (16, 16, 16)
    _say("tgv fp32", N)
    o = po.OracleSolver(N, 0.01, time_integrator="rk4", viscosity=1 / 400, dealias="2/3")
    o.initialize(po.taylor_green_3d(o.g))
    u_init = o.u_hat.copy()
    o.solve(20)
    plan = pf.Plan(N, 0.01, precision="fp32", integrator="rk4", viscosity=1 / 400, dealias="2/3")
    ud = _dev(u_init[None], plan.complex_dtype)
    plan.step(ud, 20)
    err = po.rel_l2(_host(ud)[0], o.u_hat)
    _say("  fp32 rel err", err)
    assert err < FP32_TOL


def test_diagnostics_and_spectrum(pf):
    N = (12, 8, 16)
    g = po.Grid(N)
    plan = pf.Plan(N, 0.01, precision="fp64", nonlinear=False, batch=2)
    rs = np.random.RandomState(3)
    uh = _rand_spec(rs, (2, 3) + g.shape_fourier)
    d = _host(plan.diagnostics(_dev(uh, plan.complex_dtype)))
    for b in range(2):
        assert np.isclose(d[b, 0], po.energy(g, uh[b]), rtol=1e-13)
        assert np.isclose(d[b, 1], po.enstrophy(g, uh[b]), rtol=1e-13)
    sp = _host(plan.spectrum(_dev(uh, plan.complex_dtype), 12))
    for b in range(2):
        assert np.allclose(sp[b], po.shell_spectrum(g, uh[b], 12), rtol=1e-12, atol=1e-14)


def test_rng_noise_statistics(pf):
    N = (64, 64)
    plan = pf.Plan(N, 0.01, precision="fp64", nonlinear=False, batch=2, noise_mag=1.0, seeds=[7, 8])
    out = torch.empty(plan.noise_shape(), dtype=plan.complex_dtype, device="cuda")
    plan.draw_noise(out)
    x = _host(out)
    re = x.real.ravel()
    im = x.imag.ravel()
    assert abs(re.mean()) < 0.02 and abs(im.mean()) < 0.02
    assert abs(re.var() - 1) < 0.03 and abs(im.var() - 1) < 0.03
    # members differ, and the same (seed, step) redraws the same numbers
    assert not np.allclose(x[0], x[1])
    out2 = torch.empty_like(out)
    plan.draw_noise(out2)
    assert torch.equal(out, out2)


def test_device_draws_match_restated_philox(pf):
    """FP64 device draws equal the oracle's restatement of Philox4x32-10 and
    Box-Muller (itself pinned by the generator's known-answer vectors), for
    every member, entry and mode, at a nonzero step counter."""
    N = (16, 16)
    seeds = [7, (1 << 40) + 3]
    plan = pf.Plan(N, 0.01, precision="fp64", nonlinear=False, batch=2, noise_mag=1.0, seeds=seeds)
    plan.step_counter = 5
    out = torch.empty(plan.noise_shape(), dtype=plan.complex_dtype, device="cuda")
    plan.draw_noise(out)
    nk = int(np.prod(plan.shape_fourier))
    want = po.device_noise_draws(seeds, [5, 5], nk, plan.noise_shape()[2])
    got = _host(out).reshape(want.shape)
    err = np.abs(got - want).max()
    _say("device draws vs restated Philox: max abs err", err)
    assert err < 1e-12


def test_rng_fp32_draws_match_fp64(pf):
    """The same Philox integers feed both precisions: the FP32 Box-Muller map
    (hardware log2, sin and cos) agrees with the FP64 library map to FP32
    rounding, and FP32 keeps unit variance."""
    N = (64, 64)
    kw = dict(nonlinear=False, batch=2, noise_mag=1.0, seeds=[7, 8])
    x = {}
    for prec in ("fp32", "fp64"):
        plan = pf.Plan(N, 0.01, precision=prec, **kw)
        out = torch.empty(plan.noise_shape(), dtype=plan.complex_dtype, device="cuda")
        plan.draw_noise(out)
        x[prec] = _host(out).astype(np.complex128)
    err = np.abs(x["fp32"] - x["fp64"])
    _say("fp32 vs fp64 draws: max abs err", err.max())
    assert err.max() < 2e-5
    re = x["fp32"].real.ravel()
    assert abs(re.var() - 1) < 0.03


def test_forcing_parity_fp64(pf):
    N = (8, 8, 8)
    _say("forcing")
    dt = 0.01
    o = po.OracleSolver(N, dt, time_integrator="rk3", viscosity=0.01, forcing="constant_power",
                        forcing_power=0.1)
    rs = np.random.RandomState(4)
    o.initialize(_rand_spec(rs, (3,) + o.g.shape_fourier) * 0.2, space="fourier")
    u_init = o.u_hat.copy()
    o.solve(3)
    plan = pf.Plan(N, dt, precision="fp64", integrator="rk3", viscosity=0.01, forcing="constant_power",
                   forcing_power=0.1)
    ud = _dev(u_init[None], plan.complex_dtype)
    plan.step(ud, 3)
    assert po.rel_l2(_host(ud)[0], o.u_hat) < FP64_TOL


@pytest.mark.parametrize("N,dealias,integ", [((8, 12, 16), "3/2", "rk3"), ((16, 16, 16), "2/3", "rk4"),
                                             ((24, 16), "2/3", "implicit_pc"), ((16, 24), "3/2", "explicit_pc")])
def test_fused_stage_equals_unfused(pf, N, dealias, integ):
    _say("fused vs unfused", N, dealias, integ)
    rs = np.random.RandomState(9)
    shape = (2, len(N)) + po.Grid(N).shape_fourier
    u = 0.2 * _rand_spec(rs, shape)
    out = []
    for fuse in (True, False):
        plan = pf.Plan(N, 0.01, precision="fp64", integrator=integ, viscosity=0.02, dealias=dealias, batch=2,
                       noise_mag=0.1, seeds=[3, 4])
        plan.set_fusion(fuse)
        ud = _dev(u, plan.complex_dtype)
        plan.step(ud, 3)
        out.append(_host(ud))
    assert po.rel_l2(out[0], out[1]) < 1e-14


TUNABLE = ("inv_axis0_curl", "inv_axis1", "row_c2r_cross_r2c", "fwd_axis1", "fwd_axis0", "fwd_axis0_stage")


def _radix_variant(name, v):
    """Launch variants that switch to radix plan 1 (column variant 5, row variant 3)."""
    return v == 5 or (v == 3 and name == "row_c2r_cross_r2c")


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize("N,dealias,integ", [((24, 24, 24), "3/2", "rk3"), ((32, 32, 32), "2/3", "rk4"),
                                             ((16, 24, 32), "3/2", "rk2"), ((64, 64), "3/2", "rk3"),
                                             ((48, 64), "2/3", "rk4")])
def test_launch_variants_bit_identical(pf, N, dealias, integ):
    """Every geometry or cache-policy variant of every pass (tile width,
    rows per block, nontemporal loads / stores with whole-line geometry)
    maps the same arithmetic to other threads: the right-hand side and the
    fused step agree bit for bit with the default plan, and so does the
    unfused step with the nontemporal stage kernel.  Radix-plan variants
    (another factorisation of the transform length) agree to FP32 rounding.
    Autotune leaves u_hat as it was and picks only existing variants."""
    _say("variants", N, dealias, integ)
    rs = np.random.RandomState(11)
    plan = pf.Plan(N, 1e-3, precision="fp32", integrator=integ, viscosity=0.01, dealias=dealias, batch=2,
                   noise_mag=0.05, seeds=[3, 4])
    u = _dev(0.1 * _rand_spec(rs, (2, len(N)) + plan.shape_fourier), plan.complex_dtype)
    plan.project(u)
    ref_rhs = torch.empty_like(u)
    plan.rhs_nonlinear(u, ref_rhs)
    ref_step = u.clone()
    plan.step_counter = 0
    plan.step(ref_step, 2)
    tried = 0
    radix_tried = 0
    for name in TUNABLE:
        for v in (1, 2, 3, 4, 5, 6, 7):
            try:
                plan.set_tiles({name: v})
            except NotImplementedError:
                continue
            tried += 1
            rhs = torch.empty_like(u)
            plan.rhs_nonlinear(u, rhs)
            w = u.clone()
            plan.step_counter = 0
            plan.step(w, 2)
            if _radix_variant(name, v):
                radix_tried += 1
                assert _rel(rhs, ref_rhs) < 2e-6 and _rel(w, ref_step) < 2e-6, (name, v)
            else:
                assert torch.equal(rhs, ref_rhs), (name, v)
                assert torch.equal(w, ref_step), (name, v)
    assert tried > 0
    # the separate stage kernel with nontemporal loads and stores
    plan.set_fusion(False)
    plan.set_tiles({})
    ref_unfused = u.clone()
    plan.step_counter = 0
    plan.step(ref_unfused, 2)
    plan.set_tiles({"stage": 7, "fwd_axis0": 7})
    w = u.clone()
    plan.step_counter = 0
    plan.step(w, 2)
    assert torch.equal(w, ref_unfused)
    plan.set_fusion(True)
    if dealias == "3/2" and N[0] in (24, 16):
        assert radix_tried > 0  # padded lengths 36 / 24 ... have a second radix plan
    plan.set_tiles({})
    before = u.clone()
    ms = plan.autotune(u, reps=2)
    assert torch.equal(u, before)
    tiles, fused = plan.tiles
    _say("chosen", {k: v for k, v in tiles.items() if v}, "fused", fused, {k: round(x, 4) for k, x in ms.items() if x})
    w = u.clone()
    plan.step_counter = 0
    plan.step(w, 2)
    if fused and not any(_radix_variant(k, v) for k, v in tiles.items()):
        assert torch.equal(w, ref_step)
    else:
        # the unfused sequence, or another radix plan, agrees to rounding
        err = _rel(w, ref_step)
        assert err < FP32_TOL, err


def test_set_tiles_rejects_missing_variants(pf):
    plan = pf.Plan((8, 8, 8), 1e-3, precision="fp64", integrator="rk3")
    with pytest.raises(NotImplementedError):
        plan.set_tiles({"inv_axis1": 1})
    with pytest.raises(ValueError):
        plan.set_tiles({"no_such_pass": 1})
    plan.set_tiles({"inv_axis1": 0})


def test_pass_profile_records_each_launch(pf):
    """Per-pass timing (Plan.profile / profile_read): a single-kernel pass
    hands its events to the launch, which records them at the kernel's start
    and end; the band reduction of the forcing is timed by events around it.
    Every pass of an unfused forced 3D step reports one record per launch
    with a positive duration, and their sum stays below the wall time."""
    import time

    _say("pass profile 32^3 rk3 forced, unfused")
    rs = np.random.RandomState(5)
    plan = pf.Plan((32, 32, 32), 1e-3, precision="fp32", integrator="rk3", viscosity=0.01, dealias="3/2",
                   noise_mag=0.05, forcing="constant_power", forcing_power=0.1, seeds=[7])
    plan.set_fusion(False)
    u = _dev(0.1 * _rand_spec(rs, (1, 3) + plan.shape_fourier), plan.complex_dtype)
    plan.project(u)
    plan.step(u, 1)
    torch.cuda.synchronize()
    plan.profile(True)
    t0 = time.perf_counter()
    plan.step(u, 2)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    rec = plan.profile_read()
    plan.profile(False)
    total = 0.0
    for name in ("inv_axis0_curl", "inv_axis1", "row_c2r_cross_r2c", "fwd_axis1", "fwd_axis0", "stage", "band"):
        ms, n = rec[name]
        assert n == 6, (name, n)  # 2 steps x 3 stages
        assert ms > 0.0, (name, ms)
        total += ms
    assert rec["fwd_axis0_stage"][1] == 0
    assert total < wall_ms, (total, wall_ms)
