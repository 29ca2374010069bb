"""GPU parity against the FP64 oracle at grid sizes with a factor 5.

The reference transforms any grid size through FFTW
(src/periodicflow/basis.py:62-64); these sizes run on the radix-5 butterfly
(transform lengths 20, 30, 40, 60, 80, 120, 160, 240, 320, 480: N and 3N/2
for N = 20 .. 320, and N alone under the 2/3 rule).  Bars: FP64 relative
L2 <= 1e-12, FP32 <= 1e-5 against the FP64 oracle.
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
    print("[gpu-radix5]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pf():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import periodicflow.plan as m
    return m


def _rand_spec(rs, shape):
    return rs.standard_normal(shape) + 1j * rs.standard_normal(shape)


def _dev(x, cdt):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=cdt)


def _host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("N", [(20, 40), (30, 20, 40), (160, 80), (480, 60), (40, 120, 80)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_forward_backward(pf, N, prec):
    _say("forward/backward", N, prec)
    g = po.Grid(N)
    plan = pf.Plan(N, 0.01, precision=prec, nonlinear=False, dealias="2/3")
    rs = np.random.RandomState(5)
    u = rs.standard_normal((1, len(N)) + g.shape_physical)
    uh = plan.new_state()
    plan.forward(torch.from_numpy(u).to("cuda", plan.real_dtype), uh)
    tol = FP64_TOL if prec == "fp64" else FP32_TOL
    err = po.rel_l2(_host(uh)[0], g.forward(u[0]))
    sh = _rand_spec(rs, (1, len(N)) + g.shape_fourier)
    back = torch.empty((1, len(N)) + g.shape_physical, dtype=plan.real_dtype, device="cuda")
    plan.backward(_dev(sh, plan.complex_dtype), back)
    errb = po.rel_l2(_host(back)[0], g.backward(sh[0]))
    _say("  forward", err, "backward", errb)
    assert err < tol and errb < tol


RHS = [((20, 20, 20), "3/2"), ((40, 20, 40), "2/3"), ((80, 40, 40), "3/2"), ((30, 60, 20), "2/3"),
       ((160, 320), "3/2"), ((480, 240), "2/3"), ((120, 80), "2/3")]


@pytest.mark.parametrize("N,dealias", RHS, ids=[f"{'x'.join(map(str, n))}-{d}" for n, d in RHS])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rhs_nonlinear(pf, N, dealias, prec):
    _say("rhs_nonlinear", N, dealias, prec)
    o = po.OracleSolver(N, 0.01, dealias=dealias)
    plan = pf.Plan(N, 0.01, precision=prec, dealias=dealias)
    rs = np.random.RandomState(6)
    uh = _rand_spec(rs, (1, len(N)) + o.g.shape_fourier)
    out = plan.new_state()
    plan.rhs_nonlinear(_dev(uh, plan.complex_dtype), out)
    err = po.rel_l2(_host(out)[0], o.rhs_nonlinear(uh[0]))
    _say("  rel err", err)
    assert err < (FP64_TOL if prec == "fp64" else FP32_TOL)


STEPS = [((40, 40, 40), "rk3", "3/2"), ((20, 30, 40), "rk4", "2/3"), ((160, 80), "rk2", "3/2"),
         ((80, 80), "explicit_pc", "2/3")]


@pytest.mark.parametrize("N,integ,dealias", STEPS, ids=[f"{'x'.join(map(str, n))}-{i}-{d}" for n, i, d in STEPS])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_steps_vs_oracle(pf, N, integ, dealias, prec):
    """Four steps of a random solenoidal field (no noise) scaled to max |u| = 1,
    so that the explicit schemes are well inside their stability limit
    (u k_max dt <= 0.5), fused path, against the oracle's solve loop."""
    _say("steps", N, integ, dealias, prec)
    nu, dt, n = 0.02, 2e-3, 4
    o = po.OracleSolver(N, dt, time_integrator=integ, viscosity=nu, dealias=dealias)
    rs = np.random.RandomState(7)
    uh = _rand_spec(rs, (len(N),) + o.g.shape_fourier)
    o.initialize(uh, space="fourier", project=True)
    o.initialize(o.u_hat / np.abs(o.u).max(), space="fourier", project=True)
    u0 = o.u_hat.copy()
    plan = pf.Plan(N, dt, precision=prec, integrator=integ, viscosity=nu, dealias=dealias)
    st = _dev(u0[None], plan.complex_dtype)
    plan.step(st, n)
    o.solve(n)
    assert np.linalg.norm(o.u_hat) < 2 * np.linalg.norm(u0)  # the oracle run stays bounded
    err = po.rel_l2(_host(st)[0], o.u_hat)
    _say("  rel err", err)
    assert err < (FP64_TOL if prec == "fp64" else FP32_TOL)


@pytest.mark.parametrize("N,dealias", [((40, 40, 40), "3/2"), ((80, 160), "2/3")])
def test_launch_variants_bit_identical(pf, N, dealias):
    """Every launch variant at radix-5 lengths gives the default's bits."""
    from periodicflow import _lib
    _say("variants", N, dealias)
    rs = np.random.RandomState(8)
    plan = pf.Plan(N, 1e-3, precision="fp32", integrator="rk3", viscosity=0.01, dealias=dealias)
    u = _dev(0.1 * _rand_spec(rs, (1, len(N)) + plan.shape_fourier), plan.complex_dtype)
    plan.project(u)
    ref = torch.empty_like(u)
    plan.rhs_nonlinear(u, ref)
    tried = 0
    for name in _lib.PASS_NAMES:
        for v in (1, 2, 3, 4, 5, 6, 7):
            try:
                plan.set_tiles({name: v})
            except (NotImplementedError, ValueError):
                continue
            tried += 1
            out = torch.empty_like(u)
            plan.rhs_nonlinear(u, out)
            assert torch.equal(out, ref), (name, v)
    plan.set_tiles({})
    assert tried > 0
