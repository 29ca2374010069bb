"""pf_rhs_nonlinear_noalloc (work memory in a static device array) against
pf_rhs_nonlinear of a plan (work memory the plan allocates), through the C
ABI: two calls in a row, the second on a smaller grid, in FP64 and FP32.
Bar: relative L2 <= 1e-12 (FP64) and <= 1e-5 (FP32) between the two entry
points; the plan-less call computes its twiddle tables on the device, the
plan's come from the host, so the two agree to rounding, not bit for bit."""
import ctypes as C
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _say(*a):
    print("[no-alloc]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from periodicflow import _lib, plan
    return _lib, plan


def _noalloc(_lib, p, uh, out):
    lib = _lib.load()
    rc = lib.pf_rhs_nonlinear_noalloc(C.byref(p.desc), C.c_void_p(uh.data_ptr()), C.c_void_p(out.data_ptr()),
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, lib.pf_last_error().decode()


# each case: two calls in a row, the second smaller than the first
CASES = [
    ("fp64", [((32, 32, 32), "3/2", 2), ((16, 24, 12), "2/3", 1)]),
    ("fp32", [((96, 96, 96), "3/2", 1), ((64, 64), "3/2", 3)]),
    ("fp32", [((1024, 1024), "3/2", 1), ((32, 48, 16), "3/2", 2)]),
]


@pytest.mark.parametrize("prec,calls", CASES, ids=[f"{c[0]}-{c[1][0][0]}" for c in CASES])
def test_noalloc_matches_plan(mods, prec, calls):
    _lib, plan = mods
    tol = 1e-12 if prec == "fp64" else 1e-5
    rs = np.random.RandomState(61)
    for N, dealias, B in calls:
        _say(prec, N, dealias, "batch", B)
        p = plan.Plan(N, 0.01, precision=prec, dealias=dealias, batch=B)
        shp = p.state_shape()
        uh = torch.from_numpy(rs.standard_normal(shp) + 1j * rs.standard_normal(shp)).to("cuda", p.complex_dtype)
        ref = p.new_state()
        p.rhs_nonlinear(uh, ref)
        got = torch.full_like(ref, float("nan"))
        _noalloc(_lib, p, uh, got)
        torch.cuda.synchronize()
        err = float(torch.linalg.vector_norm((got - ref).to(torch.complex128)) /
                    torch.linalg.vector_norm(ref.to(torch.complex128)))
        _say("   rel L2 vs plan", err)
        assert err < tol
        # the same call again (the scratch keeps the last call's contents)
        again = torch.full_like(ref, float("nan"))
        _noalloc(_lib, p, uh, again)
        torch.cuda.synchronize()
        assert torch.equal(again, got)
