"""The reference's Python API (Solver, integrator registry, SpaceSolver) on
the GPU, checked against the FP64 oracle on the same inputs.

Noise parity: with noise_source='numpy' the solver draws the reference's
np.random stream (same order as random_fields) from RandomState(seed), and
the oracle draws the same stream, so both see identical noise fields.
Bar: FP64 relative L2 <= 1e-12.
"""
import sys

import numpy as np
import pytest

from oracle import pf_oracle as po

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(*a):
    print("[gpu-api]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sim():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from periodicflow import simulation

    return simulation


def _np(t):
    return t.detach().cpu().numpy()


def test_tgv3d_rk3_matches_oracle(sim):
    _say("tgv3d rk3 16^3")
    N = [16, 16, 16]
    sol = sim.Solver(N, 0.01, 0.1, time_integrator="rk3", viscosity=1 / 400, seed=1)
    u0 = po.taylor_green_3d(po.Grid(N))
    sol.initialize(u0)
    assert sol.shape_local_physical == (3, 16, 16, 16)
    assert sol.shape_local_fourier == (3, 16, 16, 9)
    o = po.OracleSolver(N, 0.01, time_integrator="rk3", viscosity=1 / 400)
    o.initialize(u0)
    sol.solve()
    o.solve(10)
    assert sol.step == 10 and np.isclose(sol.t, 0.1)
    assert po.rel_l2(_np(sol.u_hat), o.u_hat) < TOL
    sol.backward()
    assert po.rel_l2(_np(sol.u), o.u) < TOL
    assert np.isclose(sol.energy(), po.energy(o.g, o.u_hat), rtol=1e-12)
    assert np.isclose(sol.dissipation(), po.dissipation(o.g, o.u_hat, 1 / 400), rtol=1e-12)


@pytest.mark.parametrize("integ", ["explicit_pc", "implicit_pc", "rk3"])
def test_thermal_stokes_with_reference_noise_stream(sim, integ):
    _say("thermal stokes", integ)
    N = [16, 24]
    nu, dt, nm, seed = 0.5, 0.02, 0.3, 77
    sol = sim.Solver(N, dt, 0.2, is_nonlinear=False, time_integrator=integ, viscosity=nu,
                     noise_type="thermal", noise_mag=nm, seed=seed, noise_source="numpy")
    o = po.OracleSolver(N, dt, is_nonlinear=False, time_integrator=integ, viscosity=nu,
                        noise_type="thermal", noise_mag=nm, seed=seed)
    rs = np.random.RandomState(3)
    u0 = rs.standard_normal(sol.shape_local_fourier) + 1j * rs.standard_normal(sol.shape_local_fourier)
    sol.initialize(u0, "fourier")
    o.initialize(u0, "fourier")
    sol.solve()
    o.solve(10)
    assert po.rel_l2(_np(sol.u_hat), o.u_hat) < TOL


def test_correlated_noise_example_setup(sim):
    """The correlated-noise example's operator set-up (examples/fluctuating_stokes_correlated_noise.py:81-90)."""
    _say("correlated noise")
    n, nu, kbt, ell = 32, 2.0, 0.5, 0.3
    N = [n, n]
    sol = sim.Solver(N, 0.01, 0.1, is_nonlinear=False, time_integrator="implicit_pc", viscosity=nu,
                     noise_type="correlated", noise_mag=(2 * nu * kbt) ** 0.5, seed=5, noise_source="numpy")
    o = po.OracleSolver(N, 0.01, is_nonlinear=False, time_integrator="implicit_pc", viscosity=nu,
                        noise_type="correlated", noise_mag=(2 * nu * kbt) ** 0.5, seed=5)
    dx = sol.x[0][1, 0] - sol.x[0][0, 0]
    kx, ky = sol.k
    Cl = np.exp(-0.5 * ((ell * kx) ** 2 + (ell * ky) ** 2)) * (np.sqrt(2 * np.pi * ell**2) / dx) ** 2
    Cl = Cl * np.where(sol.k2 == 0, 0, 1)
    rs = np.random.RandomState(4)
    u0 = 0.1 * (rs.standard_normal(sol.shape_local_fourier) + 1j * rs.standard_normal(sol.shape_local_fourier))
    for s in (sol, o):
        s.initialize(u0, "fourier")
        s.set_linear_operator(-nu * Cl)
        s.set_correlation_function(np.sqrt(Cl))
    sol.solve()
    o.solve(10)
    assert po.rel_l2(_np(sol.u_hat), o.u_hat) < TOL


def test_per_stage_path_equals_fused_path(sim):
    """A subclassed built-in (stepping overridden) runs compute_rhs + stepping
    per stage; it must agree with the fused pf_step path."""
    _say("per-stage vs fused")
    from periodicflow.time_integrator import RungeKuttaTVD3, register_integrator

    @register_integrator("rk3_per_stage")
    class RK3PerStage(RungeKuttaTVD3):
        def stepping(self, *args):
            super().stepping(*args)

    N = [8, 8, 12]
    kw = dict(viscosity=0.05, noise_mag=0.02, seed=9, noise_source="numpy")
    a = sim.Solver(N, 0.01, 0.05, time_integrator="rk3", **kw)
    b = sim.Solver(N, 0.01, 0.05, time_integrator="rk3_per_stage", **kw)
    assert a._fused() and not b._fused()
    rs = np.random.RandomState(6)
    u0 = 0.2 * (rs.standard_normal(a.shape_local_fourier) + 1j * rs.standard_normal(a.shape_local_fourier))
    for s in (a, b):
        s.initialize(u0, "fourier")
        s.solve()
    assert po.rel_l2(_np(b.u_hat), _np(a.u_hat)) < TOL


def test_user_integrator_forward_euler(sim):
    """A new integrator written against the plugin contract, in torch on the GPU."""
    _say("user integrator")
    from periodicflow.time_integrator import TimeIntegrator, register_integrator

    @register_integrator("euler_test")
    class Euler(TimeIntegrator):
        def __init__(self, dt, linear_operator=None, optimization=False):
            super().__init__(dt, linear_operator, optimization)
            self.n_stages = 1
            self.noise_factor = (0.0,)

        def stepping(self, unew, u0, stage, rhs_nonlinear, rhs_linear, forcing, noise):
            unew.copy_(u0 + self.dt * (rhs_nonlinear + rhs_linear + forcing))

    N = [16, 16]
    sol = sim.Solver(N, 0.01, 0.03, time_integrator="euler_test", viscosity=0.1, seed=2)
    o = po.OracleSolver(N, 0.01, viscosity=0.1)
    u0 = po.taylor_green_2d(o.g) + 0.1 * np.random.RandomState(8).standard_normal((2, 16, 16))
    sol.initialize(u0)
    o.initialize(u0)
    uh = o.u_hat
    for _ in range(3):
        nl, lin, f, _n = o.compute_rhs(uh, None, None, 0.0)
        uh = uh + 0.01 * (nl + lin + f)
    sol.solve()
    assert po.rel_l2(_np(sol.u_hat), uh) < TOL


def test_delegation_config_and_snapshots(sim, tmp_path):
    """read_config, attribute delegation, and HDF5 snapshots at step 0 and
    every write_interval (simulation.py:179-180, :215-217), reconfigured into
    the flat layout (io.py:142-196) and holding the physical velocity."""
    from periodicflow import io as pio

    try:
        h5 = pio.hdf5()
    except OSError as e:
        pytest.skip(f"no libhdf5: {e}")
    cfg = dict(N=[8, 8], dt=0.01, end_time=0.04, time_integrator="rk2", viscosity=0.1, write_solution=True,
               file_name=str(tmp_path / "tg2d"), write_interval=2, periodic=True, seed=3)
    sol = sim.Solver.read_config(cfg)
    g = po.Grid([8, 8])
    assert np.allclose(sol.k, g.k) and np.allclose(sol.k2, g.k2) and np.allclose(sol.x[0], g.x[0])
    assert sol.dim == 2 and sol.N == [8, 8]
    u0 = po.taylor_green_2d(g)
    sol.initialize(u0)
    sol.solve()
    u_end = _np(sol.u)
    sol.solution_writer.reconfigure_dataset()
    fid = h5.open(tmp_path / "tg2d.h5")
    try:
        assert h5.members(fid) == ["t", "u", "v", "x", "y"]
        assert h5.members(fid, "u") == ["0", "2", "4"]
        assert np.allclose(h5.read(fid, "t"), [0.0, 0.02, 0.04])
        assert np.array_equal(h5.read(fid, "v/4"), np.pad(u_end[1], [(0, 1)] * 2, mode="wrap"))
        # step 0 is the initial field (already solenoidal)
        assert np.allclose(h5.read(fid, "u/0")[:8, :8], u0[0], rtol=0, atol=1e-13)
    finally:
        h5.close(fid)
    with pytest.raises(AttributeError):
        sol.no_such_attribute


def test_snapshots_batch_fp32(sim, tmp_path):
    """With batch > 1 member m's components are the variables u_m, v_m, w_m;
    an FP32 state is written as float64 (the reference's dtype)."""
    from periodicflow import io as pio

    try:
        h5 = pio.hdf5()
    except OSError as e:
        pytest.skip(f"no libhdf5: {e}")
    N = [8, 8, 8]
    sol = sim.Solver(N, 0.01, 0.02, time_integrator="rk3", viscosity=0.1, precision="fp32", batch=2, seed=4,
                     write_solution=True, file_name=str(tmp_path / "ens"), write_interval=2)
    g = po.Grid(N)
    u0 = np.stack([po.taylor_green_3d(g), 0.5 * po.taylor_green_3d(g)])
    sol.initialize(u0)
    sol.solve()
    u_end = _np(sol.u).astype(np.float64)
    fid = h5.open(tmp_path / "ens.h5")
    try:
        assert h5.members(fid) == ["u_0", "u_1", "v_0", "v_1", "w_0", "w_1"]
        assert sorted(h5.members(fid, "v_1/3D"), key=int) == ["0", "2"]
        assert np.array_equal(h5.read(fid, "v_1/3D/2"), u_end[1, 1])
        assert np.array_equal(h5.read(fid, "w_0/3D/2"), u_end[0, 2])
        assert np.allclose(h5.read(fid, "u_1/3D/0"), u0[1, 0], rtol=0, atol=1e-6)
    finally:
        h5.close(fid)


def test_errors(sim):
    with pytest.raises(ValueError):
        sim.Solver([8, 8], 0.01, 0.1, time_integrator="rk7")
    with pytest.raises(ValueError):
        sim.Solver([8, 8], 0.01, 0.1, noise_type="pink")
    with pytest.raises(ValueError):
        sim.Solver([8, 8], 0.01, 0.1, dealias="1/2")
    sol = sim.Solver([8, 8], 0.01, 0.1)
    with pytest.raises(ValueError):
        sol.initialize(np.zeros((2, 8, 9)))
    with pytest.raises(ValueError):
        sol.initialize(np.zeros((2, 8, 8)), space="spectral")
    sol2 = sim.Solver([8, 8], 0.3, 1.0)
    sol2.initialize(np.zeros((2, 8, 8)))
    with pytest.raises(ValueError):
        sol2.solve()
