"""Solver API (src/periodicflow/simulation.py:12-225) on one GPU.

Same constructor keywords, ``read_config``, ``initialize``,
``set_linear_operator``, ``set_correlation_function``, ``solve`` and
attribute delegation to the space solver (:116-120).

Differences, all additive or forced by the single-GPU design:

* ``precision`` ('fp64' default, the reference's arithmetic; 'fp32' for the
  fast path), ``batch`` (members advanced together), ``seed`` / ``seeds``,
  ``noise_source`` ('device': counter-based Philox stream on the GPU;
  'numpy': the reference's ``np.random`` draw order, host generated, for
  bit-comparable noise), constant-power band ``forcing``.
* ``comm``/``mpi_rank``/``mpi_size`` are None/0/1; ``fft_plan``,
  ``decomposition`` and ``optimization`` are accepted and ignored.
* Errors raise Python exceptions instead of ``comm.Abort(1)``.
* ``end_time`` must be a whole number of steps from the current time (the
  reference's loop would otherwise never stop, :205).
* ``write_solution=True`` writes HDF5 snapshots through libhdf5
  (``periodicflow.io.HDF5Writer``, the reference's writer and layouts);
  with ``batch > 1`` member m's components are the variables ``u_m``,
  ``v_m``[, ``w_m``].
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import logger
from .space_solver import SpaceSolver
from .time_integrator import BUILTIN, DICT_TIME_INTEGRATORS, TimeIntegrator

_NOISE_CHUNK_BYTES = 256 << 20


class Solver:
    """Numerical solver for the incompressible (fluctuating) Navier-Stokes equations."""

    def __init__(self, N, dt, end_time, is_nonlinear=True, time_integrator="rk2", viscosity=1e-3,
                 noise_type="thermal", noise_mag=0.0, domain=2 * np.pi, check_interval=np.iinfo(np.int64).max,
                 verbose=False, write_solution=False, file_name=None, file_mode="w",
                 write_interval=np.iinfo(np.int64).max, periodic=False, dealias="3/2", mask_nyquist=False,
                 fft_plan="FFTW_MEASURE", decomposition="slab", optimization=True, *, precision="fp64",
                 batch=1, seed=None, seeds=None, noise_source="device", forcing=None, forcing_power=0.1,
                 forcing_band=(1.0, 2.5)):
        if not torch.cuda.is_available():
            raise RuntimeError("periodicflow (HIP) needs a GPU")
        self.comm = None
        self.mpi_rank = 0
        self.mpi_size = 1
        if time_integrator not in DICT_TIME_INTEGRATORS:
            raise ValueError(f"Failed to initialize time integrator: Time integrator '{time_integrator}' not found.")
        cls = DICT_TIME_INTEGRATORS[time_integrator]
        # the plan runs the nearest built-in ancestor's update (a user
        # integrator that overrides stepping only borrows the plan's terms)
        plan_integ = next((nm for base in cls.__mro__ for nm, bc in BUILTIN.items() if base is bc), "rk2")
        if seeds is None and seed is not None:
            seeds = [int(seed) + i for i in range(batch)]
        self.space_solver = SpaceSolver(
            None, 0, N, is_nonlinear=is_nonlinear, viscosity=viscosity, noise_type=noise_type,
            noise_mag=noise_mag, domain=domain, dealias=dealias, mask_nyquist=mask_nyquist, fft_plan=fft_plan,
            decomposition=decomposition, dt=dt, integrator=plan_integ,
            precision=precision, batch=batch, seeds=seeds, noise_source=noise_source, forcing=forcing,
            forcing_power=forcing_power, forcing_band=forcing_band)
        self.end_time = end_time
        self.write_solution = write_solution
        self.write_interval = write_interval
        self.check_interval = check_interval
        self.verbose = verbose
        self.file_name = file_name
        self.file_mode = file_mode
        self.periodic = periodic
        self.step = 0
        self.t = 0.0
        self.time_integrator = cls(dt, self.space_solver.linear_operator, optimization)
        self.time_integrator.bind(self.space_solver.plan)
        self._wall = 0.0
        # fft_plan (simulation.py:76) picks the planning effort: FFTW_ESTIMATE
        # keeps the default launch plans, the measuring modes time every
        # variant on the first advance (Plan.autotune)
        self.fft_plan = fft_plan
        self._tuned = str(fft_plan).upper() == "FFTW_ESTIMATE"
        self.tuning = None
        self.solution_writer = None
        self._init_writer(file_name, periodic)

    def __getattr__(self, name):
        # simulation.py:116-120
        if name != "space_solver" and "space_solver" in self.__dict__ and hasattr(self.space_solver, name):
            return getattr(self.space_solver, name)
        raise AttributeError(f"'{type(self).__name__}' object has no attribute '{name}'")

    def _init_writer(self, file_name, periodic):
        """simulation.py:133-154: the writer reads the physical velocity
        components, which ``backward()`` refreshes before every write."""
        if not self.write_solution:
            return
        if file_name is None:
            raise ValueError("write_solution=True needs a file_name")
        from .io import HDF5Writer

        ss = self.space_solver
        names = ["u", "v", "w"][: ss.dim]
        if ss.batch == 1:
            data = {nm: [ss.u[i]] for i, nm in enumerate(names)}
        else:
            data = {f"{nm}_{m}": [ss.u[m, i]] for m in range(ss.batch) for i, nm in enumerate(names)}
        self.solution_writer = HDF5Writer(None, 0, 1, file_name, solution={"space": ss, "data": data},
                                          time_points=[0.0, self.end_time],
                                          end_points=[x_end for _, x_end in ss.domain], periodic=periodic)

    @classmethod
    def read_config(cls, config: dict):
        """simulation.py:156-159."""
        return cls(**config)

    def initialize(self, u0, space="physical", project=True, mask_zero_mode=True):
        """simulation.py:161-180."""
        ss = self.space_solver
        ss.initialize_velocity(u0, space, mask_zero_mode)
        if project:
            ss.plan.project(ss._u_hat)
            ss.backward()
        if self.write_solution:
            self._write()

    def set_linear_operator(self, L):
        """simulation.py:182-184."""
        self.space_solver.set_linear_operator(L)
        self.time_integrator.update_linear_operator(self.space_solver.linear_operator)

    def set_correlation_function(self, C):
        """simulation.py:186-187."""
        self.space_solver.set_correlation_function(C)

    # ------------------------------------------------------------ stepping
    def _fused(self):
        ti = self.time_integrator
        ss = self.space_solver
        return (type(ti).pf_id is not None and type(ti).stepping is TimeIntegrator.stepping
                and type(ss).add_forcing is SpaceSolver.add_forcing)

    def _steps_left(self):
        n = int(round((self.end_time - self.t) / self.time_integrator.dt))
        if n < 0 or not np.isclose(self.t + n * self.time_integrator.dt, self.end_time, atol=1e-10):
            raise ValueError("end_time must be reached in a whole number of steps of dt")
        return n

    def advance(self, n_steps):
        """Advance n_steps (all stages each) without output or checks."""
        ss = self.space_solver
        dt = self.time_integrator.dt
        if n_steps <= 0:
            return
        if not self._tuned:
            self.tuning = ss.plan.autotune(ss._u_hat)
            self._tuned = True
        if self._fused():
            uses_noise = ss.noise_mag != 0 and any(f != 0 for f in self.time_integrator.noise_factor)
            if ss.noise_source == "numpy" and uses_noise:
                per = int(np.prod(ss.plan.noise_shape())) * (16 if ss.precision == "fp64" else 8)
                chunk = max(1, _NOISE_CHUNK_BYTES // per)
                done = 0
                while done < n_steps:
                    m = min(chunk, n_steps - done)
                    raw = torch.as_tensor(ss.host_noise(m), device="cuda", dtype=ss.plan.complex_dtype)
                    ss.plan.step(ss._u_hat, m, noise=raw)
                    done += m
            else:
                ss.plan.step(ss._u_hat, n_steps)
        else:
            ti = self.time_integrator
            u0 = torch.empty_like(ss.u_hat)
            for _ in range(n_steps):
                u0.copy_(ss.u_hat)
                ss.random_fields()
                for stage in range(ti.n_stages):
                    ss.compute_rhs(ti.noise_factor[stage])
                    ti.stepping(ss.u_hat, u0, stage, ss.rhs_nonlinear, ss.rhs_linear, ss.forcing, ss.noise)
                ss.plan.step_counter = ss.plan.step_counter + 1
        self.step += n_steps
        self.t += n_steps * dt

    def _write(self):
        # simulation.py:215-217
        self.space_solver.backward()
        self.solution_writer.write(self.step)

    def solve(self):
        """simulation.py:189-225: step until end_time, with writes every
        write_interval steps and progress logs every check_interval steps."""
        n = self._steps_left()
        t0 = time.perf_counter()
        target = self.step + n
        while self.step < target:
            nxt = target
            for iv in (self.write_interval if self.write_solution else None, self.check_interval):
                if iv is not None and iv < np.iinfo(np.int64).max:
                    nxt = min(nxt, (self.step // iv + 1) * iv)
            self.advance(nxt - self.step)
            if self.write_solution and self.step % self.write_interval == 0:
                self._write()
            if self.check_interval < np.iinfo(np.int64).max and self.step % self.check_interval == 0:
                torch.cuda.synchronize()
                if self.verbose:
                    logger.info(f"t = {self.t:.6g}, step = {self.step}, wall = {time.perf_counter() - t0:.3f} s")
        torch.cuda.synchronize()
        if self.write_solution:
            # simulation.py:222-223
            self.solution_writer.close()
        self._wall += time.perf_counter() - t0
        if self.verbose:
            logger.info(f"solve finished: {n} steps in {time.perf_counter() - t0:.3f} s")

    # ------------------------------------------------------------ diagnostics
    def energy(self):
        d = self.space_solver.diagnostics()
        return d[:, 0] if self.space_solver.batch > 1 else float(d[0, 0])

    def enstrophy(self):
        d = self.space_solver.diagnostics()
        return d[:, 1] if self.space_solver.batch > 1 else float(d[0, 1])

    def dissipation(self):
        return 2 * self.space_solver.viscosity * self.enstrophy()
