"""Spatial discretisation on the GPU (src/periodicflow/space_solver.py:67-310).

``SpaceSolver`` keeps the reference's attributes and methods.  Its arrays
are device tensors; every transform and right-hand-side term is computed by
libpfhip.so through the plan it owns.

Arrays (batch = 1 shapes; with batch > 1 a leading member axis is added):
  u (d, *N) real, u_hat (d, *Nf) complex with Nf = N[:-1] + (N[-1]//2+1,),
  rhs_linear, rhs_nonlinear, forcing, noise like u_hat,
  w1, w2 (d*d or d, *Nf): the raw standard-normal draws of the current
  step.  The reference stores them already multiplied by
  ``_noise_normal_factor``; here the kernels apply that factor, so w1/w2
  hold the unscaled draws.
"""
from __future__ import annotations

import numpy as np
import torch

from . import logger
from .basis import FourierSpace
from .plan import Plan


class SpaceSolver(FourierSpace):
    """space_solver.py:67-310 on one GPU."""

    def __init__(self, comm=None, mpi_rank=0, N=None, is_nonlinear=True, viscosity=1e-3, noise_type="thermal",
                 noise_mag=0.0, *, dt=1e-3, integrator="rk2", precision="fp64", batch=1, seeds=None,
                 noise_source="device", forcing=None, forcing_power=0.1, forcing_band=(1.0, 2.5), **kwargs):
        super().__init__(comm, mpi_rank, N, **kwargs)
        if self.dim not in (2, 3):
            raise ValueError("NS solver only supports 2D and 3D problems.")
        if noise_type not in ("thermal", "correlated"):
            raise ValueError(f"Invalid noise type '{noise_type}'. Options are 'thermal' or 'correlated'.")
        if noise_source not in ("device", "numpy"):
            raise ValueError("noise_source must be 'device' or 'numpy'")
        self._is_nonlinear = is_nonlinear
        self.viscosity = viscosity
        self.noise_type = noise_type
        self.noise_mag = noise_mag
        self.noise_source = noise_source
        self.batch = int(batch)
        self.precision = precision
        self.plan = Plan(self.N, dt, length=self.lengths, precision=precision, dealias=self.dealias,
                         mask_nyquist=self.nyquist_mask is not None, integrator=integrator,
                         nonlinear=is_nonlinear, viscosity=viscosity, noise_type=noise_type, noise_mag=noise_mag,
                         forcing=forcing, forcing_power=forcing_power, forcing_band=forcing_band, batch=batch)
        if seeds is None:
            seeds = [int(np.random.randint(0, 2**32 - 1)) + i for i in range(self.batch)]
        self.seeds = [int(s) for s in seeds]
        self.plan.set_seeds(self.seeds)
        self._rng = [np.random.RandomState(s % (2**32)) for s in self.seeds] if noise_source == "numpy" else None
        self._define_variables()
        # space_solver.py:143-148
        self._k0_mask_0 = np.where(self.k2 == 0, 0, 1)
        self._prod_n_sqrt = np.sqrt(np.prod(self.N))
        self._noise_normal_factor = self._k0_mask_0 / self._prod_n_sqrt / np.sqrt(2 - 2 / self.dim)
        self.linear_operator = -viscosity * self.k2
        self.correlation_func = np.zeros(self.shape_fourier_scalar, dtype=complex)
        self._custom_linear_operator = False

    # ------------------------------------------------------------ buffers
    def _view(self, t):
        return t[0] if self.batch == 1 else t

    def _define_variables(self):
        p = self.plan
        dev = "cuda"
        self._u = torch.zeros(p.physical_shape(), dtype=p.real_dtype, device=dev)
        self._u_hat = p.new_state(dev)
        self._rhs_linear = p.new_state(dev)
        self._rhs_nonlinear = p.new_state(dev)
        self._forcing = p.new_state(dev)
        self._noise = p.new_state(dev)
        self._raw = torch.zeros(p.noise_shape(), dtype=p.complex_dtype, device=dev)
        self.u = self._view(self._u)
        self.u_hat = self._view(self._u_hat)
        self.rhs_linear = self._view(self._rhs_linear)
        self.rhs_nonlinear = self._view(self._rhs_nonlinear)
        self.forcing = self._view(self._forcing)
        self.noise = self._view(self._noise)
        self.w1 = self._view(self._raw[:, 0])
        self.w2 = self._view(self._raw[:, 1])
        self.shape_local_physical = tuple(self.u.shape)
        self.shape_local_fourier = tuple(self.u_hat.shape)

    # ------------------------------------------------------------ initial condition
    def _to_device(self, a, dtype, shape):
        t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
        t = t.to(device="cuda", dtype=dtype)
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
        return t.contiguous()

    def initialize_velocity(self, u0, space: str = "physical", mask_zero_mode: bool = True):
        """space_solver.py:180-213."""
        if space.casefold() == "physical":
            if tuple(np.shape(u0)) != self.shape_local_physical:
                raise ValueError(
                    "SpaceSolver.initialize_velocity: Invalid shape for the input physical velocity. "
                    f"Expected {self.shape_local_physical}, got {tuple(np.shape(u0))}.")
            self.u.copy_(self._to_device(u0, self.plan.real_dtype, self.shape_local_physical))
            self.forward()
        elif space.casefold() == "fourier":
            if tuple(np.shape(u0)) != self.shape_local_fourier:
                raise ValueError(
                    "SpaceSolver.initialize_velocity: Invalid shape for the input Fourier velocity. "
                    f"Expected {self.shape_local_fourier}, got {tuple(np.shape(u0))}.")
            self.u_hat.copy_(self._to_device(u0, self.plan.complex_dtype, self.shape_local_fourier))
            if mask_zero_mode:
                self.plan.mask_zero_mode(self._u_hat)
            self.backward()
        else:
            raise ValueError("SpaceSolver.initialize_velocity: Invalid space type. Options are 'physical' or 'fourier'.")

    def forward(self):
        """u -> u_hat (space_solver.py:215-219)."""
        self.plan.forward(self._u, self._u_hat)

    def backward(self):
        """u_hat -> u (space_solver.py:221-225)."""
        self.plan.backward(self._u_hat, self._u)

    # ------------------------------------------------------------ operators
    def set_linear_operator(self, L):
        L = np.asarray(L.detach().cpu() if isinstance(L, torch.Tensor) else L)
        if np.iscomplexobj(L):
            if np.any(L.imag != 0):
                raise ValueError("the linear operator must be real")
            L = L.real
        self.linear_operator = np.broadcast_to(L, self.shape_fourier_scalar).astype(float).copy()
        self.plan.set_linear_operator(
            torch.as_tensor(self.linear_operator, device="cuda", dtype=self.plan.real_dtype).contiguous())
        self._custom_linear_operator = True

    def set_correlation_function(self, C):
        C = np.asarray(C.detach().cpu() if isinstance(C, torch.Tensor) else C)
        self.correlation_func = np.broadcast_to(C, self.shape_fourier_scalar).astype(complex).copy()
        self.plan.set_correlation(
            torch.as_tensor(self.correlation_func, device="cuda", dtype=self.plan.complex_dtype).contiguous())

    # ------------------------------------------------------------ noise
    def host_noise(self, n_steps):
        """Raw draws of the next n_steps from the NumPy streams, in the
        reference's order (space_solver.py:234-235): per member and step
        w1.real, w1.imag, w2.real, w2.imag, each a C-order randn(*shape)."""
        p = self.plan
        shape = (p.noise_components,) + p.shape_fourier
        out = np.empty((n_steps,) + p.noise_shape(), dtype=complex)
        for s in range(n_steps):
            for b, rs in enumerate(self._rng):
                out[s, b, 0] = rs.randn(*shape) + 1j * rs.randn(*shape)
                out[s, b, 1] = rs.randn(*shape) + 1j * rs.randn(*shape)
        return out

    def random_fields(self):
        """Fill w1, w2 with the raw draws of the current step (space_solver.py:227-238)."""
        if self.noise_source == "numpy":
            self._raw.copy_(torch.as_tensor(self.host_noise(1)[0], device="cuda", dtype=self.plan.complex_dtype))
        else:
            self.plan.draw_noise(self._raw)

    # ------------------------------------------------------------ right-hand side
    def compute_rhs(self, a: float = 0.0):
        """rhs_linear, rhs_nonlinear, noise(a) and forcing of the current
        u_hat (space_solver.py:302-310), from the current w1/w2."""
        self.plan.compute_rhs(self._u_hat, a, self._rhs_nonlinear, self._rhs_linear, self._forcing, self._noise,
                              noise=self._raw)
        self.add_forcing()

    def compute_rhs_linear(self):
        """space_solver.py:273-277 (stand-alone; compute_rhs fuses it)."""
        L = torch.as_tensor(self.linear_operator, device="cuda", dtype=self.plan.real_dtype)
        self._rhs_linear.copy_(L * self._u_hat)

    def compute_rhs_nonlinear(self):
        """space_solver.py:286-300."""
        self.plan.rhs_nonlinear(self._u_hat, self._rhs_nonlinear)

    def add_noise(self, a: float):
        """space_solver.py:240-252: noise(a) alone, through the fused kernel."""
        scratch = [torch.empty_like(self._u_hat) for _ in range(3)]
        self.plan.compute_rhs(self._u_hat, a, scratch[0], scratch[1], scratch[2], self._noise, noise=self._raw)

    def add_forcing(self):
        """Hook for user forcing (space_solver.py:267-271).  Built-in
        constant-power band forcing is computed inside compute_rhs."""
        pass

    # ------------------------------------------------------------ diagnostics (FP64)
    def diagnostics(self):
        """(batch, 2) NumPy: energy 1/2<|u|^2>, enstrophy 1/2<|omega|^2>."""
        return self.plan.diagnostics(self._u_hat).cpu().numpy()

    def spectrum(self, nshell=None):
        return self.plan.spectrum(self._u_hat, nshell).cpu().numpy()


__all__ = ["SpaceSolver", "logger"]
