"""Time integrators and their plugin registry (src/periodicflow/time_integrator.py).

Same registry (``DICT_TIME_INTEGRATORS``, ``@register_integrator``, :7-14),
class names, stage counts, increment tables and noise factors (:36-318).

The five built-ins carry ``pf_id``: ``Solver.solve`` then advances whole
steps inside libpfhip.so (``pf_step``), where each stage's update is fused
into the stage kernel.  Their ``stepping`` (called only on the per-stage
path) runs the same update from given terms through ``pf_stage_update``.

A user integrator registered with ``@register_integrator`` and no
``pf_id`` runs on the per-stage path: the solver computes each stage's terms
with ``pf_compute_rhs`` into device tensors and calls the user's
``stepping(unew, u0, stage, rhs_nonlinear, rhs_linear, forcing, noise)``,
which updates ``unew`` in place with torch operations on the GPU.

``noise_factor`` exists for every built-in; the reference's ``rk2`` and
``rk4`` lack it (their stepping ignores noise) and ``Solver.solve`` would
fail on them, so they carry zeros here.
"""
from __future__ import annotations

import numpy as np

DICT_TIME_INTEGRATORS = {}


def register_integrator(name):
    """time_integrator.py:10-14."""

    def decorator(cls):
        DICT_TIME_INTEGRATORS[name] = cls
        return cls

    return decorator


class TimeIntegrator:
    """Base class (time_integrator.py:17-33).

    ``linear_operator`` may be a NumPy array or a device tensor; ``_lin_op_inv``
    = 1 / (1 - dt/2 L) is kept for user integrators.  ``_plan`` is bound by
    the solver; built-in stepping goes through it.
    """

    pf_id = None
    n_stages = 1
    noise_factor = (0.0,)

    def __init__(self, dt: float, linear_operator=None, optimization: bool = False):
        self.dt = dt
        self._lin_op = linear_operator
        self._optimization = optimization
        self._plan = None
        if linear_operator is not None:
            self._lin_op_inv = 1 / (1 - self.dt / 2 * linear_operator)

    def update_linear_operator(self, L):
        """time_integrator.py:31-33; the plan is updated by the solver."""
        if self._lin_op is not None and hasattr(self._lin_op, "__setitem__"):
            self._lin_op[...] = L
        else:
            self._lin_op = L
        self._lin_op_inv = 1 / (1 - self.dt / 2 * L)

    def bind(self, plan):
        self._plan = plan

    def stepping(self, unew, u0, stage, rhs_nonlinear, rhs_linear, forcing, noise):
        if self._plan is None:
            raise RuntimeError("built-in stepping needs the solver's plan (Solver binds it)")
        sh = self._plan.state_shape()
        args = [t.reshape(sh) for t in (unew, u0, rhs_nonlinear, rhs_linear, forcing, noise)]
        self._plan.stage_update(stage, *args)


@register_integrator("explicit_pc")
class ExplicitPredictorCorrector(TimeIntegrator):
    """2-stage explicit predictor-corrector (Delong et al. 2013), :36-90.
    u = u0 + a dt (N + L + f) + sqrt(dt/2) n with a = (1/2, 1)."""

    pf_id = 0

    def __init__(self, dt, linear_operator=None, optimization=False):
        super().__init__(dt, linear_operator, optimization)
        self.n_stages = 2
        self._increment_factor = (0.5, 1.0)
        self.noise_factor = (0.0, 1.0)


@register_integrator("implicit_pc")
class ImplicitPredictorCorrector(TimeIntegrator):
    """2-stage implicit predictor-corrector (Delong et al. 2013), :93-158.
    u = (1 - dt/2 L)^-1 (u0 + a dt (N + f) + b dt L u + sqrt(dt/2) n) with
    (a, b) = (1/2, 0), (1, 1/2); the last stage uses L u0."""

    pf_id = 1

    def __init__(self, dt, linear_operator=None, optimization=False):
        super().__init__(dt, linear_operator, optimization)
        self.n_stages = 2
        self._increment_factor = [(0.5, 0.0), (1.0, 0.5)]
        self.noise_factor = (0.0, 1.0)


@register_integrator("rk2")
class RungeKuttaTVD2(TimeIntegrator):
    """TVD RK2 (Shu & Osher), :161-210: u = a u0 + b (u + dt (N + L + f))."""

    pf_id = 2

    def __init__(self, dt, linear_operator=None, optimization=False):
        super().__init__(dt, linear_operator, optimization)
        self.n_stages = 2
        self._increment_factors = [(0.0, 1.0), (0.5, 0.5)]
        self.noise_factor = (0.0, 0.0)


@register_integrator("rk3")
class RungeKuttaTVD3(TimeIntegrator):
    """Stochastic TVD RK3 (Delong et al. 2013), :213-275:
    u = a u0 + b (u + dt (N + L + f) + sqrt(dt) n)."""

    pf_id = 3

    def __init__(self, dt, linear_operator=None, optimization=False):
        super().__init__(dt, linear_operator, optimization)
        self.n_stages = 3
        self._increment_factor = [(0.0, 1.0), (3.0 / 4.0, 1.0 / 4.0), (1.0 / 3.0, 2.0 / 3.0)]
        self.noise_factor = [
            (2 * np.sqrt(2) + np.sqrt(3)) / 5,
            (-4 * np.sqrt(2) + 3 * np.sqrt(3)) / 5,
            (np.sqrt(2) - 2 * np.sqrt(3)) / 10,
        ]


@register_integrator("rk4")
class RungeKutta4(TimeIntegrator):
    """Low-storage RK4 (Bogey & Bailly 2004), :278-318: u = u0 + a dt (N + L + f)."""

    pf_id = 4

    def __init__(self, dt, linear_operator=None, optimization=False):
        super().__init__(dt, linear_operator, optimization)
        self.n_stages = 4
        self._increment_factor = [1.0 / 4.0, 1.0 / 3.0, 1.0 / 2.0, 1.0]
        self.noise_factor = (0.0, 0.0, 0.0, 0.0)


BUILTIN = {name: cls for name, cls in DICT_TIME_INTEGRATORS.items()}
