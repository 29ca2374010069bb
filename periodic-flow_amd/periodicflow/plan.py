"""Python handle of one HIP plan: a batch of same-shape members on one GPU.

Device buffers are torch tensors (torch is used for allocation, streams and
hand-off only); every computation is a call into libpfhip.so.  Shapes and
dtypes are checked here, before anything is launched.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import call


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Plan:
    """One pf_plan.  ``n``: grid per axis; ``length``: box lengths."""

    def __init__(self, n, dt, *, length=None, precision="fp64", dealias="3/2", mask_nyquist=False,
                 integrator="rk3", nonlinear=True, viscosity=1e-3, noise_type="thermal", noise_mag=0.0,
                 forcing=None, forcing_power=0.1, forcing_band=(1.0, 2.5), batch=1, seeds=None):
        lib = _lib.load()
        n = [int(v) for v in n]
        self.dim = len(n)
        if self.dim not in (2, 3):
            raise ValueError("NS solver only supports 2D and 3D problems.")
        length = [2 * np.pi] * self.dim if length is None else [float(v) for v in length]
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        dmap = {"3/2": 0, "pad": 0, "2/3": 1, "truncate": 1}
        if dealias not in dmap:
            raise ValueError(f"FourierSpace: Invalid dealiasing rule: {dealias}.")
        if integrator not in _lib.INTEGRATOR_IDS:
            raise ValueError(f"Time integrator '{integrator}' not found.")
        if noise_type not in ("thermal", "correlated"):
            raise ValueError(f"Invalid noise type '{noise_type}'. Options are 'thermal' or 'correlated'.")
        d = _lib.PfDesc()
        d.dim = self.dim
        for i in range(3):
            d.n[i] = n[i] if i < self.dim else 1
            d.length[i] = length[i] if i < self.dim else 1.0
        d.precision = _lib.PF_FP64 if precision == "fp64" else _lib.PF_FP32
        d.dealias = dmap[dealias]
        d.mask_nyquist = int(bool(mask_nyquist))
        d.integrator = _lib.INTEGRATOR_IDS[integrator]
        d.nonlinear = int(bool(nonlinear))
        d.noise_type = 0 if noise_type == "thermal" else 1
        d.forcing = 0 if forcing is None else 1
        d.batch = int(batch)
        d.viscosity = float(viscosity)
        d.dt = float(dt)
        d.noise_mag = float(noise_mag)
        d.forcing_power = float(forcing_power)
        d.forcing_kmin, d.forcing_kmax = (float(forcing_band[0]), float(forcing_band[1]))
        self.desc = d
        h = C.c_void_p()
        _lib.check(lib.pf_plan_create(C.byref(d), C.byref(h)), "pf_plan_create")
        self._h = h
        self._lib = lib
        info = _lib.PfPlanInfo()
        call("pf_plan_get_info", h, C.byref(info))
        self.n = n
        self.nk = int(info.n_fourier)
        self.nphys = int(info.n_physical)
        self.padded = [int(info.padded[i]) for i in range(self.dim)]
        self.noise_components = int(info.noise_components)
        self.workspace_bytes = int(info.workspace_bytes)
        self.batch = int(batch)
        self.precision = precision
        self.real_dtype = torch.float64 if precision == "fp64" else torch.float32
        self.complex_dtype = torch.complex128 if precision == "fp64" else torch.complex64
        self.shape_fourier = tuple(n[:-1]) + (n[-1] // 2 + 1,)
        self.shape_physical = tuple(n)
        self.integrator = integrator
        self.nonlinear = bool(nonlinear)
        self.noise_mag = float(noise_mag)
        self.noise_type = noise_type
        self._active = self.batch
        if seeds is not None:
            self.set_seeds(seeds)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.pf_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---------------------------------------------------------------- checks
    def _check(self, t, shape, dtype, name):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name} must be a CUDA (HIP) torch tensor")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
        if t.dtype != dtype:
            raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")

    def state_shape(self, ncomp=None):
        return (self.batch, self.dim if ncomp is None else ncomp) + self.shape_fourier

    def physical_shape(self, ncomp=None):
        return (self.batch, self.dim if ncomp is None else ncomp) + self.shape_physical

    def noise_shape(self):
        """Raw draws of one step for the active members."""
        return (self._active, 2, self.noise_components) + self.shape_fourier

    def new_state(self, device="cuda"):
        return torch.zeros(self.state_shape(), dtype=self.complex_dtype, device=device)

    # ---------------------------------------------------------------- calls
    def set_seeds(self, seeds):
        arr = (C.c_uint64 * self.batch)(*[int(s) & 0xFFFFFFFFFFFFFFFF for s in seeds])
        call("pf_set_seeds", self._h, arr)

    @property
    def step_counter(self):
        return int(self._lib.pf_get_step_counter(self._h))

    @step_counter.setter
    def step_counter(self, v):
        call("pf_set_step_counter", self._h, int(v))

    def add_step_counter(self, delta):
        """Add delta to the host step counter only (graph capture / replay
        bookkeeping; the device counters advance on the device)."""
        call("pf_add_step_counter", self._h, int(delta))

    @property
    def active(self):
        return self._active

    @active.setter
    def active(self, n):
        """Advance only members [0, n) (0 <= n <= batch)."""
        call("pf_set_active", self._h, int(n))
        self._active = int(n)

    def set_member(self, member, seed, step=0):
        """Seed and RNG step counter of one slot."""
        call("pf_set_member", self._h, int(member), int(seed) & 0xFFFFFFFFFFFFFFFF, int(step))

    def member_steps(self):
        arr = (C.c_int64 * self.batch)()
        call("pf_get_member_steps", self._h, arr)
        return [int(v) for v in arr]

    def set_linear_operator(self, L):
        if L is None:
            call("pf_set_linear_operator", self._h, None, _stream())
            return
        self._check(L, self.shape_fourier, self.real_dtype, "L")
        call("pf_set_linear_operator", self._h, _ptr(L), _stream())
        torch.cuda.current_stream().synchronize()

    def set_correlation(self, Cf):
        self._check(Cf, self.shape_fourier, self.complex_dtype, "C")
        call("pf_set_correlation", self._h, _ptr(Cf), _stream())
        torch.cuda.current_stream().synchronize()

    def forward(self, u, u_hat, ncomp=None):
        nc = self.dim if ncomp is None else ncomp
        self._check(u, self.physical_shape(nc), self.real_dtype, "u")
        self._check(u_hat, self.state_shape(nc), self.complex_dtype, "u_hat")
        call("pf_forward", self._h, _ptr(u), _ptr(u_hat), nc, _stream())

    def backward(self, u_hat, u, ncomp=None):
        nc = self.dim if ncomp is None else ncomp
        self._check(u_hat, self.state_shape(nc), self.complex_dtype, "u_hat")
        self._check(u, self.physical_shape(nc), self.real_dtype, "u")
        call("pf_backward", self._h, _ptr(u_hat), _ptr(u), nc, _stream())

    def project(self, u_hat):
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        call("pf_project", self._h, _ptr(u_hat), _stream())

    def mask_zero_mode(self, u_hat):
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        call("pf_mask_zero_mode", self._h, _ptr(u_hat), _stream())

    def rhs_nonlinear(self, u_hat, out):
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        self._check(out, self.state_shape(), self.complex_dtype, "rhs")
        if out.data_ptr() == u_hat.data_ptr():
            raise ValueError("rhs may not alias u_hat")
        call("pf_rhs_nonlinear", self._h, _ptr(u_hat), _ptr(out), _stream())

    def _check_noise(self, noise, n_steps):
        if noise is None:
            return None
        shp = (n_steps,) + self.noise_shape()
        self._check(noise, shp, self.complex_dtype, "noise")
        return _ptr(noise)

    def step(self, u_hat, n_steps=1, noise=None):
        """Advance n_steps; ``noise`` (n_steps, B, 2, nc, *shape_fourier) raw draws or None."""
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        call("pf_step", self._h, _ptr(u_hat), int(n_steps), self._check_noise(noise, n_steps), _stream())

    def compute_rhs(self, u_hat, a, rhs_nonlinear, rhs_linear, forcing, noise_out, noise=None):
        for t, nm in ((u_hat, "u_hat"), (rhs_nonlinear, "rhs_nonlinear"), (rhs_linear, "rhs_linear"),
                      (forcing, "forcing"), (noise_out, "noise")):
            self._check(t, self.state_shape(), self.complex_dtype, nm)
        nz = None
        if noise is not None:
            self._check(noise, self.noise_shape(), self.complex_dtype, "raw noise")
            nz = _ptr(noise)
        call("pf_compute_rhs", self._h, _ptr(u_hat), float(a), nz, _ptr(rhs_nonlinear), _ptr(rhs_linear),
             _ptr(forcing), _ptr(noise_out), _stream())

    def stage_update(self, stage, unew, u0, rhs_nonlinear, rhs_linear, forcing, noise):
        for t, nm in ((unew, "unew"), (u0, "u0"), (rhs_nonlinear, "rhs_nonlinear"), (rhs_linear, "rhs_linear"),
                      (forcing, "forcing"), (noise, "noise")):
            self._check(t, self.state_shape(), self.complex_dtype, nm)
        call("pf_stage_update", self._h, int(stage), _ptr(unew), _ptr(u0), _ptr(rhs_nonlinear),
             _ptr(rhs_linear), _ptr(forcing), _ptr(noise), _stream())

    def draw_noise(self, out):
        self._check(out, self.noise_shape(), self.complex_dtype, "raw noise")
        call("pf_draw_noise", self._h, _ptr(out), _stream())

    def diagnostics(self, u_hat, out=None):
        """(B, 2) FP64 tensor: energy 1/2<|u|^2>, enstrophy 1/2<|omega|^2>."""
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        if out is None:
            out = torch.empty((self.batch, 2), dtype=torch.float64, device=u_hat.device)
        call("pf_diagnostics", self._h, _ptr(u_hat), _ptr(out), _stream())
        return out

    def spectrum(self, u_hat, nshell=None):
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        if nshell is None:
            nshell = int(np.ceil(np.sqrt(sum((n // 2) ** 2 for n in self.n)))) + 2
        out = torch.empty((self.batch, nshell), dtype=torch.float64, device=u_hat.device)
        call("pf_spectrum", self._h, _ptr(u_hat), _ptr(out), int(nshell), _stream())
        return out

    def set_fusion(self, fuse_stage=True):
        """Fuse the axis-0 forward pass with the stage update in step() (default on)."""
        call("pf_set_fusion", self._h, int(bool(fuse_stage)))

    @property
    def tiles(self):
        """{pass name: launch variant} and whether the stage pass is fused."""
        n = len(_lib.PASS_NAMES)
        t = (C.c_int32 * n)()
        fused = C.c_int32(0)
        call("pf_get_tiles", self._h, t, C.byref(fused))
        return {nm: int(t[i]) for i, nm in enumerate(_lib.PASS_NAMES)}, bool(fused.value)

    def set_tiles(self, tiles):
        """Launch variant per pass (dict by pass name, unnamed passes 0): 0 the
        default tile width or row-pair count, 1 half of it, 2 twice it; column
        passes also 3 (two tiles per block), 4 (stage kinds: 1024-thread
        blocks), 5 (the radix-2/3 plan), 6 (nontemporal stores) and 7
        (nontemporal stores and loads); row passes also 3 (the radix-2/3
        plan) and 4 (nontemporal stores).  For Stokes flow with implicit PC,
        "stage": 1 steps one launch per step instead of many steps with the
        state in registers; for other plans "stage": 7 gives the separate
        stage kernel nontemporal loads and stores.  Results are bit-identical
        across variants except the radix plans, which change rounding only."""
        n = len(_lib.PASS_NAMES)
        unknown = set(tiles) - set(_lib.PASS_NAMES)
        if unknown:
            raise ValueError(f"unknown passes {sorted(unknown)}")
        t = (C.c_int32 * n)(*[int(tiles.get(nm, 0)) for nm in _lib.PASS_NAMES])
        call("pf_set_tiles", self._h, t)

    def autotune(self, u_hat, reps=5):
        """Time every launch variant of every nonlinear pass (and the fused
        against the unfused stage pass) on a copy of u_hat, keep the fastest;
        returns {pass name: chosen average ms}.  Every variant is bitwise the
        default except the radix-2/3 factorisation of a length with a factor
        3 (column variant 5, row variant 3), which changes rounding only."""
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        n = len(_lib.PASS_NAMES)
        ms = (C.c_double * n)()
        call("pf_autotune", self._h, _ptr(u_hat), int(reps), ms, _stream())
        return {nm: float(ms[i]) for i, nm in enumerate(_lib.PASS_NAMES)}

    def profile(self, enable=True, passes=None):
        """Start (or stop) per-pass HIP-event timing; clears earlier records.
        ``passes``: names of the passes that record (default all); each
        recorded launch adds two event records to the stream."""
        if passes is None:
            mask = 0xFFFFFFFF
        else:
            unknown = set(passes) - set(_lib.PASS_NAMES)
            if unknown:
                raise ValueError(f"unknown passes {sorted(unknown)}")
            mask = sum(1 << _lib.PASS_NAMES.index(nm) for nm in passes)
        call("pf_profile_select", self._h, mask)
        call("pf_profile_enable", self._h, int(bool(enable)))

    def profile_read(self):
        """{pass name: (total ms, launches)} of what was recorded since profile()."""
        n = len(_lib.PASS_NAMES)
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        call("pf_profile_read", self._h, ms, cnt, n)
        return {nm: (float(ms[i]), int(cnt[i])) for i, nm in enumerate(_lib.PASS_NAMES)}

    def step_accumulate_sk(self, u_hat, n_steps, acc):
        """Advance n_steps with acc += sum over members and components of
        |u_hat|^2 after every step (FP64 acc of shape shape_fourier)."""
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        self._check(acc, self.shape_fourier, torch.float64, "acc")
        call("pf_step_accumulate_sk", self._h, _ptr(u_hat), int(n_steps), _ptr(acc), _stream())

    def accumulate_sk(self, u_hat, acc):
        self._check(u_hat, self.state_shape(), self.complex_dtype, "u_hat")
        self._check(acc, self.shape_fourier, torch.float64, "acc")
        call("pf_accumulate_sk", self._h, _ptr(u_hat), _ptr(acc), _stream())


def step_multi(plans, states, n_steps=1):
    """Advance several plans (an ensemble's groups of different grids) by
    n_steps with their passes issued together, pairs of compiled transform
    lengths sharing one launch (pf_step_multi); returns the number of paired
    launches.  Results are those of each plan's own step()."""
    if len(plans) != len(states):
        raise ValueError("one state per plan")
    for p, st in zip(plans, states):
        p._check(st, p.state_shape(), p.complex_dtype, "u_hat")
    n = len(plans)
    hs = (C.c_void_p * n)(*[p._h.value for p in plans])
    us = (C.c_void_p * n)(*[st.data_ptr() for st in states])
    paired = C.c_int32(0)
    call("pf_step_multi", hs, us, n, int(n_steps), C.byref(paired), _stream())
    return int(paired.value)


def version():
    return _lib.load().pf_version().decode()
