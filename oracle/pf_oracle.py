"""FP64 NumPy restatement of periodicflow's Fourier-Galerkin RHS and time stepping.

TEST INFRASTRUCTURE ONLY.  This module is the oracle the HIP path is checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline leg may import it, and only as the checker or the timed CPU baseline;
the product path (``periodicflow`` on top of ``libpfhip.so``) never calls it.

What it restates (reference paths relative to the reference checkout):

* ``src/periodicflow/basis.py:34-92``        grid, wavenumbers, dealias table
* ``src/periodicflow/math.py:5-100``         cross, curl, linear operator, Leray
* ``src/periodicflow/space_solver.py:15-310`` noise, RHS assembly order
* ``src/periodicflow/time_integrator.py:17-318`` the five integrators
* ``src/periodicflow/simulation.py:161-225`` initialise and the step loop

The transforms and the dealiasing live in shenfun in the reference.  shenfun
is not importable here, so they are restated with ``numpy.fft`` under fixed
conventions:

* forward  u_hat_k = (1/prod M) sum_x u(x) exp(-i k.x), backward unnormalised;
  the real-to-complex axis is the last one.
* wavenumbers: integer ``fftfreq`` on complex axes (Nyquist = -N/2) and
  ``rfftfreq`` on the last axis (Nyquist = +N/2), scaled by 2 pi / L.
* 3/2 rule: backward copies the N retained modes into the M = 3N/2 spectrum;
  a complex axis splits its Nyquist coefficient half-and-half between +-N/2,
  the real-to-complex axis takes the real part of its Nyquist coefficient
  and halves it.  Forward keeps |k| < N/2, sums the +-N/2 coefficients into
  the Nyquist slot on complex axes and doubles the real part on the last axis.
* 2/3 rule: the spectral input of every padded-space backward transform is
  copied with |k_i| > floor(N_i/3) zeroed on every axis; forward outputs are
  not masked.
* c2r ignores the imaginary part of the DC and Nyquist entries of the last
  axis after the complex passes (``numpy.fft.irfft`` semantics).

Pinning (see ``tests/test_oracle.py``): brute-force DFTs on tiny grids, exact
evaluation of trigonometric polynomials on the padded grid, exact truncated
convolutions, Taylor-Green closed forms, the exact 2D Taylor-Green decay,
amplification polynomials of every integrator, projection invariants, energy
conservation of the dealiased nonlinear term, and the Crank-Nicolson
stationary variance of the fluctuating Stokes problem.  The Nyquist detail of
the 3/2 rule and the exact 2/3 boundary are shenfun conventions that nothing
available here can pin: parity unpinned for those modes only.
"""
from __future__ import annotations

import math as _m

import numpy as np

class _FFT:
    """FFT backend of the oracle: numpy.fft by default; scipy.fft with a
    worker count for the timed CPU baseline (set_fft_workers)."""

    workers = None

    def __getattr__(self, name):
        if self.workers is None:
            return getattr(np.fft, name)
        import functools

        import scipy.fft as sf
        return functools.partial(getattr(sf, name), workers=self.workers)


_F = _FFT()


def set_fft_workers(n):
    """Use scipy.fft with ``n`` threads (None: numpy.fft, single threaded)."""
    _F.workers = n


SQRT2 = np.sqrt(2.0)
SQRT3 = np.sqrt(3.0)

# integrator tables, src/periodicflow/time_integrator.py:45-47, :102-107,
# :170-174, :222-232, :287-288
RK3_NOISE = (
    (2 * SQRT2 + SQRT3) / 5,
    (-4 * SQRT2 + 3 * SQRT3) / 5,
    (SQRT2 - 2 * SQRT3) / 10,
)
INTEGRATORS = {
    "explicit_pc": dict(n_stages=2, noise_factor=(0.0, 1.0)),
    "implicit_pc": dict(n_stages=2, noise_factor=(0.0, 1.0)),
    "rk2": dict(n_stages=2, noise_factor=(0.0, 0.0)),
    "rk3": dict(n_stages=3, noise_factor=RK3_NOISE),
    "rk4": dict(n_stages=4, noise_factor=(0.0, 0.0, 0.0, 0.0)),
}


def parse_domain(domain, dim):
    """Domain parsing of src/periodicflow/basis.py:50-60 (float or list)."""
    if isinstance(domain, (float, np.floating)):
        return [(0.0, float(domain)) for _ in range(dim)]
    if isinstance(domain, (list, tuple)):
        if all(isinstance(i, (list, tuple)) for i in domain):
            return [tuple(map(float, i)) for i in domain]
        if all(isinstance(i, (float, np.floating)) for i in domain):
            return [tuple(map(float, domain)) for _ in range(dim)]
    raise ValueError("Invalid domain.")


class Grid:
    """Wavenumbers, masks and padded transforms of one periodic box.

    Follows src/periodicflow/basis.py:34-92: last axis real-to-complex,
    ``k`` broadcast and scaled (:65, :75), ``k2`` (:76), ``k_over_k2`` with
    k2 == 0 replaced by 1 (:77), dealias table (:79-87).
    """

    def __init__(self, N, domain=2 * np.pi, dealias="3/2", mask_nyquist=False):
        self.N = [int(n) for n in N]
        self.dim = d = len(self.N)
        if d not in (2, 3):
            raise ValueError("NS solver only supports 2D and 3D problems.")
        self.domain = parse_domain(domain, d)
        self.Lbox = [b - a for a, b in self.domain]
        table = {"3/2": 1.5, "pad": 1.5, "2/3": 1.0, "truncate": 1.0}
        if dealias not in table:
            raise ValueError(f"FourierSpace: Invalid dealiasing rule: {dealias}.")
        self.padding = table[dealias]
        self.dealias_direct = self.padding == 1.0
        self.M = [int(np.floor(self.padding * n)) for n in self.N]
        nh = self.N[-1] // 2 + 1
        self.shape_fourier = tuple(self.N[:-1]) + (nh,)
        self.shape_physical = tuple(self.N)
        axes = []
        for i, n in enumerate(self.N):
            scale = 2 * np.pi / self.Lbox[i]
            if i < d - 1:
                ki = np.fft.fftfreq(n, 1.0 / n) * scale
            else:
                ki = np.fft.rfftfreq(n, 1.0 / n) * scale
            shp = [1] * d
            shp[i] = ki.size
            axes.append(ki.reshape(shp))
        self.kaxes = axes
        self.k = np.array(np.broadcast_arrays(*axes))
        self.k2 = np.sum(self.k**2, axis=0)
        self.k_over_k2 = self.k / np.where(self.k2 == 0, 1, self.k2)
        # integer mode numbers (unscaled) for masks
        ints = []
        for i, n in enumerate(self.N):
            if i < d - 1:
                ii = np.fft.fftfreq(n, 1.0 / n).astype(np.int64)
            else:
                ii = np.arange(n // 2 + 1, dtype=np.int64)
            shp = [1] * d
            shp[i] = ii.size
            ints.append(ii.reshape(shp))
        self.kint = ints
        # 2/3 mask: keep |k_i| <= floor(N_i/3) on every axis
        m = np.ones(self.shape_fourier, dtype=bool)
        for i, n in enumerate(self.N):
            m = m & (np.abs(ints[i]) <= n // 3)
        self.mask23 = m
        # Nyquist mask (shenfun get_mask_nyquist): zero where any axis sits at
        # its Nyquist index (N even)
        nm = np.ones(self.shape_fourier, dtype=bool)
        for i, n in enumerate(self.N):
            if n % 2 == 0:
                nm = nm & (np.abs(ints[i]) != n // 2)
        self.nyquist_mask = nm if mask_nyquist else None
        # mesh, basis.py:74 (S.local_mesh): x_j = a + j L / N
        mesh = []
        for i, n in enumerate(self.N):
            a, _ = self.domain[i]
            xi = a + np.arange(n) * self.Lbox[i] / n
            shp = [1] * d
            shp[i] = n
            mesh.append(xi.reshape(shp))
        self.x = mesh
        # r2c weights for Parseval sums over the half spectrum
        w = np.full(self.shape_fourier, 2.0)
        w[..., 0] = 1.0
        if self.N[-1] % 2 == 0:
            w[..., -1] = 1.0
        self.r2c_weight = w

    # ---- unpadded transforms (V.forward / V.backward) ----
    def forward(self, u):
        """space_solver.py:215-219 -> shenfun V.forward: normalised rfftn."""
        axes = tuple(range(-self.dim, 0))
        return _F.rfftn(u, axes=axes) / np.prod(self.N)

    def backward(self, uh):
        """space_solver.py:221-225 -> shenfun V.backward: unnormalised irfftn."""
        axes = tuple(range(-self.dim, 0))
        return _F.irfftn(uh, s=self.N, axes=axes) * np.prod(self.N)

    # ---- padded transforms (Vp.backward / Vp.forward) ----
    def _pad_c2c(self, a, ax, N, M):
        if M == N:
            return a
        shp = list(a.shape)
        shp[ax] = M
        out = np.zeros(shp, dtype=complex)
        lo = [slice(None)] * a.ndim
        hi = [slice(None)] * a.ndim
        h = N // 2
        lo[ax] = slice(0, h)
        out[tuple(lo)] = a[tuple(lo)]
        hi[ax] = slice(N - h + 1, N)  # k = -h+1 .. -1
        dst = [slice(None)] * a.ndim
        dst[ax] = slice(M - h + 1, M)
        out[tuple(dst)] = a[tuple(hi)]
        nyq = [slice(None)] * a.ndim
        nyq[ax] = slice(h, h + 1)  # stored k = -N/2
        v = 0.5 * a[tuple(nyq)]
        p = [slice(None)] * a.ndim
        p[ax] = slice(h, h + 1)  # +N/2 in the padded spectrum
        out[tuple(p)] += v
        q = [slice(None)] * a.ndim
        q[ax] = slice(M - h, M - h + 1)  # -N/2 in the padded spectrum
        out[tuple(q)] += v
        return out

    def _trunc_c2c(self, a, ax, N, M):
        if M == N:
            return a
        shp = list(a.shape)
        shp[ax] = N
        out = np.zeros(shp, dtype=complex)
        h = N // 2
        s = [slice(None)] * a.ndim
        s[ax] = slice(0, h)
        out[tuple(s)] = a[tuple(s)]
        src = [slice(None)] * a.ndim
        src[ax] = slice(M - h + 1, M)
        dst = [slice(None)] * a.ndim
        dst[ax] = slice(N - h + 1, N)
        out[tuple(dst)] = a[tuple(src)]
        p = [slice(None)] * a.ndim
        p[ax] = slice(h, h + 1)
        q = [slice(None)] * a.ndim
        q[ax] = slice(M - h, M - h + 1)
        out[tuple(p)] = a[tuple(p)] + a[tuple(q)]
        return out

    def backward_padded(self, uh):
        """Padded backward of one or more scalar spectra (..., *shape_fourier).

        shenfun Vp/Wp.backward as used at space_solver.py:284, :291.
        """
        d = self.dim
        a = np.array(uh, dtype=complex, copy=True)
        off = a.ndim - d
        if self.dealias_direct:
            a = a * self.mask23
        for i in range(d - 1):
            a = self._pad_c2c(a, off + i, self.N[i], self.M[i])
            a = _F.ifft(a, axis=off + i) * self.M[i]
        Nl, Ml = self.N[-1], self.M[-1]
        if Ml != Nl:
            shp = list(a.shape)
            shp[-1] = Ml // 2 + 1
            b = np.zeros(shp, dtype=complex)
            h = Nl // 2
            b[..., :h] = a[..., :h]
            b[..., h] = 0.5 * a[..., h].real
            a = b
        return _F.irfft(a, n=Ml, axis=-1) * Ml

    def forward_padded(self, u):
        """Padded forward with truncation (Vp.forward, space_solver.py:296)."""
        d = self.dim
        off = u.ndim - d
        Nl, Ml = self.N[-1], self.M[-1]
        a = _F.rfft(u, axis=-1) / Ml
        if Ml != Nl:
            h = Nl // 2
            b = np.zeros(a.shape[:-1] + (h + 1,), dtype=complex)
            b[..., :h] = a[..., :h]
            b[..., h] = 2.0 * a[..., h].real
            a = b
        for i in reversed(range(d - 1)):
            a = _F.fft(a, axis=off + i) / self.M[i]
            a = self._trunc_c2c(a, off + i, self.N[i], self.M[i])
        return a

    def physical_padded_shape(self):
        return tuple(self.M)


# ---------------- point-wise kernels, src/periodicflow/math.py ----------------

def cross(a, b):
    """math.py:5-28.  3D: a x b.  2D: (a1 b, -a0 b) with scalar b."""
    if a.shape[0] == 2:
        return np.array([a[1] * b, -a[0] * b])
    return np.array([
        a[1] * b[2] - a[2] * b[1],
        a[2] * b[0] - a[0] * b[2],
        a[0] * b[1] - a[1] * b[0],
    ])


def curl(k, a):
    """math.py:32-49.  i k x a; scalar in 2D."""
    if a.shape[0] == 2:
        return 1j * (k[0] * a[1] - k[1] * a[0])
    return np.array([
        1j * (k[1] * a[2] - k[2] * a[1]),
        1j * (k[2] * a[0] - k[0] * a[2]),
        1j * (k[0] * a[1] - k[1] * a[0]),
    ])


def apply_linear_operator(L, u):
    """math.py:53-68: L broadcast over the leading component axis."""
    return L[None] * u if u.ndim == L.ndim + 1 else L * u


def leray_projection(u, k, k_over_k2):
    """math.py:72-100.  Returns (projected u, p_hat); the input is not touched."""
    p = np.sum(k_over_k2 * u, axis=0)
    return u - k * p[None], p


def symmetrize(w):
    """space_solver.py:15-42: W_sym[i,j] = (W[i,j] + W[j,i]) sqrt(1/2)."""
    d = int(round(np.sqrt(w.shape[0])))
    f = np.sqrt(0.5)
    out = np.empty_like(w)
    for i in range(d):
        for j in range(d):
            out[i * d + j] = (w[i * d + j] + w[j * d + i]) * f
    return out


def noise_divergence(w_sym, k):
    """space_solver.py:45-64: n_i = i sum_j k_j W_sym[i d + j]."""
    d = k.shape[0]
    return np.array([1j * sum(k[j] * w_sym[i * d + j] for j in range(d)) for i in range(d)])


# ---------------- noise stream ----------------

def noise_field_shape(grid, noise_type):
    """Shape of w1/w2 (space_solver.py:169-172): T space (d^2) or V space (d)."""
    nc = grid.dim**2 if noise_type == "thermal" else grid.dim
    return (nc,) + grid.shape_fourier


def draw_raw_noise(rs, shape):
    """random_fields draw order (space_solver.py:234-235), before scaling:
    w1.real, w1.imag, w2.real, w2.imag, each a C-order randn(*shape)."""
    w1 = rs.randn(*shape) + 1j * rs.randn(*shape)
    w2 = rs.randn(*shape) + 1j * rs.randn(*shape)
    return w1, w2


# ---------------------------------------------------------------- device RNG
# The on-device noise stream has no counterpart in the reference (which draws
# with np.random, restated by NoiseStream below).  It is Philox4x32-10
# (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
# SC'11) and the Box-Muller map; restated here so the device draws can be
# checked value by value.  Pinned by the published known-answer vectors.

_M32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Philox4x32 with 10 rounds.  ctr: four arrays of uint32 values, key: two
    uint32 values.  Returns the four output words as uint64 arrays."""
    x0, x1, x2, x3 = (np.asarray(v, dtype=np.uint64) & _M32 for v in ctr)
    k0, k1 = np.uint64(key[0]) & _M32, np.uint64(key[1]) & _M32
    m0, m1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    w0, w1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
    for _ in range(10):
        p0 = m0 * x0
        p1 = m1 * x2
        x0, x1, x2, x3 = (p1 >> np.uint64(32)) ^ x1 ^ k0, p1 & _M32, (p0 >> np.uint64(32)) ^ x3 ^ k1, p0 & _M32
        k0 = (k0 + w0) & _M32
        k1 = (k1 + w1) & _M32
    return x0, x1, x2, x3


def _box_muller(a, b):
    u1 = (a.astype(np.float64) + 0.5) * 2.0**-32
    u2 = b.astype(np.float64) * 2.0**-32
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)


def device_noise_draws(seeds, steps, nk, nc):
    """Raw draws (B, 2, nc, nk) of one step of the on-device stream: member b,
    entry c, mode idx use counter (idx lo, idx hi ^ c << 24, step lo, step
    hi) and key (seed lo, seed hi); w1 = Box-Muller of words 0-1, w2 of 2-3."""
    idx = np.arange(nk, dtype=np.uint64)
    out = np.zeros((len(seeds), 2, nc, nk), dtype=complex)
    for b, (seed, step) in enumerate(zip(seeds, steps)):
        seed, step = int(seed), int(step)
        for c in range(nc):
            ctr = (idx & _M32, (idx >> np.uint64(32)) ^ np.uint64(c << 24),
                   np.full(nk, step & 0xFFFFFFFF, dtype=np.uint64), np.full(nk, (step >> 32) & 0xFFFFFFFF, dtype=np.uint64))
            r = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))
            g0, g1 = _box_muller(r[0], r[1])
            g2, g3 = _box_muller(r[2], r[3])
            out[b, 0, c] = g0 + 1j * g1
            out[b, 1, c] = g2 + 1j * g3
    return out


def device_single_draws(seed, step, nk, ne):
    """Normals (ne, nk) of one step of the on-device single-draw stream (a step
    whose draws feed one stage): entry e of step s is word pair s % 2 (words
    0-1 at even s, 2-3 at odd s, through Box-Muller) of the call at counter
    (idx lo, idx hi ^ e << 24 ^ 2^31, (s // 2) lo, (s // 2) hi), key (seed lo,
    seed hi): one call serves two consecutive steps."""
    idx = np.arange(nk, dtype=np.uint64)
    seed, step = int(seed), int(step)
    n = step >> 1
    out = np.zeros((ne, nk), dtype=complex)
    for e in range(ne):
        ctr = (idx & _M32, (idx >> np.uint64(32)) ^ np.uint64((e << 24) ^ 0x80000000),
               np.full(nk, n & 0xFFFFFFFF, dtype=np.uint64), np.full(nk, (n >> 32) & 0xFFFFFFFF, dtype=np.uint64))
        r = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))
        a, b = (r[0], r[1]) if step % 2 == 0 else (r[2], r[3])
        g0, g1 = _box_muller(a, b)
        out[e] = g0 + 1j * g1
    return out


class NoiseStream:
    """Reproducible replacement for the reference's global MT19937 stream.

    ``RandomState(seed)`` draws the same numbers as ``np.random.seed(seed)``
    followed by ``np.random.randn``; simulation.py:100 reseeds from an
    unseeded draw, which the build replaces by an explicit seed.
    """

    def __init__(self, seed, shape):
        self.rs = np.random.RandomState(seed)
        self.shape = shape

    def next(self):
        return draw_raw_noise(self.rs, self.shape)


# ---------------- integrators, src/periodicflow/time_integrator.py ----------------

def stage_update(name, stage, unew, u0, rhs_nl, rhs_lin, forcing, noise, dt, L=None, inv=None):
    """Return u_hat after one stage (the ``stepping`` of each integrator).

    explicit_pc :51-56, implicit_pc :111-119, rk2 :178-180, rk3 :236-241,
    rk4 :292-293.  ``L`` and ``inv`` are the linear operator and
    1/(1 - dt/2 L) (:28-33); implicit_pc overwrites rhs_lin with L u0 at its
    last stage (:116-117).
    """
    if name == "explicit_pc":
        a = (0.5, 1.0)[stage]
        return u0 + a * dt * (rhs_nl + rhs_lin + forcing) + np.sqrt(0.5 * dt) * noise
    if name == "implicit_pc":
        a, b = ((0.5, 0.0), (1.0, 0.5))[stage]
        if stage == 1:
            rhs_lin = apply_linear_operator(L, u0)
        return inv * (u0 + a * dt * (rhs_nl + forcing) + b * dt * rhs_lin + np.sqrt(0.5 * dt) * noise)
    if name == "rk2":
        a, b = ((0.0, 1.0), (0.5, 0.5))[stage]
        return a * u0 + b * (unew + dt * (rhs_nl + rhs_lin + forcing))
    if name == "rk3":
        a, b = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))[stage]
        return a * u0 + b * (unew + dt * (rhs_nl + rhs_lin + forcing) + np.sqrt(dt) * noise)
    if name == "rk4":
        a = (0.25, 1.0 / 3.0, 0.5, 1.0)[stage]
        return u0 + a * dt * (rhs_nl + rhs_lin + forcing)
    raise ValueError(f"Time integrator '{name}' not found.")


def amplification(name, z):
    """Linear test equation u' = (z/dt) u: one-step amplification factor."""
    if name in ("explicit_pc", "rk2"):
        return 1 + z + z * z / 2
    if name == "implicit_pc":
        return (1 + z / 2) / (1 - z / 2)
    if name == "rk3":
        return 1 + z + z**2 / 2 + z**3 / 6
    if name == "rk4":
        return 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24
    raise ValueError(name)


# ---------------- the solver ----------------

class OracleSolver:
    """CPU FP64 restatement of ``Solver`` (simulation.py:56-225) for one member.

    Keyword arguments follow simulation.py:56-79.  Additions that the
    reference lacks: ``seed`` (explicit RNG seed, see NoiseStream), a
    ``noise`` argument to ``solve`` for injected streams, and constant-power
    band forcing (the reference's add_forcing, space_solver.py:267-271, is a
    stub).
    """

    def __init__(self, N, dt, end_time=None, is_nonlinear=True, time_integrator="rk2",
                 viscosity=1e-3, noise_type="thermal", noise_mag=0.0, domain=2 * np.pi,
                 dealias="3/2", mask_nyquist=False, seed=0, forcing=None,
                 forcing_power=0.1, forcing_band=(1.0, 2.5), **_ignored):
        if noise_type not in ("thermal", "correlated"):
            raise ValueError(f"Invalid noise type '{noise_type}'.")
        if time_integrator not in INTEGRATORS:
            raise ValueError(f"Time integrator '{time_integrator}' not found.")
        self.g = g = Grid(N, domain, dealias, mask_nyquist)
        self.dim = g.dim
        self.dt = float(dt)
        self.end_time = end_time
        self.is_nonlinear = is_nonlinear
        self.name = time_integrator
        self.nu = viscosity
        self.noise_type = noise_type
        self.noise_mag = float(noise_mag)
        info = INTEGRATORS[time_integrator]
        self.n_stages = info["n_stages"]
        self.noise_factor = info["noise_factor"]
        # space_solver.py:143-148
        self.k0_mask = np.where(g.k2 == 0, 0, 1)
        self.noise_normal_factor = self.k0_mask / np.sqrt(np.prod(g.N)) / np.sqrt(2 - 2 / g.dim)
        self.linear_operator = -viscosity * g.k2
        self.lin_op_inv = 1 / (1 - self.dt / 2 * self.linear_operator)
        self.correlation_func = np.zeros(g.shape_fourier, dtype=complex)
        self.u_hat = np.zeros((g.dim,) + g.shape_fourier, dtype=complex)
        self.forcing_kind = forcing
        self.forcing_power = forcing_power
        self.forcing_band = forcing_band
        self.seed = seed
        self.noise_stream = NoiseStream(seed, noise_field_shape(g, noise_type))
        self.t = 0.0
        self.step = 0
        self.rhs_linear = np.zeros_like(self.u_hat)

    # simulation.py:161-180 and space_solver.py:180-213
    def initialize(self, u0, space="physical", project=True, mask_zero_mode=True):
        g = self.g
        if space.casefold() == "physical":
            if tuple(u0.shape) != (g.dim,) + g.shape_physical:
                raise ValueError("Invalid shape for the input physical velocity.")
            self.u_hat = g.forward(np.asarray(u0, dtype=float))
        elif space.casefold() == "fourier":
            if tuple(u0.shape) != (g.dim,) + g.shape_fourier:
                raise ValueError("Invalid shape for the input Fourier velocity.")
            self.u_hat = np.array(u0, dtype=complex)
            if mask_zero_mode:
                self.u_hat = self.u_hat * self.k0_mask
        else:
            raise ValueError("Invalid space type. Options are 'physical' or 'fourier'.")
        if project:
            self.u_hat, _ = leray_projection(self.u_hat, g.k, g.k_over_k2)

    def set_linear_operator(self, L):
        """simulation.py:182-184 and time_integrator.py:31-33."""
        self.linear_operator = np.array(L, dtype=float) * np.ones(self.g.shape_fourier)
        self.lin_op_inv = 1 / (1 - self.dt / 2 * self.linear_operator)

    def set_correlation_function(self, C):
        """simulation.py:186-187."""
        self.correlation_func = np.array(C, dtype=complex) * np.ones(self.g.shape_fourier)

    @property
    def u(self):
        return self.g.backward(self.u_hat)

    # space_solver.py:286-300
    def rhs_nonlinear(self, uh):
        g = self.g
        vort_hat = curl(g.k, uh)
        vort = g.backward_padded(vort_hat)
        u = g.backward_padded(uh)
        uw = cross(u, vort)
        nl = g.forward_padded(uw)
        nl, _ = leray_projection(nl, g.k, g.k_over_k2)
        if g.nyquist_mask is not None:
            nl = nl * g.nyquist_mask
        return nl

    # space_solver.py:240-265
    def noise_term(self, w1, w2, a):
        g = self.g
        # random_fields scales each draw (space_solver.py:237-238), then
        # add_noise forms w1 + a w2 (:244)
        w = w1 * self.noise_normal_factor + a * (w2 * self.noise_normal_factor)
        if self.noise_type == "thermal":
            n = noise_divergence(symmetrize(w), g.k)
        else:
            n = apply_linear_operator(self.correlation_func, w)
        n, _ = leray_projection(n, g.k, g.k_over_k2)
        return n * self.noise_mag

    def band_energy(self, uh):
        g = self.g
        kk = np.sqrt(g.k2)
        lo, hi = self.forcing_band
        band = (kk >= lo) & (kk <= hi)
        return 0.5 * np.sum(g.r2c_weight * band * np.sum(np.abs(uh) ** 2, axis=0)), band

    def forcing_term(self, uh):
        """Constant-power band forcing f = P / (2 E_band) u_hat on lo <= |k| <= hi."""
        if self.forcing_kind is None:
            return np.zeros_like(uh)
        e, band = self.band_energy(uh)
        if e <= 0:
            return np.zeros_like(uh)
        return (self.forcing_power / (2 * e)) * band * uh

    def compute_rhs(self, uh, w1, w2, a):
        """space_solver.py:302-310: linear, [nonlinear], noise, forcing."""
        rhs_lin = apply_linear_operator(self.linear_operator, uh)
        rhs_nl = self.rhs_nonlinear(uh) if self.is_nonlinear else np.zeros_like(uh)
        noise = self.noise_term(w1, w2, a) if self.noise_mag != 0 else np.zeros_like(uh)
        forcing = self.forcing_term(uh)
        return rhs_nl, rhs_lin, forcing, noise

    def advance(self, w1=None, w2=None):
        """One time step (simulation.py:206-213)."""
        self.t += self.dt
        self.step += 1
        u0 = self.u_hat.copy()
        if w1 is None:
            w1, w2 = self.noise_stream.next()
        uh = self.u_hat
        for s in range(self.n_stages):
            rhs_nl, rhs_lin, f, n = self.compute_rhs(uh, w1, w2, self.noise_factor[s])
            uh = stage_update(self.name, s, uh, u0, rhs_nl, rhs_lin, f, n, self.dt,
                              self.linear_operator, self.lin_op_inv)
        self.u_hat = uh

    def solve(self, n_steps=None, noise=None, callback=None):
        """Step loop.  ``noise`` optional list of (w1, w2) raw draws per step."""
        if n_steps is None:
            n_steps = steps_from_end_time(self.end_time, self.dt)
        for i in range(n_steps):
            if noise is not None:
                self.advance(*noise[i])
            else:
                self.advance()
            if callback is not None:
                callback(self)


def steps_from_end_time(end_time, dt):
    """simulation.py:205 loops until isclose(t, end_time): steps = end_time/dt."""
    n = int(round(end_time / dt))
    if not np.isclose(n * dt, end_time, atol=1e-10):
        raise ValueError("end_time must be a whole number of steps")
    return n


# ---------------- FP64 diagnostics (new; the reference has none) ----------------

def energy(g, uh):
    """E = 1/2 <|u|^2> by Parseval over the half spectrum with r2c weights."""
    return 0.5 * float(np.sum(g.r2c_weight * np.sum(np.abs(uh) ** 2, axis=0)))


def enstrophy(g, uh):
    """1/2 <|omega|^2> with omega_hat = i k x u_hat."""
    w = curl(g.k, uh)
    if g.dim == 2:
        return 0.5 * float(np.sum(g.r2c_weight * np.abs(w) ** 2))
    return 0.5 * float(np.sum(g.r2c_weight * np.sum(np.abs(w) ** 2, axis=0)))


def dissipation(g, uh, nu):
    return 2.0 * nu * enstrophy(g, uh)


def shell_spectrum(g, uh, nshells=None):
    """E(k) = 1/2 sum over shells round(|k|) of weighted |u_hat|^2."""
    kk = np.sqrt(g.k2)
    sh = np.rint(kk).astype(np.int64)
    if nshells is None:
        nshells = int(sh.max()) + 1
    e = 0.5 * g.r2c_weight * np.sum(np.abs(uh) ** 2, axis=0)
    out = np.zeros(nshells)
    np.add.at(out, np.minimum(sh, nshells - 1).ravel(), e.ravel())
    return out


def taylor_green_3d(g):
    """u = sin x cos y cos z, v = -cos x sin y cos z, w = 0."""
    x, y, z = g.x
    u = np.zeros((3,) + g.shape_physical)
    u[0] = np.sin(x) * np.cos(y) * np.cos(z)
    u[1] = -np.cos(x) * np.sin(y) * np.cos(z)
    return u


def taylor_green_2d(g):
    """u = sin x cos y, v = -cos x sin y (exact NS solution decaying as e^{-2 nu t})."""
    x, y = g.x
    u = np.zeros((2,) + g.shape_physical)
    u[0] = np.sin(x) * np.cos(y)
    u[1] = -np.cos(x) * np.sin(y)
    return u


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
