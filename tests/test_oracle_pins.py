"""More pins of the FP64 oracle: the sign of the nonlinear term, the noise
coefficients of every stochastic integrator, the thermal-noise normaliser in
2D and 3D, and the shell spectrum.

Each check compares the oracle with mathematics computed here without its
code (closed forms, spectral derivatives on a finer grid, the continuous
Ornstein-Uhlenbeck variance) and is chosen so that a plausible slip fails
it.  The ``*_detects_*`` tests plant such slips in the oracle (on the CPU,
through monkeypatch) and assert that the check catches them.
"""
import numpy as np
import pytest

from oracle import pf_oracle as po


# ----------------------------------------------------------- nonlinear term
def _tgv_nonlinear_closed_form(g, kappa):
    """N(u) = P[u x omega] = -P[(u . grad) u] of u = (sin kx cos ky cos kz,
    -cos kx sin ky cos kz, 0), k = kappa.  By hand: (u . grad) u =
    kappa/4 (sin 2kx (1 + cos 2kz), sin 2ky (1 + cos 2kz), 0); its part
    kappa/4 (sin 2kx, sin 2ky, 0) is a gradient, and the rest projects to
    kappa/8 (sin 2kx cos 2kz, sin 2ky cos 2kz, -(cos 2kx + cos 2ky) sin 2kz)."""
    x, y, z = (kappa * c for c in g.x)
    c = kappa / 8
    out = np.zeros((3,) + g.shape_physical)
    out[0] = -c * np.sin(2 * x) * np.cos(2 * z)
    out[1] = -c * np.sin(2 * y) * np.cos(2 * z)
    out[2] = c * (np.cos(2 * x) + np.cos(2 * y)) * np.sin(2 * z)
    return out


@pytest.mark.parametrize("dealias", ["3/2", "2/3"])
@pytest.mark.parametrize("box", [2 * np.pi, 4 * np.pi])
def test_tgv3d_nonlinear_term_closed_form(dealias, box):
    s = po.OracleSolver((16, 16, 16), 0.01, dealias=dealias, domain=box)
    kappa = 2 * np.pi / box
    x, y, z = (kappa * c for c in s.g.x)
    u = np.zeros((3,) + s.g.shape_physical)
    u[0] = np.sin(x) * np.cos(y) * np.cos(z)
    u[1] = -np.cos(x) * np.sin(y) * np.cos(z)
    s.initialize(u)
    nl = s.g.backward(s.rhs_nonlinear(s.u_hat))
    ref = _tgv_nonlinear_closed_form(s.g, kappa)
    assert np.max(np.abs(nl - ref)) < 1e-15
    assert np.max(np.abs(ref)) > 0.1 * kappa


def test_tgv3d_closed_form_detects_flipped_cross(monkeypatch):
    """A cross product of the wrong sign (omega x u) fails the closed form."""
    s = po.OracleSolver((16, 16, 16), 0.01)
    s.initialize(po.taylor_green_3d(s.g))
    orig = po.cross
    monkeypatch.setattr(po, "cross", lambda a, b: orig(b, a))
    nl = s.g.backward(s.rhs_nonlinear(s.u_hat))
    assert np.max(np.abs(nl - _tgv_nonlinear_closed_form(s.g, 1.0))) > 0.1


def _band_limited_full(N, rs):
    """Full complex spectrum (integer modes, fftn order) of a random real
    field with every |k_i| < N_i / 2 (no Nyquist content)."""
    u = rs.standard_normal(tuple(N))
    c = np.fft.fftn(u) / np.prod(N)
    for ax, n in enumerate(N):
        k = np.fft.fftfreq(n, 1.0 / n)
        shp = [1] * len(N)
        shp[ax] = n
        c = c * (np.abs(k) < n // 2).reshape(shp)
    return c


def _advection_on_fine_grid(cs, N):
    """(u . grad) u of the fields with full spectra cs[i], evaluated on a grid
    of 2 N_i points per axis (the product of two fields of |k_i| < N_i/2 is
    then exact) and returned as full spectra on the N grid."""
    d = len(N)
    F = [2 * n for n in N]
    kf = np.meshgrid(*[np.fft.fftfreq(m, 1.0 / m) for m in F], indexing="ij")

    def to_fine(c):
        out = np.zeros(F, dtype=complex)
        src = np.meshgrid(*[np.fft.fftfreq(n, 1.0 / n).astype(int) for n in N], indexing="ij")
        dst = tuple(src[a] % F[a] for a in range(d))
        out[dst] = c
        return out

    fine = [to_fine(c) for c in cs]
    u = [np.fft.ifftn(f).real * np.prod(F) for f in fine]
    adv = []
    for i in range(d):
        acc = np.zeros(F)
        for j in range(d):
            du = np.fft.ifftn(1j * kf[j] * fine[i]).real * np.prod(F)
            acc += u[j] * du
        A = np.fft.fftn(acc) / np.prod(F)
        src = np.meshgrid(*[np.fft.fftfreq(n, 1.0 / n).astype(int) for n in N], indexing="ij")
        adv.append(A[tuple(src[a] % F[a] for a in range(d))])
    return np.array(adv)


@pytest.mark.parametrize("N", [(8, 12), (8, 8, 12), (6, 8, 8)])
def test_nonlinear_term_is_minus_projected_advection(N):
    """u x omega = grad(|u|^2 / 2) - (u . grad) u, so P[u x omega] =
    -P[(u . grad) u]; the right side is built here from spectral derivatives
    on a finer grid, without the oracle's curl, cross or padded transforms."""
    d = len(N)
    rs = np.random.RandomState(21)
    cs = [_band_limited_full(N, rs) for _ in range(d)]
    s = po.OracleSolver(N, 0.01, dealias="3/2")
    nh = N[-1] // 2 + 1
    uh = np.array([c[..., :nh] for c in cs])
    got = s.rhs_nonlinear(uh)
    adv = _advection_on_fine_grid(cs, N)
    k = np.array(np.meshgrid(*[np.fft.fftfreq(n, 1.0 / n) for n in N], indexing="ij"))
    k2 = np.sum(k**2, axis=0)
    proj = adv - k * (np.sum(k * adv, axis=0) / np.where(k2 == 0, 1, k2))[None]
    want = -proj[..., :nh]
    keep = np.ones(s.g.shape_fourier, dtype=bool)
    for i, n in enumerate(N):
        keep &= np.abs(s.g.kint[i]) != n // 2
    err = np.max(np.abs((got - want) * keep))
    assert err < 1e-14 * max(1.0, np.max(np.abs(want)))
    assert np.max(np.abs(want * keep)) > 1e-2


# ------------------------------------------------- stochastic coefficients
def _continuous_level(noise_type, nm, nu, d, Nprod, k2):
    """Stationary sum_a E|u_a(k)|^2 of du = -nu k^2 u dt + n dW, where n is
    the projected, normalised noise: thermal E|n|^2 = nm^2 k^2 d / prod N
    (from E|W_sym k|^2 of the symmetrised Gaussian tensor after projection),
    correlated with C = 1: nm^2 d / prod N; divided by 2 nu k^2."""
    if noise_type == "thermal":
        return nm**2 * d / (2 * nu * Nprod) * np.ones_like(k2)
    return nm**2 * d / (2 * nu * Nprod * np.where(k2 == 0, 1, k2))


def _stationary_level(s):
    """Exact stationary sum_a E|u_a(k)|^2 of the oracle's linear stochastic
    step: one step is u -> R(k) u + B1 w1 + B2 w2 mode by mode, so the
    variance is E|B1 w1 + B2 w2|^2 / (1 - R^2) with E|w|^2 = 2 per complex
    entry.  R and the gains B are read off single oracle steps."""
    g = s.g
    d = g.dim
    shape = s.noise_stream.shape
    zero = np.zeros(shape, dtype=complex)
    # amplification: one unit component, no noise (Stokes: no projection of u)
    s.u_hat = np.zeros((d,) + g.shape_fourier, dtype=complex)
    s.u_hat[0] = 1.0
    s.advance(zero, zero)
    R = s.u_hat[0].real
    trq = np.zeros(g.shape_fourier)
    for which in (0, 1):
        for c in range(shape[0]):
            w = np.zeros(shape, dtype=complex)
            w[c] = 1.0
            s.u_hat = np.zeros((d,) + g.shape_fourier, dtype=complex)
            s.advance(w, zero) if which == 0 else s.advance(zero, w)
            trq += 2 * np.sum(np.abs(s.u_hat) ** 2, axis=0)
    # modes the explicit schemes leave unstable or neutral (|R| >= 1 at
    # large nu k^2 dt) have no stationary level: inf / nan there
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(g.k2 > 0, trq, 0.0) / np.where(g.k2 > 0, 1 - R**2, 1.0)


def _ratio(name, N, nu, dt, noise_type="thermal", nm=0.7):
    s = po.OracleSolver(N, dt, is_nonlinear=False, time_integrator=name, viscosity=nu,
                        noise_type=noise_type, noise_mag=nm)
    if noise_type == "correlated":
        s.set_correlation_function(np.where(s.g.k2 == 0, 0.0, 1.0))
    lev = _stationary_level(s)
    ref = _continuous_level(noise_type, nm, nu, s.g.dim, np.prod(N), s.g.k2)
    sel = s.g.k2 > 0
    return lev[sel] / ref[sel], s.g.k2[sel]


@pytest.mark.parametrize("N", [(8, 8), (6, 8, 8)])
@pytest.mark.parametrize("noise_type", ["thermal", "correlated"])
def test_implicit_pc_stationary_variance_exact(N, noise_type):
    """Crank-Nicolson keeps the continuous stationary variance exactly, for
    every dt: pins sqrt(dt/2), the noise factors (0, 1) and the normaliser
    1/sqrt(prod N)/sqrt(2 - 2/d) in 2D and 3D."""
    for dt in (0.01, 0.1, 0.7):
        r, _ = _ratio("implicit_pc", N, 1.0, dt, noise_type)
        assert np.max(np.abs(r - 1)) < 1e-12, dt


def _errors(name, betas, N=(8, 8)):
    """|level / continuous level - 1| at the mode k = (1, 0) for nu k^2 dt = beta."""
    out = []
    for b in betas:
        r, k2 = _ratio(name, N, 1.0, b)
        out.append(float(np.abs(r[np.argmin(np.abs(k2 - 1))] - 1)))
    return np.array(out)


BETAS = (0.08, 0.04, 0.02, 0.01)


def _weak_order(name, N=(8, 8)):
    e = _errors(name, BETAS, N)
    return e, np.log2(e[:-1] / e[1:])


def test_explicit_pc_stationary_variance_converges():
    """Explicit PC: level R-independent to third order in nu k^2 dt
    (ratio (1 + z + z^2/2) / (1 + z + z^2/2 + z^3/8) at z = -nu k^2 dt)."""
    e, order = _weak_order("explicit_pc")
    assert np.all(np.abs(order - 3) < 0.25), order
    assert e[-1] < 2e-7


def test_rk3_stationary_variance_converges():
    """Stochastic RK3 (noise factors (2 sqrt2 + sqrt3)/5, (-4 sqrt2 + 3 sqrt3)/5,
    (sqrt2 - 2 sqrt3)/10; dt_sqrt = sqrt dt): the level tends to the
    continuous one at third order in nu k^2 dt; a perturbed factor leaves a
    first-order error."""
    e, order = _weak_order("rk3")
    assert np.all(np.abs(order - 3) < 0.25), order
    assert e[-1] < 1e-7


@pytest.mark.parametrize("N", [(8, 8), (6, 8, 8)])
@pytest.mark.parametrize("name", ["explicit_pc", "rk3"])
def test_explicit_schemes_level_every_mode(N, name):
    """At small dt every mode of both noise types sits at the continuous level."""
    for nt in ("thermal", "correlated"):
        r, k2 = _ratio(name, N, 1.0, 1e-3, nt)
        assert np.max(np.abs(r - 1)) < 5e-5, (nt, np.max(np.abs(r - 1)))


def test_rk3_check_detects_perturbed_noise_factor(monkeypatch):
    nf = po.INTEGRATORS["rk3"]["noise_factor"]
    bad = dict(po.INTEGRATORS["rk3"], noise_factor=(nf[0], nf[1] * 1.05, nf[2]))
    monkeypatch.setitem(po.INTEGRATORS, "rk3", bad)
    e, order = _weak_order("rk3")
    assert not (np.all(np.abs(order - 3) < 0.25) and e[-1] < 1e-7)


def test_explicit_pc_check_detects_swapped_noise_factor(monkeypatch):
    monkeypatch.setitem(po.INTEGRATORS, "explicit_pc", dict(po.INTEGRATORS["explicit_pc"], noise_factor=(1.0, 0.0)))
    e, order = _weak_order("explicit_pc")
    assert not (np.all(np.abs(order - 3) < 0.25) and e[-1] < 2e-7)


def test_explicit_pc_check_detects_sqrt_dt(monkeypatch):
    """The predictor-corrector's noise is scaled by sqrt(dt/2), not sqrt(dt)."""
    orig = po.stage_update

    def planted(name, stage, unew, u0, rhs_nl, rhs_lin, forcing, noise, dt, L=None, inv=None):
        return orig(name, stage, unew, u0, rhs_nl, rhs_lin, forcing, np.sqrt(2.0) * noise, dt, L, inv)

    monkeypatch.setattr(po, "stage_update", planted)
    e, _ = _weak_order("explicit_pc")
    assert e[-1] > 0.5


def test_thermal_check_detects_2d_normaliser_in_3d():
    """sqrt(2 - 2/d) with d = 3 is what the 3D level needs: the 2D constant
    sqrt(2 - 2/2) = 1 in its place moves the level by 4/3."""
    s = po.OracleSolver((6, 8, 8), 0.1, is_nonlinear=False, time_integrator="implicit_pc", viscosity=1.0,
                        noise_mag=0.7)
    s.noise_normal_factor = s.noise_normal_factor * np.sqrt(2 - 2 / 3)
    lev = _stationary_level(s)
    ref = _continuous_level("thermal", 0.7, 1.0, 3, np.prod(s.g.N), s.g.k2)
    sel = s.g.k2 > 0
    assert np.allclose(lev[sel] / ref[sel], 4 / 3, rtol=1e-12)


# ------------------------------------------------------------ shell spectrum
def test_shell_spectrum_single_mode():
    g = po.Grid((16, 16, 16))
    x, y, z = g.x
    a = 0.3
    u = np.zeros((3,) + g.shape_physical)
    u[2] = a * np.cos(3 * x + 4 * y)          # |k| = 5, on the kz = 0 plane
    u[0] = 0.5 * a * np.sin(2 * z + 1.0)      # |k| = 2, off that plane
    uh = g.forward(u)
    E = po.shell_spectrum(g, uh, nshells=12)
    want = np.zeros(12)
    want[5] = a**2 / 4
    want[2] = (0.5 * a) ** 2 / 4
    assert np.allclose(E, want, atol=1e-15)


@pytest.mark.parametrize("N", [(12, 16), (8, 12, 10)])
def test_shell_spectrum_sums_to_physical_energy(N):
    g = po.Grid(N)
    rs = np.random.RandomState(22)
    u = rs.standard_normal((len(N),) + g.shape_physical)
    uh = g.forward(u)
    E = po.shell_spectrum(g, uh)
    assert np.isclose(E.sum(), 0.5 * np.mean(np.sum(u**2, axis=0)), rtol=1e-13)
    assert np.isclose(po.energy(g, uh), 0.5 * np.mean(np.sum(u**2, axis=0)), rtol=1e-13)


# ------------------------------------- the device's thermal-noise sampling
@pytest.mark.parametrize("N", [(8, 6), (6, 4, 8)])
def test_curl_form_thermal_noise_has_the_reference_law(N):
    """The device samples thermal noise as i k x z (z: 3 independent complex
    normals per mode in 3D, one along the third axis in 2D) instead of the
    reference's P (i symmetrize(W) k) of D^2 of them.  Both are linear maps of
    i.i.d. normals, so each is Gaussian with covariance A A^H for its map A.
    Here A is built column by column from the oracle's own symmetrize,
    noise_divergence and leray_projection applied to unit draws, and the two
    covariances are compared exactly at every mode (equal to k^2 P)."""
    g = po.Grid(N)
    d = len(N)
    k = g.k
    shape = g.shape_fourier
    cols = []
    for c in range(d * d):
        w = np.zeros((d * d,) + shape, dtype=complex)
        w[c] = 1.0
        n, _ = po.leray_projection(po.noise_divergence(po.symmetrize(w), k), k, g.k_over_k2)
        cols.append(n)
    A = np.stack(cols, axis=1)  # (d, d^2, *shape)
    ref = np.einsum("ic...,jc...->ij...", A, A.conj())
    if d == 3:
        B = np.zeros((3, 3) + shape, dtype=complex)  # i k x e_c
        for c in range(3):
            e = np.zeros((3,) + shape)
            e[c] = 1.0
            B[:, c] = 1j * np.cross(k, e, axis=0)
    else:
        B = np.stack([1j * k[1], -1j * k[0]])[:, None]
    cur = np.einsum("ic...,jc...->ij...", B, B.conj())
    assert np.max(np.abs(cur - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    # and both equal k^2 P
    k2 = g.k2
    P = np.eye(d).reshape((d, d) + (1,) * d) * k2 - np.einsum("i...,j...->ij...", k, k)
    assert np.max(np.abs(ref - P)) < 1e-12 * np.max(k2)
