"""Pins the FP64 oracle against mathematics that does not depend on its code.

Every check here would fail for a plausible slip in the restatement (a wrong
sign, a dropped factor, a transposed index, a mis-split Nyquist mode):
brute-force DFTs, exact evaluation of trigonometric polynomials on the
padded grid, exact truncated convolution, Taylor-Green closed forms, the
exact 2D Taylor-Green decay, amplification polynomials, projection
invariants, energy conservation of the dealiased nonlinear term and the
closed-form stationary variance of fluctuating Stokes flow.
"""
import itertools

import numpy as np
import pytest

from oracle import pf_oracle as po


def _full_int_modes(g):
    """Integer wavenumber vectors of the full (complex) spectrum of the N grid."""
    axes = [np.fft.fftfreq(n, 1.0 / n).astype(int) for n in g.N]
    return axes


def _random_real_field(g, rs, zero_nyquist=True, ncomp=None):
    shp = g.shape_physical if ncomp is None else (ncomp,) + g.shape_physical
    u = rs.standard_normal(shp)
    uh = g.forward(u)
    if zero_nyquist:
        mask = np.ones(g.shape_fourier, dtype=bool)
        for i, n in enumerate(g.N):
            mask &= np.abs(g.kint[i]) != n // 2
        uh = uh * mask
    return uh


def _eval_trig(coeffs_full, modes, pts):
    """Directly evaluate sum_k c_k exp(i k.x) at the points ``pts`` (list of axes)."""
    d = len(modes)
    grids = np.meshgrid(*pts, indexing="ij")
    out = np.zeros(grids[0].shape, dtype=complex)
    for idx in itertools.product(*[range(len(m)) for m in modes]):
        c = coeffs_full[idx]
        if c == 0:
            continue
        ph = sum(modes[a][idx[a]] * grids[a] for a in range(d))
        out += c * np.exp(1j * ph)
    return out


@pytest.mark.parametrize("N", [(6, 8), (4, 6, 8), (8, 8)])
def test_forward_backward_vs_bruteforce_dft(N):
    g = po.Grid(N)
    rs = np.random.RandomState(1)
    u = rs.standard_normal(g.shape_physical)
    uh = g.forward(u)
    modes = _full_int_modes(g)
    pts = [np.arange(n) * 2 * np.pi / n for n in g.N]
    grids = np.meshgrid(*pts, indexing="ij")
    # brute-force forward: (1/prod N) sum_x u(x) exp(-i k.x), half spectrum
    for idx in itertools.product(*[range(s) for s in g.shape_fourier]):
        kv = [modes[a][idx[a]] if a < g.dim - 1 else idx[a] for a in range(g.dim)]
        ph = sum(kv[a] * grids[a] for a in range(g.dim))
        ref = np.sum(u * np.exp(-1j * ph)) / np.prod(g.N)
        assert abs(uh[idx] - ref) < 1e-13
    assert np.allclose(g.backward(uh), u, atol=1e-13)


def test_wavenumber_conventions():
    g = po.Grid((8, 6, 8), domain=2 * np.pi)
    # complex axes: Nyquist stored as -N/2; real axis: +N/2 (basis.py:65)
    assert g.k[0][4, 0, 0] == -4
    assert g.k[1][0, 3, 0] == -3
    assert g.k[2][0, 0, 4] == 4
    assert g.k_over_k2[:, 0, 0, 0].tolist() == [0, 0, 0]
    g2 = po.Grid((8, 8), domain=4 * np.pi)
    assert np.isclose(g2.k[0][1, 0], 0.5)
    assert np.isclose(g2.x[0][1, 0], 4 * np.pi / 8)
    with pytest.raises(ValueError):
        po.Grid((8, 8), dealias="1/2")
    with pytest.raises(ValueError):
        po.Grid((8,))


@pytest.mark.parametrize("N", [(8, 8), (8, 12), (4, 8, 8), (8, 4, 12)])
def test_padded_backward_is_exact_polynomial_evaluation(N):
    """3/2 rule backward = the trigonometric polynomial sampled on the M grid."""
    g = po.Grid(N, dealias="3/2")
    rs = np.random.RandomState(2)
    uh = _random_real_field(g, rs)
    u_full = g.backward(uh)
    c_full = np.fft.fftn(u_full) / np.prod(g.N)
    modes = _full_int_modes(g)
    pts = [np.arange(m) * 2 * np.pi / m for m in g.M]
    ref = _eval_trig(c_full, modes, pts)
    assert np.max(np.abs(ref.imag)) < 1e-12
    got = g.backward_padded(uh)
    assert got.shape == tuple(g.M)
    assert np.allclose(got, ref.real, atol=1e-12)


@pytest.mark.parametrize("N", [(9, 12), (12, 6, 9)])
def test_23_backward_masks_upper_third(N):
    g = po.Grid([n if n % 2 == 0 else n + 1 for n in N], dealias="2/3")
    rs = np.random.RandomState(3)
    uh = g.forward(rs.standard_normal(g.shape_physical))
    got = g.backward_padded(uh)
    # equivalent: zero every mode with |k_i| > floor(N_i/3) then plain backward
    keep = np.ones(g.shape_fourier, dtype=bool)
    for i, n in enumerate(g.N):
        ki = np.abs(g.kint[i])
        keep &= ki <= n // 3
    assert np.allclose(got, g.backward(uh * keep), atol=1e-12)
    # brute-force: keep mode counts on a single axis
    n0 = g.N[0]
    kept = [k for k in np.fft.fftfreq(n0, 1.0 / n0).astype(int) if abs(k) <= n0 // 3]
    assert len(kept) == 2 * (n0 // 3) + 1


@pytest.mark.parametrize("N", [(8, 8), (4, 8, 8)])
def test_padded_forward_recovers_polynomial_coefficients(N):
    g = po.Grid(N, dealias="3/2")
    rs = np.random.RandomState(4)
    uh = _random_real_field(g, rs)
    assert np.allclose(g.forward_padded(g.backward_padded(uh)), uh, atol=1e-13)


@pytest.mark.parametrize("N", [(8, 8), (4, 4, 8)])
def test_dealiased_product_is_truncated_exact_convolution(N):
    """forward_padded(a(x) b(x)) = convolution of the coefficients, truncated."""
    g = po.Grid(N, dealias="3/2")
    rs = np.random.RandomState(5)
    ah = _random_real_field(g, rs)
    bh = _random_real_field(g, rs)
    prod_h = g.forward_padded(g.backward_padded(ah) * g.backward_padded(bh))
    ca = np.fft.fftn(g.backward(ah)) / np.prod(g.N)
    cb = np.fft.fftn(g.backward(bh)) / np.prod(g.N)
    modes = _full_int_modes(g)
    ref = np.zeros(g.shape_fourier, dtype=complex)
    idxs = list(itertools.product(*[range(n) for n in g.N]))
    for ia in idxs:
        if ca[ia] == 0:
            continue
        ka = [modes[d][ia[d]] for d in range(g.dim)]
        for ib in idxs:
            if cb[ib] == 0:
                continue
            kb = [modes[d][ib[d]] for d in range(g.dim)]
            ks = [ka[d] + kb[d] for d in range(g.dim)]
            # keep |k_i| < N_i/2 on every axis (Nyquist inputs are zero)
            if any(abs(ks[d]) >= g.N[d] // 2 for d in range(g.dim)):
                continue
            if ks[-1] < 0:
                continue
            out = tuple(ks[d] % g.N[d] for d in range(g.dim - 1)) + (ks[-1],)
            ref[out] += ca[ia] * cb[ib]
    # the Nyquist slots hold the +-N/2 sums, compare everything else
    mask = np.ones(g.shape_fourier, dtype=bool)
    for i, n in enumerate(g.N):
        mask &= np.abs(g.kint[i]) != n // 2
    assert np.allclose(prod_h * mask, ref * mask, atol=1e-13)


def test_taylor_green_closed_forms():
    g = po.Grid((16, 16, 16))
    uh = g.forward(po.taylor_green_3d(g))
    mags = np.abs(uh[0])
    nz = mags[mags > 1e-14]
    # sin x cos y cos z: 8 modes of magnitude 1/8, half spectrum keeps 4
    assert len(nz) == 4 and np.allclose(nz, 1 / 8)
    assert np.isclose(po.energy(g, uh), 0.125, atol=1e-15)
    assert np.isclose(2 * po.enstrophy(g, uh), 0.75, atol=1e-15)
    # solenoidal: k . u_hat = 0
    assert np.max(np.abs(np.sum(g.k * uh, axis=0))) < 1e-15


@pytest.mark.parametrize("name", ["explicit_pc", "implicit_pc", "rk2", "rk3", "rk4"])
@pytest.mark.parametrize("dealias", ["3/2", "2/3"])
def test_2d_taylor_green_exact_decay(name, dealias):
    """The 2D TGV nonlinear term is a gradient, so u_hat(n) = R(z)^n u_hat(0)."""
    nu, dt, n = 0.05, 0.01, 6
    s = po.OracleSolver((16, 16), dt, is_nonlinear=True, time_integrator=name,
                        viscosity=nu, dealias=dealias)
    s.initialize(po.taylor_green_2d(s.g))
    u0 = s.u_hat.copy()
    s.solve(n)
    R = po.amplification(name, -2 * nu * dt)
    assert np.allclose(s.u_hat, u0 * R**n, atol=1e-14, rtol=0)
    assert abs(R**n - np.exp(-2 * nu * dt * n)) < 1e-6


@pytest.mark.parametrize("name", ["explicit_pc", "implicit_pc", "rk2", "rk3", "rk4"])
def test_linear_amplification_per_mode(name):
    nu, dt = 0.3, 0.05
    s = po.OracleSolver((8, 8, 8), dt, is_nonlinear=False, time_integrator=name, viscosity=nu)
    rs = np.random.RandomState(6)
    uh = _random_real_field(s.g, rs, ncomp=3)
    s.initialize(uh, space="fourier", project=True)
    u0 = s.u_hat.copy()
    s.solve(3)
    R = po.amplification(name, -nu * s.g.k2 * dt)
    assert np.allclose(s.u_hat, u0 * R[None] ** 3, atol=1e-14, rtol=0)


def test_leray_invariants():
    g = po.Grid((8, 6, 8))
    rs = np.random.RandomState(7)
    u = rs.standard_normal((3,) + g.shape_fourier) + 1j * rs.standard_normal((3,) + g.shape_fourier)
    pu, p = po.leray_projection(u, g.k, g.k_over_k2)
    assert np.max(np.abs(np.sum(g.k * pu, axis=0))) < 1e-12
    ppu, _ = po.leray_projection(pu, g.k, g.k_over_k2)
    assert np.allclose(ppu, pu, atol=1e-13)
    phi = rs.standard_normal(g.shape_fourier) + 0j
    grad, _ = po.leray_projection(g.k * phi, g.k, g.k_over_k2)
    assert np.allclose(grad, 0, atol=1e-13)
    # k = 0 mode untouched
    assert np.allclose(pu[:, 0, 0, 0], u[:, 0, 0, 0])


def test_symmetrize_and_divergence():
    g = po.Grid((6, 8))
    rs = np.random.RandomState(8)
    w = rs.standard_normal((4,) + g.shape_fourier) + 0j
    ws = po.symmetrize(w)
    assert np.allclose(ws[1], ws[2]) and np.allclose(ws[0], np.sqrt(2) * w[0])
    n = po.noise_divergence(ws, g.k)
    ref0 = 1j * (g.k[0] * ws[0] + g.k[1] * ws[1])
    assert np.allclose(n[0], ref0)
    g3 = po.Grid((4, 4, 6))
    w3 = rs.standard_normal((9,) + g3.shape_fourier) + 0j
    s3 = po.symmetrize(w3).reshape((3, 3) + g3.shape_fourier)
    assert np.allclose(s3, np.swapaxes(s3, 0, 1))


def test_nonlinear_term_conserves_energy():
    """<u, N(u)> = 0 exactly for band-limited u under the 3/2 rule (nu = 0)."""
    s = po.OracleSolver((8, 8, 8), 0.01, viscosity=0.0)
    rs = np.random.RandomState(9)
    uh = _random_real_field(s.g, rs, ncomp=3)
    s.initialize(uh, space="fourier", project=True)
    nl = s.rhs_nonlinear(s.u_hat)
    dEdt = float(np.sum(s.g.r2c_weight * np.sum((np.conj(s.u_hat) * nl).real, axis=0)))
    scale = float(np.sum(s.g.r2c_weight * np.sum(np.abs(nl) * np.abs(s.u_hat), axis=0)))
    assert abs(dEdt) < 1e-13 * scale
    # and the term is not trivially zero
    assert np.max(np.abs(nl)) > 1e-3


def test_noise_draw_order():
    """w1.real, w1.imag, w2.real, w2.imag are four consecutive C-order blocks."""
    shape = (4, 3, 5)
    w1, w2 = po.draw_raw_noise(np.random.RandomState(11), shape)
    flat = np.random.RandomState(11).randn(4 * int(np.prod(shape))).reshape((4,) + shape)
    assert np.array_equal(w1.real, flat[0]) and np.array_equal(w1.imag, flat[1])
    assert np.array_equal(w2.real, flat[2]) and np.array_equal(w2.imag, flat[3])


@pytest.mark.parametrize("noise_type", ["thermal", "correlated"])
def test_stokes_stationary_variance_closed_form(noise_type):
    """Implicit PC (Crank-Nicolson) on fluctuating Stokes flow.

    Stationary E sum_a |u_a(k)|^2 = nm^2 d / (2 nu prod N) for thermal noise
    (flat in k, independent of dt), and the same divided by k^2 for
    correlated noise with C(k) = 1, from the covariance of the symmetric
    noise tensor and the CN amplification (1 + z/2)/(1 - z/2).
    """
    N = (16, 16)
    nu, dt, nm = 1.0, 0.05, 0.7
    s = po.OracleSolver(N, dt, is_nonlinear=False, time_integrator="implicit_pc",
                        viscosity=nu, noise_type=noise_type, noise_mag=nm, seed=123)
    if noise_type == "correlated":
        s.set_correlation_function(np.where(s.g.k2 == 0, 0.0, 1.0))
    s.initialize(np.zeros((2,) + s.g.shape_fourier, dtype=complex), space="fourier")
    s.solve(200)
    acc = np.zeros(s.g.shape_fourier)
    nsamp = 3000
    for _ in range(nsamp):
        s.advance()
        acc += np.sum(np.abs(s.u_hat) ** 2, axis=0)
    S = acc / nsamp
    theory = nm**2 * 2 / (2 * nu * np.prod(N))
    kk = s.g.k2
    sel = (kk > 0) & (kk <= 20)
    if noise_type == "thermal":
        ratio = S[sel] / theory
    else:
        ratio = S[sel] * kk[sel] / theory
    # each mode averages 3000 correlated samples; check the mean tightly and
    # every mode loosely
    assert abs(ratio.mean() - 1) < 0.03
    assert np.all(np.abs(ratio - 1) < 0.35)
    assert S[0, 0] == 0.0


def test_constant_power_forcing_injects_P():
    s = po.OracleSolver((8, 8, 8), 0.01, forcing="constant_power", forcing_power=0.1)
    rs = np.random.RandomState(12)
    s.initialize(_random_real_field(s.g, rs, ncomp=3), space="fourier")
    f = s.forcing_term(s.u_hat)
    power = float(np.sum(s.g.r2c_weight * np.sum((np.conj(s.u_hat) * f).real, axis=0)))
    assert np.isclose(power, 0.1, rtol=1e-12)


def test_steps_from_end_time():
    assert po.steps_from_end_time(2.0, 0.01) == 200
    assert po.steps_from_end_time(1.25, 2.5e-3) == 500
    with pytest.raises(ValueError):
        po.steps_from_end_time(1.0, 0.3)


# Philox4x32-10 known-answer vectors (Random123's kat_vectors, the generator's
# reference implementation): (counter, key) -> output
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    got = po.philox4x32_10([np.array([c], dtype=np.uint64) for c in ctr], key)
    assert tuple(int(g[0]) for g in got) == want


def test_device_draw_layout_is_unit_normal():
    """The restated device stream: unit variance per real part, independent
    members and entries, and a different step gives different draws."""
    d = po.device_noise_draws([7, 8], [0, 0], 4096, 2)
    x = np.concatenate([d.real.ravel(), d.imag.ravel()])
    assert abs(x.mean()) < 0.02 and abs(x.var() - 1) < 0.03
    assert abs(np.corrcoef(d[0, 0, 0].real, d[1, 0, 0].real)[0, 1]) < 0.05
    assert abs(np.corrcoef(d[0, 0, 0].real, d[0, 0, 1].real)[0, 1]) < 0.05
    e = po.device_noise_draws([7], [1], 4096, 1)
    assert not np.allclose(e[0, 0, 0], d[0, 0, 0])


def test_device_single_draw_stream_is_unit_normal():
    """The restated single-draw stream: unit variance per real part; the two
    steps one call serves, later steps, entries and seeds uncorrelated; and
    no draw shared with the multi-stage stream of the same seed."""
    ev = po.device_single_draws(7, 10, 8192, 3)
    od = po.device_single_draws(7, 11, 8192, 3)
    nx = po.device_single_draws(7, 12, 8192, 3)
    other = po.device_single_draws(8, 10, 8192, 3)
    x = np.concatenate([ev.real.ravel(), ev.imag.ravel(), od.real.ravel(), od.imag.ravel()])
    assert abs(x.mean()) < 0.02 and abs(x.var() - 1) < 0.03
    for a, b in ((ev[0], od[0]), (ev[0], nx[0]), (ev[0], ev[1]), (ev[1], ev[2]), (ev[0], other[0]),
                 (ev[0].real, ev[0].imag)):
        assert abs(np.corrcoef(np.real(a), np.real(b))[0, 1]) < 0.05
    # step pair 5 of the single-draw stream against step 5 of the multi-stage one
    multi = po.device_noise_draws([7], [5], 8192, 3)[0].reshape(6, 8192)
    assert not np.any(np.isclose(ev[:, None, :], multi[None, :, :]))
