"""Pins of the oracle's non-periodic boundaries and of PoissonSolverMixed. CPU only.

- The r2r transforms of PoissonSolverMixed (FFTW's R2HC/HC2R and REDFT10/REDFT01) against
  scipy's DCT-II/III and numpy's FFT.
- The mixed solve: single modes come back divided by h times the symbol sum (this fixes the
  transforms' scales and the index-to-wavenumber maps), and a random RHS on a grid with
  periodic, open and wall axes against an independent scipy implementation.
- BlockLabBC's faces through the operators: the pressure RHS and the pressure gradient near
  open and wall faces against numpy with edge (open) and negated-mirror (wall) padding.
- The advection's inflow fade and the inflow correction: closed forms for a uniform field.
"""
import numpy as np
import pytest
import scipy.fft as sf

TOL = 0.01  # PoissonSolverMixed.h:50


def test_r2r_lines_match_scipy_and_numpy(oracle):
    rng = np.random.default_rng(1)
    for n in (8, 16, 24, 40):
        x = rng.standard_normal(n)
        np.testing.assert_allclose(oracle.r2r_line(x, "REDFT10"), sf.dct(x, type=2), atol=1e-12)
        np.testing.assert_allclose(oracle.r2r_line(x, "REDFT01"), sf.dct(x, type=3), atol=1e-12)
        X = np.fft.fft(x)
        hc = np.concatenate([X.real[:n // 2 + 1], X.imag[1:n // 2][::-1]])
        np.testing.assert_allclose(oracle.r2r_line(x, "R2HC"), hc, atol=1e-12)
        np.testing.assert_allclose(oracle.r2r_line(hc, "HC2R"), n * x, atol=1e-11)


def _den(n, periodic):
    """The per-axis symbol of PoissonSolverMixed::_solve over the transform index i."""
    i = np.arange(n)
    if periodic:
        k, wf, sh = np.minimum(i, n - i), 2 * np.pi / n, 0.0
    else:
        k, wf, sh = i, np.pi / n, 0.5
    return (1 - TOL) * (np.cos(2 * wf * i) - 1) / 2 - TOL * ((k + sh) * wf) ** 2


def _mode(n, periodic, k, sine=False):
    """An eigenvector of the axis transform: cos/sin(2 pi k j/n) (periodic) or
    cos(pi k (j + 1/2)/n) (the DCT-II basis)."""
    j = np.arange(n)
    if periodic:
        return np.sin(2 * np.pi * k * j / n) if sine else np.cos(2 * np.pi * k * j / n)
    return np.cos(np.pi * k * (j + 0.5) / n)


@pytest.mark.parametrize("bc", [("periodic", "dirichlet", "wall"), ("dirichlet", "dirichlet", "dirichlet"),
                                ("periodic", "periodic", "periodic")])
@pytest.mark.parametrize("ks", [((1, False), (2, False), (3, False)), ((3, True), (0, False), (5, False)),
                                ((4, False), (7, True), (1, False))])
def test_mixed_solve_of_single_modes(oracle, bc, ks):
    """RHS = product of axis eigenvectors: p = RHS / (h (denX + denY + denZ)), since the norm
    factor cancels the r2r pairs' scales (N, or 2N for the DCT pair) and h^-1 remains."""
    s = oracle.OracleSim((2, 3, 4), maxextent=1.3, bc=bc)
    Nz, Ny, Nx = s.shape
    per = [b == "periodic" for b in bc]
    fx, fy, fz = (_mode(n, p, k, sine and p) for n, p, (k, sine) in zip((Nx, Ny, Nz), per, ks))
    rhs = fz[:, None, None] * fy[None, :, None] * fx[None, None, :]
    if all(per) and all(k == 0 for k, _ in ks):
        pytest.skip("the constant mode is the zeroed one")
    s.set_rhs(rhs)
    s.poisson_mixed()
    den = (_den(Nx, per[0])[ks[0][0]] + _den(Ny, per[1])[ks[1][0]] + _den(Nz, per[2])[ks[2][0]])
    np.testing.assert_allclose(s.get("p"), rhs / (s.h * den), rtol=0, atol=1e-12 * np.abs(rhs / (s.h * den)).max())


def _scipy_mixed(rhs, h, per):
    """PoissonSolverMixed with scipy: DCT-II/III (FFTW REDFT10/01) or the real DFT in FFTW's
    halfcomplex order, per axis of a [x][y][z] array."""
    a = rhs.copy()
    N = a.shape

    def fwd(a, ax, p):
        if not p:
            return sf.dct(a, type=2, axis=ax)
        X = np.fft.fft(a, axis=ax)
        n = a.shape[ax]
        re = np.take(X.real, np.arange(n // 2 + 1), axis=ax)
        im = np.take(X.imag, np.arange(n // 2 - 1, 0, -1), axis=ax)
        return np.concatenate([re, im], axis=ax)

    def bwd(a, ax, p):
        if not p:
            return sf.dct(a, type=3, axis=ax)
        n = a.shape[ax]
        re = np.take(a, np.arange(n // 2 + 1), axis=ax)
        im = np.take(a, np.arange(n - 1, n // 2, -1), axis=ax)  # i_1 .. i_{n/2-1}
        zero = np.zeros_like(np.take(a, [0], axis=ax))
        X = re + 1j * np.concatenate([zero, im, zero], axis=ax)
        full = np.concatenate([X, np.conj(np.flip(np.take(X, np.arange(1, n // 2), axis=ax), axis=ax))], axis=ax)
        return np.real(np.fft.ifft(full, axis=ax)) * n
    for ax in range(3):
        a = fwd(a, ax, per[ax])
    norm = 1 / h
    for ax in range(3):
        norm *= (1.0 if per[ax] else 0.5) / N[ax]
    D = (_den(N[0], per[0])[:, None, None] + _den(N[1], per[1])[None, :, None] + _den(N[2], per[2])[None, None, :])
    with np.errstate(divide="ignore", invalid="ignore"):
        a = a * norm / D
    a[0, 0, 0] = 0
    for ax in range(3):
        a = bwd(a, ax, per[ax])
    return a


@pytest.mark.parametrize("bc", [("periodic", "dirichlet", "wall"), ("wall", "periodic", "dirichlet"),
                                ("periodic", "periodic", "periodic")])
def test_mixed_solve_matches_scipy(oracle, bc):
    rng = np.random.default_rng(4)
    s = oracle.OracleSim((3, 2, 4), maxextent=2.0, bc=bc)
    rhs = rng.standard_normal(s.shape)
    s.set_rhs(rhs)
    s.poisson_mixed()
    per = [b == "periodic" for b in bc]
    # the scipy version works on [x][y][z]
    ref = _scipy_mixed(np.transpose(rhs, (2, 1, 0)), s.h, per)
    np.testing.assert_allclose(s.get("p"), np.transpose(ref, (2, 1, 0)), rtol=0, atol=1e-11 * np.abs(ref).max())


def _pad(f, bc, kind):
    """f [z][y][x] padded by one cell per side as BlockLabBC fills faces: periodic wrap, open
    (the boundary cell), wall (the velocity's negated mirror, which one cell out is the negated
    boundary cell; p: the boundary cell)."""
    out = f
    for ax, b in zip((2, 1, 0), bc):
        width = [(1, 1) if a == ax else (0, 0) for a in range(3)]
        out = np.pad(out, width, mode="wrap" if b == "periodic" else "edge")
        if b == "wall" and kind == "vel":
            for sl in (slice(0, 1), slice(-1, None)):
                idx = tuple(sl if a == ax else slice(None) for a in range(3))
                out[idx] = -out[idx]
    return out


def _central_diff(fp, ax):
    """(f[+1] - f[-1]) on the padded array, back on the interior."""
    c = [slice(1, -1)] * 3
    hi, lo = list(c), list(c)
    hi[ax] = slice(2, None)
    lo[ax] = slice(0, -2)
    return fp[tuple(hi)] - fp[tuple(lo)]


@pytest.mark.parametrize("bc", [("dirichlet", "wall", "periodic"), ("wall", "dirichlet", "dirichlet"),
                                ("wall", "wall", "wall")])
def test_boundary_labs_through_rhs_and_gradient(oracle, bc):
    """KernelPressureRHS and KernelGradP with BlockLabBC faces: the RHS near open faces uses
    the boundary cell's velocity as its outside neighbour, near walls the negated mirror; the
    gradient uses the boundary cell's p outside both."""
    rng = np.random.default_rng(6)
    s = oracle.OracleSim((2, 3, 2), maxextent=1.0, bc=bc)
    h, dt = s.h, 1e-3
    u, v, w, p = (rng.standard_normal(s.shape) for _ in range(4))
    s.set_velocity(u, v, w)
    s.pressure_rhs(dt)
    up, vp, wp = (_pad(f, bc, "vel") for f in (u, v, w))
    exp = 0.5 * h * h / dt * (_central_diff(up, 2) + _central_diff(vp, 1) + _central_diff(wp, 0))
    np.testing.assert_allclose(s.rhs(), exp, rtol=0, atol=1e-12 * np.abs(exp).max())
    s.set("p", p)
    s.grad_p(dt)
    pp = _pad(p, bc, "p")
    fac = dt / (2 * h)
    for f, got, ax in ((u, s.get("u"), 2), (v, s.get("v"), 1), (w, s.get("w"), 0)):
        np.testing.assert_allclose(got, f - fac * _central_diff(pp, ax), rtol=0, atol=1e-13)


def test_inflow_fade_and_correction_closed_form(oracle):
    """A uniform field u = (c, 0, 0) with uinf = (U, 0, 0), U > 0, open x faces, periodic y, z:
    - only the west halo is faded (the inflow face for U > 0): by 1 - fac, fac = min(1, umax dt/h);
      the advection of the west boundary cells is then c + facA (c + U)(c - fade c) + facD (fade c - c);
      every other cell keeps c;
    - the inflow correction removes sumInflow / (2 (NxNy + NxNz + NyNz)) from the east face's u
      and adds it on the west face, sumInflow = sum_E u - sum_W u."""
    s = oracle.OracleSim((2, 2, 2), maxextent=1.0, bc=("dirichlet", "periodic", "periodic"))
    Nz, Ny, Nx = s.shape
    h, dt, nu, c, U = s.h, 2e-3, 0.01, 0.7, 0.4
    z = np.zeros(s.shape)
    s.set_velocity(z + c, z, z)
    umax = 2.5
    s.set_umax_measured(umax)
    s.advect_diffuse(dt, nu, (U, 0, 0))
    fac = min(1.0, umax * dt / h)
    fade = 1 - fac
    facA, facD = -dt / (2 * h), (nu / h) * (dt / h)
    west = c + facA * (c + U) * (c - fade * c) + facD * (fade * c - c)
    sum_inflow = Ny * Nz * (c - west)
    corr = sum_inflow / (2 * (Nx * Ny + Nx * Nz + Ny * Nz))
    uu = s.get("u")
    np.testing.assert_allclose(uu[:, :, 0], west + corr, rtol=1e-13)
    np.testing.assert_allclose(uu[:, :, -1], c - corr, rtol=1e-13)
    np.testing.assert_allclose(uu[:, :, 1:-1], c, rtol=1e-14)
    assert np.abs(s.get("v")).max() == 0 and np.abs(s.get("w")).max() == 0


def test_inflow_correction_closes_the_flux_with_every_face_open(oracle):
    """With all six faces open the correction removes the whole net outflow:
    sum over the faces of u.n is zero after the advection (to rounding)."""
    rng = np.random.default_rng(8)
    s = oracle.OracleSim((2, 2, 3), maxextent=1.0, bc=("dirichlet", "dirichlet", "dirichlet"))
    u, v, w = (0.3 * rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(u, v, w)
    s.calc_max_timestep(0.3, 0.01, (0.2, -0.1, 0.05))
    s.advect_diffuse(1e-3, 0.01, (0.2, -0.1, 0.05))
    uu, vv, ww = s.velocity()
    flux = (uu[:, :, -1].sum() - uu[:, :, 0].sum() + vv[:, -1, :].sum() - vv[:, 0, :].sum()
            + ww[-1].sum() - ww[0].sum())
    scale = np.abs(uu).sum()
    assert abs(flux) < 1e-13 * scale


# PoissonSolverUnbounded: the oracle's direct free-space convolution

def _g(r, h):
    """The Green's function per unit RHS: -1/(4 pi h r), -0.1924173658/h at r = 0
    (PoissonSolverUnbounded.cpp:216-218, over the h^3 of :50)."""
    return np.where(r > 0, -1.0 / (4 * np.pi * h * np.maximum(r, 1e-300)), -0.1924173658 / h)


def _cells(shape):
    z, y, x = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    return x, y, z


def test_unbounded_point_sources_give_the_green_function(oracle):
    """Two point sources: p is the sum of the Green's function around each, with the r = 0
    value at the sources (fixes the kernel, its self term and the direct sum's indexing)."""
    s = oracle.OracleSim((2, 3, 2), maxextent=0.8, bc=("freespace",) * 3)
    h = s.h
    x, y, z = _cells(s.shape)
    rhs = np.zeros(s.shape)
    src = [((3, 5, 11), 1.5), ((12, 20, 2), -0.7)]  # (x, y, z), q
    for (sx, sy, sz), q in src:
        rhs[sz, sy, sx] = q
    s.set_rhs(rhs)
    s.poisson_unbounded()
    exp = sum(q * _g(np.sqrt((x - sx) ** 2 + (y - sy) ** 2 + (z - sz) ** 2), h) for (sx, sy, sz), q in src)
    np.testing.assert_allclose(s.get("p"), exp, rtol=1e-13, atol=0)


def test_unbounded_far_field_of_a_gaussian_is_the_monopole(oracle):
    """A cubically symmetric Gaussian charge Q (sum of the RHS): far from it, p = -Q/(4 pi R)
    with R the physical distance (its dipole and quadrupole vanish; the next term is
    O((sigma/R)^4)). Fixes the sign, the 4 pi and the powers of h."""
    s = oracle.OracleSim((3, 3, 3), maxextent=2.4, bc=("freespace",) * 3)
    h = s.h
    x, y, z = _cells(s.shape)
    c = 4
    rho = np.exp(-((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) / 2.0)
    rho[np.abs(x - c) > 3] = 0
    rho[np.abs(y - c) > 3] = 0
    rho[np.abs(z - c) > 3] = 0
    Q = rho.sum()
    s.set_rhs(rho)
    s.poisson_unbounded()
    p = s.get("p")
    R = h * np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    far = R > 12 * h
    ratio = p[far] * (-4 * np.pi * R[far]) / Q
    assert np.abs(ratio - 1).max() < 2e-3
    # h = 0.1 here, so a kernel with a wrong power of h would be off by a factor of 10
    assert abs(h - 0.1) < 1e-12


def test_unbounded_solution_solves_the_poisson_equation(oracle):
    """Away from a smooth source the discrete Laplacian of p is the density (RHS / h^3) up to
    the O(h^2) error of sampling the continuous Green's function: < 3% of the peak here."""
    s = oracle.OracleSim((3, 3, 3), maxextent=2.4, bc=("freespace",) * 3)
    h = s.h
    x, y, z = _cells(s.shape)
    c, sig = 12, 3.0
    rho = np.exp(-((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) / (2 * sig ** 2))
    s.set_rhs(rho * h ** 3)  # the RHS carries h^3 (PressureRHS.cpp:54)
    s.poisson_unbounded()
    p = s.get("p")
    lap = (np.roll(p, 1, 0) + np.roll(p, -1, 0) + np.roll(p, 1, 1) + np.roll(p, -1, 1) + np.roll(p, 1, 2)
           + np.roll(p, -1, 2) - 6 * p) / h ** 2
    inner = (slice(6, -6),) * 3
    assert np.abs(lap[inner] - rho[inner]).max() < 0.03 * rho.max()


def test_freespace_is_open_for_the_labs(oracle):
    """Freespace faces fill the labs as open faces: the same advection as dirichlet."""
    rng = np.random.default_rng(11)
    a = oracle.OracleSim((2, 2, 2), maxextent=1.0, bc=("freespace",) * 3)
    b = oracle.OracleSim((2, 2, 2), maxextent=1.0, bc=("dirichlet",) * 3)
    u, v, w = (0.3 * rng.standard_normal(a.shape) for _ in range(3))
    for s in (a, b):
        s.set_velocity(u, v, w)
        s.calc_max_timestep(0.3, 0.01, (0.2, 0.0, 0.0))
        s.advect_diffuse(1e-3, 0.01, (0.2, 0.0, 0.0))
        s.pressure_rhs(1e-3)
    for f, g in zip(a.velocity(), b.velocity()):
        assert np.array_equal(f, g)
    assert np.array_equal(a.rhs(), b.rhs())
