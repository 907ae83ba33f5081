"""Pins of the oracle's core operators (halo, advection-diffusion, RHS, FFT, Poisson,
gradient, dt) against the reference's own test, closed forms, golden vectors and an
independent numpy restatement. CPU only."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cell_centres(shape, h):
    Z, Y, X = shape
    z, y, x = np.meshgrid((np.arange(Z) + 0.5) * h, (np.arange(Y) + 0.5) * h, (np.arange(X) + 0.5) * h,
                          indexing="ij")
    return x, y, z


# ---------------------------------------------------------------- halo (lab) fill

@pytest.mark.parametrize("g,shift", [(1, (-1, 0, 0)), (1, (1, 0, 0)), (2, (-1, 1, 2))])
def test_lab_periodic_shift(oracle, g, shift):
    """TestBoundaries.cpp:19-75: u = ix + 1000 iy + 1e6 iz on 128x32x32, tmpU = lab(-d).u
    must equal the periodically wrapped value exactly."""
    s = oracle.OracleSim((16, 4, 4))
    Z, Y, X = s.shape
    iz, iy, ix = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    s.set("u", ix + 1000.0 * iy + 1e6 * iz)
    dx, dy, dz = shift
    s.lab_shift(g, dx, dy, dz)
    got = s.get("tmpU")
    exp = ((ix - dx) % X) + 1000.0 * ((iy - dy) % Y) + 1e6 * ((iz - dz) % Z)
    assert np.array_equal(got, exp)


# ---------------------------------------------------------- advection-diffusion

def numpy_advdiff(u, v, w, dt, nu, h, U=(0, 0, 0)):
    """Independent restatement of KernelAdvectDiffuse with numpy rolls ([z][y][x])."""
    def d(f, ax):
        return np.roll(f, -1, ax) - np.roll(f, 1, ax)

    def lap(f):
        return sum(np.roll(f, s, a) for a in range(3) for s in (-1, 1)) - 6 * f
    facA = -dt / (2 * h)
    facD = (nu / h) * (dt / h)
    uu, vv, ww = u + U[0], v + U[1], w + U[2]
    out = []
    for f in (u, v, w):
        adv = uu * d(f, 2) + vv * d(f, 1) + ww * d(f, 0)
        out.append(f + facA * adv + facD * lap(f))
    return out


def test_advdiff_shear_wave_closed_form(oracle):
    """u = A sin(ky): (u.grad)u = 0, so one step multiplies by 1 - 4 nu dt sin^2(kh/2)/h^2."""
    s = oracle.OracleSim((2, 4, 2))
    h = s.h
    x, y, z = cell_centres(s.shape, h)
    k, A, nu, dt = 3.0, 0.7, 0.05, 1e-2
    s.set_velocity(A * np.sin(k * y), 0 * x, 0 * x)
    s.advect_diffuse(dt, nu)
    u, v, w = s.velocity()
    fac = 1 - 4 * nu * dt * np.sin(k * h / 2) ** 2 / h ** 2
    np.testing.assert_allclose(u, A * np.sin(k * y) * fac, rtol=0, atol=1e-14)
    assert np.all(v == 0) and np.all(w == 0)


def test_advdiff_advected_mode_symbol(oracle):
    """u = cos(ky) advected by U_y: one step gives cos(ky)(1 - 4 nu dt sin^2(kh/2)/h^2)
    + dt U sin(kh)/h sin(ky) (central difference symbol)."""
    s = oracle.OracleSim((2, 2, 2))
    h = s.h
    x, y, z = cell_centres(s.shape, h)
    k, U, nu, dt = 2.0, 0.8, 0.01, 5e-3
    s.set_velocity(np.cos(k * y), 0 * x, 0 * x)
    s.advect_diffuse(dt, nu, uinf=(0, U, 0))
    u, _, _ = s.velocity()
    exp = np.cos(k * y) * (1 - 4 * nu * dt * np.sin(k * h / 2) ** 2 / h ** 2) + dt * U * np.sin(k * h) / h * np.sin(k * y)
    np.testing.assert_allclose(u, exp, rtol=0, atol=1e-14)


def test_advdiff_brute_force(oracle):
    rng = np.random.default_rng(7)
    s = oracle.OracleSim((2, 2, 2))
    u, v, w = (rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(u, v, w)
    U = (0.3, -0.2, 0.1)
    s.advect_diffuse(1e-3, 0.02, uinf=U)
    got = s.velocity()
    exp = numpy_advdiff(u, v, w, 1e-3, 0.02, s.h, U)
    for a, b in zip(got, exp):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(s.get("tmpU"), got[0])


# ------------------------------------------------------------------------ RHS

def test_pressure_rhs_mode(oracle):
    """rhs = h^2/(2dt) (u_E - u_W + ...) of u = sin(kx x): h^2/dt sin(kx h) cos(kx x)."""
    s = oracle.OracleSim((2, 2, 2))
    h = s.h
    x, y, z = cell_centres(s.shape, h)
    dt = 0.01
    s.set_velocity(np.sin(2 * x), np.sin(3 * y), np.sin(z))
    s.pressure_rhs(dt)
    exp = h * h / dt * (np.sin(2 * h) * np.cos(2 * x) + np.sin(3 * h) * np.cos(3 * y) + np.sin(h) * np.cos(z))
    np.testing.assert_allclose(s.rhs(), exp, rtol=0, atol=1e-13)


# ------------------------------------------------------------------------ FFT

def test_fft_golden(oracle):
    g = np.load(os.path.join(GOLD, "fft_golden.npz"))
    np.testing.assert_allclose(oracle.fft3d(g["a"], -1), g["A"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(oracle.fft3d(g["r"], -1), g["R"], rtol=0, atol=1e-11)
    back = oracle.fft3d(g["A"], +1) / g["a"].size
    np.testing.assert_allclose(back, g["a"], rtol=0, atol=1e-13)


def test_fft_mixed_radix_golden(oracle):
    """Lengths that are not powers of two (the recursive mixed-radix path) against numpy.fft:
    lines of 12, 24, 48 and 96 points (each axis position) and a 24 x 40 x 12 real array."""
    g = np.load(os.path.join(GOLD, "fft_mixed_golden.npz"))
    for n in (12, 24, 48, 96):
        line, ref = g[f"line{n}"], g[f"LINE{n}"]
        for shape in ((n, 1, 1), (1, n, 1), (1, 1, n)):
            got = oracle.fft3d(line.reshape(shape), -1).reshape(n)
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        back = oracle.fft3d(ref.reshape(1, 1, n), +1).reshape(n) / n
        np.testing.assert_allclose(back, line, rtol=0, atol=1e-13 * np.abs(line).max())
    np.testing.assert_allclose(oracle.fft3d(g["m"], -1), g["M"], rtol=0, atol=1e-11)


def test_poisson_mixed_radix_golden(oracle):
    """The periodic solve on 40 x 16 x 24 cells (x, y, z) against numpy's."""
    g = np.load(os.path.join(GOLD, "fft_mixed_golden.npz"))
    s = oracle.OracleSim((5, 2, 3))
    assert abs(s.h - float(g["h"])) < 1e-15 and s.shape == g["rhs"].shape
    s.set_rhs(g["rhs"])
    s.poisson_solve()
    np.testing.assert_allclose(s.get("p"), g["p"], rtol=0, atol=1e-12 * np.abs(g["p"]).max())


# --------------------------------------------------------------------- Poisson

def test_poisson_golden(oracle):
    g = np.load(os.path.join(GOLD, "poisson_golden.npz"))
    s = oracle.OracleSim((2, 2, 2))
    assert abs(s.h - float(g["h"])) < 1e-15
    s.set_rhs(g["rhs"])
    s.poisson_solve()
    np.testing.assert_allclose(s.get("p"), g["p"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", [(1, 0, 0), (2, 3, 1), (5, 1, 7)])
def test_poisson_symbol_single_mode(oracle, mode):
    """A single cosine mode of the RHS is divided by (2D-6)/(4h^2) times h^3 (closed form):
    p = rhs * 4/(h (2D-6)) since the RHS carries h^3 (PoissonSolverPeriodic.cpp:23-35)."""
    s = oracle.OracleSim((2, 2, 2))
    N = s.shape[0]
    h = s.h
    kx, ky, kz = mode
    Z, Y, X = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    rhs = np.cos(2 * np.pi * (kx * X + ky * Y + kz * Z) / N)
    s.set_rhs(rhs)
    s.poisson_solve()
    D = np.cos(4 * np.pi * kx / N) + np.cos(4 * np.pi * ky / N) + np.cos(4 * np.pi * kz / N)
    np.testing.assert_allclose(s.get("p"), rhs * 4 / (h * (2 * D - 6)), rtol=1e-12, atol=1e-13)


def test_poisson_zeroed_modes(oracle):
    """The 8 modes with every index in {0, N/2} are removed (incl. the mean)."""
    s = oracle.OracleSim((2, 2, 2))
    N = s.shape[0]
    Z, Y, X = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    rhs = 1.0 + np.cos(np.pi * X) + np.cos(np.pi * Y) * np.cos(np.pi * Z)
    s.set_rhs(rhs)
    s.poisson_solve()
    np.testing.assert_allclose(s.get("p"), 0, atol=1e-13)


def test_projection_divergence_free(oracle):
    """RHS -> Poisson -> gradient leaves div_c(u) = 0 and a zero-mean p; projecting again
    changes nothing (idempotence)."""
    rng = np.random.default_rng(3)
    s = oracle.OracleSim((2, 2, 2))
    h, dt = s.h, 1e-3
    s.set_velocity(*(rng.standard_normal(s.shape) for _ in range(3)))
    s.pressure_rhs(dt); s.poisson_solve(); s.grad_p(dt)
    u, v, w = s.velocity()
    div = (np.roll(u, -1, 2) - np.roll(u, 1, 2) + np.roll(v, -1, 1) - np.roll(v, 1, 1)
           + np.roll(w, -1, 0) - np.roll(w, 1, 0)) / (2 * h)
    assert np.abs(div).max() < 1e-12 * np.abs(u).max() / h
    assert abs(s.get("p").mean()) < 1e-12
    s.pressure_rhs(dt); s.poisson_solve(); s.grad_p(dt)
    u2, _, _ = s.velocity()
    np.testing.assert_allclose(u2, u, atol=1e-12)


def test_grad_p_mode(oracle):
    s = oracle.OracleSim((2, 2, 2))
    h = s.h
    x, y, z = cell_centres(s.shape, h)
    s.set_velocity(0 * x, 0 * x, 0 * x)
    s.set("p", np.cos(2 * x) + np.sin(y))
    dt = 0.1
    s.grad_p(dt)
    u, v, _ = s.velocity()
    np.testing.assert_allclose(u, dt / h * np.sin(2 * h) * np.sin(2 * x), atol=1e-14)
    np.testing.assert_allclose(v, -dt / h * np.sin(h) * np.cos(y), atol=1e-14)


# ---------------------------------------------------------------------- dt, max

def test_calc_max_timestep(oracle):
    s = oracle.OracleSim((2, 2, 2))
    rng = np.random.default_rng(1)
    u = rng.standard_normal(s.shape)
    s.set_velocity(u, 0.5 * u, -2 * u)
    U = (0.3, 0, 0)
    umax = max(np.abs(u + 0.3).max(), np.abs(0.5 * u).max(), np.abs(2 * u).max())
    assert s.find_max_u(U) == umax
    h, nu, CFL = s.h, 0.01, 0.3
    dt, um = s.calc_max_timestep(CFL, nu, U, step=0, rampup=0)
    assert dt == pytest.approx(CFL * min(h * h / nu, h / (umax + 1e-8)), rel=1e-15)
    dt2, _ = s.calc_max_timestep(CFL, nu, U, step=9, rampup=100)
    x = 10 / 100
    ramp = np.exp(np.log(1e-3) * (1 - x) + np.log(CFL) * x)
    assert dt2 == pytest.approx(ramp * min(h * h / nu, h / (umax + 1e-8)), rel=1e-14)
