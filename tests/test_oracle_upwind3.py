"""Pins of the oracle's third-order upwind advection scheme (Euler and explicit-midpoint
RK2), which has no reference counterpart (the reference advects with the central scheme):
the exact single-mode amplification factor of linear advection on every axis and both
flow directions, the scheme's order of accuracy, and an independent numpy restatement for
fields whose velocity changes sign. CPU only."""
import numpy as np
import pytest

from test_oracle_core import cell_centres

AX = {"x": 2, "y": 1, "z": 0}  # array axis of each direction ([z][y][x])


def up3_symbol(k, h, positive):
    """6h * first-derivative symbol of the upwind-biased third-order stencil (f = e^{ikx})."""
    e = np.exp(1j * k * h)
    if positive:
        return 2 * e + 3 - 6 / e + 1 / e ** 2
    return -e ** 2 + 6 * e - 3 - 2 / e


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("U", [0.8, -0.6])
@pytest.mark.parametrize("scheme", [1, 2])
def test_linear_advection_amplification_factor(oracle, axis, U, scheme):
    """A passive component f = A sin(k s) along axis s, transported by a constant U along s
    (the other two components zero): the step is linear, so f_new = Im(G A e^{iks}) with
    z = -dt (U D(k) + 4 nu sin^2(kh/2)/h^2), G = 1 + z (Euler) or 1 + z + z^2/2 (midpoint)."""
    s = oracle.OracleSim((2, 3, 4))  # 16 x 24 x 32: different lengths per axis
    h = s.h
    x, y, z = cell_centres(s.shape, h)
    coord = {"x": x, "y": y, "z": z}[axis]
    L = s.shape[AX[axis]] * h
    k = 2 * np.pi * 3 / L
    A, nu, dt = 0.3, 0.02, 0.7 * h / abs(U)
    # transported component: one orthogonal to the axis (so its own velocity does not advect it)
    comp = {"x": "w", "y": "u", "z": "v"}[axis]
    uinf = [0.0, 0.0, 0.0]
    uinf["xyz".index(axis)] = U
    f0 = A * np.sin(k * coord)
    for c in "uvw":
        s.set(c, f0 if c == comp else np.zeros(s.shape))
    s.advect_diffuse(dt, nu, uinf, scheme=scheme)
    zz = -dt * (U * up3_symbol(k, h, U > 0) / (6 * h) + 4 * nu * np.sin(k * h / 2) ** 2 / h ** 2)
    G = 1 + zz if scheme == 1 else 1 + zz + zz * zz / 2
    exp = np.imag(G * A * np.exp(1j * k * coord))
    got = s.get(comp)
    assert np.max(np.abs(got - exp)) < 1e-14
    for c in "uvw":
        if c != comp:
            assert np.max(np.abs(s.get(c))) == 0.0
    assert np.array_equal(s.get("tmp" + comp.upper()), got)


def test_upwind3_is_third_order():
    """The derivative symbol D(k)/(6h) - ik shrinks as (kh)^4 / h, i.e. the error of the
    derivative is O(h^3), and its real part has the dissipative sign."""
    k = 1.0
    errs = []
    for h in (0.2, 0.1, 0.05):
        d = up3_symbol(k, h, True) / (6 * h)
        errs.append(abs(d - 1j * k))
        assert np.real(d) > 0  # U > 0: Re(U D) > 0 damps
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 2.9)


def numpy_up3(u, v, w, dt, nu, h, U):
    """Independent restatement with numpy rolls and np.where, one Euler stage."""
    def sh(f, ax, k):  # value at offset +k along ax
        return np.roll(f, -k, ax)

    vel = [u + U[0], v + U[1], w + U[2]]
    out = []
    for f in (u, v, w):
        adv = 0
        lap = 0
        for d, ax in ((0, 2), (1, 1), (2, 0)):
            pos = 2 * sh(f, ax, 1) + 3 * f - 6 * sh(f, ax, -1) + sh(f, ax, -2)
            neg = -sh(f, ax, 2) + 6 * sh(f, ax, 1) - 3 * f - 2 * sh(f, ax, -1)
            adv = adv + vel[d] * np.where(vel[d] > 0, pos, neg)
            lap = lap + sh(f, ax, 1) + sh(f, ax, -1) - 2 * f
        out.append(f + dt * (-adv / (6 * h) + nu * lap / h ** 2))
    return out


@pytest.mark.parametrize("scheme", [1, 2])
def test_upwind3_random_fields_match_numpy(oracle, scheme):
    rng = np.random.default_rng(7)
    s = oracle.OracleSim((2, 2, 3))
    h = s.h
    u, v, w = (rng.standard_normal(s.shape) for _ in range(3))
    U = (0.1, -0.2, 0.05)
    dt, nu = 0.2 * h, 0.01
    for c, f in zip("uvw", (u, v, w)):
        s.set(c, f)
    s.advect_diffuse(dt, nu, U, scheme=scheme)
    if scheme == 1:
        exp = numpy_up3(u, v, w, dt, nu, h, U)
    else:
        half = numpy_up3(u, v, w, dt / 2, nu, h, U)
        # second stage: u + dt L(u_half), where L(f) = (numpy_up3(f, dt) - f) / dt
        stage = numpy_up3(*half, dt, nu, h, U)
        exp = [f0 + (st - fh) for f0, st, fh in zip((u, v, w), stage, half)]
    for c, e in zip("uvw", exp):
        assert np.max(np.abs(s.get(c) - e)) < 1e-12 * max(1.0, np.max(np.abs(e)))
