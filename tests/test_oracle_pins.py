"""Independent pins of three oracle functions: the obstacle masking and perimeter correction
of the pressure RHS, the surface forces, and the 20 dissipation integrals. CPU only.

Each check comes from something other than the oracle's own formula: a conservation law,
the divergence theorem with exact ball integrals of Fourier modes, a vector-calculus
restatement in numpy (np.cross, full strain tensors), or a closed form of the central
difference symbols on single modes. Each test also plants a mistake in the oracle
(oracle.planted) and shows that its check fails for it.
"""
import numpy as np
import pytest


def centres(s):
    h = s.h
    Z, Y, X = s.shape
    return np.meshgrid((np.arange(Z) + 0.5) * h, (np.arange(Y) + 0.5) * h, (np.arange(X) + 0.5) * h,
                       indexing="ij")[::-1]  # x, y, z


def dc(f, ax, h):
    """Central difference along axis ax of a [z][y][x] array (ax 2: x, 1: y, 0: z)."""
    return (np.roll(f, -1, ax) - np.roll(f, 1, ax)) / (2 * h)


# ------------------------------------------------ pressure RHS: masking and correction

def obstacle_blocks(s):
    """The active body's obstacle blocks as a cell mask: udef is kept only on them."""
    one, zero = np.ones(s.shape), np.zeros(s.shape)
    s.set_udef(one, one, one)
    m = s.get_udef()[0] == 1
    s.set_udef(zero, zero, zero)
    return m


def plain_rhs(s, u, v, w, dt):
    """KernelPressureRHS without a body: h^2/(2 dt) times the central differences' sum."""
    d = sum(np.roll(f, -1, a) - np.roll(f, 1, a) for f, a in ((u, 2), (v, 1), (w, 0)))
    return 0.5 * s.h * s.h / dt * d


def rhs_masking_errors(s, u, v, w, dt):
    """PressureRHS with one body (PressureRHS.cpp:78-112, 215-256, 322-343) against what it
    must do, from the plain central divergence rhs0 computed here:
    - outside the obstacle blocks: rhs = rhs0;
    - inside them, off the surface: rhs = (1 - chi) rhs0;
    - on the surface: the cells whose masked value has the sign of sumRHS = -sum chi rhs0
      are scaled by one common factor; the others keep (1 - chi) rhs0;
    - the correction removes exactly what the masking added: sum rhs = sum rhs0 = 0.
    Returns (the largest violation relative to max |rhs0|, the total relative to sum |rhs0|)."""
    rhs0 = plain_rhs(s, u, v, w, dt)
    chi = s.get("chi")
    M = obstacle_blocks(s)
    s.set_velocity(u, v, w)
    s.pressure_rhs(dt)
    r = s.rhs()
    g2 = sum((np.roll(chi, -1, a) - np.roll(chi, 1, a)) ** 2 for a in range(3))
    surf = (g2 > 0) & M
    r1 = (1 - chi) * rhs0
    sum_rhs = (-chi * rhs0)[M].sum()
    adj = surf & (np.sign(r1) == np.sign(sum_rhs))
    c = (r[adj] * r1[adj]).sum() / (r1[adj] ** 2).sum()
    m = np.abs(rhs0).max()
    errs = [np.abs(r[~M] - rhs0[~M]).max(), np.abs(r[M & ~surf] - r1[M & ~surf]).max(),
            np.abs(r[surf & ~adj] - r1[surf & ~adj]).max(), np.abs(r[adj] - c * r1[adj]).max()]
    assert adj.sum() > 50 and abs(sum_rhs) > 1e-6 * np.abs(rhs0).sum()
    return max(errs) / m, abs(r.sum()) / np.abs(rhs0).sum()


@pytest.mark.parametrize("flip", [1.0, -1.0])
@pytest.mark.parametrize("R,c", [(0.2, (0.5, 0.47, 0.52)), (0.13, (0.31, 0.6, 0.44))])
def test_rhs_masking_and_perimeter_correction(oracle, R, c, flip):
    """flip = -1 reverses the velocity, and with it the sign of sumRHS: both branches of the
    correction (posRHS and negRHS) are exercised."""
    rng = np.random.default_rng(3)
    s = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    s.sphere_create(R, c)
    u, v, w = (flip * rng.standard_normal(s.shape) for _ in range(3))
    cell, total = rhs_masking_errors(s, u, v, w, 1e-3)
    assert cell < 1e-14 and total < 1e-14
    with oracle.planted("rhs_corr_sign"):
        cell, total = rhs_masking_errors(s, u, v, w, 1e-3)
    assert cell > 1e-3 and total > 1e-4


def test_rhs_masking_of_two_bodies_conserves_each(oracle):
    """Two bodies apart: outside both bodies' blocks the RHS is the plain divergence, and the
    total is still zero (each body's correction removes its own masking's sum)."""
    rng = np.random.default_rng(5)
    s = oracle.OracleSim((6, 4, 4), maxextent=1.5)
    s.sphere_create(0.12, (0.35, 0.5, 0.5))
    s.add_sphere(0.1, (1.1, 0.45, 0.55))
    u, v, w = (rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(u, v, w)
    dt = 1e-3
    rhs0 = plain_rhs(s, u, v, w, dt)
    masks = []
    for k in range(2):
        with s.body(k):
            masks.append(obstacle_blocks(s))
    assert not (masks[0] & masks[1]).any()
    s.pressure_rhs(dt)
    r = s.rhs()
    out = ~(masks[0] | masks[1])
    assert np.abs(r[out] - rhs0[out]).max() < 1e-14 * np.abs(rhs0).max()
    assert abs(r.sum()) < 1e-14 * np.abs(rhs0).sum()
    for m in masks:
        assert np.abs(r[m] - rhs0[m]).max() > 1e-3 * np.abs(rhs0).max()


# ------------------------------------------------------------------ surface forces
#
# The surface points carry n dS (outward normal times area, Towers' delta times h^3) on the
# surface one cell outside the SDF's zero (SURFDH = 1). By the divergence theorem the sums
# approach volume integrals over the ball of radius R + h:
#   pressure force  -sum p n dS            -> -int grad p dV
#   viscous force   sum 2 nu S n dS        ->  int nu (lap u + grad div u) dV = nu int lap u dV
#   viscous torque  sum r x (2 nu S n) dS  ->  nu int r x lap u dV   (S symmetric)
#   circulation     sum n x u dS           ->  int curl u dV
# For fields that are sums of Fourier modes a cos(k.x) + b sin(k.x) these integrals are
# exact: int_ball e^{ik.x} dV = e^{ik.c} g(|k|) with g(k) = 4 pi (sin kR - kR cos kR) / k^3,
# and int_ball (x - c) e^{ik.x} dV = -i e^{ik.c} g'(|k|) k/|k|.

def _g0(k, R):
    kR = k * R
    return 4 * np.pi * (np.sin(kR) - kR * np.cos(kR)) / k ** 3


def _g1(k, R):
    kR = k * R
    return 4 * np.pi * (kR * R * np.sin(kR) - 3 * np.sin(kR) + 3 * kR * np.cos(kR)) / k ** 4


def modes_eval(modes, x, y, z):
    out = np.zeros_like(x)
    for a, b, k in modes:
        ph = k[0] * x + k[1] * y + k[2] * z
        out += a * np.cos(ph) + b * np.sin(ph)
    return out


def ball_integral(modes, c, R):
    t = 0.0
    for a, b, k in modes:
        kk, ph = np.linalg.norm(k), np.dot(k, c)
        t += _g0(kk, R) * (a * np.cos(ph) + b * np.sin(ph))
    return t


def ball_moment(modes, c, R):
    """int_ball (x - c) f dV."""
    t = np.zeros(3)
    for a, b, k in modes:
        kk, ph = np.linalg.norm(k), np.dot(k, c)
        t += _g1(kk, R) * np.asarray(k) / kk * (a * np.sin(ph) - b * np.cos(ph))
    return t


def modes_grad(modes):
    return [[(b * k[d], -a * k[d], k) for a, b, k in modes] for d in range(3)]


def modes_lap(modes):
    return [(-a * np.dot(k, k), -b * np.dot(k, k), k) for a, b, k in modes]


# A divergence-free velocity in the 2 pi box: the ABC flow plus two cellular flows with
# diagonal strain (sin x cos y, -cos x sin y, 0) and (0, sin y cos z, -cos y sin z), so that
# every strain component D11..D33 is exercised.
_A, _B, _C = 0.7, -0.4, 0.9
VEL = ([(0, _A, (0, 0, 1)), (_C, 0, (0, 1, 0)), (0, 0.5, (1, 1, 0)), (0, 0.5, (1, -1, 0))],
       [(0, _B, (1, 0, 0)), (_A, 0, (0, 0, 1)), (0, -0.5, (1, 1, 0)), (0, 0.5, (1, -1, 0)), (0, 0.5, (0, 1, 1)),
        (0, 0.5, (0, 1, -1))],
       [(0, _C, (0, 1, 0)), (_B, 0, (1, 0, 0)), (0, -0.5, (0, 1, 1)), (0, 0.5, (0, 1, -1))])
PRES = [(0.8, 0, (1, 0, 0)), (0, -0.6, (0, 1, 1)), (0.3, 0.2, (1, -1, 1)), (0, 0.9, (0, 0, 2))]
CENTRE, RADIUS = np.array([3.02, 3.2, 3.1]), 1.1


def test_velocity_modes_are_solenoidal():
    gu, gv, gw = (modes_grad(m) for m in VEL)
    rng = np.random.default_rng(0)
    x, y, z = rng.uniform(0, 2 * np.pi, (3, 50))
    div = modes_eval(gu[0], x, y, z) + modes_eval(gv[1], x, y, z) + modes_eval(gw[2], x, y, z)
    assert np.abs(div).max() < 1e-14


def _viscous_errors(oracle, n, nu=0.3):
    s = oracle.OracleSim((n, n, n), maxextent=2 * np.pi)
    s.sphere_create(RADIUS, CENTRE)
    x, y, z = centres(s)
    s.set_velocity(*(modes_eval(m, x, y, z) for m in VEL))
    s.set("p", np.zeros(s.shape))
    Rs = RADIUS + s.h
    F = np.array([nu * ball_integral(modes_lap(m), CENTRE, Rs) for m in VEL])
    mom = [ball_moment(modes_lap(m), CENTRE, Rs) for m in VEL]  # mom[j][i] = int r_i lap u_j
    T = nu * np.array([mom[2][1] - mom[1][2], mom[0][2] - mom[2][0], mom[1][0] - mom[0][1]])
    g = [modes_grad(m) for m in VEL]  # g[j][i]: d u_j / d x_i
    om = np.array([ball_integral(g[2][1], CENTRE, Rs) - ball_integral(g[1][2], CENTRE, Rs),
                   ball_integral(g[0][2], CENTRE, Rs) - ball_integral(g[2][0], CENTRE, Rs),
                   ball_integral(g[1][0], CENTRE, Rs) - ball_integral(g[0][1], CENTRE, Rs)])

    def errs(q):
        return (np.abs(q[6:9] - F).max() / np.abs(F).max(), np.abs(q[9:12] - T).max() / np.abs(T).max(),
                np.abs(q[12:15] - om).max() / np.abs(om).max())
    q = s.surface_forces(nu, (0.3, -0.2, 0.1))
    assert np.all(q[3:6] == 0)
    with oracle.planted("forces_shear_half"):
        qp = s.surface_forces(nu, (0.3, -0.2, 0.1))
    return errs(q), errs(qp)


def test_surface_viscous_force_torque_and_circulation_by_the_divergence_theorem(oracle):
    """The viscous force, its torque and sum n x u against exact ball integrals, on 64^3 and
    128^3 grids (the radius is 11 and 22 cells): the force and circulation errors fall from
    3.4e-3 and 3.7e-3 to 1.2e-3 and 1.1e-3, the torque's from 3.7e-2 to 2.6e-2 of |T|. The
    strain's 1/2 on D12, D13, D23 dropped makes the force 80% wrong."""
    (f8, t8, g8), _ = _viscous_errors(oracle, 8)
    (f16, t16, g16), (fp, tp, _) = _viscous_errors(oracle, 16)
    assert f16 < 2e-3 and g16 < 2e-3 and t16 < 4e-2
    assert f16 < 0.5 * f8 and g16 < 0.5 * g8 and t16 < t8
    assert fp > 0.5 and tp > 0.2


def test_surface_pressure_force_and_power_identities(oracle):
    """A pressure made of four modes with a uniform velocity U: the pressure force against
    -int_ball grad p dV (error 1.4e-2 on 64^3, 1.2e-3 on 128^3), no viscous force, the
    pressure torque about the centre near zero (it is exactly zero for a ball: the normal
    passes through the centre), and the sums that follow from the per-point terms:
    thrust - drag = F.U/|U| with both non-negative, Pout = F.U, PoutBnd <= min(Pout, 0),
    pLocomotion = F.U + Omega.T."""
    U, Om = np.array([0.4, -0.7, 0.2]), np.array([0.3, 0.5, -0.2])
    errs = []
    for n in (8, 16):
        s = oracle.OracleSim((n, n, n), maxextent=2 * np.pi)
        s.sphere_create(RADIUS, CENTRE)
        x, y, z = centres(s)
        s.set_velocity(*(np.full(s.shape, a) for a in U))
        s.set("p", modes_eval(PRES, x, y, z))
        q = s.surface_forces(0.3, U, Om)
        F = -np.array([ball_integral(gm, CENTRE, RADIUS + s.h) for gm in modes_grad(PRES)])
        errs.append(np.abs(q[3:6] - F).max() / np.abs(F).max())
        assert np.all(q[6:9] == 0)
        assert np.abs(q[9:12]).max() < 1e-3 * np.abs(F).max() * RADIUS
        scale = np.abs(F).max()
        e = U / np.linalg.norm(U)
        assert q[15] >= 0 and q[16] >= 0
        assert q[16] - q[15] == pytest.approx(q[0:3] @ e, abs=1e-13 * scale)
        assert q[17] == pytest.approx(q[0:3] @ U, abs=1e-13 * scale)
        assert q[18] <= min(q[17], 0) + 1e-15
        assert q[21] == pytest.approx(q[0:3] @ U + Om @ q[9:12], abs=1e-13 * scale)
        with oracle.planted("forces_thrust_drag"):
            qp = s.surface_forces(0.3, U, Om)
        assert abs((qp[16] - qp[15]) - qp[0:3] @ e) > 0.5 * abs(q[0:3] @ e)
    assert errs[1] < 2e-3 and errs[1] < 0.5 * errs[0]


# ------------------------------------------------------------ dissipation integrals

def numpy_dissipation(s, nu):
    """KernelDissipation's 20 integrals (ComputeDissipation.cpp:46-115) written from their
    vector definitions: the central velocity gradient tensor G, the vorticity from it,
    impulses and momenta by np.cross about the box centre, the full strain tensor S:S."""
    h = s.h
    U = np.stack(s.velocity())
    p, chi = s.get("p"), s.get("chi")
    x, y, z = centres(s)
    ext = s.extent
    r = np.stack([x - ext[0] / 2, y - ext[1] / 2, z - ext[2] / 2])
    axis = (2, 1, 0)
    G = np.array([[dc(U[i], axis[j], h) for j in range(3)] for i in range(3)])  # G[i][j] = du_i/dx_j
    om = np.stack([G[2][1] - G[1][2], G[0][2] - G[2][0], G[1][0] - G[0][1]])
    S = 0.5 * (G + G.transpose(1, 0, 2, 3, 4))
    gp = np.stack([dc(p, axis[j], h) for j in range(3)])
    lapU = np.stack([sum(np.roll(U[i], -1, a) + np.roll(U[i], 1, a) for a in range(3)) / h ** 2 - 6 * U[i] / h ** 2
                     for i in range(3)])
    h3 = h ** 3
    out = np.zeros(20)
    out[0:3] = h3 * om.sum(axis=(1, 2, 3))
    out[3:6] = h3 / 2 * np.cross(r, om, axis=0).sum(axis=(1, 2, 3))
    out[6:9] = h3 * U.sum(axis=(1, 2, 3))
    out[9:12] = h3 / 3 * np.cross(r, np.cross(r, om, axis=0), axis=0).sum(axis=(1, 2, 3))
    out[12:15] = h3 * np.cross(r, U, axis=0).sum(axis=(1, 2, 3))
    out[15] = -h3 * ((1 - chi) * (gp * U).sum(0)).sum()
    out[16] = h3 * nu * ((1 - chi) * ((lapU * U).sum(0) + 2 * np.einsum("ij...,ij...->...", S, S))).sum()
    out[17] = h3 * (om * U).sum()
    out[18] = h3 * (U * U).sum() / 2
    out[19] = np.sqrt((om * om).sum(0)).sum()
    return out, h3 * np.sqrt((om * om).sum(0)).sum()


def test_dissipation_matches_the_vector_restatement_on_random_fields(oracle):
    """Random u, v, w, p and chi in [0, 1] on a non-cubic grid: every integral to 1e-13
    relative (the circulation, zero on a periodic grid, to 1e-13 of h^3 sum |omega|). Each
    planted mistake moves one integral by 8e-8 to 0.8 of its value."""
    rng = np.random.default_rng(11)
    s = oracle.OracleSim((2, 3, 2), maxextent=1.3)
    for n in ("u", "v", "w", "p"):
        s.set(n, rng.standard_normal(s.shape))
    s.set("chi", rng.uniform(0, 1, s.shape))
    nu = 0.07
    ref, circ_scale = numpy_dissipation(s, nu)

    def rel(got):
        return np.abs(got[3:] - ref[3:]) / np.abs(ref[3:])
    got = s.dissipation(nu)
    assert np.abs(got[0:3]).max() < 1e-13 * circ_scale and np.abs(ref[0:3]).max() < 1e-13 * circ_scale
    assert rel(got).max() < 1e-13
    for name in ("diss_shear_weight", "diss_impulse_sign", "diss_angimp_term", "diss_helicity"):
        with oracle.planted(name):
            assert rel(s.dissipation(nu)).max() > 1e-9, name


def test_dissipation_single_mode_closed_forms(oracle):
    """Closed forms of the central-difference symbols (s1 = sin(h)/h, the first derivative's;
    the 7-point Laplacian's -4 sin^2(h/2)/h^2) in the 2 pi box of volume V, chi = 0:
    - the ABC flow is a Beltrami field, curl_c u = s1 u: helicity = s1 (A^2 + B^2 + C^2) V;
    - the shear mode u = (sin y, 0, 0): viscous power nu V (-2 sin^2(h/2)/h^2 + s1^2 / 2),
      which vanishes in the continuum limit, and the pressure power of p = sin x with
      u = (cos x, 0, 0) is -s1 V / 2; the kinetic energy of the shear mode is V / 4."""
    s = oracle.OracleSim((4, 4, 4))
    h, V, nu = s.h, (2 * np.pi) ** 3, 0.02
    s1 = np.sin(h) / h
    x, y, z = centres(s)
    zero = np.zeros(s.shape)
    s.set("chi", zero)
    s.set("p", zero)
    A, B, C = 0.7, -0.4, 0.9
    s.set_velocity(A * np.sin(z) + C * np.cos(y), B * np.sin(x) + A * np.cos(z), C * np.sin(y) + B * np.cos(x))
    R = s.dissipation(nu)
    assert R[17] == pytest.approx(s1 * (A * A + B * B + C * C) * V, rel=1e-13)
    s.set_velocity(np.sin(y), zero, zero)
    R = s.dissipation(nu)
    assert R[16] == pytest.approx(nu * V * (-2 * np.sin(h / 2) ** 2 / h ** 2 + s1 ** 2 / 2), rel=1e-11)
    assert R[18] == pytest.approx(V / 4, rel=1e-13)
    s.set_velocity(np.cos(x), zero, zero)
    s.set("p", np.sin(x))
    R = s.dissipation(nu)
    assert R[15] == pytest.approx(-s1 * V / 2, rel=1e-13)
    with oracle.planted("diss_shear_weight"):
        s.set_velocity(np.sin(y), zero, zero)
        assert abs(s.dissipation(nu)[16] / (nu * V * (-2 * np.sin(h / 2) ** 2 / h ** 2 + s1 ** 2 / 2)) - 1) > 1


def _blob(x, y, z, c, sig):
    return np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * sig ** 2))


def test_dissipation_impulse_identities_on_compact_fields(oracle):
    """For a velocity that vanishes near the box faces, summation by parts with central
    differences (exact on the linear and quadratic position factors) gives the impulse
    identities exactly:
    - linear impulse 1/2 sum r x omega = linear momentum sum u (any compact u);
    - for u = curl_c A with compact A: angular momentum sum r x u = 2 sum A, and the angular
      impulse 1/3 sum r x (r x omega) equals it.
    Gaussian blobs of width 0.05-0.07 in the unit box (their tails are 1e-15 at the faces):
    agreement to 1e-8."""
    s = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    h, h3 = s.h, s.h ** 3
    x, y, z = centres(s)
    zero = np.zeros(s.shape)
    s.set("chi", zero)
    s.set("p", zero)
    s.set_velocity(0.7 * _blob(x, y, z, (0.45, 0.5, 0.55), 0.06), -0.4 * _blob(x, y, z, (0.55, 0.45, 0.5), 0.07),
                   0.9 * _blob(x, y, z, (0.5, 0.56, 0.47), 0.05))
    R = s.dissipation(0.01)
    assert np.abs(R[3:6] - R[6:9]).max() < 1e-7 * np.abs(R[6:9]).max()
    with oracle.planted("diss_impulse_sign"):
        Rp = s.dissipation(0.01)
    assert np.abs(Rp[3:6] - Rp[6:9]).max() > 1e-3 * np.abs(R[6:9]).max()
    Ax = 0.5 * _blob(x, y, z, (0.48, 0.52, 0.5), 0.06)
    Ay = -0.3 * _blob(x, y, z, (0.53, 0.47, 0.52), 0.05)
    Az = 0.8 * _blob(x, y, z, (0.5, 0.5, 0.45), 0.07)
    s.set_velocity(dc(Az, 1, h) - dc(Ay, 0, h), dc(Ax, 0, h) - dc(Az, 2, h), dc(Ay, 2, h) - dc(Ax, 1, h))
    R = s.dissipation(0.01)
    A2 = 2 * h3 * np.array([Ax.sum(), Ay.sum(), Az.sum()])
    assert np.abs(R[12:15] - A2).max() < 1e-8 * np.abs(A2).max()
    assert np.abs(R[9:12] - A2).max() < 1e-7 * np.abs(A2).max()
    assert np.abs(R[6:9]).max() < 1e-12 * np.abs(A2).max()
    with oracle.planted("diss_angimp_term"):
        Rp = s.dissipation(0.01)
    assert np.abs(Rp[9:12] - A2).max() > 1e-3 * np.abs(A2).max()
