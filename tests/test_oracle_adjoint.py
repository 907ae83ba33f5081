"""Pins of the oracle's adjoint step (no reference counterpart): the step
F(u) = P(A(u)) is quadratic in u, so the central difference (F(u + e d) - F(u - e d))/(2e)
is its exact directional derivative F'(u) d up to rounding; the adjoint must satisfy
<F'(u) d, lam> = <d, F'(u)^T lam> for random u, d, lam. The projection alone must be
symmetric and idempotent. CPU only."""
import numpy as np
import pytest


def fwd_step(oracle, bpd, u, dt, nu, U):
    s = oracle.OracleSim(bpd)
    s.set_velocity(*u)
    s.advect_diffuse(dt, nu, U)
    s.pressure_rhs(dt)
    s.poisson_solve()
    s.grad_p(dt)
    return s.velocity()


def dot(a, b):
    return sum(float(np.vdot(x, y)) for x, y in zip(a, b))


@pytest.mark.parametrize("bpd,U", [((2, 2, 2), (0.0, 0.0, 0.0)), ((4, 2, 1), (0.3, -0.2, 0.1))])
def test_adjoint_dot_product_identity(oracle, bpd, U):
    rng = np.random.default_rng(23)
    s = oracle.OracleSim(bpd)
    shape = s.shape
    u = tuple(rng.standard_normal(shape) for _ in range(3))
    d = tuple(rng.standard_normal(shape) for _ in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, eps = 0.3 * s.h, 0.02, 1e-3
    fp = fwd_step(oracle, bpd, tuple(a + eps * b for a, b in zip(u, d)), dt, nu, U)
    fm = fwd_step(oracle, bpd, tuple(a - eps * b for a, b in zip(u, d)), dt, nu, U)
    jd = tuple((a - b) / (2 * eps) for a, b in zip(fp, fm))
    s.set_velocity(*u)
    lam0 = s.adjoint_step(dt, nu, lam, U)
    lhs, rhs = dot(jd, lam), dot(d, lam0)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    # the forward state is left as it was
    for a, b in zip(s.velocity(), u):
        assert np.array_equal(a, b)


def fwd_step_scheme(oracle, bpd, u, dt, nu, U, scheme):
    s = oracle.OracleSim(bpd)
    s.set_velocity(*u)
    s.advect_diffuse(dt, nu, U, scheme)
    s.pressure_rhs(dt)
    s.poisson_solve()
    s.grad_p(dt)
    return s.velocity()


@pytest.mark.parametrize("scheme", [1, 2])
def test_upwind_adjoint_dot_product_identity(oracle, scheme):
    """The upwind step is polynomial in u (degree 2 for Euler, 4 for the midpoint form) away
    from the cells where an advecting velocity changes sign, so the five-point central
    difference with a small step is its exact directional derivative; the adjoint of
    upwind Euler and midpoint (scheme 1, 2) must satisfy the dot-product identity."""
    rng = np.random.default_rng(31)
    bpd, U = (2, 2, 2), (0.3, -0.2, 0.1)
    s = oracle.OracleSim(bpd)
    shape = s.shape
    u = tuple(rng.standard_normal(shape) for _ in range(3))
    d = tuple(rng.standard_normal(shape) for _ in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, eps = 0.3 * s.h, 0.02, 1e-5

    def f(e):
        return fwd_step_scheme(oracle, bpd, tuple(a + e * b for a, b in zip(u, d)), dt, nu, U, scheme)
    f1, f2, m1, m2 = f(eps), f(2 * eps), f(-eps), f(-2 * eps)
    jd = tuple((8 * (a - b) - (c - e)) / (12 * eps) for a, b, c, e in zip(f1, m1, f2, m2))
    s.set_velocity(*u)
    lam0 = s.adjoint_step(dt, nu, lam, U, scheme=scheme)
    lhs, rhs = dot(jd, lam), dot(d, lam0)
    assert abs(lhs - rhs) < 1e-9 * abs(lhs), (lhs, rhs)
    # not the central adjoint: the upwind scheme's transpose differs
    lam_c = s.adjoint_step(dt, nu, lam, U, scheme=0)
    assert abs(dot(d, lam_c) - rhs) > 1e-4 * abs(rhs)


def test_adjoint_of_pure_projection_is_symmetric(oracle):
    """With nu = 0, dt -> tiny advection, the adjoint reduces to P lam + O(dt); check P
    itself: <P a, b> = <a, P b> and P P = P through the oracle's own projection."""
    rng = np.random.default_rng(4)
    bpd = (2, 2, 2)

    def proj(v):
        s = oracle.OracleSim(bpd)
        s.set_velocity(*v)
        s.pressure_rhs(1e-3); s.poisson_solve(); s.grad_p(1e-3)
        return s.velocity()
    shape = oracle.OracleSim(bpd).shape
    a = tuple(rng.standard_normal(shape) for _ in range(3))
    b = tuple(rng.standard_normal(shape) for _ in range(3))
    pa, pb = proj(a), proj(b)
    assert abs(dot(pa, b) - dot(a, pb)) < 1e-12 * abs(dot(pa, b))
    ppa = proj(pa)
    assert max(np.max(np.abs(x - y)) for x, y in zip(ppa, pa)) < 1e-12 * max(np.max(np.abs(x)) for x in pa)


def fwd_step_body(oracle, bpd, u, dt, nu, U, V, lam_, ext, centre=(0.9, 0.5, 0.5), mollify=0.0, W=(0, 0, 0),
                  scheme=0, target=0.0, rate=0.0, radius=0.3):
    s = oracle.OracleSim(bpd, maxextent=ext)
    s.sphere_create(radius, centre, mollify=mollify)
    s.set_velocity(*u)
    out = s.timestep(dt, nu, uinf=U, sphere=dict(lambda_=lam_, implicit=True, trans_vel=V, ang_vel=W),
                     scheme=scheme, spectral_forcing=target > 0 or rate > 0, tke_target=target, injection_rate=rate)
    return s.velocity(), out["penal_ft"][0]


@pytest.mark.parametrize("lam_", [1e4, 1e6])
def test_adjoint_with_sphere_dot_product_identity(oracle, lam_):
    """Step with the penalized sphere, J = <lam, u_{n+1}> + w F_x: the adjoint's dJ/du_n and
    dJ/dV against central differences of the forward oracle step (the masking's correction
    makes the step rational in u: eps = 1e-5, agreement 1e-7)."""
    rng = np.random.default_rng(41)
    bpd, ext = (4, 2, 2), 2.0
    s = oracle.OracleSim(bpd, maxextent=ext)
    s.sphere_create(0.3, (0.9, 0.5, 0.5))
    shape = s.shape
    U, V = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0)
    u = tuple(0.3 * rng.standard_normal(shape) + (1.0 if q == 0 else 0.0) for q in range(3))
    d = tuple(rng.standard_normal(shape) for _ in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w, eps = 0.2 * s.h, 1e-3, 3.7, 1e-5

    def J(uu, VV):
        u1, F = fwd_step_body(oracle, bpd, uu, dt, nu, U, VV, lam_, ext)
        return dot(lam, u1) + w * F
    s.set_velocity(*u)
    lam0, dJdV = s.adjoint_step_body(dt, nu, lam, V, lam_, True, w, U)
    fd = (J(tuple(a + eps * b for a, b in zip(u, d)), V) - J(tuple(a - eps * b for a, b in zip(u, d)), V)) / (2 * eps)
    ad = dot(d, lam0)
    assert abs(fd - ad) < 1e-7 * abs(fd)
    for k in range(3):
        e = np.eye(3)[k] * eps
        fdv = (J(u, tuple(np.add(V, e))) - J(u, tuple(np.subtract(V, e)))) / (2 * eps)
        assert abs(fdv - dJdV[k]) < 1e-7 * max(abs(fdv), 1e-3)


@pytest.mark.parametrize("mollify", [1.0, 2.0])
def test_adjoint_sphere_centre_dot_product_identity(oracle, mollify):
    """The mollified chi is C1 in the sphere centre: dJ/dc of one step (J = <lam, u_{n+1}> +
    w F_x, chi entering the penalization, its force and the RHS masking) against central
    differences of the forward oracle step in the centre (eps = 1e-6 h-scale offsets, far
    below the spacing at which a cell's masking flags change), and dJ/du_n, dJ/dV as with
    Towers' chi."""
    rng = np.random.default_rng(43)
    bpd, ext = (4, 2, 2), 2.0
    c0 = np.array([0.913, 0.487, 0.502])
    s = oracle.OracleSim(bpd, maxextent=ext)
    s.sphere_create(0.3, c0, mollify=mollify)
    shape = s.shape
    U, V = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0)
    u = tuple(0.3 * rng.standard_normal(shape) + (1.0 if q == 0 else 0.0) for q in range(3))
    d = tuple(rng.standard_normal(shape) for _ in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w, lam_ = 0.2 * s.h, 1e-3, 3.7, 1e6

    def J(uu, VV, cc):
        u1, F = fwd_step_body(oracle, bpd, uu, dt, nu, U, VV, lam_, ext, centre=tuple(cc), mollify=mollify)
        return dot(lam, u1) + w * F
    s.set_velocity(*u)
    lam0, dJdV, dJdc = s.adjoint_step_body(dt, nu, lam, V, lam_, True, w, U, centre=True)
    eps = 1e-5
    fd = (J(tuple(a + eps * b for a, b in zip(u, d)), V, c0) - J(tuple(a - eps * b for a, b in zip(u, d)), V, c0)) / (2 * eps)
    assert abs(fd - dot(d, lam0)) < 1e-7 * abs(fd)
    for k in range(3):
        e = np.eye(3)[k] * eps
        fdv = (J(u, tuple(np.add(V, e)), c0) - J(u, tuple(np.subtract(V, e)), c0)) / (2 * eps)
        assert abs(fdv - dJdV[k]) < 1e-7 * max(abs(fdv), 1e-3)
    ec = 1e-6 * s.h
    for k in range(3):
        e = np.eye(3)[k] * ec
        fdc = (J(u, V, c0 + e) - J(u, V, c0 - e)) / (2 * ec)
        print(f"mollify {mollify} dJ/dc_{k}: fd {fdc:.10e} adjoint {dJdc[k]:.10e}")
        assert abs(fdc - dJdc[k]) < 1e-6 * max(abs(fdc), 1.0)


@pytest.mark.parametrize("lam_", [1e4, 1e6])
def test_adjoint_rotating_sphere_dot_product_identity(oracle, lam_):
    """The sphere rotating at W about its centre of mass (target V + W x (x - CoM)): the
    adjoint's dJ/du_n, dJ/dV and dJ/dW against central differences of the forward oracle step
    (J = <lam, u_{n+1}> + w F_x, eps = 1e-5, agreement 1e-7). A lever-arm sign or a swapped
    cross-product component fails the dJ/dW comparison."""
    rng = np.random.default_rng(47)
    bpd, ext = (4, 2, 2), 2.0
    s = oracle.OracleSim(bpd, maxextent=ext)
    s.sphere_create(0.3, (0.9, 0.5, 0.5))
    shape = s.shape
    U, V, W = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0), (0.7, -1.3, 2.1)
    u = tuple(0.3 * rng.standard_normal(shape) + (1.0 if q == 0 else 0.0) for q in range(3))
    d = tuple(rng.standard_normal(shape) for _ in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w, eps = 0.2 * s.h, 1e-3, 3.7, 1e-5

    def J(uu, VV, WW):
        u1, F = fwd_step_body(oracle, bpd, uu, dt, nu, U, VV, lam_, ext, W=WW)
        return dot(lam, u1) + w * F
    s.set_velocity(*u)
    lam0, dJdV, dJdW = s.adjoint_step_body(dt, nu, lam, V, lam_, True, w, U, ang_vel=W)
    fd = (J(tuple(a + eps * b for a, b in zip(u, d)), V, W) - J(tuple(a - eps * b for a, b in zip(u, d)), V, W)) / (2 * eps)
    assert abs(fd - dot(d, lam0)) < 1e-7 * abs(fd)
    for k in range(3):
        e = np.eye(3)[k] * eps
        fdv = (J(u, tuple(np.add(V, e)), W) - J(u, tuple(np.subtract(V, e)), W)) / (2 * eps)
        assert abs(fdv - dJdV[k]) < 1e-7 * max(abs(fdv), 1e-3)
        fdw = (J(u, V, tuple(np.add(W, e))) - J(u, V, tuple(np.subtract(W, e)))) / (2 * eps)
        assert abs(fdw - dJdW[k]) < 1e-7 * max(abs(fdw), 1e-3)
    # the lever arms are not all zero: W's gradient is not a multiple of V's
    assert np.max(np.abs(dJdW)) > 1e-3 * np.max(np.abs(dJdV))


@pytest.mark.parametrize("scheme,target,rate", [(1, 0.0, 0.0), (2, 0.0, 0.0), (0, 0.0, 0.4), (2, 1.5, 0.0)])
def test_adjoint_body_scheme_forcing_dot_product_identity(oracle, scheme, target, rate):
    """The step with the sphere, the upwind schemes (Euler, midpoint) and the linear forcing
    (constant injection rate, or the TKE target): dJ/du_n, dJ/dV and dJ/dW against central
    differences of the forward oracle step (J = <lam, u_{n+1}> + w F_x, eps = 1e-5, agreement
    1e-7). The forcing's factor must be positive at u_n (it switches the forcing on) for the
    forcing's transpose to be exercised, which the test checks. The velocity components have
    means (1, 0.8, -0.9) so that no advecting velocity comes within the differences' step of
    zero, where the upwind stencil switches (a kink: with zero means the same check misses by
    1e-5 at eps = 1e-5 and agrees to 1e-10 at eps = 1e-6). The box is 2pi x pi x pi, where the
    forcing's band (0 < |k|^2 <= 4) holds modes."""
    rng = np.random.default_rng(53 + scheme)
    bpd, ext = (4, 2, 2), 2 * np.pi
    centre, R = tuple(np.pi * np.array([0.9, 0.5, 0.5])), 0.3 * np.pi
    s = oracle.OracleSim(bpd, maxextent=ext)
    s.sphere_create(R, centre)
    shape = s.shape
    U, V, W = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0), (0.3, -0.4, 0.8)
    u = tuple(0.3 * rng.standard_normal(shape) + m for m in (1.0, 0.8, -0.9))
    d = tuple(rng.standard_normal(shape) for _ in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w, lam_, eps = 0.2 * s.h, 1e-3, 3.7, 1e5, 1e-5
    if target > 0 or rate > 0:
        f = oracle.OracleSim(bpd, maxextent=ext)
        f.set_velocity(*u)
        f.advect_diffuse(dt, nu, U, scheme)
        assert f.spectral_forcing(dt, nu, tke_target=target, injection_rate=rate)["fac"] > 0

    def J(uu, VV, WW):
        u1, F = fwd_step_body(oracle, bpd, uu, dt, nu, U, VV, lam_, ext, centre=centre, radius=R, W=WW,
                              scheme=scheme, target=target, rate=rate)
        return dot(lam, u1) + w * F
    s.set_velocity(*u)
    lam0, dJdV, dJdW = s.adjoint_step_body(dt, nu, lam, V, lam_, True, w, U, ang_vel=W, scheme=scheme,
                                           tke_target=target, injection_rate=rate)
    fd = (J(tuple(a + eps * b for a, b in zip(u, d)), V, W) - J(tuple(a - eps * b for a, b in zip(u, d)), V, W)) / (2 * eps)
    assert abs(fd - dot(d, lam0)) < 1e-7 * abs(fd)
    for k in range(3):
        e = np.eye(3)[k] * eps
        fdv = (J(u, tuple(np.add(V, e)), W) - J(u, tuple(np.subtract(V, e)), W)) / (2 * eps)
        assert abs(fdv - dJdV[k]) < 1e-7 * max(abs(fdv), 1e-3)
        fdw = (J(u, V, tuple(np.add(W, e))) - J(u, V, tuple(np.subtract(W, e)))) / (2 * eps)
        assert abs(fdw - dJdW[k]) < 1e-7 * max(abs(fdw), 1e-3)


TWO_BODIES = [(0.2 * np.pi, (0.5 * np.pi, 0.5 * np.pi, 0.5 * np.pi)),
              (0.2 * np.pi, (1.5 * np.pi, 0.45 * np.pi, 0.55 * np.pi))]


def fwd_step_bodies(oracle, bpd, u, dt, nu, U, Vs, Ws, lam_, ext, scheme=0, target=0.0, rate=0.0):
    s = oracle.OracleSim(bpd, maxextent=ext)
    for R, c in TWO_BODIES:
        s.add_sphere(R, c)
    s.set_velocity(*u)
    s.advect_diffuse(dt, nu, U, scheme)
    if target > 0 or rate > 0:
        s.spectral_forcing(dt, nu, tke_target=target, injection_rate=rate)
    s.pressure_rhs(dt)
    s.poisson()
    s.grad_p(dt)
    F = []
    for k in range(len(TWO_BODIES)):
        with s.body(k):
            F.append(s.penalize(dt, lam_, True, Vs[k], Ws[k])[0])
    return s.velocity(), F


@pytest.mark.parametrize("scheme,rate", [(0, 0.0), (2, 0.4)])
def test_adjoint_two_bodies_dot_product_identity(oracle, scheme, rate):
    """Two spheres with their own velocities and rotations, J = <lam, u_{n+1}> + sum_k w_k F_x^k:
    dJ/du_n and each body's dJ/dV, dJ/dW against central differences of the forward oracle
    step (the reference order: every body's masking in the RHS, then each body penalized;
    eps = 1e-5, agreement 1e-7). Each body's masking transpose acts on its own blocks, so a sum
    over every body's flagged cells in one body's correction fails it."""
    rng = np.random.default_rng(59 + scheme)
    bpd, ext = (4, 2, 2), 2 * np.pi
    s = oracle.OracleSim(bpd, maxextent=ext)
    for R, c in TWO_BODIES:
        s.add_sphere(R, c)
    assert s.body_count() == 2
    shape = s.shape
    U = (1.0, 0.0, 0.0)
    Vs, Ws, w = [(-1.0, 0.1, 0.0), (-0.8, 0.0, 0.2)], [(0.3, -0.4, 0.8), (-0.5, 0.2, 0.6)], [3.7, -1.9]
    u = tuple(0.3 * rng.standard_normal(shape) + m for m in (1.0, 0.8, -0.9))
    d = tuple(rng.standard_normal(shape) for _ in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, lam_, eps = 0.2 * s.h, 1e-3, 1e5, 1e-5

    def J(uu, VV, WW):
        u1, F = fwd_step_bodies(oracle, bpd, uu, dt, nu, U, VV, WW, lam_, ext, scheme=scheme, rate=rate)
        return dot(lam, u1) + sum(wk * Fk for wk, Fk in zip(w, F))
    s.set_velocity(*u)
    lam0, dV, dW = s.adjoint_step_bodies(dt, nu, lam, Vs, Ws, lam_, True, w, U, scheme=scheme, injection_rate=rate)
    fd = (J(tuple(a + eps * b for a, b in zip(u, d)), Vs, Ws) - J(tuple(a - eps * b for a, b in zip(u, d)), Vs, Ws)) / (2 * eps)
    assert abs(fd - dot(d, lam0)) < 1e-7 * abs(fd)
    for k in range(2):
        for j in range(3):
            e = np.eye(3)[j] * eps
            Vp, Vm = list(Vs), list(Vs)
            Vp[k], Vm[k] = tuple(np.add(Vs[k], e)), tuple(np.subtract(Vs[k], e))
            fdv = (J(u, Vp, Ws) - J(u, Vm, Ws)) / (2 * eps)
            assert abs(fdv - dV[k][j]) < 1e-7 * max(abs(fdv), 1e-3), (k, j, fdv, dV[k][j])
            Wp, Wm = list(Ws), list(Ws)
            Wp[k], Wm[k] = tuple(np.add(Ws[k], e)), tuple(np.subtract(Ws[k], e))
            fdw = (J(u, Vs, Wp) - J(u, Vs, Wm)) / (2 * eps)
            assert abs(fdw - dW[k][j]) < 1e-7 * max(abs(fdw), 1e-3), (k, j, fdw, dW[k][j])


def test_adjoint_rotation_zero_matches_the_translating_body(oracle):
    """W = 0 through the rotating entry point: lam_n and dJ/dV as the translating body's."""
    rng = np.random.default_rng(5)
    s = oracle.OracleSim((2, 2, 2), maxextent=1.0)
    s.sphere_create(0.2, (0.5, 0.5, 0.5))
    u = tuple(0.3 * rng.standard_normal(s.shape) for _ in range(3))
    lam = tuple(rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(*u)
    a0, av = s.adjoint_step_body(1e-3, 1e-3, lam, (0.2, 0, 0), 1e5, True, 1.5)
    s.set_velocity(*u)
    b0, bv, bw = s.adjoint_step_body(1e-3, 1e-3, lam, (0.2, 0, 0), 1e5, True, 1.5, ang_vel=(0, 0, 0))
    assert all(np.array_equal(x, y) for x, y in zip(a0, b0)) and np.array_equal(av, bv)
    assert np.all(np.isfinite(bw))


def test_adjoint_centre_is_nan_with_rotation(oracle):
    s = oracle.OracleSim((2, 2, 2), maxextent=1.0)
    s.sphere_create(0.2, (0.5, 0.5, 0.5), mollify=1.0)
    s.set_velocity(*(np.zeros(s.shape) for _ in range(3)))
    lam = tuple(np.zeros(s.shape) for _ in range(3))
    _, _, _, dc = s.adjoint_step_body(1e-3, 1e-3, lam, (0, 0, 0), 1e6, True, 1.0, centre=True, ang_vel=(0, 0, 1))
    assert np.isnan(dc).all()


def test_adjoint_centre_is_nan_with_towers_chi(oracle):
    s = oracle.OracleSim((2, 2, 2), maxextent=1.0)
    s.sphere_create(0.2, (0.5, 0.5, 0.5))
    s.set_velocity(*(np.zeros(s.shape) for _ in range(3)))
    lam = tuple(np.zeros(s.shape) for _ in range(3))
    _, _, dc = s.adjoint_step_body(1e-3, 1e-3, lam, (0, 0, 0), 1e6, True, 1.0, centre=True)
    assert np.isnan(dc).all()


def fwd_step_forced(oracle, bpd, u, dt, nu, U, scheme, target, rate):
    s = oracle.OracleSim(bpd)
    s.set_velocity(*u)
    s.advect_diffuse(dt, nu, U, scheme)
    st = s.spectral_forcing(dt, nu, tke_target=target, injection_rate=rate)
    s.pressure_rhs(dt)
    s.poisson_solve()
    s.grad_p(dt)
    return s.velocity(), st


def low_mode_field(shape, h, rng):
    """A velocity dominated by modes |k| <= 2 (so the forcing's band holds most of the energy)
    plus noise on every mode."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*(h * (np.arange(m) + 0.5) for m in shape), indexing="ij")
    u = np.sin(x) * np.cos(y) * np.cos(2 * z) + 0.3 * rng.standard_normal(shape)
    v = -np.cos(x) * np.sin(y) * np.cos(z) + 0.3 * rng.standard_normal(shape)
    w = 0.5 * np.sin(2 * y + z) + 0.3 * rng.standard_normal(shape)
    return u, v, w


@pytest.mark.parametrize("scheme,bpd,target,rate", [(0, (2, 2, 2), 1.1, 0.0), (1, (2, 2, 2), 0.0, 0.7),
                                                    (2, (4, 2, 2), 1.1, 0.0), (0, (4, 2, 2), 0.0, 0.7)])
def test_forcing_adjoint_dot_product_identity(oracle, scheme, bpd, target, rate):
    """The step with the linear forcing (to a target energy, or at a constant injection rate)
    is rational in u: the five-point central difference with a small step is its directional
    derivative to about 1e-12, and the adjoint with the forcing's transpose (the band filter,
    symmetric, and the derivative of fac through tke and tke_F) must satisfy the dot-product
    identity. Dropping either term of grad fac, or the factor fac on P lam, fails it."""
    rng = np.random.default_rng(41)
    s = oracle.OracleSim(bpd)
    u = low_mode_field(s.shape, s.h, rng)
    tke0 = 0.5 * sum(float(np.mean(a * a)) for a in u)
    target = target * tke0 if target else 0.0
    d = low_mode_field(s.shape, s.h, rng)  # a direction and a seed with energy in the band
    lam = tuple(a + 0.5 * b for a, b in zip(low_mode_field(s.shape, s.h, rng), u))
    # the upwind step's kinks (sign changes of the advecting velocity) stay out of a small step
    dt, nu, eps, U = 0.3 * s.h, 0.02, 1e-4 if scheme == 0 else 1e-5, (0.3, -0.2, 0.1)

    def f(e):
        return fwd_step_forced(oracle, bpd, tuple(a + e * b for a, b in zip(u, d)), dt, nu, U, scheme, target,
                               rate)[0]
    _, st = fwd_step_forced(oracle, bpd, u, dt, nu, U, scheme, target, rate)
    assert 1e-3 < st["fac"] < 5  # the forcing acts
    f1, f2, m1, m2 = f(eps), f(2 * eps), f(-eps), f(-2 * eps)
    jd = tuple((8 * (a - b) - (c - e)) / (12 * eps) for a, b, c, e in zip(f1, m1, f2, m2))
    s.set_velocity(*u)
    lam0 = s.adjoint_step(dt, nu, lam, U, scheme=scheme, tke_target=target, injection_rate=rate)
    lhs, rhs = dot(jd, lam), dot(d, lam0)
    tol = 1e-9 if scheme == 0 else 1e-7  # the upwind step: a.e. smooth, as above
    assert abs(lhs - rhs) < tol * abs(lhs), (lhs, rhs)
    for a, b in zip(s.velocity(), u):  # the forward state is left as it was
        assert np.array_equal(a, b)
    plain = s.adjoint_step(dt, nu, lam, U, scheme=scheme)
    assert abs(dot(d, plain) - rhs) > 1e-3 * abs(rhs)  # the forcing's transpose is not negligible


def test_forcing_adjoint_is_the_plain_one_when_the_forcing_is_off(oracle):
    """Above the target energy fac <= 0 and the forward forcing does nothing: so does its adjoint."""
    rng = np.random.default_rng(43)
    s = oracle.OracleSim((2, 2, 2))
    u = low_mode_field(s.shape, s.h, rng)
    lam = tuple(rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(*u)
    a = s.adjoint_step(0.01, 0.02, lam, tke_target=1e-3)
    b = s.adjoint_step(0.01, 0.02, lam)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
