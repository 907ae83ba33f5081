"""The free-body part of the sphere in the oracle (UpdateObstacles: KernelIntegrateFluidMomenta,
ObstaclesUpdate.cpp:15-170, and Obstacle::computeVelocities, Obstacle.cpp:153-212), pinned by
rigid motion: a velocity field u = V + Omega x (x - CoM) gives back V and Omega from the momentum
solve, with explicit and implicit penalization weights; forced translations keep their
imposed value and blocked rotations give zero."""
import numpy as np
import pytest


def rigid_field(shape, h, V, Om, com):
    Z, Y, X = shape
    z, y, x = np.meshgrid((np.arange(Z) + 0.5) * h, (np.arange(Y) + 0.5) * h, (np.arange(X) + 0.5) * h,
                          indexing="ij")
    p = (x - com[0], y - com[1], z - com[2])
    u = V[0] + Om[1] * p[2] - Om[2] * p[1]
    v = V[1] + Om[2] * p[0] - Om[0] * p[2]
    w = V[2] + Om[0] * p[1] - Om[1] * p[0]
    return u, v, w


def sphere_with_rigid_motion(oracle, V, Om):
    O = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    O.sphere_create(0.2, (0.47, 0.52, 0.55))
    info = O.sphere_info()
    O.set_velocity(*rigid_field((32, 32, 32), 1.0 / 32, V, Om, info["com"]))
    return O, info


@pytest.mark.parametrize("implicit", [False, True])
def test_rigid_motion_is_recovered(oracle, implicit):
    V, Om = np.array([0.3, -0.2, 0.1]), np.array([1.5, -0.7, 2.0])
    O, info = sphere_with_rigid_motion(oracle, V, Om)
    M = O.obstacle_momenta(1e-3, 1e6, implicit)
    assert M[0] == pytest.approx(info["mass"], rel=1e-13)
    assert M[7] > 0 and M[8] > 0 and M[9] > 0
    if implicit:
        # penalFac = dv l dt chi / (1 + chi l dt): near dv on every cell with chi > 0 when l dt >> 1
        assert M[0] < M[13] < 2 * M[0]
    else:
        assert not M[13:].any()
    Vs, Os = oracle.compute_velocities(M, implicit, [False] * 3, [0.0] * 3, [False] * 3)
    np.testing.assert_allclose(Vs, V, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Os, Om, rtol=0, atol=1e-10)


def test_forced_and_blocked_degrees_of_freedom(oracle):
    V, Om = np.array([0.3, -0.2, 0.1]), np.array([1.5, -0.7, 2.0])
    O, _ = sphere_with_rigid_motion(oracle, V, Om)
    M = O.obstacle_momenta(1e-3, 1e6, False)
    Vs, Os = oracle.compute_velocities(M, False, [True, False, False], [0.9, 0, 0], [False] * 3)
    np.testing.assert_allclose(Vs, [0.9, -0.2, 0.1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(Os, Om, rtol=0, atol=1e-10)
    Vs, Os = oracle.compute_velocities(M, False, [False] * 3, [0.0] * 3, [True] * 3)
    np.testing.assert_allclose(Vs, V, rtol=0, atol=1e-12)
    assert not Os.any()


def test_moving_one_of_two_bodies_equals_building_them_there(oracle):
    """orc_move_body rebuilds every body's chi and SDF: the pool after a move equals a pool
    built with the body at its new centre."""
    O = oracle.OracleSim((8, 4, 4), maxextent=2.0)
    O.add_sphere(0.15, (0.55, 0.5, 0.48))
    O.add_sphere(0.12, (1.45, 0.52, 0.5), hemi=True)
    chi0 = O.get("chi").copy()
    O.move_body(1, (1.47, 0.5, 0.5))
    P = oracle.OracleSim((8, 4, 4), maxextent=2.0)
    P.add_sphere(0.15, (0.55, 0.5, 0.48))
    P.add_sphere(0.12, (1.47, 0.5, 0.5), hemi=True)
    assert np.array_equal(O.get("chi"), P.get("chi")) and not np.array_equal(O.get("chi"), chi0)
    for k in range(2):
        with O.body(k):
            a = O.sphere_info()
        with P.body(k):
            b = P.sphere_info()
        assert np.array_equal(a["com"], b["com"]) and a["mass"] == b["mass"] and a["npoints"] == b["npoints"]
    with pytest.raises(IndexError):
        O.move_body(2, (1.0, 0.5, 0.5))


def test_two_bodies_superpose_single_bodies(oracle):
    """Bodies whose obstacle blocks are apart act each on its own blocks: the pool's chi is the
    sum of the single-body chis, and since the pressure RHS is linear in its per-body
    corrections and the solve is linear, p(A, B) = p(A) + p(B) - p(no body); each body's
    penalization force in the pool equals the single-body one."""
    bpd, ext = (8, 4, 4), 2.0
    A, B = ((0.55, 0.5, 0.48), 0.15), ((1.45, 0.52, 0.5), 0.12)
    rng = np.random.default_rng(11)
    u = tuple(0.3 * rng.standard_normal((32, 32, 64)) for _ in range(3))
    dt = 1e-3

    def solved(sim):
        sim.set_velocity(*u)
        sim.pressure_rhs(dt)
        sim.poisson_solve()
        return sim.get("p").copy()

    P = oracle.OracleSim(bpd, maxextent=ext)
    P.add_sphere(A[1], A[0])
    P.add_sphere(B[1], B[0], hemi=True)
    SA = oracle.OracleSim(bpd, maxextent=ext)
    SA.sphere_create(A[1], A[0])
    SB = oracle.OracleSim(bpd, maxextent=ext)
    SB.sphere_create(B[1], B[0], hemi=True)
    N = oracle.OracleSim(bpd, maxextent=ext)
    assert np.array_equal(P.get("chi"), SA.get("chi") + SB.get("chi"))
    pp, pa, pb, p0 = solved(P), solved(SA), solved(SB), solved(N)
    assert np.abs(pp - (pa + pb - p0)).max() < 1e-12 * np.abs(pp).max()
    assert np.abs(pa - p0).max() > 1e-3 * np.abs(p0).max()  # the corrections are not negligible
    for k, S in ((0, SA), (1, SB)):
        for sim in (P, S):
            sim.set_velocity(*u)
        with P.body(k):
            fp = P.penalize(dt, 1e6, True, (-1.0, 0.2, 0.0))
        fs = S.penalize(dt, 1e6, True, (-1.0, 0.2, 0.0))
        assert np.array_equal(fp, fs)


def _rigid_udef(O, V, W):
    """V + W x (x - CoM) on every cell, and the cell positions relative to the centre of mass."""
    with_info = O.sphere_info()
    nz, ny, nx = O.shape
    h = O.h
    z, y, x = np.meshgrid(*(h * (np.arange(m) + 0.5) for m in (nz, ny, nx)), indexing="ij")
    p = (x - with_info["com"][0], y - with_info["com"][1], z - with_info["com"][2])
    u = V[0] + W[1] * p[2] - W[2] * p[1]
    v = V[1] + W[2] * p[0] - W[0] * p[2]
    w = V[2] + W[0] * p[1] - W[1] * p[0]
    return (u, v, w), p


def test_udef_rigid_motion_is_removed_entirely(oracle):
    """A deformation velocity that is a rigid motion V + W x (x - CoM) is all momentum:
    F/V gives V (the centre of mass is chi's), J^-1 T gives W (T = J W about the centre of
    mass), and the removal leaves udef = 0 on every cell of the obstacle blocks. A second
    removal finds nothing (its T uses udef minus the last linear correction, whose moment
    vanishes about the centre of mass). A wrong sign in a cross product, a misplaced J entry
    or a transposed inverse fails it."""
    O = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    O.sphere_create(0.2, (0.47, 0.52, 0.55))
    V, W = np.array([0.3, -0.7, 0.2]), np.array([1.1, 0.4, -0.9])
    udef, _ = _rigid_udef(O, V, W)
    O.set_udef(*udef)
    r = O.remove_udef_momenta()
    np.testing.assert_allclose(r["trans"], V, rtol=1e-12)
    np.testing.assert_allclose(r["ang"], W, rtol=1e-9)
    assert r["mass"] == pytest.approx(O.sphere_info()["mass"], rel=1e-14)
    after = O.get_udef()
    assert max(np.abs(a).max() for a in after) < 1e-9
    r2 = O.remove_udef_momenta()
    assert np.abs(r2["trans"]).max() < 1e-9 and np.abs(r2["ang"]).max() < 1e-8


def test_udef_removal_leaves_no_momentum(oracle):
    """For any udef the removed field has zero linear and angular momentum about the centre of
    mass (sum chi udef dv and sum chi p x udef dv), the defining property of the operator."""
    rng = np.random.default_rng(17)
    O = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    O.sphere_create(0.2, (0.47, 0.52, 0.55))
    udef = tuple(rng.standard_normal(O.shape) for _ in range(3))
    O.set_udef(*udef)
    r = O.remove_udef_momenta()
    assert np.abs(r["trans"]).max() > 1e-3 and np.abs(r["ang"]).max() > 1e-3
    after = O.get_udef()
    chi = O.get("chi")
    _, p = _rigid_udef(O, np.zeros(3), np.zeros(3))
    dv = O.h ** 3
    lin = [float((chi * a).sum() * dv) for a in after]
    ang = [float((chi * (p[1] * after[2] - p[2] * after[1])).sum() * dv),
           float((chi * (p[2] * after[0] - p[0] * after[2])).sum() * dv),
           float((chi * (p[0] * after[1] - p[1] * after[0])).sum() * dv)]
    scale = float(sum((chi * np.abs(a)).sum() for a in udef) * dv)
    assert max(abs(x) for x in lin) < 1e-13 * scale and max(abs(x) for x in ang) < 1e-13 * scale
    # cells outside the body's obstacle blocks hold no udef
    far = after[0][0, 0, 0], after[0][-1, -1, -1]
    assert far == (0.0, 0.0)


def test_udef_is_the_penalization_target_in_the_stiff_limit(oracle):
    """Implicit penalization with a huge lambda drives u to the target V + udef on cells with
    chi = 1 (u_new = u + dt pf (U_tot - u), pf = lambda/(1 + lambda dt) -> 1/dt): a dropped
    udef or a wrong sign leaves u at V or V - udef there. The penalized momenta see u - udef:
    the Gu terms for (u, udef) equal those for (u - udef, 0)."""
    rng = np.random.default_rng(19)
    O = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    O.sphere_create(0.25, (0.5, 0.5, 0.5))
    udef = tuple(0.5 * rng.standard_normal(O.shape) for _ in range(3))
    u = tuple(rng.standard_normal(O.shape) for _ in range(3))
    O.set_udef(*udef)
    O.set_velocity(*u)
    V = (0.3, -0.2, 0.1)
    O.penalize(1e-3, 1e15, True, V)
    chi = O.get("chi")
    inside = chi == 1.0
    assert inside.sum() > 500
    for q, a in enumerate(O.velocity()):
        assert np.abs(a[inside] - (V[q] + udef[q][inside])).max() < 1e-9
    O.set_velocity(*u)
    m1 = O.obstacle_momenta(1e-3, 1e6, True)
    O.set_udef(*(np.zeros(O.shape),) * 3)
    O.set_velocity(*(a - b for a, b in zip(u, udef)))
    m0 = O.obstacle_momenta(1e-3, 1e6, True)
    np.testing.assert_allclose(m1[23:29], m0[23:29], rtol=1e-12, atol=1e-14 * np.abs(m0[23:29]).max())
    np.testing.assert_allclose(m1[13:23], m0[13:23], rtol=1e-14)


def test_deformation_power_of_a_uniform_udef(oracle):
    """With a uniform udef D the deformation power sum f . udef is the total force dotted with D
    (QoI 19 = QoIs 0..2 . D), and its bounded variant (QoI 20) is the sum of the negative parts,
    between it and 0; without udef both are 0."""
    rng = np.random.default_rng(29)
    O = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    O.sphere_create(0.2, (0.47, 0.52, 0.55))
    O.set_velocity(*(rng.standard_normal(O.shape) for _ in range(3)))
    O.pressure_rhs(1e-3)
    O.poisson_solve()
    q0 = O.surface_forces(1e-3, (0.3, 0, 0))
    assert q0[19] == 0 and q0[20] == 0
    D = np.array([0.7, -1.3, 0.4])
    O.set_udef(*(np.full(O.shape, d) for d in D))
    q = O.surface_forces(1e-3, (0.3, 0, 0))
    assert q[19] == pytest.approx(float(np.dot(q[0:3], D)), rel=1e-12)
    assert min(q[19], 0.0) >= q[20] and q[20] < 0
    np.testing.assert_array_equal(q[:19], q0[:19])
