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
