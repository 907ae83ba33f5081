"""Bodies that move: UpdateObstacles' momentum integrals on the device against the oracle,
and the drop-in Simulation with a free sphere (6 degrees of freedom from the penalized
momenta, chi rebuilt where the body moved) and with a sphere forced through the box, step
by step against the oracle's pipeline (ObstaclesCreate.cpp:444-468, ObstaclesUpdate.cpp:202-236,
Obstacle.cpp:153-345). FP64 to 1e-10."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PERIODIC = ["-BC_x", "periodic", "-BC_y", "periodic", "-BC_z", "periodic", "-muteAll", "1",
            "-freqDiagnostics", "0", "-precision", "double"]


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.parametrize("implicit", [False, True])
def test_obstacle_momenta_match_oracle(cup, oracle, implicit):
    print(f"obstacle momenta implicit={implicit}", flush=True)
    bpd = (4, 4, 4)
    rng = np.random.default_rng(5)
    G = cup.Grid(bpd, precision=8, maxextent=1.0)
    O = oracle.OracleSim(bpd, maxextent=1.0)
    c, R = (0.47, 0.52, 0.55), 0.2
    G.chi_from_sphere(c, R)
    O.sphere_create(R, c)
    u = tuple(rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(*u)
    O.set_velocity(*u)
    mg, mo = G.obstacle_momenta(1e-3, 1e6, implicit), O.obstacle_momenta(1e-3, 1e6, implicit)
    np.testing.assert_allclose(mg, mo, rtol=1e-11, atol=1e-13 * np.abs(mo).max())


def run_body(cup, oracle, extra, body, n, ic, nu=0.01, CFL=0.3, lam=1e6, ramp=None):
    argv = ["-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-extentx", "1.0", "-CFL", str(CFL), "-nu", str(nu),
            "-nsteps", str(n), "-rampup", "0", "-lambda", str(lam), "-implicitPenalization", "1",
            "-decomposition", "plain"] + extra + PERIODIC
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    O = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    ic(O)
    O.sphere_create(body["radius"], body["position"])
    uinf, t = [0.0] * 3, 0.0  # the first dt sees the flags' frame velocity (0)
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, uinf, step, 0)
        if ramp is not None:
            body["imposed"][0] = ramp(t)
        uinf, ft, q = O.body_timestep(dt, nu, body, lam, True)
        t += dt
    return S, O, ft, q


def test_free_sphere_in_taylor_green_flow(cup, oracle):
    """A free sphere (no forced axis, frame at rest) in a Taylor-Green flow picks up a
    velocity and a rotation from the penalized momenta and moves; 4 steps."""
    print("free sphere in TGV", flush=True)
    R, pos = 0.15, [0.41, 0.53, 0.6]
    extra = ["-initCond", "taylorGreen", "-uMax_forced", "1.0", "-sphereL", str(2 * R), "-xpos", str(pos[0]),
             "-ypos", str(pos[1]), "-zpos", str(pos[2]), "-bForcedInSimFrame", "0", "-bFixFrameOfRef", "0"]
    body = dict(radius=R, position=list(pos), transVel=[0.0] * 3, angVel=[0.0] * 3, forced=[False] * 3,
                fix=[False] * 3, block_rot=[False] * 3, imposed=[0.0] * 3)
    S, O, ft, q = run_body(cup, oracle, extra, body, 4, lambda O: O.ic_taylor_green(1.0))
    ob = S.obstacle(0)
    assert np.abs(ob["trans_vel"]).max() > 1e-3 and np.abs(ob["ang_vel"]).max() > 1e-3
    np.testing.assert_allclose(ob["trans_vel"], body["transVel"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ob["ang_vel"], body["angVel"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ob["position"], body["position"], rtol=1e-13)
    assert np.abs(np.asarray(ob["position"]) - pos).max() > 0
    assert np.linalg.norm(ob["quaternion"]) == pytest.approx(1.0, abs=1e-14)
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-10, name
    # a free body exchanges no net momentum with the fluid through the penalization: the
    # velocities solve exactly that balance, so the force is rounding on both sides
    assert np.abs(ob["force"]).max() < 1e-13 and np.abs(ft[:3]).max() < 1e-13
    np.testing.assert_allclose(ob["surface"], q, rtol=1e-8, atol=1e-10 * np.abs(q).max())


def test_forced_sphere_moving_through_the_box(cup, oracle):
    """A sphere forced at -xvel with the frame at rest crosses cells, so its chi is rebuilt
    every step (-xvel 2, 5 steps), in a fluid initially at rest."""
    print("forced moving sphere", flush=True)
    R, pos = 0.12, [0.3, 0.5, 0.5]
    extra = ["-initCond", "zero", "-sphereL", str(2 * R), "-xpos", str(pos[0]), "-ypos", str(pos[1]), "-zpos",
             str(pos[2]), "-xvel", "-2.0", "-bForcedInSimFrame", "1", "-bFixFrameOfRef", "0"]
    body = dict(radius=R, position=list(pos), transVel=[2.0, 0.0, 0.0], angVel=[0.0] * 3, forced=[True] * 3,
                fix=[False] * 3, block_rot=[True] * 3, imposed=[2.0, 0.0, 0.0])
    S, O, ft, q = run_body(cup, oracle, extra, body, 5, lambda O: None)
    ob = S.obstacle(0)
    np.testing.assert_allclose(ob["trans_vel"], [2.0, 0, 0], rtol=0, atol=0)
    np.testing.assert_allclose(ob["position"], body["position"], rtol=1e-14)
    assert ob["position"][0] - pos[0] > 2.5 / 32  # moved over more than two cells
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-10, name
    np.testing.assert_allclose(ob["force"], ft[:3], rtol=1e-9, atol=1e-11 * np.abs(ft).max())


def test_sphere_with_ramped_velocity(cup, oracle):
    """-accel: the forced x velocity rises to xvel at T and falls to 0 at 2T
    (Sphere.cpp:90-98, 147-159); the frame follows the body, so U_inf = -V changes every step."""
    print("accel/decel sphere", flush=True)
    R, pos, T, umax = 0.12, [0.5, 0.5, 0.5], 0.1, 1.5
    extra = ["-initCond", "zero", "-sphereL", str(2 * R), "-xpos", str(pos[0]), "-ypos", str(pos[1]), "-zpos",
             str(pos[2]), "-xvel", str(umax), "-accel", "1", "-T", str(T), "-bForcedInSimFrame", "1",
             "-bFixFrameOfRef", "1"]

    def ramp(t):
        return umax * t / T if t < T else (umax * (2 * T - t) / T if t < 2 * T else 0.0)

    body = dict(radius=R, position=list(pos), transVel=[-umax, 0.0, 0.0], angVel=[0.0] * 3, forced=[True] * 3,
                fix=[True] * 3, block_rot=[True] * 3, imposed=[-umax, 0.0, 0.0])
    S, O, ft, q = run_body(cup, oracle, extra, body, 6, lambda O: None, ramp=ramp)
    ob = S.obstacle(0)
    assert 0 < ob["trans_vel"][0] < umax
    np.testing.assert_allclose(ob["trans_vel"], body["transVel"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(ob["position"], pos, rtol=0, atol=0)  # the frame follows the body
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-10, name
    np.testing.assert_allclose(ob["force"], ft[:3], rtol=1e-9, atol=1e-11 * np.abs(ft).max())


def test_python_sphere_like_the_reference_test(cup):
    """tests/python/test_obstacles.py of the reference: a 64^3 box with U_inf = 0.1, a sphere
    added with the bindings' defaults (free, frame at rest), 10 steps. CreateObstacles sets
    U_inf from the bodies' fixed-frame axes, none here, so the flow stays at rest and so does
    the body."""
    print("python sphere as in the reference test", flush=True)
    SD = cup.SimulationData(cells=[64, 64, 64], CFL=0.1, uinf=[0.1, 0.0, 0.0])
    SD.nsteps = 10
    S = cup.Simulation(SD)
    S.add_obstacle(cup.Sphere(position=[0.3, 0.4, 0.5], radius=0.1))
    S.run()
    assert SD.step == 10 and SD.uinf_x == 0.0
    ob = S.obstacle(0)
    assert not np.asarray(ob["trans_vel"]).any() and not np.asarray(ob["ang_vel"]).any()
    np.testing.assert_array_equal(ob["position"], [0.3, 0.4, 0.5])
    assert ob["mass"] > 0  # a radius of about one cell (h = 2 pi / 64)
    for name in ("u", "v", "w"):
        assert not S.field(name).any(), name
