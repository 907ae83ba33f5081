"""Several bodies (ObstacleVector): two spheres, one of them a hemisphere, on the device
against the oracle: chi, centres of mass, the pressure RHS masked and corrected for each body,
the projection, each body's penalization and surface forces, and steps of the pipeline with
both; a body moved; bodies whose obstacle blocks overlap are rejected. FP64 to 1e-11."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BPD, EXT = (8, 4, 4), 2.0
BODIES = [((0.55, 0.5, 0.48), 0.15, False), ((1.45, 0.52, 0.5), 0.12, True)]


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def pair(cup, oracle):
    G = cup.Grid(BPD, precision=8, maxextent=EXT)
    O = oracle.OracleSim(BPD, maxextent=EXT)
    for c, R, hemi in BODIES:
        k, com, mass = G.add_sphere(c, R, hemi=hemi)
        assert O.add_sphere(R, c, hemi=hemi) == k
    return G, O


def test_two_bodies_chi_and_centres(cup, oracle):
    print("two bodies chi", flush=True)
    G, O = pair(cup, oracle)
    assert G.body_count() == O.body_count() == 2
    assert np.abs(G.get("chi") - O.get("chi")).max() < 1e-13
    for k, (c, R, hemi) in enumerate(BODIES):
        with O.body(k):
            info = O.sphere_info()
        m = G.move_body(k, c)  # the same centre: the body is rebuilt where it is
        np.testing.assert_allclose(m["com"], info["com"], rtol=1e-13)
        assert m["mass"] == pytest.approx(info["mass"], rel=1e-13)
    assert np.abs(G.get("chi") - O.get("chi")).max() < 1e-13


def test_two_bodies_rhs_penalization_forces(cup, oracle):
    print("two bodies rhs/penalization/forces", flush=True)
    G, O = pair(cup, oracle)
    rng = np.random.default_rng(6)
    u = tuple(0.3 * rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(*u); O.set_velocity(*u)
    dt, nu, lam = 1e-3, 1e-3, 1e6
    G.pressure_rhs(dt); O.pressure_rhs(dt)
    G.poisson_periodic(); O.poisson_solve()
    assert rel(G.get("p"), O.get("p")) < 1e-12
    G.pressure_project(dt); O.grad_p(dt)
    V = [(-1.0, 0.0, 0.0), (0.0, 0.5, 0.0)]
    for k in range(2):
        fg = G.penalize_body(k, dt, lam, True, V[k])
        with O.body(k):
            fo = O.penalize(dt, lam, True, V[k])
        np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-13)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-12
    for k in range(2):
        qg = G.surface_forces_body(k, nu, V[k])
        with O.body(k):
            qo = O.surface_forces(nu, V[k])
        np.testing.assert_allclose(qg, qo, rtol=1e-10, atol=1e-12 * np.abs(qo).max())


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_two_bodies_steps(cup, oracle, decomp):
    """Three steps of the pipeline (advection, RHS with both bodies, projection, then each
    body's penalization) in both decompositions."""
    print(f"two bodies steps decomp={decomp}", flush=True)
    G = cup.Grid(BPD, precision=8, maxextent=EXT, decomposition=decomp)
    O = oracle.OracleSim(BPD, maxextent=EXT)
    for c, R, hemi in BODIES:
        G.add_sphere(c, R, hemi=hemi)
        O.add_sphere(R, c, hemi=hemi)
    nu, dt, lam, U = 1e-3, 2e-3, 1e6, (0.5, 0.0, 0.0)
    V = [(-0.5, 0.0, 0.0), (-0.5, 0.1, 0.0)]
    for _ in range(3):
        G.advect_diffuse(dt, nu, U); O.advect_diffuse(dt, nu, U)
        G.pressure_rhs(dt); O.pressure_rhs(dt)
        G.poisson_periodic(); O.poisson_solve()
        G.pressure_project(dt); O.grad_p(dt)
        for k in range(2):
            G.penalize_body(k, dt, lam, True, V[k])
            with O.body(k):
                O.penalize(dt, lam, True, V[k])
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11


def test_overlapping_bodies_are_rejected(cup):
    G = cup.Grid(BPD, precision=8, maxextent=EXT)
    G.add_sphere((0.5, 0.5, 0.5), 0.15)
    with pytest.raises(ValueError, match="overlap"):
        G.add_sphere((0.7, 0.5, 0.5), 0.15)
    assert G.body_count() == 1


def test_rejected_move_keeps_the_body(cup, oracle):
    """A move refused because the body's obstacle blocks would overlap another's (or because it
    would touch no block) leaves every body as it was: the count, each centre of mass and mass, the
    obstacle blocks and the grid chi; a later valid move still matches the oracle."""
    print("rejected move keeps the body", flush=True)
    G, O = pair(cup, oracle)
    chi0 = G.get("chi")
    blocks = [G.obstacle_blocks(k) for k in range(2)]
    with pytest.raises(ValueError, match="overlap"):
        G.move_body(1, (0.75, 0.5, 0.48))
    with pytest.raises(ValueError, match="no block"):
        G.move_body(0, (50.0, 50.0, 50.0))
    assert G.body_count() == 2
    np.testing.assert_array_equal(G.get("chi"), chi0)
    for k in range(2):
        np.testing.assert_array_equal(G.obstacle_blocks(k), blocks[k])
    m = G.move_body(1, (1.40, 0.52, 0.5))
    O.move_body(1, (1.40, 0.52, 0.5))
    with O.body(1):
        info = O.sphere_info()
    np.testing.assert_allclose(m["com"], info["com"], rtol=1e-13)
    assert np.abs(G.get("chi") - O.get("chi")).max() < 1e-13
    # the merged obstacle tables follow the bodies: the masked and corrected RHS of both
    rng = np.random.default_rng(11)
    u = tuple(0.3 * rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(*u); O.set_velocity(*u)
    G.pressure_rhs(1e-3); O.pressure_rhs(1e-3)
    G.poisson_periodic(); O.poisson_solve()
    assert rel(G.get("p"), O.get("p")) < 1e-12



PERIODIC = ["-BC_x", "periodic", "-BC_y", "periodic", "-BC_z", "periodic", "-muteAll", "1",
            "-freqDiagnostics", "0", "-precision", "double"]


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_two_spheres_in_the_simulation(cup, oracle, decomp):
    """The drop-in Simulation with two forced spheres (the second a hemisphere) in a
    Taylor-Green flow, both fixed to the frame on x: U_inf is minus their mean x velocity
    (ObstacleVector.h:60-74), so each moves relative to the grid and is built again every
    step; 4 CFL steps against the oracle's pipeline with both bodies."""
    print(f"two spheres in the Simulation decomp={decomp}", flush=True)
    nu, CFL, n, lam = 1e-3, 0.3, 4, 1e6
    argv = ["-bpdx", str(BPD[0]), "-bpdy", str(BPD[1]), "-bpdz", str(BPD[2]), "-extentx", str(EXT), "-CFL",
            str(CFL), "-nu", str(nu), "-nsteps", str(n), "-rampup", "0", "-lambda", str(lam),
            "-implicitPenalization", "1", "-decomposition", decomp, "-initCond", "taylorGreen",
            "-uMax_forced", "1.0"] + PERIODIC
    vel = [(-1.0, 0.0, 0.0), (-0.5, 0.2, 0.0)]
    fix = [(True, True, True), (True, False, False)]
    S = cup.Simulation(cup.SimulationData(argv=argv))
    for (c, R, hemi), v, f in zip(BODIES, vel, fix):
        S.add_obstacle(cup.Sphere(radius=R, position=c, enforcedVelocity=v, bForcedInSimFrame=True,
                                  bFixFrameOfRef=f, bHemi=hemi))
    S.run()
    O = oracle.OracleSim(BPD, maxextent=EXT)
    O.ic_taylor_green(1.0)
    for c, R, hemi in BODIES:
        O.add_sphere(R, c, hemi=hemi)
    bodies = [dict(radius=R, position=list(c), transVel=list(v), angVel=[0.0] * 3, forced=[True] * 3,
                   fix=list(f), block_rot=[True] * 3, imposed=list(v))
              for (c, R, _), v, f in zip(BODIES, vel, fix)]
    uinf = [0.0] * 3  # the first dt sees the flags' frame velocity (0)
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, uinf, step, 0)
        uinf, ft, q = O.bodies_timestep(dt, nu, bodies, lam, True)
    assert uinf == pytest.approx([0.75, 0.0, 0.0])
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-10, name
    assert np.abs(S.field("chi") - O.get("chi")).max() < 1e-13
    for k, b in enumerate(bodies):
        ob = S.obstacle(k)
        np.testing.assert_allclose(ob["position"], b["position"], rtol=1e-14)
        assert np.abs(np.asarray(ob["position"]) - BODIES[k][0]).max() > 0
        np.testing.assert_allclose(ob["trans_vel"], b["transVel"], rtol=0, atol=0)
        np.testing.assert_allclose(ob["force"], ft[k][:3], rtol=1e-9, atol=1e-11 * np.abs(ft[k]).max())
        np.testing.assert_allclose(ob["surface"], q[k], rtol=1e-8, atol=1e-10 * np.abs(q[k]).max())


def test_simulation_rejects_an_overlapping_sphere(cup):
    """A sphere whose obstacle blocks would overlap another body's is refused with ValueError and
    leaves the simulation with its bodies as before."""
    print("Simulation overlapping sphere", flush=True)
    argv = ["-bpdx", str(BPD[0]), "-bpdy", str(BPD[1]), "-bpdz", str(BPD[2]), "-extentx", str(EXT), "-CFL", "0.3",
            "-nu", "1e-3", "-nsteps", "2", "-initCond", "zero"] + PERIODIC
    S = cup.Simulation(cup.SimulationData(argv=argv))
    S.add_obstacle(cup.Sphere(radius=0.15, position=(0.5, 0.5, 0.5), xvel=1.0))
    chi = S.field("chi").copy()
    with pytest.raises(ValueError, match="overlap"):
        S.add_obstacle(cup.Sphere(radius=0.15, position=(0.7, 0.5, 0.5), xvel=1.0))
    assert np.array_equal(S.field("chi"), chi)
    with pytest.raises(ValueError):
        S.obstacle(1)
    S.run()
    assert np.isfinite(S.obstacle(0)["force"]).all()


def test_two_spheres_write_their_own_files(cup, tmp_path):
    """Each body writes its own forceValues_surface_<id>, powerValues_surface_<id>,
    forceValues_<id> and computedVelocity_<id> (Obstacle.cpp:491-589), one row per step, with
    its own velocity and centre."""
    print("two spheres log files", flush=True)
    argv = ["-bpdx", str(BPD[0]), "-bpdy", str(BPD[1]), "-bpdz", str(BPD[2]), "-extentx", str(EXT), "-CFL", "0.3",
            "-nu", "1e-3", "-nsteps", "3", "-rampup", "0", "-lambda", "1e6", "-implicitPenalization", "1",
            "-initCond", "zero", "-BC_x", "periodic", "-BC_y", "periodic", "-BC_z", "periodic",
            "-freqDiagnostics", "0", "-precision", "double", "-muteAll", "0", "-serialization", str(tmp_path)]
    S = cup.Simulation(cup.SimulationData(argv=argv))
    vel = [(-1.0, 0.0, 0.0), (-0.5, 0.2, 0.0)]
    for (c, R, hemi), v in zip(BODIES, vel):
        S.add_obstacle(cup.Sphere(radius=R, position=c, enforcedVelocity=v, bForcedInSimFrame=True,
                                  bFixFrameOfRef=False, bHemi=hemi))
    S.run()
    S.flush_logs()
    for k in range(2):
        for name in (f"forceValues_surface_{k}.dat", f"powerValues_surface_{k}.dat", f"forceValues_{k}.dat",
                     f"computedVelocity_{k}.dat"):
            with open(tmp_path / name) as f:
                rows = [ln.split() for ln in f.read().splitlines() if ln.strip()]
            assert len(rows) == 4, name  # header and 3 steps
        with open(tmp_path / f"computedVelocity_{k}.dat") as f:
            rows = [ln.split() for ln in f.read().splitlines() if ln.strip()]
        head, data = rows[0], np.array([[float(x) for x in r] for r in rows[1:]])
        np.testing.assert_array_equal(data[:, head.index("vel_x")], vel[k][0])
        np.testing.assert_array_equal(data[:, head.index("vel_y")], vel[k][1])
        assert data[-1, head.index("CMx")] != BODIES[k][0][0]  # the body moved with the frame at rest
        ob = S.obstacle(k)
        with open(tmp_path / f"forceValues_{k}.dat") as f:
            rows = [ln.split() for ln in f.read().splitlines() if ln.strip()]
        np.testing.assert_allclose([float(x) for x in rows[-1][3:6]], ob["force"], rtol=1e-6, atol=1e-9)


def test_udef_momentum_removal_matches_oracle(cup, oracle):
    """CreateObstacles' udef momentum removal (ObstaclesCreate.cpp:258-440) on the device against
    the oracle, for body 1 of two (the other keeps udef = 0), twice in a row (the second call
    uses the first's linear correction), FP64; a moved body starts again from udef = 0."""
    print("udef momentum removal", flush=True)
    G, O = pair(cup, oracle)
    rng = np.random.default_rng(21)
    udef = tuple(rng.standard_normal(G.shape) for _ in range(3))
    G.set_udef(1, *udef)
    with O.body(1):
        O.set_udef(*udef)
    scale = {}
    for it in range(2):
        rg = G.remove_udef_momenta(1)
        with O.body(1):
            ro = O.remove_udef_momenta()
        assert rg["mass"] == pytest.approx(ro["mass"], rel=1e-13)
        for key in ("trans", "ang", "J"):
            if it == 0:
                scale[key] = np.abs(ro[key]).max()
            # the second call removes rounding only: compared at the first call's scale
            np.testing.assert_allclose(rg[key], ro[key], rtol=1e-11, atol=1e-13 * scale[key])
        with O.body(1):
            uo = O.get_udef()
        for a, b in zip(G.get_udef(1), uo):
            assert rel(a, b) < 1e-12
    assert np.abs(ro["trans"]).max() < 1e-12 * scale["trans"]  # nothing left to remove
    for a in G.get_udef(0):
        assert not a.any()
    r0 = G.remove_udef_momenta(0)  # a body without udef: nothing to remove
    assert not np.abs(r0["trans"]).any() and not np.abs(r0["ang"]).any()
    G.move_body(1, (1.47, 0.5, 0.5))
    for a in G.get_udef(1):
        assert not a.any()
    with pytest.raises(ValueError):
        G.set_udef(2, *udef)


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_udef_in_penalization_and_momenta(cup, oracle, decomp):
    """A sphere with a deformation velocity: the penalization's target V + Omega x p + udef and
    the penalized momenta's u - udef against the oracle, then 3 steps of the device loop with the
    body (the loop's penalization takes udef too), FP64."""
    print(f"udef penalization decomp={decomp}", flush=True)
    rng = np.random.default_rng(23)
    bpd, ext, c, R = (4, 4, 4), 1.0, (0.47, 0.52, 0.55), 0.2
    G = cup.Grid(bpd, precision=8, maxextent=ext, decomposition=decomp)
    O = oracle.OracleSim(bpd, maxextent=ext)
    G.chi_from_sphere(c, R)
    O.sphere_create(R, c)
    udef = tuple(0.3 * rng.standard_normal(G.shape) for _ in range(3))
    u = tuple(rng.standard_normal(G.shape) for _ in range(3))
    G.set_udef(0, *udef); O.set_udef(*udef)
    G.set_velocity(*u); O.set_velocity(*u)
    mg, mo = G.obstacle_momenta(1e-3, 1e6, True), O.obstacle_momenta(1e-3, 1e6, True)
    np.testing.assert_allclose(mg, mo, rtol=1e-11, atol=1e-13 * np.abs(mo).max())
    V, W = (0.3, -0.2, 0.1), (0.5, 0.0, -0.4)
    fg = G.penalize(1e-3, 1e6, True, V, W)
    fo = O.penalize(1e-3, 1e6, True, V, W)
    np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-13 * np.abs(fo).max())
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-13
    P = rng.standard_normal(G.shape)  # a pressure field on both (the fused RHS reads u*, not u)
    G.set("p", P); O.set("p", P)
    qg, qo = G.surface_forces(1e-3, V, W), O.surface_forces(1e-3, V, W)
    assert abs(qo[19]) > 0 and qo[20] < 0  # the deformation powers f . udef
    np.testing.assert_allclose(qg, qo, rtol=1e-10, atol=1e-12 * np.abs(qo).max())
    sph = dict(lambda_=1e6, implicit=True, trans_vel=V, ang_vel=W)
    G.run_steps(3, nu=1e-3, fixed_dt=2e-3, sphere=sph)
    for _ in range(3):
        O.timestep(2e-3, 1e-3, sphere=sph)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11
    with pytest.raises(cup.CupError, match="udef"):  # the body adjoint does not differentiate udef
        G.adjoint_step_body(1e-3, 1e-3, (0, 0, 0), sph)


@pytest.mark.parametrize("case", ["towed_and_resting", "two_frame_fixing"])
def test_spheres_that_move_in_the_frame_take_the_pipeline(cup, oracle, case):
    """Every sphere forced at a constant velocity, yet some move relative to the grid: a resting
    sphere beside a towed one that fixes the frame (U_inf = 1 carries it along x), or two spheres
    fixing the frame at different x velocities (U_inf = 0.75, each moves). Simulation.run in the
    fused decomposition must move and rebuild them every step as the oracle's pipeline does:
    positions, chi, fields and each body's force after 4 CFL steps."""
    print(f"moving forced spheres in the Simulation case={case}", flush=True)
    nu, CFL, n, lam = 1e-3, 0.3, 4, 1e6
    argv = ["-bpdx", str(BPD[0]), "-bpdy", str(BPD[1]), "-bpdz", str(BPD[2]), "-extentx", str(EXT), "-CFL",
            str(CFL), "-nu", str(nu), "-nsteps", str(n), "-rampup", "0", "-lambda", str(lam),
            "-implicitPenalization", "1", "-decomposition", "fused", "-initCond", "zero"] + PERIODIC
    if case == "towed_and_resting":
        vel = [(-1.0, 0.0, 0.0), (0.0, 0.0, 0.0)]
        fix = [(True, True, True), (False, False, False)]
    else:
        vel = [(-1.0, 0.0, 0.0), (-0.5, 0.0, 0.0)]
        fix = [(True, True, True), (True, True, True)]
    S = cup.Simulation(cup.SimulationData(argv=argv))
    for (c, R, hemi), v, f in zip(BODIES, vel, fix):
        S.add_obstacle(cup.Sphere(radius=R, position=c, enforcedVelocity=v, bForcedInSimFrame=True,
                                  bFixFrameOfRef=f, bHemi=hemi))
    S.run()
    O = oracle.OracleSim(BPD, maxextent=EXT)
    for c, R, hemi in BODIES:
        O.add_sphere(R, c, hemi=hemi)
    bodies = [dict(radius=R, position=list(c), transVel=list(v), angVel=[0.0] * 3, forced=[True] * 3,
                   fix=list(f), block_rot=[True] * 3, imposed=list(v))
              for (c, R, _), v, f in zip(BODIES, vel, fix)]
    uinf = [0.0] * 3
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, uinf, step, 0)
        uinf, ft, q = O.bodies_timestep(dt, nu, bodies, lam, True)
    assert uinf == pytest.approx([1.0 if case == "towed_and_resting" else 0.75, 0.0, 0.0])
    moved = 0
    for k, b in enumerate(bodies):
        ob = S.obstacle(k)
        np.testing.assert_allclose(ob["position"], b["position"], rtol=1e-14)
        moved += np.abs(np.asarray(ob["position"]) - BODIES[k][0]).max() > 0
        np.testing.assert_allclose(ob["force"], ft[k][:3], rtol=1e-9, atol=1e-11 * np.abs(ft[k]).max())
    assert moved == (1 if case == "towed_and_resting" else 2)
    assert np.abs(S.field("chi") - O.get("chi")).max() < 1e-13
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-10, name


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_two_static_spheres_in_the_device_loop(cup, oracle, decomp):
    """Two spheres held in place in a uniform inflow (both forced at -1 on x with the frame fixed
    to them, so U_inf = 1 and neither moves): Simulation.run takes the device-resident loop
    (cup_run_steps penalizing both, the second with cup_body_velocity) after the first step;
    fields, each body's penalization force and surface QoIs against the oracle's pipeline."""
    print(f"two static spheres device loop decomp={decomp}", flush=True)
    nu, CFL, n, lam = 1e-3, 0.3, 4, 1e6
    argv = ["-bpdx", str(BPD[0]), "-bpdy", str(BPD[1]), "-bpdz", str(BPD[2]), "-extentx", str(EXT), "-CFL",
            str(CFL), "-nu", str(nu), "-nsteps", str(n), "-rampup", "0", "-lambda", str(lam),
            "-implicitPenalization", "1", "-decomposition", decomp, "-initCond", "zero"] + PERIODIC
    S = cup.Simulation(cup.SimulationData(argv=argv))
    for c, R, hemi in BODIES:
        S.add_obstacle(cup.Sphere(radius=R, position=c, enforcedVelocity=(-1.0, 0.0, 0.0), bForcedInSimFrame=True,
                                  bFixFrameOfRef=True, bHemi=hemi))
    S.run()
    O = oracle.OracleSim(BPD, maxextent=EXT)
    for c, R, hemi in BODIES:
        O.add_sphere(R, c, hemi=hemi)
    bodies = [dict(radius=R, position=list(c), transVel=[-1.0, 0.0, 0.0], angVel=[0.0] * 3, forced=[True] * 3,
                   fix=[True] * 3, block_rot=[True] * 3, imposed=[-1.0, 0.0, 0.0]) for c, R, _ in BODIES]
    uinf = [0.0] * 3
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, uinf, step, 0)
        uinf, ft, q = O.bodies_timestep(dt, nu, bodies, lam, True)
    assert uinf == [1.0, 0.0, 0.0]
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-10, name
    for k in range(2):
        ob = S.obstacle(k)
        np.testing.assert_array_equal(ob["position"], BODIES[k][0])
        np.testing.assert_allclose(ob["force"], ft[k][:3], rtol=1e-9, atol=1e-11 * np.abs(ft[k]).max())
        np.testing.assert_allclose(ob["surface"], q[k], rtol=1e-8, atol=1e-10 * np.abs(q[k]).max())
