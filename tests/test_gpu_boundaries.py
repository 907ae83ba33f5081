"""Non-periodic boxes on the device against the CPU oracle (FP64, the PLAIN decomposition):
BlockLabBC's open and wall faces (Definitions.h:261-378) in every stencil of the pipeline, the
advection's inflow fade and UpdateAndCorrectInflow (AdvectionDiffusion.cpp:20-104, 209-296),
PoissonSolverMixed (PoissonSolverMixed.h:27-80), and whole steps with a forced sphere in a box
with the reference sphere preset's boundaries (launch/settingsSphere.sh: -BC_x/y/z dirichlet).
The oracle's boundaries and mixed solve are pinned in tests/test_oracle_boundaries.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BCS = [("dirichlet", "wall", "periodic"), ("wall", "dirichlet", "dirichlet"), ("dirichlet",) * 3, ("wall",) * 3]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def pair(cup, oracle, bpd, bc, maxextent=1.0):
    return (cup.Grid(bpd, maxextent=maxextent, precision=8, decomposition="plain", bc=bc),
            oracle.OracleSim(bpd, maxextent=maxextent, bc=bc))


@pytest.mark.parametrize("bc", BCS)
def test_projection_with_boundaries(cup, oracle, bc):
    """PressureRHS with boundary labs, PoissonSolverMixed and KernelGradP."""
    print(f"mixed projection bc={bc}", flush=True)
    rng = np.random.default_rng(3)
    G, O = pair(cup, oracle, (3, 2, 4), bc)
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w)
    O.set_velocity(u, v, w)
    dt = 1e-3
    G.pressure_rhs(dt); G.poisson_mixed(); G.pressure_project(dt)
    O.pressure_rhs(dt); O.poisson_mixed(); O.grad_p(dt)
    ep = rel(G.get("p"), O.get("p"))
    print(f"  p rel L2 {ep:.2e}", flush=True)
    assert ep < 1e-11
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-12


def test_mixed_solver_on_a_periodic_box(cup, oracle):
    """The Python path's solver of periodic boxes (its SimulationData never sets
    bUseFourierBC): the blended symbol on 24 x 16 x 32 cells."""
    rng = np.random.default_rng(9)
    G, O = pair(cup, oracle, (3, 2, 4), ("periodic",) * 3, maxextent=2 * np.pi)
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w)
    O.set_velocity(u, v, w)
    G.pressure_rhs(1e-3); G.poisson_mixed()
    O.pressure_rhs(1e-3); O.poisson_mixed()
    assert rel(G.get("p"), O.get("p")) < 1e-11


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("uinf", [(0.3, -0.2, 0.1), (-0.25, 0.0, 0.4)])
def test_advection_with_boundaries(cup, oracle, bc, uinf):
    """KernelAdvectDiffuse with the open faces' inflow fade (uMax_measured from the CFL
    rule) and the inflow correction after the copy."""
    print(f"advection bc={bc} uinf={uinf}", flush=True)
    rng = np.random.default_rng(5)
    G, O = pair(cup, oracle, (2, 3, 2), bc)
    u, v, w = (0.5 * rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w)
    O.set_velocity(u, v, w)
    dg, ug = G.calc_max_timestep(0.3, 0.01, uinf)
    do, uo = O.calc_max_timestep(0.3, 0.01, uinf)
    assert dg == do and ug == uo
    dt = 4 * dg  # a larger step makes the fade factor of order one
    G.advect_diffuse(dt, 0.01, uinf)
    O.advect_diffuse(dt, 0.01, uinf)
    for a, b in zip(G.velocity(), O.velocity()):
        e = rel(a, b)
        print(f"  rel L2 {e:.2e}", flush=True)
        assert e < 1e-13


@pytest.mark.parametrize("bc", BCS)
def test_diagnostics_with_boundaries(cup, oracle, bc):
    rng = np.random.default_rng(2)
    G, O = pair(cup, oracle, (2, 2, 3), bc)
    for n in ("u", "v", "w", "p"):
        f = rng.standard_normal(G.shape)
        G.set(n, f)
        O.set(n, f)
    chi = rng.uniform(0, 1, G.shape)
    G.set("chi", chi)
    O.set("chi", chi)
    a, b = G.diagnostics(0.02), O.dissipation(0.02)
    scale = np.abs(b).max()
    assert np.all(np.abs(a - b) <= 1e-12 * np.maximum(np.abs(b), 1e-3 * scale))


@pytest.mark.parametrize("bc", [("dirichlet",) * 3, ("periodic", "wall", "dirichlet")])
def test_sphere_steps_in_a_bounded_box(cup, oracle, bc):
    """The sphere preset's setting on a small grid: a sphere held at V = (-0.5, 0, 0) in a box
    with open faces, uinf = (0.5, 0, 0) flowing past it; 4 CFL steps (advection with fade and
    inflow correction, RHS masking, PoissonSolverMixed, gradient, implicit penalization, surface
    forces) against the oracle's steps."""
    print(f"sphere in a box bc={bc}", flush=True)
    bpd, nu, lam, R = (4, 2, 2), 0.005, 1e6, 0.12
    G, O = pair(cup, oracle, bpd, bc)
    centre = (0.35, 0.25, 0.27)
    G.chi_from_sphere(centre, R)
    O.sphere_create(R, centre)
    uinf, V = (0.5, 0.0, 0.0), (-0.5, 0.0, 0.0)
    z = np.zeros(G.shape)
    G.set_velocity(z, z, z)
    O.set_velocity(z, z, z)
    for step in range(4):
        dg, _ = G.calc_max_timestep(0.3, nu, uinf, step, 0)
        do, _ = O.calc_max_timestep(0.3, nu, uinf, step, 0)
        assert abs(dg - do) < 1e-14 * do
        G.advect_diffuse(dg, nu, uinf); O.advect_diffuse(do, nu, uinf)
        G.pressure_rhs(dg); O.pressure_rhs(do)
        G.poisson_mixed(); O.poisson_mixed()
        G.pressure_project(dg); O.grad_p(do)
        fg = G.penalize(dg, lam, True, V)
        fo = O.penalize(do, lam, True, V)
        assert np.allclose(fg, fo, rtol=1e-10, atol=1e-14)
    for n in ("u", "v", "w", "p"):
        e = rel(G.get(n), O.get(n))
        print(f"  {n} rel L2 {e:.2e}", flush=True)
        assert e < 1e-11
    qg, qo = G.surface_forces(nu, V), O.surface_forces(nu, V)
    assert np.allclose(qg, qo, rtol=1e-9, atol=1e-12 * np.abs(qo).max())


@pytest.mark.parametrize("bc", BCS)
def test_refusals_on_non_periodic_boxes(cup, bc):
    """The periodic-only operations refuse a non-periodic box; FUSED needs a periodic box."""
    with pytest.raises(ValueError):
        cup.Grid((2, 2, 2), decomposition="fused", bc=bc)
    G = cup.Grid((2, 2, 2), bc=bc)
    for call in (G.poisson_periodic, lambda: G.ic_hit(4.0, 1.0, 1), lambda: G.hit_analysis(0.01),
                 lambda: G.advect_diffuse(1e-3, 0.01, (0, 0, 0), cup._lib.CUP_UPWIND3_EULER)):
        with pytest.raises(ValueError):
            call()


@pytest.mark.parametrize("bc", ["dirichlet", "wall"])
def test_simulation_sphere_preset_boundaries(cup, oracle, bc):
    """The drop-in Simulation with the sphere preset's flags (launch/settingsSphere.sh: -BC_x/y/z
    dirichlet, a sphere forced at xvel with the frame fixed to it), on 64 x 32 x 32 cells, and
    with walls instead: 4 CFL steps through the pipeline (PoissonSolverMixed chosen by
    PressureProjection, PressureProjection.cpp:86-105) against the oracle's steps."""
    print(f"Simulation sphere preset BC {bc}", flush=True)
    nu, CFL, n, R = 1.5625e-4 * 8, 0.1, 4, 0.0625
    argv = ["-bpdx", "8", "-bpdy", "4", "-bpdz", "4", "-extentx", "1.0", "-CFL", str(CFL), "-nu", str(nu),
            "-nsteps", str(n), "-rampup", "0", "-sphereL", str(2 * R), "-xpos", "0.35", "-xvel", "0.125",
            "-bForcedInSimFrame", "1", "-bFixFrameOfRef", "1", "-lambda", "1e5", "-implicitPenalization", "1",
            "-BC_x", bc, "-BC_y", bc, "-BC_z", bc, "-muteAll", "1", "-freqDiagnostics", "0"]
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    O = oracle.OracleSim((8, 4, 4), maxextent=1.0, bc=(bc, bc, bc))
    O.sphere_create(R, (0.35, 0.25, 0.25))
    U = (0.125, 0.0, 0.0)
    sph = dict(lambda_=1e5, implicit=True, trans_vel=(-0.125, 0.0, 0.0))
    u_cfl = (0.0, 0.0, 0.0)
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, u_cfl, step, 0)
        out = O.timestep(dt, nu, uinf=U, sphere=sph)
        u_cfl = U
    for name in ("u", "v", "w", "p"):
        e = rel(S.field(name), O.get(name))
        print(f"  {name} rel L2 {e:.2e}", flush=True)
        assert e < 1e-10, name
    ob = S.obstacle(0)
    np.testing.assert_allclose(ob["force"], out["penal_ft"][:3], rtol=1e-9, atol=1e-12 * np.abs(out["penal_ft"]).max())


@pytest.mark.parametrize("bpd", [(2, 2, 2), (3, 2, 4)])
@pytest.mark.parametrize("precision,tol", [(8, 1e-11), (4, 2e-6)])
def test_unbounded_solve_matches_the_direct_sum(cup, oracle, bpd, precision, tol):
    """PoissonSolverUnbounded on the device (FFTs of the grid padded to 2N per axis, 48 cells
    for the 24-cell axis: radix 3) against the oracle's direct free-space sum, from the pressure
    RHS of a random velocity."""
    print(f"unbounded solve bpd={bpd} precision={precision}", flush=True)
    rng = np.random.default_rng(12)
    bc = ("freespace",) * 3
    G = cup.Grid(bpd, maxextent=1.0, precision=precision, decomposition="plain", bc=bc)
    O = oracle.OracleSim(bpd, maxextent=1.0, bc=bc)
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w)
    O.set_velocity(u, v, w)
    G.pressure_rhs(1e-3); G.poisson_unbounded(); G.poisson_unbounded()  # the second call reuses the kernel
    O.pressure_rhs(1e-3); O.poisson_unbounded()
    e = rel(G.get("p"), O.get("p"))
    print(f"  p rel L2 {e:.2e}", flush=True)
    assert e < tol


def test_freespace_refusals(cup):
    with pytest.raises(ValueError):  # all or none (SimulationData.cpp:136-143)
        cup.Grid((2, 2, 2), bc=("freespace", "dirichlet", "freespace"))
    G = cup.Grid((2, 2, 2), bc=("freespace",) * 3, decomposition="plain")
    with pytest.raises(ValueError):
        G.poisson_periodic()


def test_simulation_sphere_in_freespace(cup, oracle):
    """The drop-in Simulation with -BC_x/y/z freespace (PressureProjection picks
    PoissonSolverUnbounded, PressureProjection.cpp:86-105): a forced sphere, 3 CFL steps on
    32 x 16 x 16 cells against the oracle's steps."""
    print("Simulation sphere freespace", flush=True)
    nu, CFL, n, R = 1e-3, 0.1, 3, 0.1
    argv = ["-bpdx", "4", "-bpdy", "2", "-bpdz", "2", "-extentx", "1.0", "-CFL", str(CFL), "-nu", str(nu),
            "-nsteps", str(n), "-rampup", "0", "-sphereL", str(2 * R), "-xpos", "0.35", "-xvel", "0.125",
            "-bForcedInSimFrame", "1", "-bFixFrameOfRef", "1", "-lambda", "1e5", "-implicitPenalization", "1",
            "-BC_x", "freespace", "-BC_y", "freespace", "-BC_z", "freespace", "-muteAll", "1",
            "-freqDiagnostics", "0"]
    S = cup.Simulation(cup.SimulationData(argv=argv))
    S.run()
    O = oracle.OracleSim((4, 2, 2), maxextent=1.0, bc=("freespace",) * 3)
    O.sphere_create(R, (0.35, 0.25, 0.25))
    U = (0.125, 0.0, 0.0)
    sph = dict(lambda_=1e5, implicit=True, trans_vel=(-0.125, 0.0, 0.0))
    u_cfl = (0.0, 0.0, 0.0)
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, u_cfl, step, 0)
        O.timestep(dt, nu, uinf=U, sphere=sph)
        u_cfl = U
    for name in ("u", "v", "w", "p"):
        e = rel(S.field(name), O.get(name))
        print(f"  {name} rel L2 {e:.2e}", flush=True)
        assert e < 1e-10, name
