"""HIP kernels through the C-ABI against the oracle: halo gather, advection-diffusion,
pressure RHS + FFT Poisson + gradient, reductions and dt, full Taylor-Green steps (C1).
FP64 device results must match the FP64 oracle to 1e-11 relative; FP32 device results
match to FP32 rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def cell_centres(shape, h):
    Z, Y, X = shape
    z, y, x = np.meshgrid((np.arange(Z) + 0.5) * h, (np.arange(Y) + 0.5) * h, (np.arange(X) + 0.5) * h,
                          indexing="ij")
    return x, y, z


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("g,shift", [(1, (-1, 0, 0)), (1, (1, 0, 0)), (2, (-1, 1, 2))])
def test_lab_periodic_shift_device(cup, prec, g, shift):
    """TestBoundaries.cpp:19-75 on the device block store (exact)."""
    print(f"lab shift prec={prec} g={g} shift={shift}", flush=True)
    G = cup.Grid((16, 4, 4), precision=prec)
    Z, Y, X = G.shape
    iz, iy, ix = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    val = ix + 1000.0 * iy + (1e6 if prec == 8 else 1e4) * iz
    G.set("u", val)
    dx, dy, dz = shift
    G.lab_shift(g, dx, dy, dz)
    exp = ((ix - dx) % X) + 1000.0 * ((iy - dy) % Y) + (1e6 if prec == 8 else 1e4) * ((iz - dz) % Z)
    assert np.array_equal(G.get("tmpU"), exp)


@pytest.mark.parametrize("bpd", [(2, 2, 2), (4, 2, 3)])
def test_advdiff_fp64_parity(cup, oracle, bpd):
    print(f"advdiff parity bpd={bpd}", flush=True)
    rng = np.random.default_rng(11)
    G = cup.Grid(bpd, precision=8)
    O = oracle.OracleSim(bpd)
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w); O.set_velocity(u, v, w)
    U = (0.2, -0.1, 0.05)
    G.advect_diffuse(2e-3, 0.01, U); O.advect_diffuse(2e-3, 0.01, U)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-14


@pytest.mark.parametrize("decomp", ["plain", "fused"])
@pytest.mark.parametrize("bpd", [(2, 2, 2), (4, 2, 1), (1, 2, 4), (8, 8, 8), (4, 16, 16), (2, 32, 32)])
def test_projection_fp64_parity(cup, oracle, decomp, bpd):
    """RHS, Poisson and gradient against the oracle; y/z FFTs of 8 to 256 points (64 and up
    take the register four-step form: 4, 8, 16 threads per column)."""
    print(f"projection parity decomp={decomp} bpd={bpd}", flush=True)
    rng = np.random.default_rng(5)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    dt = 1e-3
    O.set_velocity(u, v, w)
    if decomp == "fused":   # FUSED reads the advected velocity from tmp
        G.set("tmpU", u); G.set("tmpV", v); G.set("tmpW", w)
    else:
        G.set_velocity(u, v, w)
    G.pressure_rhs(dt); G.poisson_periodic(); G.pressure_project(dt)
    O.pressure_rhs(dt); O.poisson_solve(); O.grad_p(dt)
    assert rel(G.get("p"), O.get("p")) < 1e-12
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-13


def test_max_velocity_and_dt(cup, oracle):
    rng = np.random.default_rng(2)
    G = cup.Grid((2, 2, 2), precision=8)
    O = oracle.OracleSim((2, 2, 2))
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w); O.set_velocity(u, v, w)
    U = (0.5, 0, -0.25)
    assert G.max_velocity(U) == O.find_max_u(U)
    for step, ramp in [(0, 0), (3, 100), (150, 100)]:
        dg, ug = G.calc_max_timestep(0.3, 0.01, U, step, ramp)
        do, uo = O.calc_max_timestep(0.3, 0.01, U, step, ramp)
        assert ug == uo and dg == pytest.approx(do, rel=1e-15)


def c1_taylor_green(G_or_O, shape, h):
    """C1: u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 on [0,2pi]^3."""
    x, y, z = cell_centres(shape, h)
    return np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z), 0 * x


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_c1_taylor_green_fp64(cup, oracle, decomp):
    """C1 (64^3 TGV, nu=0.01, dt=1e-3, 20 steps) FP64 device vs FP64 oracle <= 1e-11 rel;
    kinetic energy decay agrees."""
    print(f"C1 TGV decomp={decomp}", flush=True)
    bpd = (8, 8, 8)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    u, v, w = c1_taylor_green(None, G.shape, G.h)
    G.set_velocity(u, v, w); O.set_velocity(u, v, w)
    nu, dt = 0.01, 1e-3
    G.run_steps(20, nu=nu, fixed_dt=dt)
    for _ in range(20):
        O.timestep(dt, nu)
    st = G.step_state()
    assert st["step"] == 20 and st["time"] == pytest.approx(20 * dt, rel=1e-14)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11
    assert rel(G.get("p"), O.get("p")) < 1e-10
    kg = G.diagnostics(nu)[18]
    ko = O.dissipation(nu)[18]
    assert kg == pytest.approx(ko, rel=1e-12)


def test_c1_reference_ic_is_a_16_cell_roll(cup, oracle):
    """The reference's TGV kernel with uMax = -1 rolled by 16 cells is C1's field."""
    G = cup.Grid((8, 8, 8), precision=8)
    G.ic_taylor_green(-1.0)
    u, v, w = c1_taylor_green(None, G.shape, G.h)
    gu = np.roll(G.get("u"), (-16, -16, -16), axis=(0, 1, 2))
    gv = np.roll(G.get("v"), (-16, -16, -16), axis=(0, 1, 2))
    assert np.abs(gu - u).max() < 1e-14 and np.abs(gv - v).max() < 1e-14
    O = oracle.OracleSim((8, 8, 8))
    O.ic_taylor_green(-1.0)
    assert np.abs(O.get("u") - G.get("u")).max() < 1e-15


def test_c1_fp32_vs_fp64(cup):
    """FP32 form vs the FP64 device oracle on C1: relative L2 within FP32 accumulation."""
    bpd = (8, 8, 8)
    G8 = cup.Grid(bpd, precision=8, decomposition="plain")
    G4 = cup.Grid(bpd, precision=4, decomposition="fused")
    u, v, w = c1_taylor_green(None, G8.shape, G8.h)
    G8.set_velocity(u, v, w); G4.set_velocity(u, v, w)
    G8.run_steps(20, nu=0.01, fixed_dt=1e-3)
    G4.run_steps(20, nu=0.01, fixed_dt=1e-3)
    for a, b in zip(G4.velocity()[:2], G8.velocity()[:2]):
        assert rel(a, b) < 1e-6


def test_cfl_driven_steps_match(cup, oracle):
    """Device-resident dt (FUSED) follows calcMaxTimestep with ramp like the host loop."""
    bpd = (4, 4, 4)
    rng = np.random.default_rng(9)
    G = cup.Grid(bpd, precision=8, decomposition="fused")
    O = oracle.OracleSim(bpd)
    u, v, w = (0.3 * rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w); O.set_velocity(u, v, w)
    O.pressure_rhs(1.0); O.poisson_solve(); O.grad_p(1.0)   # start divergence free
    G.set_velocity(*O.velocity())
    nu, CFL = 0.02, 0.3
    G.run_steps(6, CFL=CFL, nu=nu, rampup=3)
    t = 0.0
    for step in range(6):
        dt, _ = O.calc_max_timestep(CFL, nu, (0, 0, 0), step, 3)
        O.timestep(dt, nu)
        t += dt
    st = G.step_state()
    assert st["time"] == pytest.approx(t, rel=1e-12)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11
