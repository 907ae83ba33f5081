"""Grids whose cells per axis are not powers of two (8 x blocks with factors 3 and 5, as the
reference's presets have them, e.g. launch/settingsHIT_DNS.sh:7 with BPDX=12, 96^3): the FFT
passes take the run-time radix plan (radix 8, 4, 2, 3, 5 Stockham stages in LDS), which FFTW
gives the reference for any size (PoissonSolverPeriodic.cpp:108-121). Each case against the
CPU oracle, whose mixed-radix transform is pinned by numpy golden vectors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("decomp", ["plain", "fused"])
@pytest.mark.parametrize("bpd", [(3, 2, 5), (6, 5, 3), (12, 3, 2), (5, 12, 4), (2, 4, 10)])
def test_projection_mixed_radix_fp64(cup, oracle, decomp, bpd):
    """RHS, Poisson and gradient on 24 to 96 cells per axis against the oracle."""
    print(f"mixed-radix projection decomp={decomp} bpd={bpd}", flush=True)
    rng = np.random.default_rng(7)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    dt = 1e-3
    O.set_velocity(u, v, w)
    if decomp == "fused":
        G.set("tmpU", u); G.set("tmpV", v); G.set("tmpW", w)
    else:
        G.set_velocity(u, v, w)
    G.pressure_rhs(dt); G.poisson_periodic(); G.pressure_project(dt)
    O.pressure_rhs(dt); O.poisson_solve(); O.grad_p(dt)
    ep = rel(G.get("p"), O.get("p"))
    print(f"  p rel L2 {ep:.2e}", flush=True)
    assert ep < 1e-12
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-13


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_taylor_green_96_matches_oracle(cup, oracle, decomp):
    """A C1-style run on 96^3 (12^3 blocks): the Taylor-Green IC and 20 steps of dt = 1e-3,
    nu = 0.01, against the oracle at 1e-11."""
    bpd, nsteps, dt, nu = (12, 12, 12), 20, 1e-3, 0.01
    print(f"96^3 Taylor-Green {decomp}", flush=True)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    G.ic_taylor_green(-1.0)
    O.ic_taylor_green(-1.0)
    G.run_steps(nsteps, nu=nu, fixed_dt=dt)
    for _ in range(nsteps):
        O.timestep(dt, nu)
    for a, b in zip(G.velocity(), O.velocity()):
        e = rel(a, b)
        print(f"  rel L2 {e:.2e}", flush=True)
        assert e < 1e-11


@pytest.mark.parametrize("prec,tol", [(8, 1e-11), (4, 2e-6)])
def test_forced_turbulence_mixed_radix(cup, oracle, prec, tol):
    """Forced isotropic turbulence on 40 x 48 x 24 cells: the seeded IC and 3 forced CFL steps
    (fused forcing, the z pass with the division on 24 points) against the oracle."""
    bpd = (5, 6, 3)
    print(f"forced mixed-radix precision {prec}", flush=True)
    O = oracle.OracleSim(bpd)
    O.ic_hit(4.0, 1.0, 20260101)
    G = cup.Grid(bpd, precision=prec, decomposition="fused")
    G.ic_hit(4.0, 1.0, 20260101)
    ref = O.velocity()
    if prec == 8:
        for a, b in zip(G.velocity(), ref):
            assert rel(a, b) < 1e-12
    G.set_velocity(*ref)
    tke = 0.5 * float(sum((a * a).mean() for a in ref))
    for step in range(3):
        dt, _ = O.calc_max_timestep(0.3, 0.005, (0, 0, 0), step, 0)
        O.timestep(dt, 0.005, spectral_forcing=True, tke_target=tke)
    G.run_steps(3, CFL=0.3, nu=0.005, rampup=0, spectral_forcing=True, tke_target=tke)
    for n in ("u", "v", "w"):
        e = rel(G.get(n), O.get(n))
        print(f"  {n} rel L2 {e:.2e}", flush=True)
        assert e < tol
