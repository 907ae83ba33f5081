"""The third-order upwind advection scheme (Euler and explicit-midpoint RK2) on the device
against the oracle (whose scheme is pinned by the exact single-mode amplification factor):
one advection-diffusion call, CFL-driven steps with forcing in both decompositions, steps
with a penalized sphere, and the FP32 form against the FP64 form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CUP_UPWIND3_EULER, CUP_UPWIND3_RK2 = 1, 2


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.parametrize("decomp", ["plain", "fused"])
@pytest.mark.parametrize("scheme", [CUP_UPWIND3_EULER, CUP_UPWIND3_RK2])
def test_upwind3_advdiff_matches_oracle(cup, oracle, decomp, scheme):
    print(f"upwind3 advdiff decomp={decomp} scheme={scheme}", flush=True)
    rng = np.random.default_rng(19)
    bpd = (4, 2, 1)  # one block along z: the z halo wraps onto the block itself
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    u, v, w = (rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(u, v, w); O.set_velocity(u, v, w)
    U = (0.3, -0.2, 0.1)
    dt, nu = 0.1 * G.h, 0.01
    G.advect_diffuse(dt, nu, U, scheme=scheme)
    O.advect_diffuse(dt, nu, U, scheme=scheme)
    ustar = ("u", "v", "w") if decomp == "plain" or scheme == CUP_UPWIND3_RK2 else ("tmpU", "tmpV", "tmpW")
    for n, b in zip(ustar, O.velocity()):
        assert rel(G.get(n), b) < 1e-14
    if decomp == "plain":
        for n, b in zip(("tmpU", "tmpV", "tmpW"), O.velocity()):
            assert rel(G.get(n), b) < 1e-14


@pytest.mark.parametrize("decomp", ["plain", "fused"])
@pytest.mark.parametrize("scheme", [CUP_UPWIND3_EULER, CUP_UPWIND3_RK2])
def test_upwind3_forced_steps_match_oracle(cup, oracle, decomp, scheme):
    """Forced isotropic turbulence, CFL-driven, 6 steps on 32^3."""
    print(f"upwind3 forced steps decomp={decomp} scheme={scheme}", flush=True)
    bpd = (4, 4, 4)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    O.ic_hit(4.0, 1.0, 20260101)
    G.set_velocity(*O.velocity())
    u, v, w = O.velocity()
    tgt = 1.05 * 0.5 * (u ** 2 + v ** 2 + w ** 2).mean()
    nu, CFL = 0.005, 0.3
    G.run_steps(6, CFL=CFL, nu=nu, spectral_forcing=True, tke_target=tgt, scheme=scheme)
    t = 0
    for step in range(6):
        dt, _ = O.calc_max_timestep(CFL, nu, (0, 0, 0), step, 0)
        O.timestep(dt, nu, spectral_forcing=True, tke_target=tgt, scheme=scheme)
        t += dt
    assert G.step_state()["time"] == pytest.approx(t, rel=1e-11)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_upwind3_rk2_sphere_steps_match_oracle(cup, oracle, decomp):
    print(f"upwind3 rk2 sphere steps decomp={decomp}", flush=True)
    bpd = (8, 4, 4)
    G = cup.Grid(bpd, precision=8, maxextent=2.0, decomposition=decomp)
    O = oracle.OracleSim(bpd, maxextent=2.0)
    c, R = (0.5, 0.5, 0.5), 0.15
    G.chi_from_sphere(c, R); O.sphere_create(R, c)
    nu, dt, U = 1e-3, 2e-3, (1.0, 0.0, 0.0)
    sph = dict(lambda_=1e6, implicit=True, trans_vel=(-1.0, 0.0, 0.0))
    for _ in range(4):
        O.timestep(dt, nu, uinf=U, sphere=sph, scheme=CUP_UPWIND3_RK2)
    G.run_steps(4, nu=nu, uinf=U, fixed_dt=dt, sphere=sph, scheme=CUP_UPWIND3_RK2)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11
    qg = G.surface_forces(nu, sph["trans_vel"])
    qo = O.surface_forces(nu, sph["trans_vel"])
    assert qg[15] == pytest.approx(qo[15], rel=1e-9)


def test_upwind3_rk2_fp32_vs_fp64(cup):
    bpd = (4, 4, 4)
    G8 = cup.Grid(bpd, precision=8, decomposition="fused")
    G4 = cup.Grid(bpd, precision=4, decomposition="fused")
    G8.ic_hit(4.0, 1.0, 3)
    G4.set_velocity(*G8.velocity())
    tgt = G8.diagnostics(0.005)[18] / (2 * np.pi) ** 3
    for G in (G8, G4):
        G.run_steps(10, CFL=0.3, nu=0.005, spectral_forcing=True, tke_target=tgt, scheme=CUP_UPWIND3_RK2)
    for a, b in zip(G4.velocity(), G8.velocity()):
        assert rel(a, b) < 2e-6
