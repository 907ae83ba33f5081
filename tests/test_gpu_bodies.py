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
    with pytest.raises(cup.CupError):  # the device-resident loop takes one body
        G.add_sphere((1.5, 0.5, 0.5), 0.1)
        G.run_steps(1, nu=1e-3, fixed_dt=1e-3, sphere=dict(lambda_=1e6, implicit=True, trans_vel=(0, 0, 0)))
