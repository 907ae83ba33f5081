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
