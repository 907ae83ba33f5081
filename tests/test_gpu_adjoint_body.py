"""The adjoint through penalization (cup_adjoint_step_body and its general form
cup_adjoint_step_body_full, the PenalizedRollout built on them) against the oracle's (pinned
by the dot-product identity with central differences of the forward step with the sphere),
and, at the 256 x 128 x 128 sphere case, the gradient of the time-averaged drag with respect
to the body velocity and centre (at 128 x 64 x 64, its rotation) against finite differences
of the FP64 forward."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_adjoint_step_body_matches_oracle(cup, oracle, decomp):
    print(f"adjoint step with sphere decomp={decomp}", flush=True)
    rng = np.random.default_rng(41)
    bpd, ext = (4, 2, 2), 2.0
    G = cup.Grid(bpd, maxextent=ext, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd, maxextent=ext)
    G.chi_from_sphere((0.9, 0.5, 0.5), 0.3)
    O.sphere_create(0.3, (0.9, 0.5, 0.5))
    shape = G.shape
    U, V = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0)
    u = tuple(0.3 * rng.standard_normal(shape) + (1.0 if q == 0 else 0.0) for q in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w = 0.2 * G.h, 1e-3, 3.7
    sph = dict(lambda_=1e6, implicit=True, trans_vel=V)
    G.set_velocity(*u); O.set_velocity(*u)
    for n, f in zip(("adjU", "adjV", "adjW"), lam):
        G.set(n, f)
    dv = G.adjoint_step_body(dt, nu, U, sph, w)
    lam0, dvo = O.adjoint_step_body(dt, nu, lam, V, 1e6, True, w, U)
    for n, b in zip(("adjU", "adjV", "adjW"), lam0):
        assert rel(G.get(n), b) < 1e-11
    np.testing.assert_allclose(dv, dvo, rtol=1e-11, atol=1e-12 * np.abs(dvo).max())


@pytest.mark.parametrize("decomp,mollify", [("plain", 1.0), ("fused", 1.0), ("fused", 2.0)])
def test_adjoint_step_body_centre_matches_oracle(cup, oracle, decomp, mollify):
    print(f"adjoint step with the mollified sphere decomp={decomp} mollify={mollify}", flush=True)
    rng = np.random.default_rng(43)
    bpd, ext, c0 = (4, 2, 2), 2.0, (0.913, 0.487, 0.502)
    G = cup.Grid(bpd, maxextent=ext, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd, maxextent=ext)
    G.chi_from_sphere(c0, 0.3, mollify=mollify)
    O.sphere_create(0.3, c0, mollify=mollify)
    np.testing.assert_allclose(G.get("chi"), O.get("chi"), rtol=0, atol=1e-15)
    shape = G.shape
    U, V = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0)
    u = tuple(0.3 * rng.standard_normal(shape) + (1.0 if q == 0 else 0.0) for q in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w = 0.2 * G.h, 1e-3, 3.7
    sph = dict(lambda_=1e6, implicit=True, trans_vel=V)
    G.set_velocity(*u); O.set_velocity(*u)
    for n, f in zip(("adjU", "adjV", "adjW"), lam):
        G.set(n, f)
    dv, dc = G.adjoint_step_body(dt, nu, U, sph, w, centre=True)
    lam0, dvo, dco = O.adjoint_step_body(dt, nu, lam, V, 1e6, True, w, U, centre=True)
    for n, b in zip(("adjU", "adjV", "adjW"), lam0):
        assert rel(G.get(n), b) < 1e-11
    np.testing.assert_allclose(dv, dvo, rtol=1e-11, atol=1e-12 * np.abs(dvo).max())
    print(f"dJ/dc gpu {dc} oracle {dco}", flush=True)
    np.testing.assert_allclose(dc, dco, rtol=1e-10, atol=1e-11 * np.abs(dco).max())
    # the velocity step is unchanged by asking for the centre
    G.set_velocity(*u)
    for n, f in zip(("adjU", "adjV", "adjW"), lam):
        G.set(n, f)
    dv2 = G.adjoint_step_body(dt, nu, U, sph, w)
    for n, b in zip(("adjU", "adjV", "adjW"), lam0):
        assert rel(G.get(n), b) < 1e-11
    np.testing.assert_allclose(dv2, dvo, rtol=1e-11, atol=1e-12 * np.abs(dvo).max())


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_adjoint_step_rotating_body_matches_oracle(cup, oracle, decomp):
    """The sphere rotating at W about its centre of mass: lam_n, dJ/dV and dJ/dW against the
    oracle's (pinned by central differences of its forward step in u, V and W)."""
    print(f"adjoint step with a rotating sphere decomp={decomp}", flush=True)
    rng = np.random.default_rng(47)
    bpd, ext = (4, 2, 2), 2.0
    G = cup.Grid(bpd, maxextent=ext, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd, maxextent=ext)
    G.chi_from_sphere((0.9, 0.5, 0.5), 0.3)
    O.sphere_create(0.3, (0.9, 0.5, 0.5))
    shape = G.shape
    U, V, W = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0), (0.7, -1.3, 2.1)
    u = tuple(0.3 * rng.standard_normal(shape) + (1.0 if q == 0 else 0.0) for q in range(3))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w = 0.2 * G.h, 1e-3, 3.7
    sph = dict(lambda_=1e6, implicit=True, trans_vel=V, ang_vel=W)
    G.set_velocity(*u); O.set_velocity(*u)
    for n, f in zip(("adjU", "adjV", "adjW"), lam):
        G.set(n, f)
    dv, dw = G.adjoint_step_body(dt, nu, U, sph, w, rotation=True)
    lam0, dvo, dwo = O.adjoint_step_body(dt, nu, lam, V, 1e6, True, w, U, ang_vel=W)
    for n, b in zip(("adjU", "adjV", "adjW"), lam0):
        assert rel(G.get(n), b) < 1e-11
    np.testing.assert_allclose(dv, dvo, rtol=1e-11, atol=1e-12 * np.abs(dvo).max())
    print(f"dJ/dW gpu {dw} oracle {dwo}", flush=True)
    np.testing.assert_allclose(dw, dwo, rtol=1e-10, atol=1e-11 * np.abs(dwo).max())
    with pytest.raises(ValueError, match="without rotation"):
        G.adjoint_step_body(dt, nu, U, sph, w, centre=True)


@pytest.mark.parametrize("decomp,scheme,target,rate", [("fused", 1, 0.0, 0.0), ("fused", 2, 0.0, 0.0),
                                                        ("plain", 2, 0.0, 0.0), ("fused", 0, 0.0, 0.4),
                                                        ("plain", 0, 0.0, 0.4), ("fused", 2, 1.5, 0.0)])
def test_step_and_adjoint_body_scheme_forcing_match_oracle(cup, oracle, decomp, scheme, target, rate):
    """The step with the rotating sphere, the upwind schemes and the linear forcing: the forward
    step (run_steps) and its adjoint (lam_n, dJ/dV, dJ/dW) against the oracle's, which central
    differences of its forward step pin (2pi x pi x pi box, where the forcing's band holds
    modes; velocity means (1, 0.8, -0.9) as in the oracle's check)."""
    print(f"body step and adjoint decomp={decomp} scheme={scheme} target={target} rate={rate}", flush=True)
    rng = np.random.default_rng(53 + scheme)
    bpd, ext = (4, 2, 2), 2 * np.pi
    centre, R = tuple(np.pi * np.array([0.9, 0.5, 0.5])), 0.3 * np.pi
    G = cup.Grid(bpd, maxextent=ext, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd, maxextent=ext)
    G.chi_from_sphere(centre, R)
    O.sphere_create(R, centre)
    shape = G.shape
    U, V, W = (1.0, 0.0, 0.0), (-1.0, 0.1, 0.0), (0.3, -0.4, 0.8)
    u = tuple(0.3 * rng.standard_normal(shape) + m for m in (1.0, 0.8, -0.9))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, w, lam_ = 0.2 * G.h, 1e-3, 3.7, 1e5
    sph = dict(lambda_=lam_, implicit=True, trans_vel=V, ang_vel=W)
    fkw = dict(spectral_forcing=target > 0 or rate > 0, tke_target=target, injection_rate=rate)
    # the forward step
    G.set_velocity(*u); O.set_velocity(*u)
    G.run_steps(1, nu=nu, uinf=U, fixed_dt=dt, sphere=sph, scheme=scheme, **fkw)
    out = O.timestep(dt, nu, uinf=U, sphere=sph, scheme=scheme, **fkw)
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11
    np.testing.assert_allclose(G.last_body_force(), out["penal_ft"], rtol=1e-10, atol=1e-12)
    # its adjoint
    G.set_velocity(*u); O.set_velocity(*u)
    for n, f in zip(("adjU", "adjV", "adjW"), lam):
        G.set(n, f)
    dv, dw = G.adjoint_step_body(dt, nu, U, sph, w, rotation=True, scheme=scheme, tke_target=target,
                                 injection_rate=rate)
    lam0, dvo, dwo = O.adjoint_step_body(dt, nu, lam, V, lam_, True, w, U, ang_vel=W, scheme=scheme,
                                         tke_target=target, injection_rate=rate)
    for n, b in zip(("adjU", "adjV", "adjW"), lam0):
        assert rel(G.get(n), b) < 1e-11
    for a, b in zip(G.velocity(), u):  # u_n is left in place
        assert np.array_equal(a, b)
    np.testing.assert_allclose(dv, dvo, rtol=1e-11, atol=1e-12 * np.abs(dvo).max())
    np.testing.assert_allclose(dw, dwo, rtol=1e-10, atol=1e-11 * np.abs(dwo).max())


TWO_BODIES = [(0.2 * np.pi, (0.5 * np.pi, 0.5 * np.pi, 0.5 * np.pi)),
              (0.2 * np.pi, (1.5 * np.pi, 0.45 * np.pi, 0.55 * np.pi))]


@pytest.mark.parametrize("decomp,scheme,rate", [("plain", 0, 0.0), ("fused", 0, 0.0), ("fused", 2, 0.4)])
def test_step_and_adjoint_two_bodies_match_oracle(cup, oracle, decomp, scheme, rate):
    """Two spheres with their own velocities and rotations (body 1's through cup_body_velocity):
    the forward step (run_steps), each body's penalization force, and the adjoint with both
    (lam_n, each body's dJ/dV and dJ/dW) against the oracle's, which central differences of its
    forward step pin."""
    print(f"two-body step and adjoint decomp={decomp} scheme={scheme} rate={rate}", flush=True)
    rng = np.random.default_rng(59 + scheme)
    bpd, ext = (4, 2, 2), 2 * np.pi
    G = cup.Grid(bpd, maxextent=ext, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd, maxextent=ext)
    for R, c in TWO_BODIES:
        G.add_sphere(c, R)
        O.add_sphere(R, c)
    shape = G.shape
    U = (1.0, 0.0, 0.0)
    Vs, Ws, w = [(-1.0, 0.1, 0.0), (-0.8, 0.0, 0.2)], [(0.3, -0.4, 0.8), (-0.5, 0.2, 0.6)], [3.7, -1.9]
    u = tuple(0.3 * rng.standard_normal(shape) + m for m in (1.0, 0.8, -0.9))
    lam = tuple(rng.standard_normal(shape) for _ in range(3))
    dt, nu, lam_ = 0.2 * G.h, 1e-3, 1e5
    fkw = dict(spectral_forcing=rate > 0, injection_rate=rate)
    sph = dict(lambda_=lam_, implicit=True, trans_vel=Vs[0], ang_vel=Ws[0])
    G.body_velocity(1, Vs[1], Ws[1])
    G.set_velocity(*u); O.set_velocity(*u)
    G.run_steps(1, nu=nu, uinf=U, fixed_dt=dt, sphere=sph, scheme=scheme, **fkw)
    O.advect_diffuse(dt, nu, U, scheme)
    if rate > 0:
        O.spectral_forcing(dt, nu, injection_rate=rate)
    O.pressure_rhs(dt)
    O.poisson()
    O.grad_p(dt)
    for k in range(2):
        with O.body(k):
            ft = O.penalize(dt, lam_, True, Vs[k], Ws[k])
        np.testing.assert_allclose(G.last_body_force_body(k), ft, rtol=1e-10, atol=1e-12 * np.abs(ft).max())
    for a, b in zip(G.velocity(), O.velocity()):
        assert rel(a, b) < 1e-11
    G.set_velocity(*u); O.set_velocity(*u)
    for n, f in zip(("adjU", "adjV", "adjW"), lam):
        G.set(n, f)
    bodies = [dict(trans_vel=Vs[k], ang_vel=Ws[k]) for k in range(2)]
    dv, dw = G.adjoint_step_bodies(dt, nu, U, bodies, lam_, True, w, scheme=scheme, injection_rate=rate)
    lam0, dvo, dwo = O.adjoint_step_bodies(dt, nu, lam, Vs, Ws, lam_, True, w, U, scheme=scheme, injection_rate=rate)
    for n, b in zip(("adjU", "adjV", "adjW"), lam0):
        assert rel(G.get(n), b) < 1e-11
    print(f"dJ/dV gpu {dv.tolist()} oracle {dvo.tolist()}", flush=True)
    np.testing.assert_allclose(dv, dvo, rtol=1e-11, atol=1e-12 * np.abs(dvo).max())
    np.testing.assert_allclose(dw, dwo, rtol=1e-10, atol=1e-11 * np.abs(dwo).max())
    with pytest.raises(RuntimeError, match="one body only"):
        G.adjoint_step_body(dt, nu, U, sph, 1.0)
    with pytest.raises(ValueError, match="number of bodies"):
        G.adjoint_step_bodies(dt, nu, U, bodies[:1], lam_, True, w[:1])


def test_bodies_drag_gradient_vs_finite_differences(cup):
    """Two spheres D = 0.2 in a periodic 2 x 1 x 1 box at 128 x 64 x 64 (the second behind and
    beside the first), inflow U = 1, Re = 200, lambda = 1e6, 30 steps of dt = 2e-3 from rest,
    checkpoints every 10, J = sum_n (dt/T) (F_x^0 + 0.5 F_x^1): every body's dJ/dV and dJ/dW,
    FP32 and FP64 adjoints, against central differences (eps = 1e-4, below the masking's kinks)
    of the FP64 forward."""
    import torch
    from cubismup_3d_amd.adjoint import bodies_drag
    bpd, ext, R = (16, 8, 8), 2.0, 0.1
    cs = [(0.5, 0.4988, 0.4999), (1.2, 0.62, 0.47)]
    U, nu, nsteps, every = (1.0, 0.0, 0.0), 1.0 * 2 * R / 200, 30, 10
    V0 = np.array([[-1.0, 0.0, 0.0], [-0.9, 0.1, 0.0]])
    W0 = np.array([[0.5, -1.0, 4.0], [0.0, 2.0, -1.0]])
    weights = [1.0, 0.5]
    dts = [2e-3] * nsteps

    def grid(prec):
        G = cup.Grid(bpd, maxextent=ext, precision=prec, decomposition="fused")
        for c in cs:
            G.add_sphere(c, R)
        return G
    gv, gw = {}, {}
    for prec in (4, 8):
        G = grid(prec)
        dtype = torch.float32 if prec == 4 else torch.float64
        u0 = torch.zeros((3,) + G.shape, dtype=dtype, device="cuda", requires_grad=True)
        V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
        W = torch.tensor(W0, dtype=torch.float64, requires_grad=True)
        J = bodies_drag(u0, V, W, G, nsteps, nu, U, dts, weights=weights, every=every)
        J.backward()
        gv[prec], gw[prec] = V.grad.numpy().copy(), W.grad.numpy().copy()
        print(f"precision {prec}: J {J.item():.8e}, dJ/dV {gv[prec].tolist()}, dJ/dW {gw[prec].tolist()}", flush=True)
    G8 = grid(8)

    def J64(V, W):
        G8.set_velocity(*(np.zeros(G8.shape) for _ in range(3)))
        G8.body_velocity(1, V[1], W[1])
        sph = dict(lambda_=1e6, implicit=True, trans_vel=tuple(V[0]), ang_vel=tuple(W[0]))
        J = 0.0
        for dt in dts:
            G8.run_steps(1, nu=nu, uinf=U, fixed_dt=dt, sphere=sph)
            J += dt / sum(dts) * sum(weights[k] * G8.last_body_force_body(k)[0] for k in range(2))
        return J
    eps = 1e-4
    for name, P0, grads in (("V", V0, gv), ("W", W0, gw)):
        scale = np.linalg.norm(grads[8])
        assert scale > 0
        for k in range(2):
            for j in range(3):
                e = np.zeros((2, 3))
                e[k, j] = eps
                fd = ((J64(V0 + e, W0) - J64(V0 - e, W0)) if name == "V" else
                      (J64(V0, W0 + e) - J64(V0, W0 - e))) / (2 * eps)
                print(f"dJ/d{name}[{k}][{j}]: fd {fd:.10e} fp64 {grads[8][k, j]:.10e} "
                      f"({abs(grads[8][k, j] - fd) / scale:.1e}) fp32 {grads[4][k, j]:.10e} "
                      f"({abs(grads[4][k, j] - fd) / scale:.1e} of |dJ/d{name}|)", flush=True)
                assert abs(grads[8][k, j] - fd) < 1e-6 * scale
                assert abs(grads[4][k, j] - fd) < 1e-4 * scale


def test_mean_drag_rotation_gradient_vs_finite_differences(cup):
    """Sphere D = 0.2 rotating at W = (0.5, -1, 4) about its centre of mass in a periodic
    2 x 1 x 1 box at 128 x 64 x 64, inflow U = 1, Re = 200, lambda = 1e6, 40 steps of
    dt = 2e-3 from rest, checkpoints every 10: dJ/dW of the time-averaged drag, FP32 and FP64
    adjoints against central differences (eps = 1e-4) of the FP64 forward. The x drag is
    nearly even in W (a sphere's spin changes it at second order), so dJ/dW is small, about
    3e-4 of dJ/dV, and J has kinks where the masking's sign pattern changes along the
    trajectory: steps of 1e-2 and above cross them and miss by up to 1e-3 of |dJ/dW|, steps
    of 3e-4 and below agree to 1e-6, rounding keeping 1e-4's error near 1e-8."""
    import torch
    from cubismup_3d_amd.adjoint import mean_drag
    bpd, ext, R = (16, 8, 8), 2.0, 0.1
    c0 = (0.4982, 0.4988, 0.4999)
    U, nu, nsteps, every = (1.0, 0.0, 0.0), 1.0 * 2 * R / 200, 40, 10
    V0, W0 = np.array([-1.0, 0.0, 0.0]), np.array([0.5, -1.0, 4.0])
    dts = [2e-3] * nsteps
    gradw, gradv = {}, {}
    for prec in (4, 8):
        G = cup.Grid(bpd, maxextent=ext, precision=prec, decomposition="fused")
        G.chi_from_sphere(c0, R)
        dtype = torch.float32 if prec == 4 else torch.float64
        u0 = torch.zeros((3,) + G.shape, dtype=dtype, device="cuda", requires_grad=True)
        V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
        W = torch.tensor(W0, dtype=torch.float64, requires_grad=True)
        J = mean_drag(u0, V, G, nsteps, nu, U, dts, every=every, angvel=W)
        J.backward()
        gradw[prec], gradv[prec] = W.grad.numpy().copy(), V.grad.numpy().copy()
        print(f"precision {prec}: mean drag {J.item():.8e}, dJ/dV {gradv[prec]}, dJ/dW {gradw[prec]}", flush=True)
    G8 = cup.Grid(bpd, maxextent=ext, precision=8, decomposition="fused")
    G8.chi_from_sphere(c0, R)

    def J64(Wv):
        G8.set_velocity(*(np.zeros(G8.shape) for _ in range(3)))
        sph = dict(lambda_=1e6, implicit=True, trans_vel=tuple(V0), ang_vel=tuple(Wv))
        J = 0.0
        for dt in dts:
            G8.run_steps(1, nu=nu, uinf=U, fixed_dt=dt, sphere=sph)
            J += dt / sum(dts) * G8.last_body_force()[0]
        return J
    eps = 1e-4
    wscale = np.linalg.norm(gradw[8])
    assert wscale > 0
    for k in range(3):
        e = np.eye(3)[k] * eps
        fd = (J64(W0 + e) - J64(W0 - e)) / (2 * eps)
        print(f"dJ/dW_{k}: fd {fd:.10e} fp64 adjoint {gradw[8][k]:.10e} ({abs(gradw[8][k] - fd) / wscale:.1e}) "
              f"fp32 adjoint {gradw[4][k]:.10e} ({abs(gradw[4][k] - fd) / wscale:.1e} of |dJ/dW|)", flush=True)
        assert abs(gradw[8][k] - fd) < 1e-6 * wscale
        assert abs(gradw[4][k] - fd) < 1e-4 * wscale


def test_centre_gradient_needs_mollified_chi(cup):
    G = cup.Grid((2, 2, 2), maxextent=1.0, precision=8, decomposition="fused")
    G.chi_from_sphere((0.5, 0.5, 0.5), 0.2)
    G.set_velocity(*(np.zeros(G.shape) for _ in range(3)))
    for n in ("adjU", "adjV", "adjW"):
        G.set(n, np.zeros(G.shape))
    with pytest.raises(RuntimeError, match="mollified"):
        G.adjoint_step_body(1e-3, 1e-3, (1, 0, 0), dict(lambda_=1e6, implicit=True, trans_vel=(0, 0, 0)), 1.0,
                            centre=True)


def test_last_body_force_matches_oracle(cup, oracle):
    rng = np.random.default_rng(5)
    bpd, ext = (4, 2, 2), 2.0
    G = cup.Grid(bpd, maxextent=ext, precision=8, decomposition="fused")
    O = oracle.OracleSim(bpd, maxextent=ext)
    G.chi_from_sphere((0.9, 0.5, 0.5), 0.3)
    O.sphere_create(0.3, (0.9, 0.5, 0.5))
    u = tuple(0.2 * rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(*u); O.set_velocity(*u)
    sph = dict(lambda_=1e6, implicit=True, trans_vel=(-1.0, 0.0, 0.0))
    dt = 0.2 * G.h
    G.run_steps(1, nu=1e-3, uinf=(1, 0, 0), fixed_dt=dt, sphere=sph)
    out = O.timestep(dt, 1e-3, uinf=(1, 0, 0), sphere=sph)
    np.testing.assert_allclose(G.last_body_force(), out["penal_ft"], rtol=1e-11, atol=1e-13)


def test_mean_drag_gradient_vs_finite_differences(cup):
    """Sphere D = 0.2 in a periodic 2 x 1 x 1 box at 256 x 128 x 128, mollified chi (half-width
    one cell), inflow U = 1, Re = 200, implicit penalization lambda = 1e6, 100 steps of
    dt = 1e-3 from rest, checkpoints every 10: the gradient of the time-averaged drag with
    respect to the body velocity and the sphere centre, FP32 and FP64 adjoints against
    central differences of the FP64 forward (velocity: fourth order, eps = 1e-3; centre:
    second order, eps = 5e-7, every cell of the chi band staying more than 10 eps away from
    its edges |x - c| = R +- h, where a cell's masking flags change)."""
    import torch
    from cubismup_3d_amd.adjoint import mean_drag, recompute_memory_bytes
    bpd, ext, R, moll = (32, 16, 16), 2.0, 0.1, 1.0
    c0 = np.array([0.4982, 0.4988, 0.4999])
    U, nu, nsteps, every = (1.0, 0.0, 0.0), 1.0 * 2 * R / 200, 100, 10
    V0 = np.array([-1.0, 0.0, 0.0])
    dts = [1e-3] * nsteps
    h = ext / 256
    ax = [(np.arange(n) + 0.5) * h for n in (256, 128, 128)]
    sl = [a[(a > ci - R - 3 * h) & (a < ci + R + 3 * h)] for a, ci in zip(ax, c0)]
    X, Y, Z = np.meshgrid(*sl, indexing="ij")
    r = np.sqrt((X - c0[0]) ** 2 + (Y - c0[1]) ** 2 + (Z - c0[2]) ** 2)
    ec = 5e-7
    assert min(np.abs(r - (R + moll * h)).min(), np.abs(r - (R - moll * h)).min()) > 10 * ec
    grads, gradc, vals = {}, {}, {}
    for prec in (4, 8):
        G = cup.Grid(bpd, maxextent=ext, precision=prec, decomposition="fused")
        dtype = torch.float32 if prec == 4 else torch.float64
        u0 = torch.zeros((3,) + G.shape, dtype=dtype, device="cuda", requires_grad=True)
        V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
        C = torch.tensor(c0, dtype=torch.float64, requires_grad=True)
        J = mean_drag(u0, V, G, nsteps, nu, U, dts, every=every, centre=C, radius=R, mollify=moll)
        J.backward()
        grads[prec], gradc[prec], vals[prec] = V.grad.numpy().copy(), C.grad.numpy().copy(), J.item()
        print(f"precision {prec}: mean drag {J.item():.8e} (Cd {J.item() / (0.5 * np.pi * R * R):.5f}), "
              f"dJ/dV {grads[prec]}, dJ/dc {gradc[prec]}, "
              f"recompute memory {recompute_memory_bytes(G, nsteps, every) / 2**20:.0f} MiB", flush=True)
    G8 = cup.Grid(bpd, maxextent=ext, precision=8, decomposition="fused")

    def J64(V, c):
        G8.chi_from_sphere(tuple(c), R, mollify=moll)
        G8.set_velocity(*(np.zeros(G8.shape) for _ in range(3)))
        sph = dict(lambda_=1e6, implicit=True, trans_vel=tuple(V))
        J = 0.0
        for dt in dts:
            G8.run_steps(1, nu=nu, uinf=U, fixed_dt=dt, sphere=sph)
            J += dt / sum(dts) * G8.last_body_force()[0]
        return J
    assert abs(J64(V0, c0) - vals[8]) < 1e-12 * abs(vals[8])
    # each component against the gradient's norm: the finite differences carry an absolute
    # error of about 1e-8 (the masking's sign pattern changes along the trajectory, which
    # puts kinks in J), the size of the small components themselves
    eps = 1e-3
    vscale = np.linalg.norm(grads[8])
    for k in range(3):
        e = np.eye(3)[k] * eps
        fd = (8 * (J64(V0 + e, c0) - J64(V0 - e, c0)) - (J64(V0 + 2 * e, c0) - J64(V0 - 2 * e, c0))) / (12 * eps)
        print(f"dJ/dV_{k}: fd {fd:.10e} fp64 adjoint {grads[8][k]:.10e} ({abs(grads[8][k] - fd) / vscale:.1e}) "
              f"fp32 adjoint {grads[4][k]:.10e} ({abs(grads[4][k] - fd) / vscale:.1e} of |dJ/dV|)", flush=True)
        assert abs(grads[8][k] - fd) < 1e-6 * vscale
        assert abs(grads[4][k] - fd) < 1e-4 * vscale
    cscale = np.linalg.norm(gradc[8])
    for k in range(3):
        e = np.eye(3)[k] * ec
        fd = (J64(V0, c0 + e) - J64(V0, c0 - e)) / (2 * ec)
        print(f"dJ/dc_{k}: fd {fd:.10e} fp64 adjoint {gradc[8][k]:.10e} ({abs(gradc[8][k] - fd) / cscale:.1e}) "
              f"fp32 adjoint {gradc[4][k]:.10e} ({abs(gradc[4][k] - fd) / cscale:.1e} of |dJ/dc|)", flush=True)
        assert abs(gradc[8][k] - fd) < 1e-5 * cscale
        assert abs(gradc[4][k] - fd) < 1e-4 * cscale
