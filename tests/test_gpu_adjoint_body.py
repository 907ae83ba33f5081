"""The adjoint through penalization (cup_adjoint_step_body, the PenalizedRollout built on
it) against the oracle's (pinned by the dot-product identity with central differences of
the forward step with the sphere), and, at the 256 x 128 x 128 sphere case, the gradient of
the time-averaged drag with respect to the body velocity against finite differences of the
FP64 forward."""
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


def test_mean_drag_velocity_gradient_vs_finite_differences(cup):
    """Sphere D = 0.2 at (0.5, 0.5, 0.5) in a periodic 2 x 1 x 1 box at 256 x 128 x 128, inflow
    U = 1, Re = 200, implicit penalization lambda = 1e6, 100 steps of dt = 1e-3 from rest,
    checkpoints every 10: the gradient of the time-averaged drag with respect to the body
    velocity, FP32 and FP64 adjoints against fourth-order central differences (eps = 1e-3)
    of the FP64 forward."""
    import torch
    from cubismup_3d_amd.adjoint import mean_drag, recompute_memory_bytes
    bpd, ext, R, c = (32, 16, 16), 2.0, 0.1, (0.5, 0.5, 0.5)
    U, nu, nsteps, every = (1.0, 0.0, 0.0), 1.0 * 2 * R / 200, 100, 10
    V0 = np.array([-1.0, 0.0, 0.0])
    dts = [1e-3] * nsteps
    grads, vals = {}, {}
    for prec in (4, 8):
        G = cup.Grid(bpd, maxextent=ext, precision=prec, decomposition="fused")
        G.chi_from_sphere(c, R)
        dtype = torch.float32 if prec == 4 else torch.float64
        u0 = torch.zeros((3,) + G.shape, dtype=dtype, device="cuda", requires_grad=True)
        V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
        J = mean_drag(u0, V, G, nsteps, nu, U, dts, every=every)
        J.backward()
        grads[prec], vals[prec] = V.grad.numpy().copy(), J.item()
        print(f"precision {prec}: mean drag {J.item():.8e} (Cd {J.item() / (0.5 * np.pi * R * R):.5f}), "
              f"dJ/dV {grads[prec]}, recompute memory {recompute_memory_bytes(G, nsteps, every) / 2**20:.0f} MiB",
              flush=True)
    G8 = cup.Grid(bpd, maxextent=ext, precision=8, decomposition="fused")
    G8.chi_from_sphere(c, R)

    def J64(V):
        G8.set_velocity(*(np.zeros(G8.shape) for _ in range(3)))
        sph = dict(lambda_=1e6, implicit=True, trans_vel=tuple(V))
        J = 0.0
        for dt in dts:
            G8.run_steps(1, nu=nu, uinf=U, fixed_dt=dt, sphere=sph)
            J += dt / sum(dts) * G8.last_body_force()[0]
        return J
    eps = 1e-3
    for k in range(3):
        e = np.eye(3)[k] * eps
        fd = (8 * (J64(V0 + e) - J64(V0 - e)) - (J64(V0 + 2 * e) - J64(V0 - 2 * e))) / (12 * eps)
        scale = max(abs(fd), 1e-3 * abs(vals[8]))
        print(f"dJ/dV_{k}: fd {fd:.10e} fp64 adjoint {grads[8][k]:.10e} fp32 adjoint {grads[4][k]:.10e}", flush=True)
        assert abs(grads[8][k] - fd) < 1e-6 * scale
        assert abs(grads[4][k] - fd) < 1e-3 * scale
