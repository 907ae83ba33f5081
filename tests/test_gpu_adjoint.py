"""The hand-written adjoint of the step (cup_adjoint_step) and the torch rollout built on
it, against the oracle's adjoint (itself pinned by the dot-product identity with the exact
central difference of the quadratic step) and, at the 128^3 isotropic-turbulence size,
against central finite differences of the FP64 forward along seeded solenoidal directions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def set_adjoint(G, lam):
    for n, f in zip(("adjU", "adjV", "adjW"), lam):
        G.set(n, f)


def get_adjoint(G):
    return tuple(G.get(n) for n in ("adjU", "adjV", "adjW"))


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_adjoint_step_matches_oracle(cup, oracle, decomp):
    print(f"adjoint step decomp={decomp}", flush=True)
    rng = np.random.default_rng(31)
    bpd = (4, 2, 2)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    u = tuple(rng.standard_normal(G.shape) for _ in range(3))
    lam = tuple(rng.standard_normal(G.shape) for _ in range(3))
    U, dt, nu = (0.3, -0.2, 0.1), 0.3 * G.h, 0.02
    G.set_velocity(*u); O.set_velocity(*u)
    set_adjoint(G, lam)
    G.adjoint_step(dt, nu, U)
    exp = O.adjoint_step(dt, nu, lam, U)
    for a, b in zip(get_adjoint(G), exp):
        assert rel(a, b) < 1e-12
    for a, b in zip(G.velocity(), u):
        assert np.array_equal(a, b)


def test_adjoint_step_fp32_vs_fp64(cup):
    rng = np.random.default_rng(8)
    bpd = (4, 4, 4)
    u = tuple(rng.standard_normal((32, 32, 32)) for _ in range(3))
    lam = tuple(rng.standard_normal((32, 32, 32)) for _ in range(3))
    out = []
    for prec in (8, 4):
        G = cup.Grid(bpd, precision=prec, decomposition="fused")
        G.set_velocity(*u)
        set_adjoint(G, lam)
        G.adjoint_step(0.2 * G.h, 0.01, (0.1, 0.0, 0.0))
        out.append(get_adjoint(G))
    for a, b in zip(out[1], out[0]):
        assert rel(a, b) < 1e-6


def test_adjoint_rejects_other_schemes_and_bodies(cup):
    G = cup.Grid((2, 2, 2), precision=8)
    with pytest.raises(ValueError):
        G.adjoint_step(1e-3, 0.01, scheme=2)
    G.chi_from_sphere((np.pi, np.pi, np.pi), 1.0)
    with pytest.raises(cup.CupError):
        G.adjoint_step(1e-3, 0.01)


def test_rollout_gradient_matches_oracle(cup, oracle):
    """torch rollout (6 steps, checkpoints every 4) vs the oracle's adjoint chain, FP64."""
    import torch
    rng = np.random.default_rng(12)
    bpd = (4, 4, 2)
    G = cup.Grid(bpd, precision=8, decomposition="fused")
    shape = G.shape
    u0 = np.stack([rng.standard_normal(shape) for _ in range(3)])
    wgt = np.stack([rng.standard_normal(shape) for _ in range(3)])
    nu, U = 0.01, (0.2, 0.0, -0.1)
    dts = [0.2 * G.h * (1 + 0.1 * k) for k in range(6)]
    from cubismup_3d_amd.adjoint import rollout
    ut = torch.tensor(u0, dtype=torch.float64, device="cuda", requires_grad=True)
    uN = rollout(ut, G, 6, nu, U, dts=dts, every=4)
    J = (uN * torch.tensor(wgt, device="cuda")).sum()
    J.backward()
    grad = ut.grad.cpu().numpy()
    # oracle: forward states, then the adjoint steps backwards
    O = oracle.OracleSim(bpd)
    O.set_velocity(*u0)
    states = []
    for dt in dts:
        states.append(O.velocity())
        O.timestep(dt, nu, uinf=U)
    for a, b in zip(O.velocity(), uN.detach().cpu().numpy()):
        assert rel(b, a) < 1e-11
    lam = tuple(wgt)
    for n in reversed(range(6)):
        O.set_velocity(*states[n])
        lam = O.adjoint_step(dts[n], nu, lam, U)
    for q in range(3):
        assert rel(grad[q], lam[q]) < 1e-10


def test_dissipated_energy_gradient_vs_finite_differences(cup):
    """128^3 seeded isotropic turbulence (k0 = 4, u_rms = 1), 50 CFL steps (nu = 0.005,
    CFL 0.3), checkpoints every 10: the adjoint gradient of the dissipated kinetic energy
    J = KE(u0) - KE(u50) against FP64 central differences with eps = 1e-3 along 8 seeded
    random solenoidal directions d. The fourth-order central stencil is used: the
    second-order one's O(eps^2) truncation is about 4e-5 of dJ.d here.
    - FP64 forward and adjoint: the same map as the differences, 1e-6 (measured about 1e-8).
    - FP32 forward and adjoint: dJ.d is the difference of h^3 <u0, d> and <dKE(u50)/du0, d>,
      terms 7 to 300 times larger than it, and the FP32 trajectory (which also seeds the
      backward through u50) and adjoint carry an error of a few 1e-6 of those terms after 50
      steps each way: measured 1e-5 to 2e-4 of dJ.d, asserted 5e-4."""
    import torch
    from cubismup_3d_amd.adjoint import kinetic_energy, recompute_memory_bytes, rollout
    bpd, nsteps, every, nu = (16, 16, 16), 50, 10, 0.005
    G4 = cup.Grid(bpd, precision=4, decomposition="fused")
    G4.ic_hit(4.0, 1.0, 20260101)
    h = G4.h
    u0 = G4.get_torch_velocity().clone().requires_grad_(True)
    dts = []
    uN = rollout(u0, G4, nsteps, nu, dts=None, CFL=0.3, every=every, dts_out=dts)
    J = kinetic_energy(u0, h) - kinetic_energy(uN, h)
    J.backward()
    grad32 = u0.grad.double()
    print(f"J={J.item():.6e} recompute memory {recompute_memory_bytes(G4, nsteps, every) / 2**20:.0f} MiB",
          flush=True)
    G8 = cup.Grid(bpd, precision=8, decomposition="fused")
    base = u0.detach().double()
    v0 = base.clone().requires_grad_(True)
    vN = rollout(v0, G8, nsteps, nu, dts=dts, every=every)
    (kinetic_energy(v0, h) - kinetic_energy(vN, h)).backward()
    grad64 = v0.grad
    P = cup.Grid(bpd, precision=8, decomposition="plain")

    def J64(v):
        G8.set_torch_velocity(v)
        for dt in dts:
            G8.run_steps(1, nu=nu, fixed_dt=dt)
        return (kinetic_energy(v, h) - kinetic_energy(G8.get_torch_velocity(), h)).item()
    rng = np.random.default_rng(2026)
    eps = 1e-3
    for k in range(8):
        r = [rng.standard_normal(G8.shape) for _ in range(3)]
        P.set_velocity(*r)
        P.pressure_rhs(1e-3); P.poisson_periodic(); P.pressure_project(1e-3)
        d = torch.tensor(np.stack(P.velocity()), device="cuda")
        d = d / d.pow(2).mean().sqrt()
        jp1, jm1 = J64(base + eps * d), J64(base - eps * d)
        jp2, jm2 = J64(base + 2 * eps * d), J64(base - 2 * eps * d)
        fd = (8 * (jp1 - jm1) - (jp2 - jm2)) / (12 * eps)
        fd2 = (jp1 - jm1) / (2 * eps)
        a64 = (grad64 * d).sum().item()
        a32 = (grad32 * d).sum().item()
        e64, e32 = abs(a64 - fd) / abs(fd), abs(a32 - fd) / abs(fd)
        term = abs(h ** 3 * (base * d).sum().item())
        print(f"direction {k}: fd4 {fd:.10e} (fd2 rel {abs(fd2 - fd) / abs(fd):.1e}, |dJ.d|/term "
              f"{abs(fd) / term:.1e}) fp64 adjoint rel {e64:.1e}, fp32 adjoint rel {e32:.1e} "
              f"(of the term {abs(a32 - fd) / term:.1e})", flush=True)
        assert e64 < 1e-6
        assert e32 < 5e-4
