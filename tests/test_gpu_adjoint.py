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


@pytest.mark.parametrize("scheme", [0, 1, 2])
@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_adjoint_step_matches_oracle(cup, oracle, decomp, scheme):
    """Central Euler (0), upwind Euler (1) and upwind midpoint (2)."""
    print(f"adjoint step decomp={decomp} scheme={scheme}", flush=True)
    rng = np.random.default_rng(31)
    bpd = (4, 2, 2)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    u = tuple(rng.standard_normal(G.shape) for _ in range(3))
    lam = tuple(rng.standard_normal(G.shape) for _ in range(3))
    U, dt, nu = (0.3, -0.2, 0.1), 0.3 * G.h, 0.02
    G.set_velocity(*u); O.set_velocity(*u)
    set_adjoint(G, lam)
    G.adjoint_step(dt, nu, U, scheme=scheme)
    exp = O.adjoint_step(dt, nu, lam, U, scheme=scheme)
    for a, b in zip(get_adjoint(G), exp):
        assert rel(a, b) < 1e-12
    for a, b in zip(G.velocity(), u):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("scheme", [0, 2])
def test_adjoint_step_fp32_vs_fp64(cup, scheme):
    rng = np.random.default_rng(8)
    bpd = (4, 4, 4)
    u = tuple(rng.standard_normal((32, 32, 32)) for _ in range(3))
    lam = tuple(rng.standard_normal((32, 32, 32)) for _ in range(3))
    out = []
    for prec in (8, 4):
        G = cup.Grid(bpd, precision=prec, decomposition="fused")
        G.set_velocity(*u)
        set_adjoint(G, lam)
        G.adjoint_step(0.2 * G.h, 0.01, (0.1, 0.0, 0.0), scheme=scheme)
        out.append(get_adjoint(G))
    for a, b in zip(out[1], out[0]):
        assert rel(a, b) < 1e-6


def test_adjoint_rejects_other_schemes_and_bodies(cup):
    G = cup.Grid((2, 2, 2), precision=8)
    with pytest.raises(ValueError):
        G.adjoint_step(1e-3, 0.01, scheme=7)
    G.chi_from_sphere((np.pi, np.pi, np.pi), 1.0)
    with pytest.raises(cup.CupError):
        G.adjoint_step(1e-3, 0.01)


@pytest.mark.parametrize("scheme", [0, 2])
def test_rollout_gradient_matches_oracle(cup, oracle, scheme):
    """torch rollout (6 steps, checkpoints every 4) vs the oracle's adjoint chain, FP64;
    central Euler and upwind midpoint."""
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
    uN = rollout(ut, G, 6, nu, U, dts=dts, every=4, scheme=scheme)
    J = (uN * torch.tensor(wgt, device="cuda")).sum()
    J.backward()
    grad = ut.grad.cpu().numpy()
    # oracle: forward states, then the adjoint steps backwards
    O = oracle.OracleSim(bpd)
    O.set_velocity(*u0)
    states = []
    for dt in dts:
        states.append(O.velocity())
        O.timestep(dt, nu, uinf=U, scheme=scheme)
    for a, b in zip(O.velocity(), uN.detach().cpu().numpy()):
        assert rel(b, a) < 1e-11
    lam = tuple(wgt)
    for n in reversed(range(6)):
        O.set_velocity(*states[n])
        lam = O.adjoint_step(dts[n], nu, lam, U, scheme=scheme)
    for q in range(3):
        assert rel(grad[q], lam[q]) < 1e-10


def test_upwind_midpoint_gradient_vs_finite_differences(cup):
    """64^3 seeded isotropic turbulence advected by a uniform flow U = (6, -6, 6), 20 CFL steps
    of the upwind midpoint form (nu = 0.005, CFL 0.3), checkpoints every 5: the adjoint
    gradient of J = KE(u0) - KE(u20) against FP64 central differences (fourth-order stencil,
    eps = 1e-4) along 4 seeded solenoidal directions, to 1e-6. The upwind map is smooth only
    away from cells where an advecting velocity changes sign: there its derivative jumps by
    the fourth difference of the advected field, and a difference over +-eps averages the
    jumps of every cell that crosses zero inside the interval (for U = 0 at this size, 1e-4 to
    2e-4 of dJ.d at eps = 1e-5, erratic in eps). With |U| above every velocity of the run no
    cell crosses zero, the x and z axes take the a > 0 stencil and y the other one, and the
    difference is the derivative; the per-step adjoint also matches the oracle's, which the
    dot-product identity pins, to 1e-12."""
    import torch
    from cubismup_3d_amd.adjoint import kinetic_energy, rollout
    bpd, nsteps, every, nu, scheme, U = (8, 8, 8), 20, 5, 0.005, 2, (6.0, -6.0, 6.0)
    G8 = cup.Grid(bpd, precision=8, decomposition="fused")
    G8.ic_hit(4.0, 1.0, 20260101)
    h = G8.h
    base = G8.get_torch_velocity().clone()
    assert base.abs().max().item() < 5.0
    v0 = base.clone().requires_grad_(True)
    dts = []
    vN = rollout(v0, G8, nsteps, nu, U, dts=None, CFL=0.3, every=every, dts_out=dts, scheme=scheme)
    assert vN.abs().max().item() < 5.0  # every advecting velocity keeps its sign
    (kinetic_energy(v0, h) - kinetic_energy(vN, h)).backward()
    grad = v0.grad
    P = cup.Grid(bpd, precision=8, decomposition="plain")

    def J64(v):
        G8.set_torch_velocity(v)
        for dt in dts:
            G8.run_steps(1, nu=nu, uinf=U, fixed_dt=dt, scheme=scheme)
        return (kinetic_energy(v, h) - kinetic_energy(G8.get_torch_velocity(), h)).item()
    rng = np.random.default_rng(7)
    eps = 1e-4
    for k in range(4):
        r = [rng.standard_normal(G8.shape) for _ in range(3)]
        P.set_velocity(*r)
        P.pressure_rhs(1e-3); P.poisson_periodic(); P.pressure_project(1e-3)
        d = torch.tensor(np.stack(P.velocity()), device="cuda")
        d = d / d.pow(2).mean().sqrt()
        fd = (8 * (J64(base + eps * d) - J64(base - eps * d)) - (J64(base + 2 * eps * d) - J64(base - 2 * eps * d))) / (
            12 * eps)
        a = (grad * d).sum().item()
        print(f"direction {k}: fd4 {fd:.10e} adjoint rel {abs(a - fd) / abs(fd):.1e}", flush=True)
        assert abs(a - fd) < 1e-6 * abs(fd)


def low_mode_field(shape, h, rng):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*(h * (np.arange(m) + 0.5) for m in shape), indexing="ij")
    u = np.sin(x) * np.cos(y) * np.cos(2 * z) + 0.3 * rng.standard_normal(shape)
    v = -np.cos(x) * np.sin(y) * np.cos(z) + 0.3 * rng.standard_normal(shape)
    w = 0.5 * np.sin(2 * y + z) + 0.3 * rng.standard_normal(shape)
    return u, v, w


@pytest.mark.parametrize("scheme,target,rate", [(0, 1.1, 0.0), (1, 0.0, 0.7), (2, 1.1, 0.0), (2, 0.0, 0.7)])
@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_forcing_adjoint_step_matches_oracle(cup, oracle, decomp, scheme, target, rate):
    """cup_adjoint_step_forcing against the oracle's forced adjoint (pinned by the dot-product
    identity), FP64, to a target energy (1.1 x the field's) or at an injection rate; the
    forward state is left as it was (the midpoint form keeps u_n aside)."""
    print(f"forcing adjoint decomp={decomp} scheme={scheme} target={target} rate={rate}", flush=True)
    rng = np.random.default_rng(41)
    bpd = (4, 2, 2)
    G = cup.Grid(bpd, precision=8, decomposition=decomp)
    O = oracle.OracleSim(bpd)
    u = low_mode_field(G.shape, G.h, rng)
    target = target * 0.5 * sum(float(np.mean(a * a)) for a in u) if target else 0.0
    lam = tuple(a + 0.5 * b for a, b in zip(low_mode_field(G.shape, G.h, rng), u))
    U, dt, nu = (0.3, -0.2, 0.1), 0.3 * G.h, 0.02
    G.set_velocity(*u); O.set_velocity(*u)
    set_adjoint(G, lam)
    G.adjoint_step(dt, nu, U, scheme=scheme, tke_target=target, injection_rate=rate)
    exp = O.adjoint_step(dt, nu, lam, U, scheme=scheme, tke_target=target, injection_rate=rate)
    plain = O.adjoint_step(dt, nu, lam, U, scheme=scheme)
    for a, b, c in zip(get_adjoint(G), exp, plain):
        assert rel(a, b) < 1e-11
        assert rel(c, b) > 1e-4  # the forcing's transpose shows
    for a, b in zip(G.velocity(), u):
        assert np.array_equal(a, b)


def test_forced_rollout_gradient_vs_finite_differences(cup):
    """A forced rollout (64^3 seeded isotropic turbulence, linear forcing to 1.2 x its energy,
    central Euler, 10 fixed steps, checkpoints every 4): the adjoint gradient of J = <w, u_10>
    against FP64 central differences (fourth-order stencil, eps = 1e-4) along 3 seeded
    directions. The step is smooth (rational in u through fac), so the agreement is at the
    stencil's truncation and rounding, asserted to 1e-6 of |dJ.d|."""
    import torch
    from cubismup_3d_amd.adjoint import rollout
    print("forced rollout FD", flush=True)
    G = cup.Grid((8, 8, 8), precision=8, decomposition="fused")
    G.ic_hit(4.0, 1.0, 20260101)
    base = G.get_torch_velocity().clone()
    tke = 0.5 * (base * base).sum(0).mean().item()
    nu, n, target = 0.005, 10, 1.2 * tke
    dts = [0.3 * G.h / 4.0] * n
    gen = torch.Generator(device="cpu").manual_seed(5)
    wgt = torch.randn(base.shape, generator=gen, dtype=torch.float64).cuda()
    v0 = base.clone().requires_grad_(True)
    vN = rollout(v0, G, n, nu, dts=dts, every=4, tke_target=target)
    (wgt * vN).sum().backward()
    grad = v0.grad

    def J(v):
        G.set_torch_velocity(v)
        G.run_steps(n, nu=nu, fixed_dt=dts[0], spectral_forcing=True, tke_target=target)
        return (wgt * G.get_torch_velocity()).sum().item()
    eps = 1e-4
    for k in range(3):
        d = torch.randn(base.shape, generator=gen, dtype=torch.float64).cuda()
        fd = (8 * (J(base + eps * d) - J(base - eps * d)) - (J(base + 2 * eps * d) - J(base - 2 * eps * d))) / (
            12 * eps)
        ad = (grad * d).sum().item()
        print(f"direction {k}: adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
        assert abs(ad - fd) < 1e-6 * abs(fd)


def test_forcing_adjoint_step_fp32_vs_fp64(cup):
    """The forced adjoint step in FP32 against FP64 on a 32^3 field (fused), to a target energy."""
    print("forcing adjoint fp32 vs fp64", flush=True)
    rng = np.random.default_rng(9)
    bpd = (4, 4, 4)
    G8 = cup.Grid(bpd, precision=8, decomposition="fused")
    u = low_mode_field(G8.shape, G8.h, rng)
    lam = tuple(a + 0.5 * b for a, b in zip(low_mode_field(G8.shape, G8.h, rng), u))
    target = 1.1 * 0.5 * sum(float(np.mean(a * a)) for a in u)
    out = []
    for prec in (8, 4):
        G = cup.Grid(bpd, precision=prec, decomposition="fused")
        G.set_velocity(*u)
        set_adjoint(G, lam)
        G.adjoint_step(0.2 * G.h, 0.01, (0.1, 0.0, 0.0), tke_target=target)
        out.append(get_adjoint(G))
    for a, b in zip(out[1], out[0]):
        assert rel(a, b) < 1e-5
