"""The sphere-in-uniform-flow case at its full size: periodic 4 x 2 x 2 box at 512 x 256 x 256
cells (64 x 32 x 32 blocks), sphere of diameter 0.25 at (1, 1, 1), inflow U = 1, Re = 300
(nu = U D / Re), implicit penalization with lambda = 1e6, CFL 0.1 without ramp. The FP32
fused form's drag-coefficient series over 200 steps against the FP64 form (the GPU oracle),
and the FP64 form's first steps against the CPU oracle at the same size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BPD, EXTENT = (64, 32, 32), 4.0
CENTER, RADIUS = (1.0, 1.0, 1.0), 0.125
U, RE, CFL = 1.0, 300.0, 0.1
NU = U * 2 * RADIUS / RE
SPHERE = dict(lambda_=1e6, implicit=True, trans_vel=(-U, 0.0, 0.0))
QDRAG = 15  # the drag column of the surface-force QoIs (FluidSolidForces.cpp:136-138)


def cd(q):
    return q[QDRAG] / (0.5 * U * U * np.pi * RADIUS ** 2)


def drag_series(cup, precision, decomposition, nsteps):
    G = cup.Grid(BPD, maxextent=EXTENT, precision=precision, decomposition=decomposition)
    G.chi_from_sphere(CENTER, RADIUS)
    out = []
    for _ in range(nsteps):
        G.run_steps(1, CFL=CFL, nu=NU, uinf=(U, 0, 0), sphere=SPHERE)
        st = G.step_state()
        out.append((st["time"], cd(G.surface_forces(NU, SPHERE["trans_vel"]))))
    return np.array(out)


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_small_sphere_fp32_vs_fp64(cup, decomp):
    """The FP32 form with a body (chi, masked RHS, penalization, forces) against FP64 on a
    small grid: 20 CFL steps, velocity to 1e-5 and drag to 1e-4."""
    bpd = (16, 8, 8)
    out = []
    for prec in (8, 4):
        G = cup.Grid(bpd, maxextent=EXTENT, precision=prec, decomposition=decomp)
        G.chi_from_sphere(CENTER, RADIUS)
        G.run_steps(20, CFL=CFL, nu=NU, uinf=(U, 0, 0), sphere=SPHERE)
        out.append((G.velocity(), cd(G.surface_forces(NU, SPHERE["trans_vel"]))))
    (v8, c8), (v4, c4) = out
    num = sum(np.linalg.norm(a - b) ** 2 for a, b in zip(v4, v8))
    den = sum(np.linalg.norm(b) ** 2 for b in v8)
    print(f"{decomp}: rel L2 {np.sqrt(num / den):.2e}, Cd fp64 {c8:.6f} fp32 {c4:.6f}", flush=True)
    assert np.sqrt(num / den) < 1e-5
    assert c4 == pytest.approx(c8, rel=1e-4)


def test_drag_series_fp32_vs_fp64(cup):
    n = 200
    s32 = drag_series(cup, 4, "fused", n)
    s64 = drag_series(cup, 8, "plain", n)
    rel = np.abs(s32[:, 1] - s64[:, 1]) / np.abs(s64[:, 1])
    for k in (0, 9, 49, 99, 149, 199):
        print(f"step {k + 1:3d} t={s64[k, 0]:.5f} Cd fp64 {s64[k, 1]:.6f} fp32 {s32[k, 1]:.6f} rel {rel[k]:.1e}",
              flush=True)
    print(f"max rel difference over {n} steps: {rel.max():.2e}; time rel difference "
          f"{abs(s32[-1, 0] - s64[-1, 0]) / s64[-1, 0]:.1e}", flush=True)
    assert np.all(np.isfinite(s32)) and np.all(s64[:, 1] > 0)
    assert rel.max() < 1e-3
    assert abs(s32[-1, 0] - s64[-1, 0]) < 1e-5 * s64[-1, 0]


def test_fp64_first_steps_match_oracle_full_size(cup, oracle):
    G = cup.Grid(BPD, maxextent=EXTENT, precision=8, decomposition="fused")
    G.chi_from_sphere(CENTER, RADIUS)
    O = oracle.OracleSim(BPD, maxextent=EXTENT)
    O.sphere_create(RADIUS, CENTER)
    assert np.array_equal(G.get("chi"), O.get("chi")) or np.max(np.abs(G.get("chi") - O.get("chi"))) < 1e-13
    for step in range(2):
        dt, _ = O.calc_max_timestep(CFL, NU, (U, 0, 0), step, 0)
        O.timestep(dt, NU, uinf=(U, 0, 0), sphere=SPHERE)
    G.run_steps(2, CFL=CFL, nu=NU, uinf=(U, 0, 0), sphere=SPHERE)
    for a, b in zip(G.velocity(), O.velocity()):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-11
    qg = G.surface_forces(NU, SPHERE["trans_vel"])
    qo = O.surface_forces(NU, SPHERE["trans_vel"])
    print(f"Cd after 2 steps: device {cd(qg):.12f} oracle {cd(qo):.12f}", flush=True)
    assert cd(qg) == pytest.approx(cd(qo), rel=1e-9)
