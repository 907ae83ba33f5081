"""The drop-in Simulation mirror on the device: SimulationData built from the reference's
command-line flags, Simulation(SD).run() through the operator pipeline (plain) and the
device-resident loop (fused), against the oracle's step loop; a sphere through
CreateObstacles/Penalization/ComputeForces; the HIT analysis files of the Analysis operator
(Analysis.cpp:96-196, SpectralAnalysis.cpp:124-205)."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PERIODIC = ["-BC_x", "periodic", "-BC_y", "periodic", "-BC_z", "periodic", "-muteAll", "1",
            "-freqDiagnostics", "0", "-precision", "double"]


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_run_taylor_green_matches_oracle(cup, oracle, decomp):
    print(f"Simulation.run TGV decomp={decomp}", flush=True)
    nu, CFL, n = 0.02, 0.3, 6
    argv = ["-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-extentx", repr(2 * math.pi), "-decomposition", decomp,
            "-initCond", "taylorGreen", "-uMax_forced", "1.0", "-CFL", str(CFL), "-nu", str(nu), "-nsteps", str(n),
            "-rampup", "3"] + PERIODIC
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    O = oracle.OracleSim((4, 4, 4))
    O.ic_taylor_green(1.0)
    t = 0.0
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, (0, 0, 0), step, 3)
        O.timestep(dt, nu)
        t += dt
    assert SD.step == n and SD.time == pytest.approx(t, rel=1e-12)
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-11, name
    assert rel(S.field("p"), O.get("p")) < 1e-10


@pytest.mark.parametrize("decomp,hemi", [("plain", False), ("fused", False), ("plain", True)])
def test_run_sphere_matches_oracle(cup, oracle, decomp, hemi):
    """A sphere at rest in a uniform inflow (the frame follows the body, -xvel 1), implicit
    penalization, CFL-driven steps: the pipeline's velocity, penalization force and surface
    QoIs against the oracle's step."""
    print(f"Simulation.run sphere decomp={decomp} hemisphere={hemi}", flush=True)
    nu, CFL, n, R = 1e-3, 0.3, 4, 0.15
    argv = ["-bpdx", "8", "-bpdy", "4", "-bpdz", "4", "-extentx", "2.0", "-CFL", str(CFL), "-nu", str(nu),
            "-nsteps", str(n), "-rampup", "0", "-sphereL", str(2 * R), "-xpos", "0.5", "-ypos", "0.5", "-zpos",
            "0.5", "-xvel", "1.0", "-bForcedInSimFrame", "1", "-bFixFrameOfRef", "1", "-lambda", "1e6",
            "-implicitPenalization", "1", "-decomposition", decomp, "-hemisphere", str(int(hemi))] + PERIODIC
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    O = oracle.OracleSim((8, 4, 4), maxextent=2.0)
    O.sphere_create(R, (0.5, 0.5, 0.5), hemi=hemi)
    U = (1.0, 0.0, 0.0)
    sph = dict(lambda_=1e6, implicit=True, trans_vel=(-1.0, 0.0, 0.0))
    # the first dt sees the frame velocity of the flags (0): CreateObstacles sets U after it
    u_cfl = (0.0, 0.0, 0.0)
    for step in range(n):
        dt, _ = O.calc_max_timestep(CFL, nu, u_cfl, step, 0)
        out = O.timestep(dt, nu, uinf=U, sphere=sph)
        u_cfl = U
    for name in ("u", "v", "w"):
        assert rel(S.field(name), O.get(name)) < 1e-11, name
    ob = S.obstacle(0)
    np.testing.assert_allclose(ob["force"], out["penal_ft"][:3], rtol=1e-10, atol=1e-12 * np.abs(out["penal_ft"]).max())
    qo = np.asarray(out["surface"])
    np.testing.assert_allclose(ob["surface"], qo, rtol=1e-9, atol=1e-11 * np.abs(qo).max())


def read_analysis(path):
    """Rows 'name value #comment' and the two-column spectrum of a spectralAnalysis file."""
    rows, spec = {}, []
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0].startswith("Spectral Analysis")
    i = 1
    while lines[i].strip():
        name, value = lines[i].split()[:2]
        rows[name] = float(value)
        i += 1
    assert lines[i + 1].split()[:2] == ["k", "*"]
    for ln in lines[i + 2:]:
        if ln.strip():
            k, e = ln.split()
            spec.append((float(k), float(e)))
    return rows, np.array(spec)


def test_analysis_files_every_f_steps(cup, oracle, tmp_path, monkeypatch):
    """-analysis HIT -fAnalysis 2 over 4 forced steps writes files 1 and 3 (steps whose
    step+1 is a multiple of 2); the last one describes the final field."""
    print("Simulation analysis fAnalysis", flush=True)
    monkeypatch.chdir(tmp_path)
    nu = 0.01
    argv = ["-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-extentx", repr(2 * math.pi), "-initCond", "HITurbulence",
            "-k0", "4", "-spectralForcing", "1", "-CFL", "0.3", "-nu", str(nu), "-nsteps", "4", "-rampup", "0",
            "-analysis", "HIT", "-fAnalysis", "2", "-decomposition", "plain"] + PERIODIC
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    assert sorted(os.listdir("analysis")) == ["spectralAnalysis_000000001", "spectralAnalysis_000000003"]
    rows, spec = read_analysis("analysis/spectralAnalysis_000000003")
    O = oracle.OracleSim((4, 4, 4))
    O.set_velocity(*(S.field(n) for n in "uvw"))
    ro = O.hit_analysis(nu)
    for k in ("tke", "eps", "lambda", "Re_lambda", "l_integral", "tau_integral"):
        assert rows[k] == pytest.approx(ro[k], rel=2e-5), k
    assert rows["eta"] == pytest.approx((nu ** 3 / ro["eps"]) ** 0.25, rel=2e-5)
    assert rows["tau_eta"] == pytest.approx(math.sqrt(nu / ro["eps"]), rel=2e-5)
    assert rows["mean_grad"] == pytest.approx(ro["grad_mean"], rel=2e-5)
    assert rows["std_grad"] == pytest.approx(ro["grad_std"], rel=2e-5)
    assert rows["lBox"] == pytest.approx(2 * math.pi, rel=1e-5)
    assert rows["nu_sgs"] == 0 and rows["cs2_avg"] == 0
    assert 0 < rows["time"] < SD.time
    assert spec.shape == (15, 2)
    np.testing.assert_allclose(spec[:, 0], np.arange(1, 16), rtol=1e-5)  # (i+1) 2 pi / L with L = 2 pi
    np.testing.assert_allclose(spec[:, 1], ro["E"], rtol=2e-5, atol=1e-9 * ro["E"].max())
    assert SD.grad_mean == pytest.approx(ro["grad_mean"], rel=1e-11)


def test_analysis_files_every_t_time(cup, tmp_path, monkeypatch):
    """-tAnalysis T: a file whenever time + dt passes the next analysis time, numbered
    0 at time 0 and 1 + floor(time / T) after (Analysis.cpp:99-112)."""
    print("Simulation analysis tAnalysis", flush=True)
    monkeypatch.chdir(tmp_path)
    argv = ["-bpdx", "2", "-bpdy", "2", "-bpdz", "2", "-extentx", repr(2 * math.pi), "-initCond", "HITurbulence",
            "-k0", "2", "-CFL", "0.4", "-nu", "0.02", "-nsteps", "0", "-rampup", "0", "-analysis", "HIT",
            "-tAnalysis", "0.05", "-decomposition", "plain"] + PERIODIC
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    expect, nxt, T = set(), 0.0, 0.05
    for _ in range(12):
        t = SD.time
        dt = S.calcMaxTimestep()
        if t + dt > nxt:
            nxt += T
            expect.add(0 if t <= 0 else 1 + int(t / T))
        S.timestep(dt)
    assert SD.nextAnalysisTime == pytest.approx(nxt, rel=1e-12)
    assert sorted(os.listdir("analysis")) == sorted(f"spectralAnalysis_{i:09d}" for i in expect)
    assert len(expect) >= 3


def test_body_log_files(cup, tmp_path):
    """The per-body files of Obstacle::computeForces and Obstacle::update
    (Obstacle.cpp:491-589): forceValues_surface_0, powerValues_surface_0, forceValues_0 and
    computedVelocity_0, one row per step after the header, consistent with each other."""
    print("body log files", flush=True)
    argv = ["-bpdx", "8", "-bpdy", "4", "-bpdz", "4", "-extentx", "2.0", "-CFL", "0.3", "-nu", "1e-3", "-nsteps", "3",
            "-rampup", "0", "-sphereL", "0.3", "-xpos", "0.5", "-ypos", "0.5", "-zpos", "0.5", "-xvel", "1.0",
            "-bForcedInSimFrame", "1", "-bFixFrameOfRef", "1", "-lambda", "1e6", "-implicitPenalization", "1",
            "-decomposition", "plain", "-BC_x", "periodic", "-BC_y", "periodic", "-BC_z", "periodic",
            "-freqDiagnostics", "0", "-precision", "double", "-muteAll", "0", "-serialization", str(tmp_path)]
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    S.flush_logs()

    def table(name):
        with open(tmp_path / name) as f:
            rows = [ln.split() for ln in f.read().splitlines() if ln.strip()]
        head = rows[0]
        data = np.array([[float(x) for x in r] for r in rows[1:]])
        return head, data

    hs, fs = table("forceValues_surface_0.dat")
    hp, fp = table("powerValues_surface_0.dat")
    hd, fd = table("forceValues_0.dat")
    hv, fv = table("computedVelocity_0.dat")
    assert hs[:3] == ["step", "time", "mass"] and hs[-2:] == ["drag", "thrust"] and len(hs) == 20
    assert hp == ["step", "time", "Pthrust", "Pdrag", "Pout", "pDef", "etaPDef", "pLocom", "PoutBnd", "defPowerBnd",
                  "etaPDefBnd"]
    assert len(hd) == 31 and hd[-1] == "penalJ5" and len(hv) == 22 and hv[-1] == "J5"
    assert fs.shape[0] == fp.shape[0] == fd.shape[0] == 3 and fv.shape[0] == 3
    np.testing.assert_array_equal(fs[:, 0], [0, 1, 2])
    # |V| = 1: Pdrag = drag, Pthrust = thrust
    np.testing.assert_allclose(fp[:, 3], fs[:, hs.index("drag")], rtol=1e-6)
    np.testing.assert_allclose(fp[:, 2], fs[:, hs.index("thrust")], rtol=1e-6, atol=1e-12)
    # the forced axis solves to the imposed velocity; the penalization force is the body's
    assert (fd[:, hd.index("linVel_comp_x")] == -1.0).all() and (fd[:, hd.index("penalM")] > 0).all()
    ob = S.obstacle(0)
    np.testing.assert_allclose(fd[-1, 3:6], ob["force"], rtol=1e-6, atol=1e-9)
    assert (fv[:, hv.index("vel_x")] == -1.0).all() and (fv[:, hv.index("CMx")] == 0.5).all()


def test_command_line_program(cup, oracle, tmp_path):
    """cup_simulation (the reference's source/main.cpp): a Taylor-Green run from the flags
    writes diagnostics.dat every step; its kinetic energy column against the oracle's steps
    with the same CFL timesteps (the file's 6 significant digits)."""
    import subprocess
    print("cup_simulation", flush=True)
    exe = os.path.join(os.path.dirname(cup.__file__), "cup_simulation")
    flags = ["-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-extentx", repr(2 * math.pi), "-BC_x", "periodic", "-BC_y",
             "periodic", "-BC_z", "periodic", "-initCond", "taylorGreen", "-uMax_forced", "-1", "-nu", "0.01",
             "-CFL", "0.3", "-rampup", "0", "-nsteps", "5", "-freqDiagnostics", "1", "-muteAll", "0",
             "-serialization", str(tmp_path)]
    r = subprocess.run([exe] + flags, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(tmp_path / "diagnostics.dat") as f:
        rows = [ln.split() for ln in f.read().splitlines()]
    assert rows[0][:2] == ["step_id", "time"] and rows[0][-2:] == ["kineticEn", "enstrophy"]
    data = np.array([[float(x) for x in r_] for r_ in rows[1:]])
    assert data.shape == (5, 22)
    O = oracle.OracleSim((4, 4, 4))
    O.ic_taylor_green(-1.0)
    t = 0.0
    for step in range(5):  # row k: ComputeDissipation inside step k, after its update, at time t_k
        dt, _ = O.calc_max_timestep(0.3, 0.01, (0, 0, 0), step, 0)
        O.timestep(dt, 0.01)
        assert data[step, 0] == step and data[step, 1] == pytest.approx(t, rel=1e-5, abs=1e-12)
        assert data[step, 2 + 18] == pytest.approx(O.dissipation(0.01)[18], rel=1e-5)
        t += dt


@pytest.mark.parametrize("decomp", ["plain", "fused"])
def test_external_forcing_uniform_acceleration(cup, decomp):
    """ExternalForcing (ExternalForcing.cpp:35-48, added by Simulation.cpp:336-339 when
    uMax_forced > 0 and the IC is not the Taylor-Green vortex): from rest, a uniform field is
    neither advected, diffused nor projected, so after the run u = 8 uMax nu t / H^2 with
    H = extent_z, exactly, and v = w = 0."""
    print(f"external forcing decomp={decomp}", flush=True)
    argv = ["-bpdx", "4", "-bpdy", "4", "-bpdz", "2", "-extentx", "2.0", "-initCond", "zero", "-uMax_forced", "0.5",
            "-nu", "0.01", "-CFL", "0.3", "-rampup", "0", "-nsteps", "3", "-decomposition", decomp] + PERIODIC
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    H = SD.extent_z
    assert H == pytest.approx(1.0)
    u = S.field("u")
    np.testing.assert_allclose(u, 8 * 0.5 * 0.01 / H ** 2 * SD.time, rtol=1e-13)
    assert not S.field("v").any() and not S.field("w").any()
