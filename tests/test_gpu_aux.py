"""The reference's debugging auxiliaries on the device: Operator::check's NaN scan
(operators/Operator.h:23-50) as cup_check_nan and as the Simulation's -checkNaN runs, and the
profiler regions of SimulationData::startProfiler / stopProfiler (SimulationData.cpp:293-305)
with their summary every 50 verbose steps (Simulation.cpp:614)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,decomposition", [(8, "plain"), (4, "fused"), (4, "plain")])
def test_check_nan_counts(cup, precision, decomposition):
    """Counts per field; infinities and large values are not NaN."""
    print(f"check_nan precision={precision} {decomposition}", flush=True)
    G = cup.Grid((2, 3, 2), precision=precision, decomposition=decomposition)
    rng = np.random.default_rng(1)
    n = int(np.prod(G.shape))
    want = {"u": 3, "v": 0, "w": 1, "p": 7}
    for name, k in want.items():
        f = rng.standard_normal(n)
        idx = rng.choice(n, size=k + 2, replace=False)
        f[idx[:k]] = np.nan
        f[idx[k]] = np.inf
        f[idx[k + 1]] = -1e30 if precision == 8 else -3e38
        G.set(name, f.reshape(G.shape))
    assert list(G.check_nan()) == [want[k] for k in ("u", "v", "w", "p")]
    for name in want:
        G.set(name, np.zeros(G.shape))
    assert list(G.check_nan()) == [0, 0, 0, 0]


def _sd(cup, extra):
    argv = ["-bpdx", "2", "-bpdy", "2", "-bpdz", "2", "-extentx", str(2 * np.pi), "-nu", "0.01", "-CFL", "0.3",
            "-initCond", "taylorGreen", "-uMax_forced", "1.0", "-BC_x", "periodic", "-BC_y", "periodic",
            "-BC_z", "periodic", "-muteAll", "1", "-freqDiagnostics", "0", "-rampup", "0"] + extra
    return cup.SimulationData(argv=argv)


def test_simulation_check_names_the_operator(cup):
    """With -checkNaN 1 a NaN in the velocity stops the step at PressureRHS's check with the
    reference's message; without the flag the step runs."""
    print("Simulation -checkNaN plain FP64", flush=True)
    S = cup.Simulation(_sd(cup, ["-checkNaN", "1", "-precision", "double", "-decomposition", "plain"]))
    S.timestep(1e-3)  # a clean step passes every check
    u = S.field("u")
    u[3, 4, 5] = np.nan
    S.set_field("u", u)
    with pytest.raises(RuntimeError, match="GenericCoordinator::check isnan PressureRHS"):
        S.timestep(1e-3)
    T = cup.Simulation(_sd(cup, ["-precision", "double", "-decomposition", "plain"]))
    T.set_field("u", u)
    T.timestep(1e-3)
    assert np.isnan(T.field("u")).any()


def test_fused_advance_checks_after_the_loop(cup):
    """The device-resident loop is checked once after its steps."""
    print("Simulation -checkNaN fused FP32", flush=True)
    S = cup.Simulation(_sd(cup, ["-checkNaN", "1", "-precision", "float", "-decomposition", "fused"]))
    S.advance(2)
    u = S.field("u")
    u[0, 1, 2] = np.nan
    S.set_field("u", u)
    with pytest.raises(RuntimeError, match="GenericCoordinator::check isnan Fused steps"):
        S.advance(1)


def test_verbose_profiler_summary(cup, capfd):
    """A verbose run prints the regions' table after step 50, with the reference's region names."""
    print("profiler summary, 50 verbose plain steps", flush=True)
    S = cup.Simulation(_sd(cup, ["-verbose", "1", "-precision", "double", "-decomposition", "plain",
                                 "-nsteps", "50"]))
    S.run()
    out = capfd.readouterr().out
    lines = out[out.index("region"):].splitlines()[1:]
    rows = {ln[:30].strip(): ln[30:].split() for ln in lines if len(ln) > 30}
    for name in ("DT", "AdvDiff Kernel", "PresRHS Kernel", "FFTW solve", "GradP", "Save"):
        assert name in rows, (name, list(rows))
        total, mean, share, count = rows[name]
        assert int(count) == 50  # one region per step
        assert float(total) > 0
