"""Saves and restarts of the Simulation mirror (Simulation::_serialize / _deserialize,
Simulation.cpp:426-580; Obstacle::save / restart, Obstacle.cpp:402-450): the files a run with
-fsave writes, their contents against the run's state, and a restarted run that continues
exactly as an uninterrupted one (FP64, a free sphere moving in a Taylor-Green flow). A
missing restart.status starts afresh; a corrupt one is an error."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def flags(path, n, extra=()):
    return ["-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-extentx", "1.0", "-BC_x", "periodic", "-BC_y", "periodic",
            "-BC_z", "periodic", "-muteAll", "1", "-freqDiagnostics", "0", "-precision", "double",
            "-decomposition", "plain", "-initCond", "taylorGreen", "-uMax_forced", "1.0", "-nu", "0.01", "-CFL",
            "0.3", "-rampup", "2", "-nsteps", str(n), "-sphereL", "0.24", "-xpos", "0.45", "-ypos", "0.5", "-zpos",
            "0.55", "-bForcedInSimFrame", "0", "-bFixFrameOfRef", "0", "-implicitPenalization", "1", "-lambda", "1e6",
            "-serialization", str(path)] + list(extra)


def run(cup, argv):
    SD = cup.SimulationData(argv=argv)
    S = cup.Simulation(SD)
    S.run()
    return SD, S


def test_saves_and_restart_continue_exactly(cup, tmp_path):
    print("save every 3 steps, restart at 6, compare with 9 uninterrupted steps", flush=True)
    a, c = tmp_path / "a", tmp_path / "c"
    SDa, Sa = run(cup, flags(a, 6, ["-fsave", "3"]))
    names = sorted(os.listdir(a))
    for stem in ("init000000000", "restart_000000003", "restart_000000006"):
        for f in (f"{stem}.status", f"{stem}_0.txt", f"vel_{stem}.cupf", f"pres_{stem}.cupf", f"chi_{stem}.cupf"):
            assert f in names, f
    assert "restart.status" in names
    with open(a / "restart.status") as f:
        lines = f.read().splitlines()
    assert [ln.split(":")[0] for ln in lines] == ["time", "stepid", "uinfx", "uinfy", "uinfz"]
    assert lines[1] == "stepid: 6" and float(lines[0].split()[1]) == SDa.time
    vel = cup.read_field_file(str(a / "vel_restart_000000006.cupf"))
    assert vel["step"] == 6 and vel["time"] == SDa.time
    for q, name in enumerate("uvw"):
        np.testing.assert_array_equal(vel["data"][q], Sa.field(name))
    chi = cup.read_field_file(str(a / "chi_restart_000000006.cupf"))
    np.testing.assert_array_equal(chi["data"][0], Sa.field("chi"))
    # the XDMF descriptors open the same arrays on the grid of cell corners (Simulation.cpp:475-505)
    h = 1.0 / Sa.field("u").shape[2]  # -extentx 1.0 along the longest side
    for stem, names in (("vel", ("u", "v", "w")), ("pres", ("p",)), ("chi", ("chi",))):
        x = cup.read_xdmf(str(a / f"{stem}_restart_000000006.xmf"))
        assert x["time"] == SDa.time and x["cells"] == Sa.field("u").shape[::-1]
        np.testing.assert_allclose(x["spacing"], [h, h, h], rtol=1e-15)
        assert not x["origin"].any()
        for name in names:
            key = {"p": "pressure"}.get(name, name)
            np.testing.assert_array_equal(x["attributes"][key], Sa.field(name))
        if stem == "vel":
            np.testing.assert_array_equal(x["attributes"]["velocity"],
                                          np.stack([Sa.field(n) for n in "uvw"], axis=-1))
    ob = Sa.obstacle(0)
    with open(a / "restart_000000006_0.txt") as f:
        rows = [[float(x) for x in ln.split()] for ln in f.read().splitlines()]
    assert rows[0] == [SDa.time]
    np.testing.assert_array_equal(rows[1], ob["position"])
    np.testing.assert_array_equal(rows[3], ob["quaternion"])
    np.testing.assert_array_equal(rows[4], ob["trans_vel"])
    np.testing.assert_array_equal(rows[5], ob["ang_vel"])
    del Sa, SDa

    SDb, Sb = run(cup, flags(a, 9, ["-restart", "1"]))
    SDc, Sc = run(cup, flags(c, 9))
    assert SDb.step == SDc.step == 9 and SDb.time == SDc.time
    for name in ("u", "v", "w", "p"):
        np.testing.assert_array_equal(Sb.field(name), Sc.field(name), err_msg=name)
    ob_b, ob_c = Sb.obstacle(0), Sc.obstacle(0)
    for k in ("position", "trans_vel", "ang_vel", "force", "surface"):
        np.testing.assert_array_equal(ob_b[k], ob_c[k], err_msg=k)


def test_restart_without_status_starts_afresh(cup, tmp_path):
    print("restart without restart.status", flush=True)
    SDa, Sa = run(cup, flags(tmp_path / "empty", 2, ["-restart", "1"]))
    SDb, Sb = run(cup, flags(tmp_path / "fresh", 2))
    assert SDa.step == 2 and SDa.time == SDb.time
    np.testing.assert_array_equal(Sa.field("u"), Sb.field("u"))


def test_corrupt_restart_status_is_an_error(cup, tmp_path):
    print("corrupt restart.status", flush=True)
    d = tmp_path / "bad"
    d.mkdir()
    (d / "restart.status").write_text("time: not-a-number\nstepid: 3\n")
    SD = cup.SimulationData(argv=flags(d, 2, ["-restart", "1"]))
    with pytest.raises(OSError, match="restart"):
        cup.Simulation(SD)
    (d / "restart.status").write_text("time: 1.0\nstepid: 3\nuinfx: 0\nuinfy: 0\nuinfz: 0\n")
    SD = cup.SimulationData(argv=flags(d, 2, ["-restart", "1"]))
    with pytest.raises(OSError, match="vel_restart_000000003"):
        cup.Simulation(SD)
