"""Python surface of the reference's pybind11 module ``cubismup3d``
(source/bindings/Simulation.cpp:59-94, source/bindings/Obstacles.cpp:79-128) over the
Simulation / SimulationData host C++ classes of libcup.so.

    SD = SimulationData(cells=[64, 64, 64], CFL=0.1, uinf=[0.1, 0, 0])
    SD.nsteps = 10
    S = Simulation(SD)
    S.add_obstacle(Sphere(position=[0.3, 0.4, 0.5], radius=0.1))
    S.run()
"""
from __future__ import annotations

import ctypes as C
import math

from . import _lib as L
from .grid import Grid, from_blocks, to_blocks


def read_field_file(path):
    """A field file of the saves (vel_/pres_/chi_<name>.cupf, written by Simulation::_serialize):
    a 64-byte header ("CUPFIELD", int32 version 1, ncomp, cells x, y, z, step, 0, FP64 time)
    then ncomp FP64 arrays in [comp][z][y][x] order. Returns dict(time, step, data)."""
    import struct
    import numpy as np
    with open(path, "rb") as f:
        head = f.read(64)
        if len(head) != 64 or head[:8] != b"CUPFIELD":
            raise ValueError(f"{path}: not a field file")
        version, ncomp, nx, ny, nz, step, _ = struct.unpack("<7i", head[8:36])
        if version != 1:
            raise ValueError(f"{path}: version {version}")
        (time,) = struct.unpack("<d", head[36:44])
        data = np.fromfile(f, dtype="<f8", count=ncomp * nx * ny * nz)
    if data.size != ncomp * nx * ny * nz:
        raise ValueError(f"{path}: truncated")
    return dict(time=time, step=step, data=data.reshape(ncomp, nz, ny, nx))


def read_xdmf(path):
    """The XDMF descriptor written beside each field file of the saves (vel_/pres_/chi_<name>.xmf):
    parses the 3DCoRectMesh topology, the ORIGIN_DXDYDZ geometry and each cell-centred
    attribute's binary DataItems (file, byte offset, dimensions, FP64 little endian), and reads
    the arrays the way a visualisation tool does. Returns dict(time, cells (x, y, z), origin,
    spacing (both z, y, x as written), attributes {name: array [z][y][x], or [z][y][x][3] for a
    vector joined from three items})."""
    import os
    import xml.etree.ElementTree as ET
    import numpy as np
    root = ET.parse(path).getroot()
    grid = root.find("Domain/Grid")
    topo = grid.find("Topology")
    if topo.get("TopologyType") != "3DCoRectMesh":
        raise ValueError(f"{path}: topology {topo.get('TopologyType')}")
    nz, ny, nx = (int(v) - 1 for v in topo.get("Dimensions").split())
    geo = {d.get("Name"): np.array([float(v) for v in d.text.split()]) for d in grid.find("Geometry")}
    here = os.path.dirname(os.path.abspath(path))

    def item(d):
        dims = tuple(int(v) for v in d.get("Dimensions").split())
        if d.get("Format") != "Binary" or d.get("NumberType") != "Float" or d.get("Precision") != "8":
            raise ValueError(f"{path}: unsupported DataItem")
        dt = "<f8" if d.get("Endian", "Little") == "Little" else ">f8"
        a = np.fromfile(os.path.join(here, d.text.strip()), dtype=dt, count=int(np.prod(dims)),
                        offset=int(d.get("Seek", "0")))
        if a.size != np.prod(dims):
            raise ValueError(f"{path}: {d.text.strip()} is truncated")
        return a.reshape(dims)
    attrs = {}
    for at in grid.findall("Attribute"):
        d = at.find("DataItem")
        if d.get("ItemType") == "Function":
            attrs[at.get("Name")] = np.stack([item(s) for s in d.findall("DataItem")], axis=-1)
        else:
            attrs[at.get("Name")] = item(d)
    t = grid.find("Time")
    return dict(time=float(t.get("Value")) if t is not None else None, cells=(nx, ny, nz),
                origin=geo["Origin"], spacing=geo["Spacing"], attributes=attrs)


class Sphere:
    """ObstacleAndSphereArguments of the bindings (bindings/Obstacles.cpp:12-71, 79-128):
    ``radius`` or ``length``; ``position``; ``enforcedVelocity``, the velocity of the axes
    forced in the simulation frame; ``bForcedInSimFrame`` and ``bFixFrameOfRef``, a bool or one
    per axis (default: free, frame at rest); ``umax``, ``tmax``, ``accel_decel``, the ramped x
    velocity of Sphere::computeVelocities (Sphere.cpp:147-159); ``bHemi``, the hemisphere
    (Sphere.cpp:53-84). ``xvel`` is the command-line shorthand: every axis forced, the frame
    fixed to the body, enforcedVelocity = (-xvel, 0, 0) (Obstacle.cpp:64). ``bFixToPlanar``, a
    non-identity ``quaternion`` and ``bComputeForces=False`` are not supported and raise
    ValueError, as unknown keywords do."""

    def __init__(self, radius=0.0, umax=0.0, tmax=1.0, accel_decel=False, bHemi=False, **kwargs):
        length = float(kwargs.pop("length", 0.0))
        radius = float(radius)
        if radius > 0 and length > 0 and radius != 0.5 * length:
            raise ValueError("Cannot specify both `radius` and `length`.")
        if radius <= 0 and length <= 0:
            raise ValueError("Expected a `radius` or `length`.")
        self.radius = radius if radius > 0 else 0.5 * length
        self.position = [float(x) for x in kwargs.pop("position", (0.0, 0.0, 0.0))]
        xvel = kwargs.pop("xvel", None)
        forced, fix = kwargs.pop("bForcedInSimFrame", None), kwargs.pop("bFixFrameOfRef", None)
        vel = kwargs.pop("enforcedVelocity", None)
        if xvel is not None:
            if vel is not None:
                raise ValueError("give xvel or enforcedVelocity, not both")
            vel = (-float(xvel), 0.0, 0.0)
            forced = True if forced is None else forced
            fix = True if fix is None else fix

        def axes(v):
            v = False if v is None else v
            return [bool(v)] * 3 if isinstance(v, (bool, int)) else [bool(a) for a in v]

        self.forced, self.fix_frame = axes(forced), axes(fix)
        self.trans_vel = [float(x) for x in (vel if vel is not None else (0.0, 0.0, 0.0))]
        if len(self.position) != 3 or len(self.trans_vel) != 3 or len(self.forced) != 3 or len(self.fix_frame) != 3:
            raise ValueError("position, enforcedVelocity and the axis flags take three values")
        q = [float(x) for x in kwargs.pop("quaternion", (1.0, 0.0, 0.0, 0.0))]
        if kwargs.pop("bFixToPlanar", False) or q != [1.0, 0.0, 0.0, 0.0] or not kwargs.pop("bComputeForces", True):
            raise ValueError("bFixToPlanar, a rotated quaternion and bComputeForces=False are not supported")
        self.bHemi = bool(bHemi)
        if kwargs:
            raise ValueError(f"unknown sphere arguments: {sorted(kwargs)}")
        self.umax, self.tmax, self.accel_decel = float(umax), float(tmax), bool(accel_decel)

    def masks(self):
        return (sum(1 << d for d in range(3) if self.forced[d]), sum(1 << d for d in range(3) if self.fix_frame[d]))


class SimulationData:
    """SimulationData with the reference's command-line vocabulary
    (SimulationData.cpp:32-172). ``argv`` takes precedence over keyword settings."""

    _INT = {"nsteps", "rampup", "freqDiagnostics", "freqAnalysis", "bpdx", "bpdy", "bpdz"}
    _BOOL = {"verbose", "muteAll", "bCheckNaN", "bImplicitPenalization", "spectralForcing"}

    def __init__(self, cells=None, CFL=None, uinf=(0.0, 0.0, 0.0), argv=None, **kw):
        if argv is None:
            if cells is None:
                raise TypeError("cells are required")
            if CFL is None:
                raise TypeError("CFL is required")
            cells = [int(c) for c in cells]
            if len(cells) != 3 or any(c <= 0 or c % 8 for c in cells):
                raise ValueError("Number of cells must be multiple of block size.")
            argv = ["-bpdx", str(cells[0] // 8), "-bpdy", str(cells[1] // 8), "-bpdz", str(cells[2] // 8),
                    "-BC_x", "periodic", "-BC_y", "periodic", "-BC_z", "periodic",
                    "-CFL", repr(float(CFL)), "-nu", repr(float(kw.pop("nu", 0.0))),
                    "-uinfx", repr(float(uinf[0])), "-uinfy", repr(float(uinf[1])),
                    "-uinfz", repr(float(uinf[2])), "-extentx", repr(float(kw.pop("extent", 2 * math.pi))),
                    "-freqDiagnostics", str(kw.pop("freqDiagnostics", 0)), "-muteAll",
                    str(int(kw.pop("muteAll", True)))]
            for k, v in kw.items():
                argv += [f"-{k}", str(v)]
        self.argv = [str(a) for a in argv]
        arr = (C.c_char_p * len(self.argv))(*[a.encode() for a in self.argv])
        h = C.c_void_p()
        rc = L.lib().cup_sim_create(len(self.argv), arr, C.byref(h))
        self._h = h
        L.check(rc, sim=self._h)
        self._owned = True

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        v = C.c_double()
        rc = L.lib().cup_sim_get(self.__dict__["_h"], name.encode(), C.byref(v))
        if rc != L.CUP_OK:
            raise AttributeError(name)
        if name in self._BOOL:
            return bool(v.value)
        return int(v.value) if name in self._INT or name == "step" else v.value

    def __setattr__(self, name, value):
        if name.startswith("_") or name == "argv":
            object.__setattr__(self, name, value)
            return
        L.check(L.lib().cup_sim_set(self._h, name.encode(), float(value)), sim=self._h)

    def setCells(self, nx, ny, nz):
        L.check(L.lib().cup_sim_set_cells(self._h, int(nx), int(ny), int(nz)), sim=self._h)

    def __del__(self):
        if self.__dict__.get("_owned") and self.__dict__.get("_h"):
            L.lib().cup_sim_destroy(self._h)
            object.__setattr__(self, "_h", None)


class Simulation:
    """Simulation(SD): grid, operator pipeline and initial condition on the device."""

    def __init__(self, SD: SimulationData):
        if not isinstance(SD, SimulationData):
            raise TypeError("Simulation needs a SimulationData")
        self.sim = SD
        self._h = SD._h
        L.check(L.lib().cup_sim_init(self._h), sim=self._h)
        ctx = C.c_void_p()
        L.check(L.lib().cup_sim_context(self._h, C.byref(ctx)), sim=self._h)
        self._ctx = ctx
        bpd = (SD.bpdx, SD.bpdy, SD.bpdz)
        self.shape = (8 * bpd[2], 8 * bpd[1], 8 * bpd[0])

    def add_obstacle(self, s: Sphere):
        if not isinstance(s, Sphere):  # Simulation_addObstacle (bindings/Obstacles.cpp:120-128)
            raise ValueError(str(s))
        fm, xm = s.masks()
        L.check(L.lib().cup_sim_add_sphere(self._h, s.radius, L.vec3(s.position), L.vec3(s.trans_vel), fm, xm,
                                           int(s.accel_decel), s.umax, s.tmax, int(s.bHemi)), sim=self._h)

    def run(self):
        L.check(L.lib().cup_sim_run(self._h), sim=self._h)

    def calcMaxTimestep(self) -> float:
        dt = C.c_double()
        L.check(L.lib().cup_sim_calc_max_timestep(self._h, C.byref(dt)), sim=self._h)
        return dt.value

    def timestep(self, dt) -> bool:
        f = C.c_int()
        L.check(L.lib().cup_sim_timestep(self._h, float(dt), C.byref(f)), sim=self._h)
        return bool(f.value)

    def advance(self, nsteps) -> bool:
        f = C.c_int()
        L.check(L.lib().cup_sim_advance(self._h, int(nsteps), C.byref(f)), sim=self._h)
        return bool(f.value)

    def field(self, name):
        import numpy as np
        b = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        L.check(L.lib().cup_get_field(self._ctx, L.FIELD_IDS[name], L.dptr(b)), ctx=self._ctx)
        return from_blocks(b, self.shape)

    def set_field(self, name, g):
        """Write a [z][y][x] field (chi, u, v, w, p) into the simulation's device context."""
        import numpy as np
        g = np.asarray(g, dtype=np.float64)
        if g.shape != self.shape:
            raise ValueError(f"field shape {g.shape} != grid {self.shape}")
        b = to_blocks(g)
        L.check(L.lib().cup_set_field(self._ctx, L.FIELD_IDS[name], L.dptr(b)), ctx=self._ctx)

    def obstacle(self, idx=0):
        import numpy as np
        o = np.zeros(34)
        L.check(L.lib().cup_sim_obstacle(self._h, int(idx), L.dptr(o)), sim=self._h)
        st = np.zeros(13)
        L.check(L.lib().cup_sim_obstacle_state(self._h, int(idx), L.dptr(st)), sim=self._h)
        return dict(com=o[0:3].copy(), mass=o[3], force=o[4:7].copy(), torque=o[7:10].copy(),
                    surface=o[10:32].copy(), position=st[0:3].copy(), quaternion=st[3:7].copy(),
                    trans_vel=st[7:10].copy(), ang_vel=st[10:13].copy())

    def diagnostics(self):
        import numpy as np
        o = np.zeros(20)
        L.check(L.lib().cup_sim_diagnostics(self._h, L.dptr(o)), sim=self._h)
        return o

    def flush_logs(self):
        L.check(L.lib().cup_sim_flush_logs(self._h), sim=self._h)
