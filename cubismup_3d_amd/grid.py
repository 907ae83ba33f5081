"""Operator-level access to the device block store (the C-ABI of include/cup.h).

``Grid`` owns one device context: a periodic box of bpd[0] x bpd[1] x bpd[2] blocks of
8^3 cells. Fields cross the boundary as numpy float64 arrays indexed [z][y][x]
(x fastest); on the device they live block-major in the context precision.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

BS = 8


def to_blocks(g: np.ndarray) -> np.ndarray:
    """[Z][Y][X] field -> block-major [block][z][y][x], block = bx + nbx*(by + nby*bz)."""
    Z, Y, X = g.shape
    nbz, nby, nbx = Z // BS, Y // BS, X // BS
    return np.ascontiguousarray(
        g.reshape(nbz, BS, nby, BS, nbx, BS).transpose(0, 2, 4, 1, 3, 5).reshape(-1), dtype=np.float64)


def from_blocks(b: np.ndarray, shape) -> np.ndarray:
    Z, Y, X = shape
    nbz, nby, nbx = Z // BS, Y // BS, X // BS
    return np.ascontiguousarray(
        b.reshape(nbz, nby, nbx, BS, BS, BS).transpose(0, 3, 1, 4, 2, 5).reshape(Z, Y, X))


class Grid:
    """A device context of the per-timestep fluid update (cup_create / cup_destroy)."""

    def __init__(self, bpd, maxextent=2 * np.pi, precision=8, decomposition="plain", device=0,
                 bc=("periodic", "periodic", "periodic")):
        self.bpd = tuple(int(b) for b in bpd)
        self.bc = tuple(bc)
        d = L.GridDesc()
        for a in range(3):
            d.bpd[a] = self.bpd[a]
            d.bc[a] = L.BC_IDS[self.bc[a]]
        d.maxextent = float(maxextent)
        d.precision = int(precision)
        d.decomposition = L.CUP_FUSED if decomposition == "fused" else L.CUP_PLAIN
        d.device = int(device)
        self.precision = int(precision)
        self.decomposition = decomposition
        h = C.c_void_p()
        rc = L.lib().cup_create(C.byref(d), C.byref(h))
        if rc != L.CUP_OK:
            msg = (L.lib().cup_last_error(None) or b"").decode()
            if rc == L.CUP_E_INVALID:
                raise ValueError(f"invalid grid {bpd} / precision {precision}: {msg}")
            raise L.CupError(rc, msg or "cup_create failed")
        self._h = h
        cells = (C.c_int * 3)()
        hh = C.c_double()
        ext = (C.c_double * 3)()
        L.lib().cup_get_geometry(self._h, cells, C.byref(hh), ext)
        self.cells = tuple(cells)
        self.shape = (self.cells[2], self.cells[1], self.cells[0])
        self.h = hh.value
        self.extent = np.array(list(ext))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().cup_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        L.check(rc, ctx=self._h)

    # ----------------------------------------------------------------- fields
    def set(self, name: str, g: np.ndarray):
        g = np.asarray(g, dtype=np.float64)
        if g.shape != self.shape:
            raise ValueError(f"field shape {g.shape} != grid {self.shape}")
        b = to_blocks(g)
        self._c(L.lib().cup_set_field(self._h, L.FIELD_IDS[name], L.dptr(b)))

    def get(self, name: str) -> np.ndarray:
        b = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        self._c(L.lib().cup_get_field(self._h, L.FIELD_IDS[name], L.dptr(b)))
        return from_blocks(b, self.shape)

    def set_velocity(self, u, v, w):
        self.set("u", u); self.set("v", v); self.set("w", w)

    def velocity(self):
        return self.get("u"), self.get("v"), self.get("w")

    def device_ptr(self, name: str) -> int:
        p = C.c_void_p()
        self._c(L.lib().cup_field_device_ptr(self._h, L.FIELD_IDS[name], C.byref(p)))
        return p.value

    def synchronize(self):
        self._c(L.lib().cup_synchronize(self._h))

    # ---------------------------------------------------- torch device tensors
    # A field as a torch tensor [Z][Y][X] on the GPU in the context's precision. Copies are
    # device to device; both streams are synchronized around them, so the tensors may be
    # used on torch's current stream right away.
    def _torch_dtype(self):
        import torch
        return torch.float32 if self.precision == 4 else torch.float64

    def set_torch(self, name: str, t):
        import torch
        if tuple(t.shape) != self.shape or t.dtype != self._torch_dtype() or t.device.type != "cuda":
            raise ValueError(f"expected a cuda {self._torch_dtype()} tensor of shape {self.shape}")
        Z, Y, X = self.shape
        blk = t.detach().reshape(Z // BS, BS, Y // BS, BS, X // BS, BS).permute(0, 2, 4, 1, 3, 5).contiguous()
        torch.cuda.current_stream().synchronize()
        self._c(L.lib().cup_set_field_device(self._h, L.FIELD_IDS[name], C.c_void_p(blk.data_ptr())))
        self.synchronize()

    def get_torch(self, name: str):
        import torch
        Z, Y, X = self.shape
        blk = torch.empty((Z // BS, Y // BS, X // BS, BS, BS, BS), dtype=self._torch_dtype(), device="cuda")
        torch.cuda.current_stream().synchronize()
        self._c(L.lib().cup_get_field_device(self._h, L.FIELD_IDS[name], C.c_void_p(blk.data_ptr())))
        self.synchronize()
        return blk.permute(0, 3, 1, 4, 2, 5).reshape(Z, Y, X)

    def set_torch_velocity(self, u):
        """u: [3][Z][Y][X] tensor -> (u, v, w)."""
        for q, n in enumerate("uvw"):
            self.set_torch(n, u[q])

    def get_torch_velocity(self):
        import torch
        return torch.stack([self.get_torch(n) for n in "uvw"])

    def set_torch_adjoint(self, lam):
        for q, n in enumerate(("adjU", "adjV", "adjW")):
            self.set_torch(n, lam[q])

    def get_torch_adjoint(self):
        import torch
        return torch.stack([self.get_torch(n) for n in ("adjU", "adjV", "adjW")])

    def adjoint_step(self, dt, nu, uinf=(0, 0, 0), scheme=L.CUP_CENTRAL_EULER, tke_target=0.0, injection_rate=0.0):
        """adj <- dJ/du_n from adj = dJ/du_{n+1}, with u_n the current velocity; with the linear
        forcing's transpose when tke_target or injection_rate > 0."""
        self._c(L.lib().cup_adjoint_step_forcing(self._h, float(dt), float(nu), L.vec3(uinf), int(scheme),
                                                 float(tke_target), float(injection_rate)))

    def adjoint_step_body(self, dt, nu, uinf, sphere, wdrag=0.0, centre=False, rotation=False, scheme=0,
                          tke_target=0.0, injection_rate=0.0):
        """With the sphere (rotating at sphere["ang_vel"] about its centre of mass when that is
        given): adj <- dJ/du_n for the step with scheme's advection and, with tke_target or
        injection_rate > 0, the linear forcing after it; returns this step's dJ/d(body
        velocity), J holding wdrag times the step's drag (penalization force, x).
        rotation=True returns (dJ/dV, dJ/d(ang_vel)); centre=True (mollified chi, no rotation)
        appends dJ/d(sphere centre)."""
        out, dw, dc = np.zeros(3), np.zeros(3), np.zeros(3)
        self._c(L.lib().cup_adjoint_step_body_full(
            self._h, float(dt), float(nu), L.vec3(uinf), int(scheme), float(tke_target), float(injection_rate),
            float(sphere["lambda_"]), int(bool(sphere["implicit"])), L.vec3(sphere["trans_vel"]),
            L.vec3(sphere.get("ang_vel", (0.0, 0.0, 0.0))), float(wdrag), L.dptr(out), L.dptr(dw),
            L.dptr(dc) if centre else None))
        res = (out, dw) if rotation else (out,)
        res = res + (dc,) if centre else res
        return res if len(res) > 1 else out

    def adjoint_step_bodies(self, dt, nu, uinf, bodies, lambda_, implicit, wdrag, scheme=0, tke_target=0.0,
                            injection_rate=0.0):
        """With every body (bodies[k] = dict(trans_vel, ang_vel) of body k, as run_steps
        penalized them): adj <- dJ/du_n for J holding wdrag[k] times body k's drag of this step;
        returns (dJ/dV, dJ/d(ang_vel)), each [nbody][3]."""
        n = len(bodies)
        vel = np.array([b["trans_vel"] for b in bodies], dtype=np.float64).reshape(n, 3)
        om = np.array([b.get("ang_vel", (0.0, 0.0, 0.0)) for b in bodies], dtype=np.float64).reshape(n, 3)
        w = np.ascontiguousarray(np.asarray(wdrag, dtype=np.float64).reshape(n))
        dv, dw = np.zeros((n, 3)), np.zeros((n, 3))
        self._c(L.lib().cup_adjoint_step_bodies(self._h, float(dt), float(nu), L.vec3(uinf), int(scheme),
                                                float(tke_target), float(injection_rate), float(lambda_),
                                                int(bool(implicit)), n, L.dptr(vel), L.dptr(om), L.dptr(w),
                                                L.dptr(dv), L.dptr(dw)))
        return dv, dw

    def last_body_force(self) -> np.ndarray:
        """Penalization force and torque of the last run_steps step with the sphere."""
        out = np.zeros(6)
        self._c(L.lib().cup_last_body_force(self._h, L.dptr(out)))
        return out

    def stream(self) -> int:
        s = C.c_void_p()
        self._c(L.lib().cup_stream(self._h, C.byref(s)))
        return s.value or 0

    # -------------------------------------------------------------- operators
    def lab_shift(self, g, dx, dy, dz):
        self._c(L.lib().cup_lab_shift(self._h, int(g), int(dx), int(dy), int(dz)))

    def max_velocity(self, uinf=(0, 0, 0)) -> float:
        m = C.c_double()
        self._c(L.lib().cup_max_velocity(self._h, L.vec3(uinf), C.byref(m)))
        return m.value

    def calc_max_timestep(self, CFL, nu, uinf=(0, 0, 0), step=0, rampup=0):
        dt, um = C.c_double(), C.c_double()
        self._c(L.lib().cup_calc_max_timestep(self._h, float(CFL), float(nu), L.vec3(uinf), int(step),
                                              int(rampup), C.byref(dt), C.byref(um)))
        return dt.value, um.value

    def advect_diffuse(self, dt, nu, uinf=(0, 0, 0), scheme=L.CUP_CENTRAL_EULER):
        self._c(L.lib().cup_advect_diffuse(self._h, float(dt), float(nu), L.vec3(uinf), int(scheme)))

    def pressure_rhs(self, dt):
        self._c(L.lib().cup_pressure_rhs(self._h, float(dt)))

    def poisson_periodic(self):
        self._c(L.lib().cup_poisson_periodic(self._h))

    def poisson_mixed(self):
        self._c(L.lib().cup_poisson_mixed(self._h))

    def poisson_unbounded(self):
        """PoissonSolverUnbounded: the free-space solution of the RHS into p (freespace boxes)."""
        self._c(L.lib().cup_poisson_unbounded(self._h))

    def pressure_project(self, dt):
        self._c(L.lib().cup_pressure_project(self._h, float(dt)))

    def diagnostics(self, nu) -> np.ndarray:
        out = np.zeros(20)
        self._c(L.lib().cup_diagnostics(self._h, float(nu), L.dptr(out)))
        return out

    def check_nan(self) -> np.ndarray:
        """Operator::check's scan (Operator.h:23-50): the NaN cells of u, v, w, p."""
        out = np.zeros(4, dtype=np.int64)
        self._c(L.lib().cup_check_nan(self._h, out.ctypes.data_as(C.POINTER(C.c_longlong))))
        return out

    def hit_analysis(self, nu):
        """The HIT analysis of Analysis.cpp (gradient statistics and the spectral analysis of
        SpectralManipFFTW::_compute_analysis): dict of tke, eps, uprime, lambda, Re_lambda,
        l_integral, tau_integral, grad_mean, grad_std, and E (the shell spectrum, max(N)/2 - 1
        values) with k (its wavenumbers (i+1) 2 pi / max extent)."""
        out = np.zeros(9)
        nb = max(self.cells) // 2 - 1
        E = np.zeros(nb)
        self._c(L.lib().cup_hit_analysis(self._h, float(nu), L.dptr(out), L.dptr(E), int(nb)))
        keys = ("tke", "eps", "uprime", "lambda", "Re_lambda", "l_integral", "tau_integral", "grad_mean",
                "grad_std")
        r = dict(zip(keys, out.tolist()))
        r["E"] = E
        r["k"] = (np.arange(nb) + 1) * 2 * np.pi / float(max(self.extent))
        return r

    def ic_taylor_green(self, uMax):
        self._c(L.lib().cup_ic_taylor_green(self._h, float(uMax)))

    def ic_hit(self, k0=4.0, u_rms=1.0, seed=20260101):
        self._c(L.lib().cup_ic_hit(self._h, float(k0), float(u_rms), int(seed)))

    def ic_hit_spectrum(self, K, E, u_rms=0.0, seed=20260101):
        """The seeded isotropic IC with E(|k|) interpolated from the target table (K, E);
        u_rms <= 0 keeps the table's energy."""
        K = np.ascontiguousarray(K, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        if K.shape != E.shape:
            raise ValueError("K and E differ in length")
        self._c(L.lib().cup_ic_hit_spectrum(self._h, L.dptr(K), L.dptr(E), int(K.size), float(u_rms), int(seed)))

    def spectral_forcing(self, dt, nu, tke_target=0.0, injection_rate=0.0, sim_dt=None):
        st = np.zeros(5)
        self._c(L.lib().cup_spectral_forcing(self._h, float(dt), float(dt if sim_dt is None else sim_dt),
                                             float(nu), float(tke_target), float(injection_rate), L.dptr(st)))
        return dict(tke=st[0], eps=st[1], tke_filtered=st[2], l_integral=st[3], fac=st[4])

    def chi_from_sphere(self, center, radius, mollify=0.0, hemi=False):
        """The sphere's chi by Towers' formula, or with mollify (a half-width in cells, in
        (0, 2]) by the mollified Heaviside of its SDF, which is C1 in the centre. hemi: the
        hemisphere (the ball's part below the centre in x)."""
        out = np.zeros(4)
        if hemi:
            if mollify:
                raise ValueError("the hemisphere takes Towers' chi")
            self._c(L.lib().cup_chi_from_hemisphere(self._h, L.vec3(center), float(radius), L.dptr(out)))
        elif mollify:
            self._c(L.lib().cup_chi_from_sphere_mollified(self._h, L.vec3(center), float(radius), float(mollify),
                                                          L.dptr(out)))
        else:
            self._c(L.lib().cup_chi_from_sphere(self._h, L.vec3(center), float(radius), L.dptr(out)))
        return dict(com=out[:3].copy(), mass=out[3])

    def penalize(self, dt, lam, implicit, trans_vel, ang_vel=(0, 0, 0)) -> np.ndarray:
        ft = np.zeros(6)
        self._c(L.lib().cup_penalize(self._h, float(dt), float(lam), int(bool(implicit)), L.vec3(trans_vel),
                                     L.vec3(ang_vel), L.dptr(ft)))
        return ft

    def add_sphere(self, center, radius, hemi=False):
        """Another body (its obstacle blocks apart from the others'); returns (index, com, mass)."""
        out = np.zeros(4)
        k = C.c_int()
        self._c(L.lib().cup_add_sphere(self._h, L.vec3(center), float(radius), int(bool(hemi)), L.dptr(out),
                                       C.byref(k)))
        return k.value, out[:3].copy(), out[3]

    def body_count(self) -> int:
        n = C.c_int()
        self._c(L.lib().cup_body_count(self._h, C.byref(n)))
        return n.value

    def move_body(self, k, center):
        out = np.zeros(4)
        self._c(L.lib().cup_move_body(self._h, int(k), L.vec3(center), L.dptr(out)))
        return dict(com=out[:3].copy(), mass=out[3])

    def body_velocity(self, k, trans_vel, ang_vel=(0, 0, 0)):
        """The velocity and rotation run_steps penalizes body k >= 1 with (body 0 takes sphere=)."""
        self._c(L.lib().cup_body_velocity(self._h, int(k), L.vec3(trans_vel), L.vec3(ang_vel)))

    def last_body_force_body(self, k) -> np.ndarray:
        out = np.zeros(6)
        self._c(L.lib().cup_last_body_force_body(self._h, int(k), L.dptr(out)))
        return out

    def obstacle_blocks(self, k=0) -> np.ndarray:
        """Block ids of body k's obstacle blocks, in the order its udef uses."""
        n = C.c_int()
        self._c(L.lib().cup_body_obstacle_blocks(self._h, int(k), None, 0, C.byref(n)))
        ids = (C.c_int * max(n.value, 1))()
        self._c(L.lib().cup_body_obstacle_blocks(self._h, int(k), ids, n.value, C.byref(n)))
        return np.array(ids[:n.value], dtype=np.int64)

    def set_udef(self, k, ux, uy, uz):
        """Body k's deformation velocity from three [Z][Y][X] arrays (its obstacle blocks are kept)."""
        ids = self.obstacle_blocks(k)
        blk = np.stack([to_blocks(np.asarray(f, dtype=np.float64)).reshape(-1, 512)[ids] for f in (ux, uy, uz)], 1)
        buf = np.ascontiguousarray(blk)
        self._c(L.lib().cup_body_set_udef(self._h, int(k), L.dptr(buf), len(ids)))

    def get_udef(self, k=0):
        """Body k's deformation velocity as three [Z][Y][X] arrays, 0 outside its obstacle blocks."""
        ids = self.obstacle_blocks(k)
        buf = np.zeros((len(ids), 3, 512))
        self._c(L.lib().cup_body_get_udef(self._h, int(k), L.dptr(buf), len(ids)))
        out = []
        for q in range(3):
            full = np.zeros((int(np.prod(self.shape)) // 512, 512))
            full[ids] = buf[:, q]
            out.append(from_blocks(full.ravel(), self.shape))
        return tuple(out)

    def remove_udef_momenta(self, k=0):
        """CreateObstacles' udef momentum removal for body k: dict(mass, trans, ang, J)."""
        o = np.zeros(13)
        self._c(L.lib().cup_body_remove_udef_momenta(self._h, int(k), L.dptr(o)))
        return dict(mass=o[0], trans=o[1:4].copy(), ang=o[4:7].copy(), J=o[7:13].copy())

    def penalize_body(self, k, dt, lam, implicit, trans_vel, ang_vel=(0, 0, 0)) -> np.ndarray:
        ft = np.zeros(6)
        self._c(L.lib().cup_penalize_body(self._h, int(k), float(dt), float(lam), int(bool(implicit)),
                                          L.vec3(trans_vel), L.vec3(ang_vel), L.dptr(ft)))
        return ft

    def surface_forces_body(self, k, nu, trans_vel, ang_vel=(0, 0, 0)) -> np.ndarray:
        q = np.zeros(22)
        self._c(L.lib().cup_surface_forces_body(self._h, int(k), float(nu), L.vec3(trans_vel), L.vec3(ang_vel),
                                                L.dptr(q)))
        return q

    def obstacle_momenta(self, dt, lam, implicit) -> np.ndarray:
        """UpdateObstacles' 29 momentum sums over the sphere's blocks (ObstaclesUpdate.cpp:15-170)."""
        m = np.zeros(29)
        self._c(L.lib().cup_obstacle_momenta(self._h, float(dt), float(lam), int(bool(implicit)), L.dptr(m)))
        return m

    def surface_forces(self, nu, trans_vel, ang_vel=(0, 0, 0)) -> np.ndarray:
        q = np.zeros(22)
        self._c(L.lib().cup_surface_forces(self._h, float(nu), L.vec3(trans_vel), L.vec3(ang_vel), L.dptr(q)))
        return q

    def timestep(self, dt, nu, uinf=(0, 0, 0), spectral_forcing=False, tke_target=0.0, sphere=None):
        """One step through the operator entry points in the reference pipeline order."""
        out = {}
        self.advect_diffuse(dt, nu, uinf)
        if spectral_forcing:
            out["forcing"] = self.spectral_forcing(dt, nu, tke_target=tke_target)
        self.pressure_rhs(dt)
        self.poisson_periodic()
        self.pressure_project(dt)
        if sphere is not None:
            out["penal_ft"] = self.penalize(dt, sphere["lambda_"], sphere["implicit"], sphere["trans_vel"],
                                            sphere.get("ang_vel", (0, 0, 0)))
            out["surface"] = self.surface_forces(nu, sphere["trans_vel"], sphere.get("ang_vel", (0, 0, 0)))
        return out

    def run_steps(self, nsteps, CFL=0.3, nu=0.0, uinf=(0, 0, 0), rampup=0, fixed_dt=0.0,
                  spectral_forcing=False, tke_target=0.0, injection_rate=0.0, sphere=None,
                  scheme=L.CUP_CENTRAL_EULER):
        p = L.StepParams()
        p.CFL = float(CFL); p.nu = float(nu)
        for a in range(3):
            p.uinf[a] = float(uinf[a])
        p.rampup = int(rampup); p.scheme = int(scheme)
        p.spectral_forcing = int(bool(spectral_forcing)); p.tke_target = float(tke_target)
        p.injection_rate = float(injection_rate); p.fixed_dt = float(fixed_dt)
        if sphere is not None:
            p.has_body = 1
            p.lambda_ = float(sphere["lambda_"]); p.implicit_penalization = int(bool(sphere["implicit"]))
            for a in range(3):
                p.body_vel[a] = float(sphere["trans_vel"][a])
                p.body_angvel[a] = float(sphere.get("ang_vel", (0, 0, 0))[a])
        self._c(L.lib().cup_run_steps(self._h, C.byref(p), int(nsteps)))

    def step_state(self):
        s = C.c_int(); t, dt, um, ke = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self._c(L.lib().cup_step_state(self._h, C.byref(s), C.byref(t), C.byref(dt), C.byref(um), C.byref(ke)))
        return dict(step=s.value, time=t.value, last_dt=dt.value, umax=um.value, ke=ke.value)

    STAGES = {"advdiff": 0, "xfwd": 1, "forcing_reduce": 2, "yfwd": 3, "zpass": 4, "yinv": 5, "xinv": 6,
              "gradp": 7}

    def stage_timing_begin(self, stage, max_steps):
        self._c(L.lib().cup_stage_timing_begin(self._h, int(self.STAGES.get(stage, stage)), int(max_steps)))

    def stage_timing_end(self):
        ms, n = C.c_double(), C.c_int()
        self._c(L.lib().cup_stage_timing_end(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_step_state(self, step, time):
        self._c(L.lib().cup_set_step_state(self._h, int(step), float(time)))
