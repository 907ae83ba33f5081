"""ctypes front end of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
bench.py's cpu_baseline leg, never by the product package. The oracle is an FP64 CPU
restatement of CubismUP_3D's per-timestep update (see cup_oracle.c for the reference
file:line each function follows and for what pins it).

Arrays handed to and from this module are numpy float64 "global" fields indexed
[z][y][x] (x fastest). The C side stores blocks; ``to_blocks``/``from_blocks`` convert.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
BS = 8
FIELDS = {"chi": 0, "u": 1, "v": 2, "w": 3, "p": 4, "tmpU": 5, "tmpV": 6, "tmpW": 7}

_lib = None


def build() -> str:
    """Compile oracle/liboracle.so from cup_oracle.c with the Makefile beside it."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    d, i, p = C.c_double, C.c_int, C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "orc_create": (p, [i, i, i, d]),
        "orc_destroy": (None, [p]),
        "orc_h": (d, [p]),
        "orc_extent": (None, [p, dp]),
        "orc_set_threads": (None, [i]),
        "orc_set_bc": (None, [p, C.POINTER(C.c_int)]),
        "orc_poisson_unbounded": (None, [p]),
        "orc_set_umax_measured": (None, [p, d]),
        "orc_poisson_mixed": (None, [p]),
        "orc_r2r_line": (None, [dp, i, i]),
        "orc_set_plant": (None, [i]),
        "orc_set_field": (None, [p, i, dp]),
        "orc_get_field": (None, [p, i, dp]),
        "orc_lab_shift": (None, [p, i, i, i, i]),
        "orc_find_max_u": (d, [p, dp]),
        "orc_calc_max_timestep": (d, [p, d, d, dp, i, i, dp]),
        "orc_advect_diffuse": (None, [p, d, d, dp]),
        "orc_advect_diffuse_scheme": (None, [p, d, d, dp, i]),
        "orc_adjoint_step": (None, [p, d, d, dp, dp]),
        "orc_adjoint_step_scheme": (None, [p, d, d, dp, i, dp]),
        "orc_adjoint_step_forcing": (None, [p, d, d, dp, i, d, d, dp]),
        "orc_adjoint_step_body": (None, [p, d, d, dp, d, i, dp, d, dp, dp, dp]),
        "orc_adjoint_step_body_full": (None, [p, d, d, dp, i, d, d, d, i, dp, dp, d, dp, dp, dp, dp]),
        "orc_adjoint_step_bodies": (None, [p, d, d, dp, i, d, d, d, i, i, dp, dp, dp, dp, dp, dp]),
        "orc_pressure_rhs": (None, [p, d]),
        "orc_poisson_solve": (None, [p]),
        "orc_grad_p": (None, [p, d]),
        "orc_get_rhs": (None, [p, dp]),
        "orc_set_rhs": (None, [p, dp]),
        "orc_dissipation": (None, [p, d, dp]),
        "orc_hit_analysis": (None, [p, d, dp, dp]),
        "orc_spectral_forcing": (None, [p, d, d, d, d, d, dp]),
        "orc_ic_taylor_green": (None, [p, d]),
        "orc_ic_hit": (None, [p, d, d, C.c_uint64]),
        "orc_ic_hit_spectrum": (None, [p, dp, dp, i, d, C.c_uint64]),
        "orc_spectrum_interp": (d, [dp, dp, i, d]),
        "orc_sphere_create": (None, [p, d, dp]),
        "orc_sphere_create_mollified": (None, [p, d, dp, d]),
        "orc_sphere_info": (None, [p, dp]),
        "orc_penalize": (None, [p, d, d, i, dp, dp, dp]),
        "orc_surface_forces": (None, [p, d, dp, dp, dp]),
        "orc_obstacle_momenta": (None, [p, d, d, i, dp]),
        "orc_hemisphere_create": (None, [p, d, dp]),
        "orc_add_sphere": (i, [p, d, dp, i]),
        "orc_body_count": (i, [p]),
        "orc_move_body": (i, [p, i, dp]),
        "orc_set_udef": (None, [p, dp]),
        "orc_get_udef": (None, [p, dp]),
        "orc_remove_udef_momenta": (None, [p, dp]),
        "orc_select_body": (None, [p, i]),
        "orc_fft3d": (None, [i, i, i, dp, dp, i]),
        "orc_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vec(x, n=3):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(n))
    return a


def to_blocks(g: np.ndarray) -> np.ndarray:
    """[Z][Y][X] global field -> block-major [block][z][y][x], block = bx + nbx*(by + nby*bz)."""
    Z, Y, X = g.shape
    nbz, nby, nbx = Z // BS, Y // BS, X // BS
    return np.ascontiguousarray(
        g.reshape(nbz, BS, nby, BS, nbx, BS).transpose(0, 2, 4, 1, 3, 5).reshape(-1))


def from_blocks(b: np.ndarray, shape) -> np.ndarray:
    Z, Y, X = shape
    nbz, nby, nbx = Z // BS, Y // BS, X // BS
    return np.ascontiguousarray(
        b.reshape(nbz, nby, nbx, BS, BS, BS).transpose(0, 3, 1, 4, 2, 5).reshape(Z, Y, X))


PLANTS = {"rhs_corr_sign": 1, "forces_shear_half": 2, "forces_thrust_drag": 3, "diss_shear_weight": 4,
          "diss_impulse_sign": 5, "diss_angimp_term": 6, "diss_helicity": 7}


def planted(name):
    """Context manager: the oracle runs with the named mistake planted inside (the pin tests
    show with it that their checks fail for that mistake)."""
    import contextlib

    @contextlib.contextmanager
    def scope():
        lib().orc_set_plant(PLANTS[name])
        try:
            yield
        finally:
            lib().orc_set_plant(0)
    return scope()


BC = {"periodic": 0, "dirichlet": 1, "open": 1, "absorbing": 1, "wall": 2, "freespace": 3, "unbounded": 3}


def r2r_line(a, kind):
    """One r2r transform as FFTW defines it: kind 'R2HC', 'HC2R', 'REDFT10' or 'REDFT01'."""
    b = np.array(a, dtype=np.float64)
    lib().orc_r2r_line(_dp(b), b.size, ("R2HC", "HC2R", "REDFT10", "REDFT01").index(kind))
    return b


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return [int(v) for v in o]


def fft3d(a: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalized 3-D complex DFT of a (A,B,C) array: sign -1 forward, +1 backward."""
    re = np.ascontiguousarray(np.real(a), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(a), dtype=np.float64)
    A, B, Cc = a.shape
    lib().orc_fft3d(A, B, Cc, _dp(re), _dp(im), int(sign))
    return re + 1j * im


def spectrum_interp(K, E, k):
    """EnergySpectrum::interpE (SpectralManip.cpp:96-116) of the oracle."""
    K = np.ascontiguousarray(K, dtype=np.float64)
    E = np.ascontiguousarray(E, dtype=np.float64)
    return lib().orc_spectrum_interp(_dp(K), _dp(E), int(K.size), float(k))


def compute_velocities(M, implicit, forced, imposed, block_rot):
    """Obstacle::computeVelocities (Obstacle.cpp:153-212) on the sums of obstacle_momenta:
    solve the 6x6 system of the penalized momenta for (V, Omega). A forced translation keeps
    only its diagonal with right-hand side penalM * imposed; a blocked rotation keeps only
    its diagonal with right-hand side 0 (numpy's LU with partial pivoting in place of GSL's)."""
    M = np.asarray(M, dtype=np.float64)
    if implicit:
        pm, cm, J, L, A_ = M[13], M[14:17], M[17:23], M[23:26], M[26:29]
    else:
        pm, cm, J, L, A_ = M[0], np.zeros(3), M[7:13], M[1:4], M[4:7]
    A = np.array([[pm, 0, 0, 0, cm[2], -cm[1]],
                  [0, pm, 0, -cm[2], 0, cm[0]],
                  [0, 0, pm, cm[1], -cm[0], 0],
                  [0, -cm[2], cm[1], J[0], J[3], J[4]],
                  [cm[2], 0, -cm[0], J[3], J[1], J[5]],
                  [-cm[1], cm[0], 0, J[4], J[5], J[2]]], dtype=np.float64)
    b = np.concatenate([L, A_]).astype(np.float64)
    for d in range(3):
        if forced[d]:
            diag = A[d, d]
            A[d, :] = 0
            A[d, d] = diag
            b[d] = pm * imposed[d]
        if block_rot[d]:
            diag = A[3 + d, 3 + d]
            A[3 + d, :] = 0
            A[3 + d, 3 + d] = diag
            b[3 + d] = 0
    x = np.linalg.solve(A, b)
    V, Om = x[:3].copy(), x[3:].copy()
    for d in range(3):
        if forced[d]:
            V[d] = imposed[d]
        if block_rot[d]:
            Om[d] = 0.0
    return V, Om


class OracleSim:
    """FP64 CPU restatement of the reference's periodic block grid and operators."""

    def __init__(self, bpd, maxextent: float = 2 * np.pi, bc=("periodic", "periodic", "periodic")):
        self.bpd = tuple(int(b) for b in bpd)
        self.cells = tuple(BS * b for b in self.bpd)
        self.shape = (self.cells[2], self.cells[1], self.cells[0])  # [z][y][x]
        self._h = lib().orc_create(self.bpd[0], self.bpd[1], self.bpd[2], float(maxextent))
        if not self._h:
            raise ValueError(f"invalid blocks per dimension {bpd}")
        self.bc = tuple(bc)
        lib().orc_set_bc(self._h, (C.c_int * 3)(*[BC[b] for b in self.bc]))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.orc_destroy(h)
            self._h = None

    @property
    def h(self) -> float:
        return lib().orc_h(self._h)

    @property
    def extent(self):
        e = np.zeros(3)
        lib().orc_extent(self._h, _dp(e))
        return e

    def set(self, name: str, g: np.ndarray):
        b = to_blocks(np.asarray(g, dtype=np.float64))
        lib().orc_set_field(self._h, FIELDS[name], _dp(b))

    def get(self, name: str) -> np.ndarray:
        b = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        lib().orc_get_field(self._h, FIELDS[name], _dp(b))
        return from_blocks(b, self.shape)

    def set_velocity(self, u, v, w):
        self.set("u", u); self.set("v", v); self.set("w", w)

    def velocity(self):
        return self.get("u"), self.get("v"), self.get("w")

    def lab_shift(self, g, dx, dy, dz):
        lib().orc_lab_shift(self._h, int(g), int(dx), int(dy), int(dz))

    def find_max_u(self, uinf=(0, 0, 0)) -> float:
        return lib().orc_find_max_u(self._h, _dp(_vec(uinf)))

    def calc_max_timestep(self, CFL, nu, uinf=(0, 0, 0), step=0, rampup=0):
        um = C.c_double(0)
        dt = lib().orc_calc_max_timestep(self._h, float(CFL), float(nu), _dp(_vec(uinf)),
                                         int(step), int(rampup), C.byref(um))
        return dt, um.value

    def advect_diffuse(self, dt, nu, uinf=(0, 0, 0), scheme=0):
        """scheme 0: central/Euler (the reference); 1: upwind3/Euler; 2: upwind3/RK2 midpoint."""
        lib().orc_advect_diffuse_scheme(self._h, float(dt), float(nu), _dp(_vec(uinf)), int(scheme))

    def adjoint_step(self, dt, nu, lam, uinf=(0, 0, 0), scheme=0, tke_target=0.0, injection_rate=0.0):
        """dJ/du_n from dJ/du_{n+1} = lam (a tuple of three [Z][Y][X] arrays), with the forward
        state u_n in this sim's velocity; central Euler (scheme 0) or upwind Euler/midpoint (1, 2);
        with the linear forcing after the advection when tke_target or injection_rate > 0; no body."""
        buf = np.concatenate([to_blocks(np.asarray(f, dtype=np.float64)) for f in lam])
        lib().orc_adjoint_step_forcing(self._h, float(dt), float(nu), _dp(_vec(uinf)), int(scheme),
                                       float(tke_target), float(injection_rate), _dp(buf))
        n = buf.size // 3
        return tuple(from_blocks(buf[q * n:(q + 1) * n], self.shape) for q in range(3))

    def adjoint_step_body(self, dt, nu, lam, vel, lambda_, implicit, wdrag, uinf=(0, 0, 0), centre=False,
                          ang_vel=None, scheme=0, tke_target=0.0, injection_rate=0.0):
        """With the sphere: (dJ/du_n, dJ/dV of this step) from lam = dJ/du_{n+1}, J containing
        wdrag * F_x of this step (F the penalization force); u_n is this sim's velocity. The
        step advects with scheme and, with tke_target or injection_rate > 0, applies the linear
        forcing after the advection. With ang_vel (the sphere rotating about its centre of mass)
        this step's dJ/d(ang_vel) follows dJ/dV. With centre=True (mollified chi) also this
        step's dJ/d(sphere centre) as the last value (NaN with a nonzero ang_vel)."""
        buf = np.concatenate([to_blocks(np.asarray(f, dtype=np.float64)) for f in lam])
        dv, dw, dc = np.zeros(3), np.zeros(3), np.zeros(3)
        lib().orc_adjoint_step_body_full(self._h, float(dt), float(nu), _dp(_vec(uinf)), int(scheme),
                                         float(tke_target), float(injection_rate), float(lambda_),
                                         int(bool(implicit)), _dp(_vec(vel)),
                                         _dp(_vec(ang_vel if ang_vel is not None else (0, 0, 0))), float(wdrag),
                                         _dp(buf), _dp(dv), _dp(dw), _dp(dc) if centre else None)
        n = buf.size // 3
        lam0 = tuple(from_blocks(buf[q * n:(q + 1) * n], self.shape) for q in range(3))
        out = (lam0, dv) if ang_vel is None else (lam0, dv, dw)
        return out + (dc,) if centre else out

    def adjoint_step_bodies(self, dt, nu, lam, vels, ang_vels, lambda_, implicit, wdrag, uinf=(0, 0, 0), scheme=0,
                            tke_target=0.0, injection_rate=0.0):
        """With every body (vels[k], ang_vels[k] those of body k): (dJ/du_n, dJ/dV, dJ/dW) from
        lam = dJ/du_{n+1}, J containing wdrag[k] times body k's drag of this step; dJ/dV and
        dJ/dW are [nbody][3]."""
        n = len(wdrag)
        buf = np.concatenate([to_blocks(np.asarray(f, dtype=np.float64)) for f in lam])
        V = np.ascontiguousarray(np.asarray(vels, dtype=np.float64).reshape(n, 3))
        W = np.ascontiguousarray(np.asarray(ang_vels, dtype=np.float64).reshape(n, 3))
        w = np.ascontiguousarray(np.asarray(wdrag, dtype=np.float64).reshape(n))
        dv, dw = np.zeros((n, 3)), np.zeros((n, 3))
        lib().orc_adjoint_step_bodies(self._h, float(dt), float(nu), _dp(_vec(uinf)), int(scheme), float(tke_target),
                                      float(injection_rate), float(lambda_), int(bool(implicit)), n, _dp(V), _dp(W),
                                      _dp(w), _dp(buf), _dp(dv), _dp(dw))
        m = buf.size // 3
        return tuple(from_blocks(buf[q * m:(q + 1) * m], self.shape) for q in range(3)), dv, dw

    def hit_analysis(self, nu):
        """HIT analysis: dict of tke, eps, uprime, lambda, Re_lambda, l_integral, tau_integral,
        grad_mean, grad_std and E (the shell spectrum, maxN/2 - 1 bins)."""
        out = np.zeros(9)
        E = np.zeros(max(self.shape) // 2 - 1)
        lib().orc_hit_analysis(self._h, float(nu), _dp(out), _dp(E))
        keys = ("tke", "eps", "uprime", "lambda", "Re_lambda", "l_integral", "tau_integral", "grad_mean", "grad_std")
        r = dict(zip(keys, out.tolist()))
        r["E"] = E
        return r

    def pressure_rhs(self, dt):
        lib().orc_pressure_rhs(self._h, float(dt))

    def rhs(self) -> np.ndarray:
        b = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        lib().orc_get_rhs(self._h, _dp(b))
        return from_blocks(b, self.shape)

    def set_rhs(self, g):
        b = to_blocks(np.asarray(g, dtype=np.float64))
        lib().orc_set_rhs(self._h, _dp(b))

    def poisson_solve(self):
        lib().orc_poisson_solve(self._h)

    def poisson(self):
        """The solver PressureProjection picks (PressureProjection.cpp:86-105): FFT for a fully
        periodic box, PoissonSolverUnbounded with freespace faces, PoissonSolverMixed otherwise."""
        if all(b == "periodic" for b in self.bc):
            self.poisson_solve()
        elif all(BC[b] == 3 for b in self.bc):
            self.poisson_unbounded()
        else:
            self.poisson_mixed()

    def poisson_unbounded(self):
        """PoissonSolverUnbounded's solve (the free-space convolution) of the RHS into p."""
        lib().orc_poisson_unbounded(self._h)

    def poisson_mixed(self):
        """PoissonSolverMixed's solve (DFT on periodic axes, DCT on the others) of the RHS into p."""
        lib().orc_poisson_mixed(self._h)

    def set_umax_measured(self, umax):
        lib().orc_set_umax_measured(self._h, float(umax))

    def grad_p(self, dt):
        lib().orc_grad_p(self._h, float(dt))

    def dissipation(self, nu) -> np.ndarray:
        out = np.zeros(20)
        lib().orc_dissipation(self._h, float(nu), _dp(out))
        return out

    def spectral_forcing(self, dt, nu, tke_target=0.0, injection_rate=0.0, sim_dt=None):
        st = np.zeros(5)
        lib().orc_spectral_forcing(self._h, float(dt), float(dt if sim_dt is None else sim_dt),
                                   float(nu), float(tke_target), float(injection_rate), _dp(st))
        return dict(tke=st[0], eps=st[1], tke_filtered=st[2], l_integral=st[3], fac=st[4])

    def ic_taylor_green(self, uMax):
        lib().orc_ic_taylor_green(self._h, float(uMax))

    def ic_hit(self, k0=4.0, u_rms=1.0, seed=20260101):
        lib().orc_ic_hit(self._h, float(k0), float(u_rms), int(seed))

    def ic_hit_spectrum(self, K, E, u_rms=0.0, seed=20260101):
        """The seeded IC with E(|k|) interpolated from the table (K, E); u_rms <= 0 keeps its energy."""
        K = np.ascontiguousarray(K, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        lib().orc_ic_hit_spectrum(self._h, _dp(K), _dp(E), int(K.size), float(u_rms), int(seed))

    def sphere_create(self, radius, pos, mollify=0.0, hemi=False):
        """The sphere's chi by Towers' formula, or (mollify = half-width in cells, in (0, 2])
        by the mollified Heaviside of the SDF; hemi: the hemisphere (x below the centre)."""
        if hemi:
            if mollify:
                raise ValueError("the hemisphere takes Towers' chi")
            lib().orc_hemisphere_create(self._h, float(radius), _dp(_vec(pos)))
        elif mollify:
            if not 0 < mollify <= 2:
                raise ValueError("mollify: the half-width must lie in (0, 2] cells")
            lib().orc_sphere_create_mollified(self._h, float(radius), _dp(_vec(pos)), float(mollify))
        else:
            lib().orc_sphere_create(self._h, float(radius), _dp(_vec(pos)))

    def add_sphere(self, radius, pos, hemi=False):
        """Another sphere (its obstacle blocks apart from the others'); returns its index."""
        return lib().orc_add_sphere(self._h, float(radius), _dp(_vec(pos)), int(bool(hemi)))

    def body_count(self):
        return lib().orc_body_count(self._h)

    def set_udef(self, ux, uy, uz):
        """The active body's deformation velocity (three [Z][Y][X] arrays; only its obstacle
        blocks are kept)."""
        buf = np.concatenate([to_blocks(np.asarray(f, dtype=np.float64)) for f in (ux, uy, uz)])
        lib().orc_set_udef(self._h, _dp(buf))

    def get_udef(self):
        n = int(np.prod(self.shape))
        buf = np.zeros(3 * n)
        lib().orc_get_udef(self._h, _dp(buf))
        return tuple(from_blocks(buf[q * n:(q + 1) * n], self.shape) for q in range(3))

    def remove_udef_momenta(self):
        """The active body's udef momentum removal; returns dict(mass, trans, ang, J)."""
        o = np.zeros(13)
        lib().orc_remove_udef_momenta(self._h, _dp(o))
        return dict(mass=o[0], trans=o[1:4].copy(), ang=o[4:7].copy(), J=o[7:13].copy())

    def move_body(self, k, pos):
        """Body k at centre pos; chi and the SDF of every body are built again."""
        if lib().orc_move_body(self._h, int(k), _dp(_vec(pos))) != 0:
            raise IndexError(f"no body {k}")

    def body(self, k):
        """Context manager: body k is the active body inside (penalize, surface_forces,
        obstacle_momenta and sphere_info act on the active body)."""
        import contextlib

        @contextlib.contextmanager
        def scope():
            lib().orc_select_body(self._h, int(k))
            try:
                yield self
            finally:
                lib().orc_select_body(self._h, int(k))
        return scope()

    def sphere_info(self):
        o = np.zeros(8)
        lib().orc_sphere_info(self._h, _dp(o))
        return dict(com=o[0:3].copy(), mass=o[3], nblocks=int(o[4]), npoints=int(o[5]))

    def penalize(self, dt, lam, implicit, trans_vel, ang_vel=(0, 0, 0)):
        ft = np.zeros(6)
        lib().orc_penalize(self._h, float(dt), float(lam), int(bool(implicit)),
                           _dp(_vec(trans_vel)), _dp(_vec(ang_vel)), _dp(ft))
        return ft

    def obstacle_momenta(self, dt, lam, implicit):
        """The 29 sums of KernelIntegrateFluidMomenta / KernelFinalizeObstacleVel
        (ObstaclesUpdate.cpp:15-170) over the sphere's obstacle blocks."""
        m = np.zeros(29)
        lib().orc_obstacle_momenta(self._h, float(dt), float(lam), int(bool(implicit)), _dp(m))
        return m

    def surface_forces(self, nu, trans_vel, ang_vel=(0, 0, 0)):
        q = np.zeros(22)
        lib().orc_surface_forces(self._h, float(nu), _dp(_vec(trans_vel)), _dp(_vec(ang_vel)), _dp(q))
        return q

    def body_timestep(self, dt, nu, body, lambda_, implicit):
        """One step with a sphere that may move, in the reference pipeline order
        (Simulation.cpp:307-401): CreateObstacles (frame velocity from the fixed-frame axes,
        ObstacleVector.h:60-74; position += dt (V + U_inf), Obstacle.cpp:314-319; chi at the new
        position), advection-diffusion, RHS, Poisson, gradient, UpdateObstacles (momenta and
        computeVelocities), Penalization, ComputeForces. ``body``: dict with radius,
        position, transVel, angVel, forced, fix, block_rot, imposed; position and velocities
        are updated in place. Returns (U_inf, penalization force/torque, surface QoIs)."""
        uinf = [-body["transVel"][d] if body["fix"][d] else 0.0 for d in range(3)]
        body["position"] = [body["position"][d] + dt * (body["transVel"][d] + uinf[d]) for d in range(3)]
        self.sphere_create(body["radius"], body["position"])
        self.advect_diffuse(dt, nu, uinf)
        self.pressure_rhs(dt)
        self.poisson()
        self.grad_p(dt)
        M = self.obstacle_momenta(dt, lambda_, implicit)
        V, Om = compute_velocities(M, implicit, body["forced"], body["imposed"], body["block_rot"])
        body["transVel"], body["angVel"] = V, Om
        ft = self.penalize(dt, lambda_, implicit, V, Om)
        q = self.surface_forces(nu, V, Om)
        return uinf, ft, q

    def bodies_timestep(self, dt, nu, bodies, lambda_, implicit):
        """body_timestep for the bodies added with add_sphere (bodies[k] is body k): the frame
        velocity is minus the mean velocity of the bodies fixed to the frame on each axis
        (ObstacleVector::updateUinf, ObstacleVector.h:60-74), every body moves and is built
        again where it moved, then one step; the momenta of every body are taken before any
        is penalized (UpdateObstacles, then Penalization). Returns (U_inf, [force/torque of
        each body], [surface QoIs of each body])."""
        uinf = [0.0] * 3
        for d in range(3):
            fixed = [b["transVel"][d] for b in bodies if b["fix"][d]]
            uinf[d] = -sum(fixed) / len(fixed) if fixed else 0.0
        for k, b in enumerate(bodies):
            new = [b["position"][d] + dt * (b["transVel"][d] + uinf[d]) for d in range(3)]
            if new != list(b["position"]):
                self.move_body(k, new)
            b["position"] = new
        self.advect_diffuse(dt, nu, uinf)
        self.pressure_rhs(dt)
        self.poisson()
        self.grad_p(dt)
        for k, b in enumerate(bodies):
            with self.body(k):
                M = self.obstacle_momenta(dt, lambda_, implicit)
            b["transVel"], b["angVel"] = compute_velocities(M, implicit, b["forced"], b["imposed"], b["block_rot"])
        ft, q = [], []
        for k, b in enumerate(bodies):
            with self.body(k):
                ft.append(self.penalize(dt, lambda_, implicit, b["transVel"], b["angVel"]))
        for k, b in enumerate(bodies):
            with self.body(k):
                q.append(self.surface_forces(nu, b["transVel"], b["angVel"]))
        return uinf, ft, q

    def timestep(self, dt, nu, uinf=(0, 0, 0), spectral_forcing=False, tke_target=0.0,
                 sphere=None, scheme=0, injection_rate=0.0):
        """One step in the reference pipeline order. ``sphere``: dict(lambda_, implicit,
        trans_vel, ang_vel) when a sphere was created with sphere_create."""
        out = {}
        self.advect_diffuse(dt, nu, uinf, scheme)
        if spectral_forcing:
            out["forcing"] = self.spectral_forcing(dt, nu, tke_target=tke_target, injection_rate=injection_rate)
        self.pressure_rhs(dt)
        self.poisson()
        self.grad_p(dt)
        if sphere is not None:
            out["penal_ft"] = self.penalize(dt, sphere["lambda_"], sphere["implicit"],
                                            sphere["trans_vel"], sphere.get("ang_vel", (0, 0, 0)))
            out["surface"] = self.surface_forces(nu, sphere["trans_vel"], sphere.get("ang_vel", (0, 0, 0)))
        return out
