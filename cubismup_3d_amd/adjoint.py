"""Differentiable rollouts of the periodic timestep for torch.

``rollout(u0, grid, nsteps, nu, ...)`` advances a [3][Z][Y][X] velocity tensor through
nsteps steps of the device path (central advection with Euler, or the third-order upwind
scheme with Euler or the midpoint rule; optionally the linear forcing, to a target energy or at
a constant injection rate; FFT projection; no obstacle) and is a
``torch.autograd.Function``: its backward is the hand-written adjoint of the step
(cup_adjoint_step: the transposed linearised advection-diffusion stencil and the
symmetric projection), run backwards over the steps with
checkpoint-and-recompute: the forward keeps the velocity every ``every`` steps, and the
backward recomputes each segment's states from its checkpoint before stepping the adjoint
through it. Timesteps are those of the forward pass, held fixed (the CFL's max over the
velocity is not differentiated). No counterpart in the reference.
"""
from __future__ import annotations

import torch

from .grid import Grid


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, grid: Grid, nsteps: int, nu: float, uinf, dts, every: int, CFL: float, dts_out: list,
                scheme: int, forcing: tuple):
        """dts: fixed timesteps, or None for CFL timesteps (appended to dts_out); forcing:
        (tke_target, injection_rate), both 0 for none."""
        grid.set_torch_velocity(u0.detach())
        ckpts = {0: u0.detach().clone()}
        used = []
        fkw = dict(spectral_forcing=any(f > 0 for f in forcing), tke_target=forcing[0], injection_rate=forcing[1])
        for n in range(nsteps):
            if dts is None:
                grid.run_steps(1, CFL=CFL, nu=nu, uinf=uinf, scheme=scheme, **fkw)
                used.append(grid.step_state()["last_dt"])
            else:
                grid.run_steps(1, nu=nu, uinf=uinf, fixed_dt=dts[n], scheme=scheme, **fkw)
                used.append(float(dts[n]))
            if (n + 1) % every == 0 and n + 1 < nsteps:
                ckpts[n + 1] = grid.get_torch_velocity()
        dts_out.extend(used)
        ctx.grid, ctx.nsteps, ctx.nu, ctx.uinf, ctx.dts, ctx.every = grid, nsteps, nu, uinf, used, every
        ctx.scheme = scheme
        ctx.forcing, ctx.fkw = forcing, fkw
        ctx.ckpts = ckpts
        ctx.u0_dtype = u0.dtype
        return grid.get_torch_velocity()

    @staticmethod
    def backward(ctx, g):
        grid, every = ctx.grid, ctx.every
        lam = g.detach().contiguous().to(grid._torch_dtype())
        for s0 in sorted(ctx.ckpts, reverse=True):
            s1 = min(s0 + every, ctx.nsteps)
            states = [ctx.ckpts[s0]]
            grid.set_torch_velocity(ctx.ckpts[s0])
            for n in range(s0, s1 - 1):
                grid.run_steps(1, nu=ctx.nu, uinf=ctx.uinf, fixed_dt=ctx.dts[n], scheme=ctx.scheme, **ctx.fkw)
                states.append(grid.get_torch_velocity())
            for n in reversed(range(s0, s1)):
                grid.set_torch_velocity(states[n - s0])
                grid.set_torch_adjoint(lam)
                grid.adjoint_step(ctx.dts[n], ctx.nu, ctx.uinf, scheme=ctx.scheme, tke_target=ctx.forcing[0],
                                  injection_rate=ctx.forcing[1])
                lam = grid.get_torch_adjoint()
        ctx.ckpts = None
        return lam.to(ctx.u0_dtype), None, None, None, None, None, None, None, None, None, None


def rollout(u0, grid: Grid, nsteps: int, nu: float, uinf=(0.0, 0.0, 0.0), dts=None, CFL=0.3, every=10,
            dts_out=None, scheme=0, tke_target=0.0, injection_rate=0.0):
    """u_N of a rollout, differentiable with respect to u0 ([3][Z][Y][X], cuda). With dts
    None the steps take CFL timesteps (rampup 0), which are appended to dts_out and held
    fixed by the backward. scheme: 0 central Euler, 1 upwind Euler, 2 upwind midpoint.
    tke_target or injection_rate > 0: each step applies the linear forcing after the
    advection (SpectralForcing), differentiated through its energy statistics too."""
    out = dts_out if dts_out is not None else []
    if tke_target < 0 or injection_rate < 0:
        raise ValueError("tke_target and injection_rate must be >= 0")
    return Rollout.apply(u0, grid, int(nsteps), float(nu), tuple(float(x) for x in uinf),
                         None if dts is None else list(dts), int(every), float(CFL), out, int(scheme),
                         (float(tke_target), float(injection_rate)))


class PenalizedRollout(torch.autograd.Function):
    """Time-averaged drag of a rollout with the penalized sphere (forced at the body velocity
    V and, with angvel, rotating at W about its centre of mass), J = sum_n (dt_n / T) F_x^n
    with F^n the step's penalization force; differentiable with respect to u0, V and W, and,
    when a centre is given (the sphere is then rebuilt there with the mollified chi) and W is
    zero, with respect to the sphere centre. Without a centre the grid's chi from
    Grid.chi_from_sphere is used as it stands. The steps advect with any of the three schemes
    and may apply the linear forcing after the advection. Fixed timesteps."""

    @staticmethod
    def forward(ctx, u0, vel, centre, angvel, grid: Grid, nsteps: int, nu: float, uinf, dts, every: int,
                lambda_: float, implicit: bool, radius: float, mollify: float, scheme: int, forcing: tuple):
        """forcing: (tke_target, injection_rate), both 0 for none."""
        sphere = dict(lambda_=lambda_, implicit=implicit, trans_vel=tuple(float(x) for x in vel.tolist()))
        fkw = dict(spectral_forcing=any(f > 0 for f in forcing), tke_target=forcing[0], injection_rate=forcing[1],
                   scheme=scheme)
        if angvel is not None:
            sphere["ang_vel"] = tuple(float(x) for x in angvel.tolist())
        body = None
        if centre is not None:
            body = (tuple(float(x) for x in centre.tolist()), float(radius), float(mollify))
            grid.chi_from_sphere(body[0], body[1], mollify=body[2])
        grid.set_torch_velocity(u0.detach())
        T = float(sum(dts))
        ckpts = {0: u0.detach().clone()}
        J = 0.0
        for n in range(nsteps):
            grid.run_steps(1, nu=nu, uinf=uinf, fixed_dt=dts[n], sphere=sphere, **fkw)
            J += dts[n] / T * grid.last_body_force()[0]
            if (n + 1) % every == 0 and n + 1 < nsteps:
                ckpts[n + 1] = grid.get_torch_velocity()
        ctx.grid, ctx.nsteps, ctx.nu, ctx.uinf, ctx.dts, ctx.every = grid, nsteps, nu, uinf, list(dts), every
        ctx.sphere, ctx.T, ctx.ckpts, ctx.u0_dtype, ctx.vdev, ctx.body = sphere, T, ckpts, u0.dtype, vel.device, body
        ctx.cdev = centre.device if centre is not None else None
        ctx.wdev = angvel.device if angvel is not None else None
        ctx.fkw, ctx.forcing, ctx.scheme = fkw, forcing, scheme
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, gJ):
        grid, every, sphere = ctx.grid, ctx.every, ctx.sphere
        want_c = ctx.body is not None and ctx.needs_input_grad[2]
        want_w = ctx.wdev is not None and ctx.needs_input_grad[3]
        if ctx.body is not None:
            grid.chi_from_sphere(ctx.body[0], ctx.body[1], mollify=ctx.body[2])
        g = float(gJ)
        lam = torch.zeros((3,) + grid.shape, dtype=grid._torch_dtype(), device="cuda")
        dV = torch.zeros(3, dtype=torch.float64)
        dc = torch.zeros(3, dtype=torch.float64)
        dW = torch.zeros(3, dtype=torch.float64)
        for s0 in sorted(ctx.ckpts, reverse=True):
            s1 = min(s0 + every, ctx.nsteps)
            states = [ctx.ckpts[s0]]
            grid.set_torch_velocity(ctx.ckpts[s0])
            for n in range(s0, s1 - 1):
                grid.run_steps(1, nu=ctx.nu, uinf=ctx.uinf, fixed_dt=ctx.dts[n], sphere=sphere, **ctx.fkw)
                states.append(grid.get_torch_velocity())
            for n in reversed(range(s0, s1)):
                grid.set_torch_velocity(states[n - s0])
                grid.set_torch_adjoint(lam)
                out = grid.adjoint_step_body(ctx.dts[n], ctx.nu, ctx.uinf, sphere, g * ctx.dts[n] / ctx.T,
                                             centre=want_c, rotation=True, scheme=ctx.scheme,
                                             tke_target=ctx.forcing[0], injection_rate=ctx.forcing[1])
                dV += torch.from_numpy(out[0])
                dW += torch.from_numpy(out[1])
                if want_c:
                    dc += torch.from_numpy(out[2])
                lam = grid.get_torch_adjoint()
        ctx.ckpts = None
        return (lam.to(ctx.u0_dtype), dV.to(ctx.vdev), dc.to(ctx.cdev) if want_c else None,
                dW.to(ctx.wdev) if want_w else None, None, None, None, None, None, None, None, None, None, None,
                None, None)


def mean_drag(u0, vel, grid: Grid, nsteps: int, nu: float, uinf, dts, every=10, lambda_=1e6, implicit=True,
              centre=None, radius=None, mollify=1.0, angvel=None, scheme=0, tke_target=0.0, injection_rate=0.0):
    """Time-averaged drag (penalization force, x) of a rollout with the sphere; differentiable
    with respect to u0 ([3][Z][Y][X], cuda), vel (the body velocity, 3 values) and angvel (its
    rotation about the centre of mass, 3 values; None: no rotation). With centre (3 values)
    and radius the sphere is rebuilt at centre with the mollified chi of half-width mollify
    cells, and J is differentiable with respect to the centre as well (a body without
    rotation only). scheme: 0 central Euler, 1 upwind Euler, 2 upwind midpoint; tke_target or
    injection_rate > 0: each step applies the linear forcing after the advection."""
    if centre is not None and radius is None:
        raise ValueError("mean_drag: a centre needs the sphere's radius")
    if tke_target < 0 or injection_rate < 0:
        raise ValueError("tke_target and injection_rate must be >= 0")
    if (centre is not None and centre.requires_grad and angvel is not None
            and bool(torch.any(torch.as_tensor(angvel) != 0))):
        raise ValueError("mean_drag: the centre gradient is formed for a body without rotation only")
    return PenalizedRollout.apply(u0, vel, centre, angvel, grid, int(nsteps), float(nu), tuple(float(x) for x in uinf),
                                  list(dts), int(every), float(lambda_), bool(implicit),
                                  float(radius) if radius is not None else 0.0, float(mollify), int(scheme),
                                  (float(tke_target), float(injection_rate)))


class BodiesRollout(torch.autograd.Function):
    """J = sum_n (dt_n / T) sum_k w_k F_x^{k,n} for the grid's bodies (added with Grid.add_sphere
    or built with Grid.chi_from_sphere), body k forced at vels[k] and rotating at angvels[k]
    about its centre of mass; differentiable with respect to u0, vels and angvels ([nbody][3]
    float64). Fixed timesteps, checkpoints every ``every`` steps."""

    @staticmethod
    def forward(ctx, u0, vels, angvels, grid: Grid, nsteps: int, nu: float, uinf, dts, every: int, weights,
                lambda_: float, implicit: bool, scheme: int, forcing: tuple):
        n = grid.body_count()
        bodies = [dict(trans_vel=tuple(float(x) for x in vels[k].tolist()),
                       ang_vel=tuple(float(x) for x in angvels[k].tolist())) for k in range(n)]
        sphere = dict(lambda_=lambda_, implicit=implicit, **bodies[0])
        fkw = dict(spectral_forcing=any(f > 0 for f in forcing), tke_target=forcing[0], injection_rate=forcing[1],
                   scheme=scheme)
        for k in range(1, n):
            grid.body_velocity(k, bodies[k]["trans_vel"], bodies[k]["ang_vel"])
        grid.set_torch_velocity(u0.detach())
        T = float(sum(dts))
        ckpts = {0: u0.detach().clone()}
        J = 0.0
        for s in range(nsteps):
            grid.run_steps(1, nu=nu, uinf=uinf, fixed_dt=dts[s], sphere=sphere, **fkw)
            J += dts[s] / T * sum(weights[k] * grid.last_body_force_body(k)[0] for k in range(n))
            if (s + 1) % every == 0 and s + 1 < nsteps:
                ckpts[s + 1] = grid.get_torch_velocity()
        ctx.grid, ctx.nsteps, ctx.nu, ctx.uinf, ctx.dts, ctx.every = grid, nsteps, nu, uinf, list(dts), every
        ctx.sphere, ctx.bodies, ctx.weights, ctx.fkw, ctx.forcing, ctx.scheme = \
            sphere, bodies, list(weights), fkw, forcing, scheme
        ctx.T, ctx.ckpts, ctx.u0_dtype, ctx.vdev = T, ckpts, u0.dtype, vels.device
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, gJ):
        grid, every, n = ctx.grid, ctx.every, len(ctx.bodies)
        g = float(gJ)
        lam = torch.zeros((3,) + grid.shape, dtype=grid._torch_dtype(), device="cuda")
        dV = torch.zeros((n, 3), dtype=torch.float64)
        dW = torch.zeros((n, 3), dtype=torch.float64)
        for s0 in sorted(ctx.ckpts, reverse=True):
            s1 = min(s0 + every, ctx.nsteps)
            states = [ctx.ckpts[s0]]
            grid.set_torch_velocity(ctx.ckpts[s0])
            for s in range(s0, s1 - 1):
                grid.run_steps(1, nu=ctx.nu, uinf=ctx.uinf, fixed_dt=ctx.dts[s], sphere=ctx.sphere, **ctx.fkw)
                states.append(grid.get_torch_velocity())
            for s in reversed(range(s0, s1)):
                grid.set_torch_velocity(states[s - s0])
                grid.set_torch_adjoint(lam)
                w = [g * ctx.dts[s] / ctx.T * wk for wk in ctx.weights]
                dv, dw = grid.adjoint_step_bodies(ctx.dts[s], ctx.nu, ctx.uinf, ctx.bodies, ctx.sphere["lambda_"],
                                                  ctx.sphere["implicit"], w, scheme=ctx.scheme,
                                                  tke_target=ctx.forcing[0], injection_rate=ctx.forcing[1])
                dV += torch.from_numpy(dv)
                dW += torch.from_numpy(dw)
                lam = grid.get_torch_adjoint()
        ctx.ckpts = None
        return (lam.to(ctx.u0_dtype), dV.to(ctx.vdev), dW.to(ctx.vdev), None, None, None, None, None, None, None,
                None, None, None, None)


def bodies_drag(u0, vels, angvels, grid: Grid, nsteps: int, nu: float, uinf, dts, weights=None, every=10,
                lambda_=1e6, implicit=True, scheme=0, tke_target=0.0, injection_rate=0.0):
    """Time-averaged weighted drag sum_k w_k F_x^k of a rollout with every body of the grid
    (weights default to 1); differentiable with respect to u0 ([3][Z][Y][X], cuda), vels and
    angvels ([nbody][3], float64: each body's velocity and rotation about its centre of mass)."""
    n = grid.body_count()
    if n < 1:
        raise ValueError("bodies_drag: the grid has no body")
    if tuple(vels.shape) != (n, 3) or tuple(angvels.shape) != (n, 3):
        raise ValueError(f"bodies_drag: vels and angvels must be [{n}][3]")
    if tke_target < 0 or injection_rate < 0:
        raise ValueError("tke_target and injection_rate must be >= 0")
    w = [1.0] * n if weights is None else [float(x) for x in weights]
    if len(w) != n:
        raise ValueError(f"bodies_drag: {n} weights needed")
    return BodiesRollout.apply(u0, vels, angvels, grid, int(nsteps), float(nu), tuple(float(x) for x in uinf),
                               list(dts), int(every), w, float(lambda_), bool(implicit), int(scheme),
                               (float(tke_target), float(injection_rate)))


def kinetic_energy(u, h):
    """KE = h^3/2 sum |u|^2 (ComputeDissipation.cpp:109)."""
    return 0.5 * h ** 3 * (u * u).sum()


def recompute_memory_bytes(grid: Grid, nsteps: int, every: int) -> int:
    """Peak device bytes the rollout keeps for its backward: the checkpoints plus one
    recomputed segment of states, three components each."""
    ncell = grid.cells[0] * grid.cells[1] * grid.cells[2]
    nck = (nsteps + every - 1) // every
    return (nck + every) * 3 * ncell * grid.precision
