/*
 * cup.h -- the C-ABI of cubismup_3d_amd: the per-timestep periodic fluid update of
 * CubismUP_3D on a device-resident grid of 8^3-cell blocks, for AMD Instinct MI355X
 * (gfx950). Plain pointers and sizes only; every entry point returns 0 on success or a
 * negative CUP_E* code, and cup_last_error(ctx) holds the message (no exception crosses
 * this boundary).
 *
 * Each entry point names the reference code it replaces (paths relative to the
 * CubismUP_3D source tree). Work is stream-ordered on the context's HIP stream; scalar
 * results (maxima, integrals) are returned synchronously as doubles.
 *
 * Host field format: one double per cell in block-major order
 * [block][iz][iy][ix], block = bx + bpdx*(by + bpdy*bz) -- the reference's
 * FluidBlock::data[z][y][x] (source/Definitions.h:115) walked over the grid's blocks.
 */
#ifndef CUP_H
#define CUP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUP_OK 0
#define CUP_E_INVALID -1   /* bad argument (the reference throws std::invalid_argument / aborts) */
#define CUP_E_DEVICE -2    /* HIP runtime error */
#define CUP_E_STATE -3     /* call out of order (e.g. no obstacle created) */
#define CUP_E_NOMEM -4
#define CUP_E_IO -5        /* a file could not be read or written (saves, restarts, analysis files) */

/* Field ids, same order as the reference's FluidElement (Definitions.h:39), then the
 * adjoint velocity of cup_adjoint_step (no reference counterpart; allocated on first use). */
enum cup_field {
  CUP_CHI = 0, CUP_U = 1, CUP_V = 2, CUP_W = 3, CUP_P = 4, CUP_TMPU = 5, CUP_TMPV = 6, CUP_TMPW = 7,
  CUP_ADJ_U = 8, CUP_ADJ_V = 9, CUP_ADJ_W = 10
};

/* Decomposition of one step. PLAIN follows the reference's operator decomposition
 * kernel by kernel (AdvectionDiffusion.cpp:402-428 kernel + copy, PressureRHS.cpp:273-346
 * memset + kernel, PoissonSolverPeriodic.cpp:129-146 r2c / solve / c2r as separate passes,
 * PoissonSolver.cpp:116-130 copy, PressureProjection.cpp:107-129 gradient). FUSED keeps
 * the arithmetic but fuses passes (RHS into the first FFT pass, the spectral division
 * into the z pass, reductions into the gradient epilogue, dt kept on the device). */
enum cup_decomposition { CUP_PLAIN = 0, CUP_FUSED = 1 };

/* Advection scheme. CENTRAL_EULER is the reference's (AdvectionDiffusion.cpp:106-145).
 * UPWIND3_* (no reference counterpart) advect with the third-order upwind-biased stencil
 * (6h f' = 2f(+1) + 3f(0) - 6f(-1) + f(-2) for positive advecting velocity, mirrored
 * otherwise) and diffuse with the 7-point Laplacian, stepped by Euler or by the explicit
 * midpoint rule u* = u + dt L(u + dt/2 L(u)). */
enum cup_scheme { CUP_CENTRAL_EULER = 0, CUP_UPWIND3_EULER = 1, CUP_UPWIND3_RK2 = 2 };

/* Boundary condition of an axis (SimulationData's BCx/y/z_flag, SimulationData.cpp:123-166):
 * periodic; open, the reference's dirichlet (its labs treat dirichlet, absorbing and
 * freespace alike: the boundary cell is copied outward, BlockLabBC::applyBCfaceOpen,
 * Definitions.h:278-301); wall (no slip: velocities mirrored with their sign flipped, p
 * copied, chi zero, applyBCfaceWall :302-334); freespace, open faces for every stencil whose
 * pressure solve is PoissonSolverUnbounded (all three axes or none, SimulationData.cpp:136-143). */
enum cup_bc { CUP_BC_PERIODIC = 0, CUP_BC_OPEN = 1, CUP_BC_WALL = 2, CUP_BC_FREESPACE = 3 };

typedef struct cup_grid_desc {
  int bpd[3];          /* blocks per dimension (SimulationData::bpdx/y/z, SimulationData.h:88) */
  double maxextent;    /* longest side; h = maxextent / max cells (SimulationData.cpp:184-195) */
  int precision;       /* 4: FP32 bulk fields, 8: FP64 */
  int decomposition;   /* cup_decomposition */
  int device;          /* HIP device ordinal */
  int bc[3];           /* cup_bc per axis; non-periodic axes need CUP_PLAIN */
} cup_grid_desc;

typedef struct cup_ctx cup_ctx;

/* Grid and block store. Replaces the Cubism FluidGridMPI allocation in
 * Simulation::setupGrid (Simulation.cpp:221-243) and the PoissonSolverPeriodic buffers and
 * plans (PoissonSolverPeriodic.cpp:90-127). Cells per axis (8 x blocks) from 8 to 1024 with no
 * prime factor above 5; other sizes are refused with CUP_E_INVALID (FFTW plans any size).
 * Non-periodic axes (desc->bc) take the PLAIN decomposition, the central scheme and
 * cup_poisson_mixed (cup_poisson_unbounded with freespace faces); the spectral operations (forcing, HIT IC and analysis), the upwind schemes
 * and the adjoint need a fully periodic box (CUP_E_INVALID otherwise). */
int cup_create(const cup_grid_desc* desc, cup_ctx** out);
int cup_destroy(cup_ctx* ctx);
/* The context's last error message; with ctx == NULL, the message of the calling thread's
 * last failed cup_create. */
const char* cup_last_error(const cup_ctx* ctx);
/* cells[3], h, extent[3] */
int cup_get_geometry(const cup_ctx* ctx, int cells[3], double* h, double extent[3]);
int cup_synchronize(cup_ctx* ctx);

/* Field upload/download in host block-major doubles (converted to the context precision),
 * replacing direct FluidBlock access through BlockInfo::ptrBlock (Definitions.h:103-167). */
int cup_set_field(cup_ctx* ctx, int field, const double* host);
int cup_get_field(cup_ctx* ctx, int field, double* host);
/* Raw device pointer of a field (context precision), for zero-copy callers. */
int cup_field_device_ptr(cup_ctx* ctx, int field, void** dptr);
/* Device-to-device copies of one field in the context's precision and block-major order,
 * stream-ordered on the context's stream (the caller orders its own stream before, and
 * calls cup_synchronize before reading dst elsewhere). */
int cup_set_field_device(cup_ctx* ctx, int field, const void* src);
int cup_get_field_device(cup_ctx* ctx, int field, void* dst);

/* TestBoundaries.cpp:19-75 restated on the device: tmpU = lab(-dx,-dy,-dz).u through the
 * LDS-staged periodic halo gather with halo width g (1 or 2). */
int cup_lab_shift(cup_ctx* ctx, int g, int dx, int dy, int dz);

/* findMaxU (operators/ProcessHelpers.h:74-99): max over cells of max_i |u_i + uinf_i|. */
int cup_max_velocity(cup_ctx* ctx, const double uinf[3], double* umax);

/* Simulation::calcMaxTimestep (Simulation.cpp:403-424) with the CFL ramp. */
int cup_calc_max_timestep(cup_ctx* ctx, double CFL, double nu, const double uinf[3], int step,
                          int rampup, double* dt_out, double* umax_out);

/* AdvectionDiffusion::operator() (AdvectionDiffusion.cpp:402-428): KernelAdvectDiffuse
 * (:106-145) into tmpU/V/W and the copy tmp -> u with UpdateAndCorrectInflow's correction on
 * the open faces (:209-296, none on periodic boxes). On open inflow faces the halo velocity is
 * faded by 1 - min(1, uMax_measured dt/h) (uinf_d/max|uinf|)^2 (:20-104), uMax_measured being
 * the last cup_calc_max_timestep's maximum. With CUP_FUSED the copy is skipped: u keeps the
 * old velocity until cup_pressure_project writes u = tmp - dt grad p. */
int cup_advect_diffuse(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme);

/* PressureRHS::operator() (PressureRHS.cpp:273-346): h^2/(2dt) * central divergence of
 * the velocity (tmp in FUSED, u in PLAIN), with the obstacle (1-chi) masking and perimeter
 * correction when a sphere exists. */
int cup_pressure_rhs(cup_ctx* ctx, double dt);

/* PoissonSolverPeriodic::solve (PoissonSolverPeriodic.cpp:129-146) and _fftw2cub
 * (PoissonSolver.cpp:116-130): hand-written r2c FFT, grid-consistent spectral division
 * (:15-61), c2r, result in p. */
int cup_poisson_periodic(cup_ctx* ctx);

/* PoissonSolverMixed::solve (PoissonSolverMixed.cpp:66-103, _solve PoissonSolverMixed.h:27-80)
 * and _fftw2cub: r2r transforms per axis (the real DFT on periodic axes, DCT-II / DCT-III on
 * the others), division by the blended symbol (1 - tol) grid-consistent + tol spectral,
 * tol = 0.01, the first coefficient zeroed; result in p. The solver of non-periodic boxes
 * (PressureProjection.cpp:86-105), and of periodic boxes on the reference's Python path
 * (its SimulationData never sets bUseFourierBC). CUP_PLAIN decomposition only (CUP_E_STATE):
 * it solves the RHS of cup_pressure_rhs. */
int cup_poisson_mixed(cup_ctx* ctx);

/* PoissonSolverUnbounded::solve (PoissonSolverUnbounded.cpp:126-157, the Green's function
 * :195-246): the free-space solution of the RHS of cup_pressure_rhs, Eastwood and Brownrigg's
 * convolution with the uniform mesh's Green's function -h^2/(4 pi r) (r = 0: -0.1924173658 h^2)
 * over h^3, by FFTs of a zero-padded grid of twice the cells per axis (at most 512 cells per
 * axis, CUP_E_INVALID otherwise; the padded buffers and the Green's function's spectrum are
 * made on the first call); result in p. CUP_PLAIN decomposition only (CUP_E_STATE). */
int cup_poisson_unbounded(cup_ctx* ctx);

/* KernelGradP (PressureProjection.cpp:27-51): u -= dt/(2h) (p_E - p_W), etc. */
int cup_pressure_project(cup_ctx* ctx, double dt);

/* A sphere (Sphere.cpp:17-50 SDF, ObstacleLibrary.h:41-69 block fill, Obstacle.h:186-207
 * block selection) turned into chi by CreateObstacles' KernelCharacteristicFunction
 * (ObstaclesCreate.cpp:22-137), with surface points and the grid centre of mass
 * (:230-256). A body that moved is built again; its buffers are kept while it covers the
 * same blocks. com_mass_out = CoM[3], mass (may be NULL). */
int cup_chi_from_sphere(cup_ctx* ctx, const double center[3], double radius, double com_mass_out[4]);
/* The hemisphere (Sphere with bHemi, HemiSphereObstacle::FillBlocks, Sphere.cpp:53-84): the
 * ball's part with x < center[0], SDF min(center[0] - x, R - |x - c|); otherwise as
 * cup_chi_from_sphere. */
int cup_chi_from_hemisphere(cup_ctx* ctx, const double center[3], double radius, double com_mass_out[4]);
/* The same sphere with a mollified chi, no reference counterpart (it replaces Towers' formula
 * of ObstaclesCreate.cpp:51-137 so that chi is C1 in the centre, for cup_adjoint_step_body's
 * centre gradient): chi = H(R - |x - c|) with H(d) = 0 below -e, 1 above e,
 * (1 + d/e + sin(pi d/e)/pi)/2 between, e = width*h, 0 < width <= 2. The SDF, surface points,
 * block selection and centre of mass are as cup_chi_from_sphere forms them. */
int cup_chi_from_sphere_mollified(cup_ctx* ctx, const double center[3], double radius, double width,
                                  double com_mass_out[4]);

/* Several bodies (ObstacleVector, ObstacleVector.h:20-80): the calls above and below act on
 * body 0 (cup_chi_from_sphere and its variants replace every body by one); cup_add_sphere
 * adds a sphere (hemi: the hemisphere) whose obstacle blocks must be apart from the others'
 * (CUP_E_INVALID otherwise) and returns its index; cup_move_body builds body id again at a
 * new centre. PressureRHS masks and corrects for every body; the _body variants penalize,
 * integrate forces and momenta for one. The device-resident loop (cup_run_steps with a body)
 * penalizes every body (cup_body_velocity); the adjoint with a body takes a single body. */
int cup_body_count(cup_ctx* ctx, int* n);
int cup_add_sphere(cup_ctx* ctx, const double center[3], double radius, int hemi, double com_mass_out[4], int* id);
int cup_move_body(cup_ctx* ctx, int id, const double center[3], double com_mass_out[4]);
/* Body id's deformation velocity udef (ObstacleBlock::udef, ObstacleBlock.h:41-45) on its
 * obstacle blocks: cup_body_obstacle_blocks gives their number (*n_out) and, with ids_out of
 * cap >= n entries, their block ids in the order udef uses; udef is n x 3 x 512 FP64,
 * [block][component][cell] with cell = ix + 8 (iy + 8 iz). A body built again (cup_move_body,
 * cup_chi_from_sphere) has udef = 0 until it is set again. cup_body_remove_udef_momenta is
 * CreateObstacles' momentum removal (ObstaclesCreate.cpp:258-440): over cells with chi > 0,
 * V = sum chi dv, F = sum chi udef dv, T = sum chi (x - CoM) x (udef - c_old) dv with c_old the
 * body's previous linear correction, J the inertia tensor; corrections F/V and J^-1 T;
 * udef -= F/V + (J^-1 T) x (x - CoM) on every cell of the obstacle blocks. out: mass,
 * the linear correction (3), the angular correction (3), J0..J5 (xx, yy, zz, xy, xz, yz).
 * Once set, udef enters the penalization's target velocity V + Omega x p + udef
 * (Penalization.cpp:74-80, also in cup_run_steps) and the penalized momenta's u - udef
 * (ObstaclesUpdate.cpp:117-126) and the surface forces' deformation powers f . udef
 * (FluidSolidForces.cpp:110-112,145-147); cup_adjoint_step_body refuses a body that has one
 * (CUP_E_STATE). */
int cup_body_obstacle_blocks(cup_ctx* ctx, int id, int* ids_out, int cap, int* n_out);
int cup_body_set_udef(cup_ctx* ctx, int id, const double* udef, int n);
int cup_body_get_udef(cup_ctx* ctx, int id, double* udef, int n);
int cup_body_remove_udef_momenta(cup_ctx* ctx, int id, double out[13]);
int cup_penalize_body(cup_ctx* ctx, int id, double dt, double lambda, int implicit, const double trans_vel[3],
                      const double ang_vel[3], double ft_out[6]);
int cup_surface_forces_body(cup_ctx* ctx, int id, double nu, const double trans_vel[3], const double ang_vel[3],
                            double out[22]);
int cup_obstacle_momenta_body(cup_ctx* ctx, int id, double dt, double lambda, int implicit, double out[29]);

/* Penalization::operator() (Penalization.cpp:129-156): u += dt*pf*(U_body - u) with
 * pf = chi*lambda/(1+lambda*dt*chi) (implicit) or chi/dt (explicit), and the force/torque
 * integrals ft_out[6] (KernelFinalizePenalizationForce, :101-124). */
int cup_penalize(cup_ctx* ctx, double dt, double lambda, int implicit, const double trans_vel[3],
                 const double ang_vel[3], double ft_out[6]);

/* ComputeForces (FluidSolidForces.cpp:17-158, 205-223) reduced like Obstacle::computeForces
 * (Obstacle.cpp:266-286): out[22] in ObstacleBlock::sumQoI order (ObstacleBlock.h:81-93). */
int cup_surface_forces(cup_ctx* ctx, double nu, const double trans_vel[3], const double ang_vel[3],
                       double out[22]);

/* UpdateObstacles' momentum integrals (KernelIntegrateFluidMomenta, ObstaclesUpdate.cpp:15-131,
 * summed over the obstacle blocks as KernelFinalizeObstacleVel does, :133-170) for the sphere
 * (udef = 0), positions relative to its grid centre of mass: out[29] = V, sum chi u dv (3),
 * sum chi (p x u) dv (3), J0..J5, then for implicit penalization GfX, Gp (3), Gj (6), Gu (3),
 * Ga (3) with penalFac = dv lambda dt chi / (1 + chi lambda dt), zero otherwise. The input of
 * Obstacle::computeVelocities (Obstacle.cpp:153-264). */
int cup_obstacle_momenta(cup_ctx* ctx, double dt, double lambda, int implicit, double out[29]);

/* ExternalForcing (ExternalForcing.cpp:35-48): the uniform velocity increment du (the
 * reference's 8 uMax_forced nu dt / H^2 along x) added to every cell of the step's velocity:
 * u, and u* when the advection left it apart from u. */
int cup_add_uniform_velocity(cup_ctx* ctx, const double du[3]);

/* KernelDissipation (ComputeDissipation.cpp:17-116): out[20] in the RDX order (:136-155). */
int cup_diagnostics(cup_ctx* ctx, double nu, double out[20]);

/* cup_diagnostics without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). No context: fields[5] are the caller's device arrays chi,
 * u, v, w, p (cup_field order), each nb*512 cells of desc->precision in the block-major layout
 * of desc's grid (bpd, maxextent, device; decomposition ignored); out_dev receives the same 20
 * integrals as cup_diagnostics, as device doubles. Stream-ordered on hip_stream (NULL: the
 * null stream); nothing is copied to the host and nothing is synchronized. The per-workgroup
 * partial sums live in a static device array of CUP_NOALLOC_MAX_PARTS x 20 doubles (5 MiB on
 * each GPU the call runs on, held until the process ends); a grid that needs more partial
 * sums (more than 262144 blocks when bpdx is a multiple of 8, 32768 when bpdx is odd) is
 * refused with CUP_E_INVALID. Every call on a GPU shares that array: calls that overlap on
 * one GPU (two streams, two threads) race. Errors: cup_last_error(NULL). */
#define CUP_NOALLOC_MAX_PARTS 32768
int cup_diagnostics_noalloc(const cup_grid_desc* desc, const void* const fields[5], double nu, void* hip_stream,
                            double* out_dev);

/* Operator::check (operators/Operator.h:23-50): the scan for NaN in u, v, w and p that the
 * reference's debug builds run after PressureRHS, PressureProjection and ComputeDissipation.
 * counts[4] receives the number of NaN cells of u, v, w and p (cup_field's velocity and
 * pressure, whatever the last operation left there); infinities are not counted, as isnan
 * does not count them. Synchronizes the context's stream. */
int cup_check_nan(cup_ctx* ctx, long long counts[4]);

/* The HIT branch of Analysis::operator() (Analysis.cpp:159-196): KernelAnalysis_gradStats
 * (:18-52) and SpectralAnalysis::run, i.e. SpectralManipFFTW::_compute_analysis
 * (SpectralManipFFTW.cpp:110-207) on the velocity's forward FFT. out[9] = tke, eps (viscous
 * dissipation from the central-difference vorticity), uprime, lambda (Taylor microscale),
 * Re_lambda, l_integral, tau_integral, mean_grad, std_grad; spectrum[max(N)/2 - 1] the shell
 * energy spectrum E_k of HITstatistics (SpectralManip.h:51-110; shell i at k = (i+1) 2 pi / L).
 * nbin is the capacity of spectrum. Deterministic. */
int cup_hit_analysis(cup_ctx* ctx, double nu, double out[9], double* spectrum, int nbin);

/* Initial conditions: KernelIC_taylorGreen (InitialConditions.cpp:32-52). */
int cup_ic_taylor_green(cup_ctx* ctx, double uMax);
/* Seeded isotropic turbulence (construction of SpectralManipFFTW::_compute_IC,
 * SpectralManipFFTW.cpp:217-288, with the "art" spectrum, SpectralIcGenerator.cpp:63-81):
 * Philox-4x32-10 phases from seed, exact Hermitian symmetry, DC and Nyquist zeroed,
 * scaled to sqrt(<|u|^2>/3) = u_rms. Generated in FP64 on the device. */
int cup_ic_hit(cup_ctx* ctx, double k0, double u_rms, uint64_t seed);
/* The same construction with E(|k|) from a target table (SpectralIcGenerator::_generateTarget,
 * SpectralIcGenerator.cpp:21-98, read through EnergySpectrum::interpE, SpectralManip.cpp:96-116,
 * in SpectralManipFFTW::_compute_IC): K[i], E[i], i < n, n >= 2, K increasing. u_rms > 0 scales
 * the field to it; u_rms <= 0 keeps the table's energy, as the reference does. */
int cup_ic_hit_spectrum(cup_ctx* ctx, const double* K, const double* E, int n, double u_rms, uint64_t seed);

/* SpectralForcing::operator() (SpectralForcing.cpp:23-82) with
 * SpectralManipFFTW::_compute_largeModesForcing (SpectralManipFFTW.cpp:25-108):
 * u += fac * P_{0<k^2<=4} u, fac = sim_dt*injection/(2 tke_low), injection =
 * (tke_target - tke)/dt or injection_rate. stats_out = tke, eps, tke_filtered, l_integral, fac. */
int cup_spectral_forcing(cup_ctx* ctx, double dt, double sim_dt, double nu, double tke_target,
                         double injection_rate, double stats_out[5]);

/* ------------------------------------------------------------------------------------
 * Whole-step driver with the step scalars resident on the device (no host round trip
 * per step): dt from the CFL rule (with ramp) of the maximum velocity measured in the
 * previous step's gradient epilogue, the reference pipeline order of
 * Simulation::timestep (Simulation.cpp:593-623), step counter and time kept on the
 * device. It is the FUSED decomposition's form of Simulation::run (Simulation.cpp:582-591).
 */
typedef struct cup_step_params {
  double CFL;
  double nu;
  double uinf[3];
  int rampup;
  int scheme;             /* cup_scheme */
  int spectral_forcing;   /* band-limited linear forcing on/off */
  double tke_target;      /* forcing target (<=0: use injection_rate) */
  double injection_rate;
  double fixed_dt;        /* > 0: use this dt instead of the CFL rule */
  int has_body;           /* penalize the sphere created by cup_chi_from_sphere */
  double lambda;
  int implicit_penalization;
  double body_vel[3];
  double body_angvel[3];
} cup_step_params;

int cup_run_steps(cup_ctx* ctx, const cup_step_params* prm, int nsteps);
/* step, time, dt of the last step, umax measured after it, tke after it. */
int cup_step_state(cup_ctx* ctx, int* step, double* time, double* last_dt, double* umax, double* tke);
int cup_set_step_state(cup_ctx* ctx, int step, double time);

/* Adjoint of one step (no reference counterpart): with the forward state u_n in
 * (CUP_U, CUP_V, CUP_W) and dJ/du_{n+1} in (CUP_ADJ_U, _V, _W), replaces the adjoint
 * fields by dJ/du_n for the step u_{n+1} = P(A(u_n)) of the given dt and scheme (central
 * Euler, upwind Euler, upwind midpoint), no forcing, no obstacle (CUP_E_STATE if one
 * exists). P (PressureRHS + PoissonSolverPeriodic + KernelGradP) is symmetric and is
 * applied to the adjoint; A'(u_n)^T is the transposed linearised advection-diffusion
 * stencil, for the midpoint form applied at the recomputed half step and at u_n. Overwrites
 * p and tmp; the device pointers of tmp and the adjoint fields may change
 * (cup_field_device_ptr again after this call). */
int cup_adjoint_step(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme);
/* The same for the step with the linear forcing after the advection (SpectralForcing.cpp:37-77,
 * cup_spectral_forcing: to the energy tke_target when > 0, else at the constant injection
 * rate injection_rate), no reference counterpart. Between the projection's transpose and the
 * advection's it applies the forcing's: lam += fac P lam + <P u*, lam> grad fac at u* = A(u_n),
 * P the forcing's band filter, grad fac = -u* / (2 n tke_F) (target only) - fac P u* / (n tke_F)
 * (n cells, tke_F = <u*, P u*>/2); nothing when fac <= 0, as the forward. Both zero: exactly
 * cup_adjoint_step. Synchronous (two scalar read-backs). */
int cup_adjoint_step_forcing(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme,
                             double tke_target, double injection_rate);
/* The same with the sphere (forced at body_vel; chi from cup_chi_from_sphere),
 * no reference counterpart. The step adds the implicit (or explicit) penalization of
 * Penalization.cpp:40-98 after the projection, whose RHS is masked and corrected on the
 * body as in PressureRHS.cpp:78-112,322-343. The objective may hold wdrag times this step's
 * drag F_x = sum h^3 pf (V_x - u_x) (the penalization force, Penalization.cpp:93-96). The
 * masking's sign pattern is held at the forward state. Returns this step's dJ/dV in dJdV.
 * With dJdc non-NULL (the chi from cup_chi_from_sphere_mollified, else CUP_E_STATE) it also
 * returns this step's dJ/d(sphere centre): chi enters the penalization, its force and the
 * RHS masking, whose surface flags are held; this re-solves the forward pressure once. */
int cup_adjoint_step_body(cup_ctx* ctx, double dt, double nu, const double uinf[3], double lambda, int implicit,
                          const double body_vel[3], double wdrag, double dJdV[3], double dJdc[3]);
/* cup_adjoint_step_body for the step with the scheme's advection (CUP_CENTRAL_EULER,
 * CUP_UPWIND3_EULER, CUP_UPWIND3_RK2), the linear forcing after it (tke_target or
 * injection_rate > 0, as cup_adjoint_step_forcing; the masking is then linearised at the forced
 * u*) and a sphere that also rotates at body_angvel about its centre of mass (the
 * penalization's target V + Omega x (x - CoM), Penalization.cpp:74-80); no reference
 * counterpart. dJdW (may be NULL) returns this step's dJ/dOmega = sum over penalized cells of
 * (x - CoM) x g, g_d = dt pf lam_d + wdrag h^3 pf [d = x] the cell's dJ/d(target velocity).
 * dJdc must be NULL when body_angvel is nonzero (CUP_E_INVALID: the lever arms move with the
 * centre, which this gradient does not differentiate). cup_adjoint_step_body is this call with
 * the central scheme, no forcing, body_angvel = 0 and dJdW = NULL. */
int cup_adjoint_step_body_full(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme,
                               double tke_target, double injection_rate, double lambda, int implicit,
                               const double body_vel[3], const double body_angvel[3], double wdrag, double dJdV[3],
                               double dJdW[3], double dJdc[3]);
/* cup_adjoint_step_body_full with several bodies (nbody == cup_body_count, CUP_E_INVALID
 * otherwise; their obstacle blocks are apart, so each body's transposes act on its own blocks),
 * no reference counterpart. Body k is forced at body_vel[3k..3k+2] and rotates at
 * body_angvel[3k..3k+2] (the velocities cup_run_steps was given: body 0's in the step
 * parameters, the others' through cup_body_velocity); J holds wdrag[k] times body k's drag.
 * dJdV[3k..3k+2] and dJdW[3k..3k+2] (may be NULL) return body k's gradients of this step. No
 * centre gradient. Synchronous (one scalar read-back per body). */
int cup_adjoint_step_bodies(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme, double tke_target,
                            double injection_rate, double lambda, int implicit, int nbody, const double* body_vel,
                            const double* body_angvel, const double* wdrag, double* dJdV, double* dJdW);
/* Penalization force and torque (6 values) of the last step cup_run_steps made with a body. */
int cup_last_body_force(cup_ctx* ctx, double out[6]);
/* Several bodies in cup_run_steps (has_body): body 0 is penalized with the params' body_vel and
 * body_angvel, every other body with the velocity and rotation last given here (default 0),
 * each after the projection (Penalization::operator(), Penalization.cpp:129-156, for each
 * obstacle). cup_last_body_force_body gives body id's force and torque of the loop's last
 * step (id 0: cup_last_body_force). */
int cup_body_velocity(cup_ctx* ctx, int id, const double vel[3], const double angvel[3]);
int cup_last_body_force_body(cup_ctx* ctx, int id, double out[6]);

/* Measurement: the stream the context launches on, and HIP events recorded on it around
 * one stage of the FUSED step (0 advection-diffusion, 1 x-forward FFT with the RHS,
 * 2 forcing reduction, 3 y forward, 4 z pass with the division, 5 y inverse, 6 x inverse,
 * 7 gradient) for up to max_steps steps of the following cup_run_steps calls, leaving out the
 * first step of each call (it follows an idle stream, so its events would hold the host's
 * launch latency, and it runs without the previous step's folded bookkeeping); _end
 * synchronizes and returns the average duration (ms) and the number of launches timed.
 * No counterpart in the reference (its cubism::Profiler regions time host phases). */
int cup_stream(cup_ctx* ctx, void** hip_stream);
int cup_stage_timing_begin(cup_ctx* ctx, int stage, int max_steps);
int cup_stage_timing_end(cup_ctx* ctx, double* avg_ms, int* count);

/* ------------------------------------------------------------------------------------
 * Simulation-level entry points: the reference's Simulation / SimulationData pair
 * (source/Simulation.h:24-71, source/SimulationData.h:46-179) built by the host C++
 * layer of this library on top of the calls above, for bindings (the reference's pybind11
 * module, source/bindings/Simulation.cpp:59-94, binds these same objects).
 */
typedef struct cup_sim cup_sim;

/* SimulationData(comm, parser) (SimulationData.cpp:32-172): "-bpdx 8 -nu 0.01 ..."; adds
 * -precision float|double, -decomposition plain|fused, -device, -icSeed, -icUrms, and a
 * sphere through -sphereL -xpos -ypos -zpos -xvel -yvel -zvel -bForcedInSimFrame
 * -bFixFrameOfRef -accel -T (Sphere.cpp:87-99). */
int cup_sim_create(int argc, const char* const* argv, cup_sim** out);
/* SimulationData::setCells (SimulationData.cpp:280-291); before cup_sim_init. */
int cup_sim_set_cells(cup_sim* sim, int nx, int ny, int nz);
/* Simulation(const SimulationData&) (Simulation.cpp:53-64): grid, pipeline, IC. */
int cup_sim_init(cup_sim* sim);
int cup_sim_destroy(cup_sim* sim);
const char* cup_sim_last_error(const cup_sim* sim);
/* Simulation_addObstacle with a Sphere (bindings/Obstacles.cpp:12-128): forced_mask and
 * fix_mask hold bForcedInSimFrame and bFixFrameOfRef per axis (bit d for axis d);
 * enforced_vel is the velocity of the forced axes; accel_decel, umax, tmax select the
 * ramped x velocity of Sphere::computeVelocities (Sphere.cpp:147-159); hemi the hemisphere. */
int cup_sim_add_sphere(cup_sim* sim, double radius, const double position[3], const double enforced_vel[3],
                       int forced_mask, int fix_mask, int accel_decel, double umax, double tmax, int hemi);
int cup_sim_run(cup_sim* sim);                                   /* Simulation::run (Simulation.cpp:582-591) */
int cup_sim_calc_max_timestep(cup_sim* sim, double* dt);         /* Simulation.cpp:403-424 */
int cup_sim_timestep(cup_sim* sim, double dt, int* finished);    /* Simulation.cpp:593-623 */
int cup_sim_advance(cup_sim* sim, int nsteps, int* finished);    /* FUSED device-resident loop */
int cup_sim_context(cup_sim* sim, cup_ctx** ctx);
int cup_sim_get(cup_sim* sim, const char* name, double* value);  /* SimulationData members by name */
int cup_sim_set(cup_sim* sim, const char* name, double value);
int cup_sim_obstacle(cup_sim* sim, int idx, double out[34]);     /* CoM, mass, force, torque, 22 QoIs */
int cup_sim_diagnostics(cup_sim* sim, double out[20]);
/* An obstacle's kinematic state (Obstacle.h:99-104): position[3], quaternion[4], transVel[3], angVel[3]. */
int cup_sim_obstacle_state(cup_sim* sim, int idx, double out[13]);
int cup_sim_flush_logs(cup_sim* sim);

#ifdef __cplusplus
}
#endif
#endif /* CUP_H */
