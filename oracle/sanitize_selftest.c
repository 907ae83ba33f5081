/*
 * sanitize_selftest.c -- drives every entry point of the oracle on small grids, for the
 * AddressSanitizer / UndefinedBehaviorSanitizer build of `make -C oracle sanitize` (the
 * reference's debug configurations segf and nans, makefiles/make.gcc.flags:14-20). A
 * sanitizer report aborts the program with a non-zero status; the few value checks here only
 * make sure the calls did their work (the oracle's numbers are pinned by the tests).
 *
 * TEST INFRASTRUCTURE ONLY, like the oracle it drives.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cup_oracle.h"

static int failures = 0;

#define EXPECT(cond)                                                       \
  do {                                                                     \
    if (!(cond)) {                                                         \
      fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

static unsigned long long rng_state = 0x9e3779b97f4a7c15ull;

static double urand(void)
{
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

static size_t cells(int bx, int by, int bz) { return (size_t)bx * by * bz * 512; }

static void random_fields(orc_sim* s, size_t n, double amp)
{
  double* f = malloc(n * sizeof(double));
  for (int k = ORC_U; k <= ORC_W; ++k) {
    for (size_t i = 0; i < n; ++i) f[i] = amp * urand();
    orc_set_field(s, k, f);
  }
  free(f);
}

static int finite_field(const orc_sim* s, int k, size_t n)
{
  double* f = malloc(n * sizeof(double));
  orc_get_field(s, k, f);
  int ok = 1;
  for (size_t i = 0; i < n; ++i) ok = ok && isfinite(f[i]);
  free(f);
  return ok;
}

/* the periodic pipeline: ICs, labs, dt, each operator, full steps with forcing and every scheme */
static void periodic(int bx, int by, int bz)
{
  const size_t n = cells(bx, by, bz);
  orc_sim* s = orc_create(bx, by, bz, 2 * M_PI);
  double ext[3];
  orc_extent(s, ext);
  EXPECT(orc_h(s) > 0 && ext[0] > 0);
  const double uinf[3] = {0.1, -0.05, 0.02};
  orc_ic_taylor_green(s, 1.0);
  orc_ic_hit(s, 4.0, 1.0, 7);
  const double K[4] = {1, 2, 4, 8}, E[4] = {1, 0.5, 0.1, 0.01};
  EXPECT(orc_spectrum_interp(K, E, 4, 3.0) > 0);
  orc_ic_hit_spectrum(s, K, E, 4, 0.8, 11);
  for (int g = 1; g <= 3; ++g) orc_lab_shift(s, g, g, -1, 1);
  double umax = 0;
  const double dt = orc_calc_max_timestep(s, 0.3, 0.01, uinf, 0, 10, &umax);
  EXPECT(dt > 0 && umax > 0 && orc_find_max_u(s, uinf) > 0);
  for (int scheme = 0; scheme < 3; ++scheme) orc_advect_diffuse_scheme(s, dt, 0.01, uinf, scheme);
  orc_advect_diffuse(s, dt, 0.01, uinf);
  orc_pressure_rhs(s, dt);
  double* rhs = malloc(n * sizeof(double));
  orc_get_rhs(s, rhs);
  orc_set_rhs(s, rhs);
  orc_poisson_solve(s);
  orc_set_rhs(s, rhs);
  orc_poisson_mixed(s);
  orc_grad_p(s, dt);
  double d[20], a[9], st[5];
  orc_dissipation(s, 0.01, d);
  const int maxN = 8 * (bx > by ? (bx > bz ? bx : bz) : (by > bz ? by : bz));
  double* Ek = calloc((size_t)(maxN / 2 - 1), sizeof(double));
  orc_hit_analysis(s, 0.01, a, Ek);
  EXPECT(a[0] > 0);
  orc_spectral_forcing(s, dt, dt, 0.01, 0.5, 0, st);
  orc_spectral_forcing(s, dt, dt, 0.01, 0, 0.1, st);
  orc_params p;
  memset(&p, 0, sizeof(p));
  p.nu = 0.01;
  p.spectral_forcing = 1;
  p.tke_target = 0.5;
  double fs[5], ft[6], q[22];
  for (int scheme = 0; scheme < 3; ++scheme) {
    p.scheme = scheme;
    orc_timestep(s, dt, &p, fs, ft, q);
  }
  EXPECT(finite_field(s, ORC_U, n) && finite_field(s, ORC_P, n));
  /* adjoints of the plain step, of each scheme and of the forcing */
  double* lam = malloc(3 * n * sizeof(double));
  for (size_t i = 0; i < 3 * n; ++i) lam[i] = urand();
  orc_adjoint_step(s, dt, 0.01, uinf, lam);
  for (int scheme = 0; scheme < 3; ++scheme) orc_adjoint_step_scheme(s, dt, 0.01, uinf, scheme, lam);
  orc_adjoint_step_forcing(s, dt, 0.01, uinf, 0, 0.5, 0, lam);
  orc_adjoint_step_forcing(s, dt, 0.01, uinf, 2, 0, 0.1, lam);
  int ok = 1;
  for (size_t i = 0; i < 3 * n; ++i) ok = ok && isfinite(lam[i]);
  EXPECT(ok);
  free(lam);
  free(Ek);
  free(rhs);
  orc_destroy(s);
}

/* bodies: spheres, hemispheres, several bodies, udef, penalization, forces, the body adjoint */
static void bodies(void)
{
  const int bx = 4, by = 2, bz = 2;
  const size_t n = cells(bx, by, bz);
  orc_sim* s = orc_create(bx, by, bz, 1.0);
  random_fields(s, n, 0.2);
  const double c0[3] = {0.3, 0.25, 0.26}, c1[3] = {0.7, 0.24, 0.25};
  orc_sphere_create(s, 0.1, c0);
  double info[8];
  orc_sphere_info(s, info);
  EXPECT(info[3] > 0);
  EXPECT(orc_add_sphere(s, 0.08, c1, 1) == 1 && orc_body_count(s) == 2);
  const double c1b[3] = {0.72, 0.25, 0.25};
  EXPECT(orc_move_body(s, 1, c1b) == 0);
  const double V[3] = {-0.2, 0, 0}, W[3] = {0, 0, 0.1}, uinf[3] = {0.2, 0, 0};
  double ft[6], M[29], q[22], rm[13];
  for (int k = 0; k < 2; ++k) {
    orc_select_body(s, k);
    orc_penalize(s, 1e-3, 1e5, k, V, W, ft);
    orc_obstacle_momenta(s, 1e-3, 1e5, 1, M);
    orc_surface_forces(s, 0.01, V, W, q);
  }
  orc_select_body(s, 0);
  double* udef = malloc(3 * n * sizeof(double));
  for (size_t i = 0; i < 3 * n; ++i) udef[i] = 0.01 * urand();
  orc_set_udef(s, udef);
  orc_remove_udef_momenta(s, rm);
  orc_get_udef(s, udef);
  free(udef);
  orc_hemisphere_create(s, 0.1, c0);
  orc_sphere_create_mollified(s, 0.1, c0, 1.5);
  orc_pressure_rhs(s, 1e-3);
  double* lam = malloc(3 * n * sizeof(double));
  for (size_t i = 0; i < 3 * n; ++i) lam[i] = urand();
  double dJdV[3] = {0, 0, 0}, dJdc[3] = {0, 0, 0};
  orc_adjoint_step_body(s, 1e-3, 0.01, uinf, 1e5, 1, V, 1.0, lam, dJdV, dJdc);
  EXPECT(isfinite(dJdV[0]) && isfinite(dJdc[0]));
  const double Wr[3] = {0.1, -0.2, 0.3};
  double dJdW[3] = {0, 0, 0};
  orc_adjoint_step_body_full(s, 1e-3, 0.01, uinf, 0, 0.0, 0.0, 1e5, 1, V, Wr, 1.0, lam, dJdV, dJdW, dJdc);
  EXPECT(isfinite(dJdW[0]) && isfinite(dJdW[2]) && isnan(dJdc[0]));
  for (int scheme = 1; scheme < 3; ++scheme)
    orc_adjoint_step_body_full(s, 1e-3, 0.01, uinf, scheme, 0.5, 0.0, 1e5, 1, V, Wr, 1.0, lam, dJdV, dJdW, NULL);
  orc_adjoint_step_body_full(s, 1e-3, 0.01, uinf, 0, 0.0, 0.1, 1e5, 1, V, Wr, 1.0, lam, dJdV, dJdW, NULL);
  orc_params p;
  memset(&p, 0, sizeof(p));
  p.nu = 0.01;
  p.has_sphere = 1;
  p.lambda = 1e5;
  p.implicit_penalization = 1;
  p.trans_vel[0] = -0.2;
  double fs[5], pf[6];
  orc_timestep(s, 1e-3, &p, fs, pf, q);
  EXPECT(finite_field(s, ORC_U, n));
  free(lam);
  orc_destroy(s);
}

/* non-periodic boxes: open and wall labs, the inflow fade and correction, PoissonSolverMixed */
static void boundaries(void)
{
  const int kinds[4][3] = {{ORC_BC_OPEN, ORC_BC_WALL, ORC_BC_PERIODIC},
                           {ORC_BC_WALL, ORC_BC_OPEN, ORC_BC_OPEN},
                           {ORC_BC_OPEN, ORC_BC_OPEN, ORC_BC_OPEN},
                           {ORC_BC_WALL, ORC_BC_WALL, ORC_BC_WALL}};
  for (int c = 0; c < 4; ++c) {
    const int bx = 2, by = 3, bz = 2;
    const size_t n = cells(bx, by, bz);
    orc_sim* s = orc_create(bx, by, bz, 1.0);
    orc_set_bc(s, kinds[c]);
    random_fields(s, n, 0.3);
    const double uinf[3] = {0.2, -0.1, 0.05}, c0[3] = {0.5, 0.37, 0.25};
    double umax = 0;
    const double dt = orc_calc_max_timestep(s, 0.3, 0.01, uinf, 0, 0, &umax);
    orc_set_umax_measured(s, umax);
    orc_advect_diffuse(s, dt, 0.01, uinf);
    orc_pressure_rhs(s, dt);
    orc_poisson_mixed(s);
    orc_grad_p(s, dt);
    double d[20];
    orc_dissipation(s, 0.01, d);
    orc_sphere_create(s, 0.1, c0);
    orc_params p;
    memset(&p, 0, sizeof(p));
    p.nu = 0.01;
    p.uinf[0] = 0.2;
    p.has_sphere = 1;
    p.lambda = 1e5;
    p.implicit_penalization = 1;
    p.trans_vel[0] = -0.2;
    double fs[5], ft[6], q[22];
    orc_timestep(s, dt, &p, fs, ft, q);
    EXPECT(finite_field(s, ORC_U, n) && finite_field(s, ORC_P, n));
    orc_destroy(s);
  }
}

/* the r2r lines and the complex FFT, with their round trips */
static void transforms(void)
{
  for (int n = 8; n <= 40; n += 8) {
    double* a = malloc((size_t)n * sizeof(double));
    double* b = malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; ++i) a[i] = b[i] = urand();
    for (int kind = 0; kind < 4; ++kind) orc_r2r_line(a, n, kind);
    free(a);
    free(b);
  }
  const int nx = 8, ny = 24, nz = 40, m = nx * ny * nz;
  double* re = malloc(m * sizeof(double));
  double* im = malloc(m * sizeof(double));
  double* r0 = malloc(m * sizeof(double));
  for (int i = 0; i < m; ++i) re[i] = r0[i] = urand(), im[i] = 0;
  orc_fft3d(nx, ny, nz, re, im, -1);
  orc_fft3d(nx, ny, nz, re, im, +1);
  double err = 0;
  for (int i = 0; i < m; ++i) err = fmax(err, fabs(re[i] / m - r0[i]) + fabs(im[i] / m));
  EXPECT(err < 1e-12);
  free(re);
  free(im);
  free(r0);
}

int main(void)
{
  orc_set_threads(2);
  periodic(2, 2, 2);
  periodic(3, 2, 5); /* lengths with radix 3 and 5 */
  bodies();
  boundaries();
  transforms();
  for (int plant = 1; plant <= 7; ++plant) { /* the planted mistakes' branches */
    orc_set_plant(plant);
    bodies();
  }
  orc_set_plant(0);
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle sanitizer self-test: ok\n");
  return 0;
}
