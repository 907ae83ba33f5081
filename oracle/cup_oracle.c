/*
 * cup_oracle.c -- the oracle: an FP64 CPU restatement of CubismUP_3D's per-timestep fluid
 * update on a uniform, periodic grid of 8^3-cell blocks, with OpenMP over blocks.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker. It is built into oracle/liboracle.so
 * and is loaded only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg; nothing in the product links or calls it.
 *
 * It keeps the reference CPU path's structure on purpose, because it doubles as the CPU
 * baseline: an AoS cell element of 8 doubles (source/Definitions.h:40-52), blocks stored
 * data[z][y][x] (Definitions.h:115), a halo'd "lab" copy of every block filled by periodic
 * wrap before each stencil kernel (Cubism BlockLab + BlockLabBC::_apply_bc,
 * Definitions.h:349-377, where periodic faces are a no-op), and one OpenMP loop over blocks
 * per operator (source/operators/Operator.h:113-170). The FFT is written here (radix-2,
 * full complex, unnormalized like FFTW's r2c/c2r pair) because FFTW is not available.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   lab fill          -- TestBoundaries shift test restated, exact (tests/cubismup3d/TestBoundaries.cpp:19-75)
 *   advect_diffuse    -- shear-wave closed form (exact one-step factor), advected-mode symbol,
 *                        brute force against an independent numpy restatement on 16^3
 *   pressure_rhs      -- single-mode closed form (2i sin(kh) h^2/(2dt) symbol); with bodies, the
 *                        masking and perimeter correction: plain divergence outside the obstacle
 *                        blocks, (1-chi) rhs off the surface, one common factor on the surface cells
 *                        of sumRHS's sign, the total restored to zero, for both signs and two bodies
 *   fft3d             -- numpy.fft golden vectors (tests/golden/fft_*.npz), round trip
 *   poisson_solve     -- closed-form symbol per mode, the 8 zeroed modes, div_c(u_new)=0
 *   grad_p            -- single-mode closed form, projection idempotence
 *   find_max_u / dt   -- closed form incl. ramp
 *   dissipation       -- a numpy restatement from vector definitions (np.cross, full strain
 *                        tensor) on random fields; single-mode closed forms of helicity (ABC flow),
 *                        viscous power, pressure power and KE; the impulse identities (linear
 *                        impulse = momentum, angular momentum = 2 sum A = angular impulse) on
 *                        compact fields
 *   spectral_forcing  -- Parseval (tke vs physical-space mean), 32-mode count, injected energy
 *   ic_hit            -- Philox-4x32-10 known-answer vectors, solenoidality, u_rms, spectrum shape
 *   sphere chi        -- Towers formula on a planar SDF (closed form), sphere volume vs 4/3 pi R^3
 *   penalize          -- per-cell closed form; force = sum of per-cell terms
 *   surface_forces    -- the divergence theorem against exact ball integrals of Fourier modes:
 *                        pressure force, viscous force, viscous torque and circulation, converging
 *                        from 64^3 to 128^3; thrust - drag, Pout and pLocomotion against the force
 *                        and torque; zero force for uniform pressure
 *   The pins of pressure_rhs's masking, surface_forces and dissipation plant mistakes here
 *   (orc_set_plant) and show that their checks fail for them.
 */
#include "cup_oracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define BS 8
#define SURFDH 1 /* ObstacleBlock.h:16 */

/* Planted mistakes (orc_set_plant): the pin tests switch one on to show that their checks
 * fail for it, and switch it off again. 0, the default, is the restatement as written. */
static int g_plant = 0;
void orc_set_plant(int code) { g_plant = code; }

typedef struct {
  double chi, u, v, w, p, tmpU, tmpV, tmpW;
} FE;

typedef struct {
  FE d[BS][BS][BS];
} Block;

/* Per-block obstacle data (ObstacleBlock.h:34-67), sphere only. */
typedef struct {
  double chi[BS][BS][BS];
  double sdf[BS][BS][BS];
  double udef[3][BS][BS][BS]; /* the deformation velocity (0 until orc_set_udef) */
  double CoM[3], mass;
  int npts, cap;
  int* pix;
  double *dchidx, *dchidy, *dchidz, *delta;
} OBlock;

/* A body kept aside while another is the sim's active one (orc_select_body swaps them). */
typedef struct {
  int has_sphere;
  double radius, spos[3];
  double moll;
  int hemi;
  OBlock** ob;
  double CoM[3], mass;
  double corr[6]; /* transVel_correction, angVel_correction of the last udef removal */
} OBody;

struct orc_sim {
  int bpd[3], N[3], nb;
  OBody* more; /* bodies 1.. (more[k-1] holds body k while body 0 is active) */
  int nmore;
  double h, maxextent, extent[3];
  Block* blk;
  double* rhs; /* solver buffer, global [x][y][z], z fastest (PoissonSolverPeriodic.cpp:112-116) */
  int has_sphere;
  double radius, spos[3];
  double moll; /* half-width in cells of the mollified chi, 0 for Towers' chi */
  int hemi;    /* the hemisphere of HemiSphereObstacle (flat face at x = c_x, body at x < c_x) */
  OBlock** ob;
  double CoM[3], mass;
  double corr[6];
  int bcf[3];        /* per axis: ORC_BC_PERIODIC, ORC_BC_OPEN, ORC_BC_WALL */
  double umax_meas;  /* SimulationData::uMax_measured, set by orc_calc_max_timestep */
  int unbounded;     /* freespace faces on every axis: PoissonSolverUnbounded */
};

static const double EPS = DBL_EPSILON;

/* ------------------------------------------------------------------------ grid */

orc_sim* orc_create(int bpdx, int bpdy, int bpdz, double maxextent)
{
  if (bpdx < 1 || bpdy < 1 || bpdz < 1) return NULL;
  orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
  s->bpd[0] = bpdx; s->bpd[1] = bpdy; s->bpd[2] = bpdz;
  for (int d = 0; d < 3; d++) s->N[d] = s->bpd[d] * BS;
  s->nb = bpdx * bpdy * bpdz;
  int maxN = s->N[0];
  if (s->N[1] > maxN) maxN = s->N[1];
  if (s->N[2] > maxN) maxN = s->N[2];
  s->maxextent = maxextent;
  s->h = maxextent / maxN;
  for (int d = 0; d < 3; d++) s->extent[d] = ((double)s->N[d] / maxN) * maxextent;
  s->blk = (Block*)calloc((size_t)s->nb, sizeof(Block));
  s->rhs = (double*)calloc((size_t)s->N[0] * s->N[1] * s->N[2], sizeof(double));
  return s;
}

/* Boundary conditions per axis (SimulationData's BCx/y/z_flag, SimulationData.cpp:123-166). */
void orc_set_bc(orc_sim* s, const int bc[3])
{
  /* freespace faces are open faces for the labs (Definitions.h:355-373); all three make the
     pressure solve PoissonSolverUnbounded (SimulationData.cpp:136-143) */
  int nfree = 0;
  for (int d = 0; d < 3; d++) {
    nfree += bc[d] == ORC_BC_FREESPACE;
    s->bcf[d] = bc[d] == ORC_BC_FREESPACE ? ORC_BC_OPEN : bc[d];
  }
  s->unbounded = nfree == 3;
}

void orc_set_umax_measured(orc_sim* s, double umax) { s->umax_meas = umax; }

static void free_sphere(orc_sim* s)
{
  if (!s->ob) return;
  for (int b = 0; b < s->nb; b++) {
    OBlock* o = s->ob[b];
    if (!o) continue;
    free(o->pix); free(o->dchidx); free(o->dchidy); free(o->dchidz); free(o->delta);
    free(o);
  }
  free(s->ob);
  s->ob = NULL;
  s->has_sphere = 0;
}

static void body_swap(orc_sim* s, OBody* b)
{
  OBody t = {s->has_sphere, s->radius, {s->spos[0], s->spos[1], s->spos[2]}, s->moll, s->hemi, s->ob,
             {s->CoM[0], s->CoM[1], s->CoM[2]}, s->mass};
  s->has_sphere = b->has_sphere; s->radius = b->radius; s->moll = b->moll; s->hemi = b->hemi; s->ob = b->ob;
  s->mass = b->mass;
  for (int d = 0; d < 3; d++) { s->spos[d] = b->spos[d]; s->CoM[d] = b->CoM[d]; }
  for (int q = 0; q < 6; q++) { t.corr[q] = s->corr[q]; s->corr[q] = b->corr[q]; }
  *b = t;
}

/* Body k becomes the active one; calling again with the same k swaps it back. */
void orc_select_body(orc_sim* s, int k)
{
  if (k > 0 && k <= s->nmore) body_swap(s, &s->more[k - 1]);
}

static void free_other_bodies(orc_sim* s)
{
  for (int k = 1; k <= s->nmore; k++) {
    orc_select_body(s, k);
    free_sphere(s);
    orc_select_body(s, k);
  }
  free(s->more);
  s->more = NULL;
  s->nmore = 0;
}

void orc_destroy(orc_sim* s)
{
  free_other_bodies(s);
  if (!s) return;
  free_sphere(s);
  free(s->blk);
  free(s->rhs);
  free(s);
}

double orc_h(const orc_sim* s) { return s->h; }
void orc_extent(const orc_sim* s, double out[3]) { for (int d = 0; d < 3; d++) out[d] = s->extent[d]; }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

static inline double* fe_field(FE* e, int f) { return &e->chi + f; }

void orc_set_field(orc_sim* s, int f, const double* src)
{
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    FE* e = &s->blk[b].d[0][0][0];
    const double* q = src + (size_t)b * BS * BS * BS;
    for (int i = 0; i < BS * BS * BS; i++) *fe_field(&e[i], f) = q[i];
  }
}

void orc_get_field(const orc_sim* s, int f, double* dst)
{
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    FE* e = &s->blk[b].d[0][0][0];
    double* q = dst + (size_t)b * BS * BS * BS;
    for (int i = 0; i < BS * BS * BS; i++) q[i] = *fe_field(&e[i], f);
  }
}

/* Block coordinates of block id b (block = bx + bpdx*(by + bpdy*bz)). */
static inline void bcoords(const orc_sim* s, int b, int bc[3])
{
  bc[0] = b % s->bpd[0];
  bc[1] = (b / s->bpd[0]) % s->bpd[1];
  bc[2] = b / (s->bpd[0] * s->bpd[1]);
}

/* Cell-centre position x = (index*BS + i + 1/2) h (LinearInterpolation.h:155-166). */
static inline void cell_pos(const orc_sim* s, const int bc[3], int ix, int iy, int iz, double p[3])
{
  p[0] = (bc[0] * BS + ix + 0.5) * s->h;
  p[1] = (bc[1] * BS + iy + 0.5) * s->h;
  p[2] = (bc[2] * BS + iz + 0.5) * s->h;
}

static inline int wrap(int i, int n) { i %= n; return i < 0 ? i + n : i; }

static inline const FE* gcell(const orc_sim* s, int gx, int gy, int gz)
{
  gx = wrap(gx, s->N[0]); gy = wrap(gy, s->N[1]); gz = wrap(gz, s->N[2]);
  const int b = gx / BS + s->bpd[0] * (gy / BS + s->bpd[1] * (gz / BS));
  return &s->blk[b].d[gz % BS][gy % BS][gx % BS];
}

/* Lab: a (BS+2g)^3 halo'd copy of block b, periodic wrap everywhere (tensorial). */
typedef struct {
  FE* c;
  int g, L;
} Lab;

static Lab lab_alloc(int g)
{
  Lab l;
  l.g = g;
  l.L = BS + 2 * g;
  l.c = (FE*)malloc(sizeof(FE) * l.L * l.L * l.L);
  return l;
}

static void lab_load(const orc_sim* s, int b, Lab* l)
{
  int bc[3];
  bcoords(s, b, bc);
  const int g = l->g, L = l->L;
  for (int iz = -g; iz < BS + g; iz++)
    for (int iy = -g; iy < BS + g; iy++)
      for (int ix = -g; ix < BS + g; ix++)
        l->c[((iz + g) * L + (iy + g)) * L + (ix + g)] =
            *gcell(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz);
  /* BlockLabBC::_apply_bc (Definitions.h:349-377) on the domain faces of the non-periodic
   * axes, x then y then z, over the face's own extent [0, BS) in the other two axes:
   * open (the reference's dirichlet == absorbing == freespace, applyBCfaceOpen :278-301)
   * copies the boundary cell into every halo layer; wall (applyBCfaceWall :302-334) mirrors
   * the cells across the face with the velocities (u, v, w and tmp) negated, copies the
   * boundary cell's p and sets chi = 0. */
  for (int d = 0; d < 3; d++) {
    if (s->bcf[d] == ORC_BC_PERIODIC) continue;
    for (int side = 0; side < 2; side++) {
      if (bc[d] != (side ? s->bpd[d] - 1 : 0)) continue;
      for (int q = 1; q <= g; q++) {
        const int dst = side ? BS - 1 + q : -q;       /* halo index along d */
        const int bnd = side ? BS - 1 : 0;            /* the boundary cell */
        const int mir = side ? 2 * BS - 1 - dst : -1 - dst;  /* its mirror image */
        for (int a = 0; a < BS; a++)
          for (int e = 0; e < BS; e++) {
            int id[3], ib[3], im[3];
            const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
            id[d] = dst; ib[d] = bnd; im[d] = mir;
            id[o1] = ib[o1] = im[o1] = a;
            id[o2] = ib[o2] = im[o2] = e;
            FE* D = &l->c[((id[2] + g) * L + (id[1] + g)) * L + (id[0] + g)];
            const FE* B = &l->c[((ib[2] + g) * L + (ib[1] + g)) * L + (ib[0] + g)];
            const FE* M = &l->c[((im[2] + g) * L + (im[1] + g)) * L + (im[0] + g)];
            if (s->bcf[d] == ORC_BC_OPEN) {
              *D = *B;
            } else {
              D->p = B->p;
              D->chi = 0;
              D->u = -M->u; D->tmpU = -M->tmpU;
              D->v = -M->v; D->tmpV = -M->tmpV;
              D->w = -M->w; D->tmpW = -M->tmpW;
            }
          }
      }
    }
  }
}

static inline const FE* LAB(const Lab* l, int ix, int iy, int iz)
{
  return &l->c[((iz + l->g) * l->L + (iy + l->g)) * l->L + (ix + l->g)];
}

void orc_lab_shift(orc_sim* s, int g, int dx, int dy, int dz)
{
  #pragma omp parallel
  {
    Lab l = lab_alloc(g);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++)
            s->blk[b].d[iz][iy][ix].tmpU = LAB(&l, ix - dx, iy - dy, iz - dz)->u;
    }
    free(l.c);
  }
}

/* ------------------------------------------------------------ reductions, dt */

/* findMaxU (ProcessHelpers.h:74-99). */
double orc_find_max_u(const orc_sim* s, const double uinf[3])
{
  double maxU = 0;
  #pragma omp parallel for schedule(static) reduction(max : maxU)
  for (int b = 0; b < s->nb; b++) {
    const FE* e = &s->blk[b].d[0][0][0];
    for (int i = 0; i < BS * BS * BS; i++) {
      const double au = fabs(e[i].u + uinf[0]);
      const double av = fabs(e[i].v + uinf[1]);
      const double aw = fabs(e[i].w + uinf[2]);
      double m = au > av ? au : av;
      m = m > aw ? m : aw;
      maxU = maxU > m ? maxU : m;
    }
  }
  return maxU;
}

/* calcMaxTimestep (Simulation.cpp:403-424), uniform grid so hMin = h. */
double orc_calc_max_timestep(orc_sim* s, double CFL, double nu, const double uinf[3],
                             int step, int rampup, double* umax_out)
{
  const double hMin = s->h;
  const double umax = orc_find_max_u(s, uinf);
  s->umax_meas = umax;  /* sim.uMax_measured (Simulation.cpp:407) */
  if (umax_out) *umax_out = umax;
  const double dtDif = hMin * hMin / nu;
  const double dtAdv = hMin / (umax + 1e-8);
  const double m = dtDif < dtAdv ? dtDif : dtAdv;
  double dt = CFL * m;
  if (step < rampup) {
    const double x = (step + 1.0) / rampup;
    const double rampCFL = exp(log(1e-3) * (1 - x) + log(CFL) * x);
    dt = rampCFL * m;
  }
  return dt;
}

/* ---------------------------------------------------------------- operators */

/* KernelAdvectDiffuse (AdvectionDiffusion.cpp:106-145) into tmpU/V/W, then the copy
 * tmp -> u of UpdateAndCorrectInflow_nonUniform::operate (:315-328); with every face
 * periodic the inflow sums are zero and the correction returns early (:363-365). */
/* KernelAdvectDiffuseBase::applyBC* (AdvectionDiffusion.cpp:20-104): on an open domain face
 * (neither periodic nor wall) the halo layer's velocity, edges included, is scaled by the
 * face's fade 1 - fac (max or min of uinf_d over |uinf|_max)^2, fac = min(1, uMax_measured dt/h),
 * when the fade is below 1 (inflow faces). */
static void lab_fade(const orc_sim* s, int b, Lab* l, double dt, const double uinf[3])
{
  int bc[3];
  bcoords(s, b, bc);
  double fac = s->umax_meas * dt / s->h;
  fac = fac < 1 ? fac : 1;
  double nor = EPS;
  for (int d = 0; d < 3; d++) nor = fabs(uinf[d]) > nor ? fabs(uinf[d]) : nor;
  for (int d = 0; d < 3; d++) {
    if (s->bcf[d] != ORC_BC_OPEN) continue;
    for (int side = 0; side < 2; side++) {
      if (bc[d] != (side ? s->bpd[d] - 1 : 0)) continue;
      const double q = side ? (uinf[d] < 0 ? uinf[d] : 0) : (uinf[d] > 0 ? uinf[d] : 0);
      const double fade = 1 - fac * pow(q / nor, 2);
      if (fade >= 1) continue;
      const int dst = side ? BS : -1;
      for (int a = -1; a <= BS; a++)
        for (int e = -1; e <= BS; e++) {
          int id[3];
          id[d] = dst; id[(d + 1) % 3] = a; id[(d + 2) % 3] = e;
          FE* D = &l->c[((id[2] + 1) * l->L + (id[1] + 1)) * l->L + (id[0] + 1)];
          D->u *= fade; D->v *= fade; D->w *= fade;
        }
    }
  }
}

/* UpdateAndCorrectInflow::operate (AdvectionDiffusion.cpp:209-296) after the copy tmp -> u:
 * the net outflow through the open faces (not wall, not periodic), divided by the cell count
 * of all six faces, is removed from the face-normal velocity of their boundary cells. */
static void inflow_correction(orc_sim* s)
{
  if (s->bcf[0] != ORC_BC_OPEN && s->bcf[1] != ORC_BC_OPEN && s->bcf[2] != ORC_BC_OPEN) return;
  double* part = (double*)calloc((size_t)s->nb, sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    double sum = 0;
    for (int d = 0; d < 3; d++) {
      if (s->bcf[d] != ORC_BC_OPEN) continue;
      for (int side = 0; side < 2; side++) {
        if (bc[d] != (side ? s->bpd[d] - 1 : 0)) continue;
        for (int a = 0; a < BS; a++)
          for (int e = 0; e < BS; e++) {
            int id[3];
            id[d] = side ? BS - 1 : 0; id[(d + 1) % 3] = a; id[(d + 2) % 3] = e;
            const FE* c = &s->blk[b].d[id[2]][id[1]][id[0]];
            const double un = d == 0 ? c->u : (d == 1 ? c->v : c->w);
            sum += side ? un : -un;
          }
      }
    }
    part[b] = sum;
  }
  double sumInflow = 0;
  for (int b = 0; b < s->nb; b++) sumInflow += part[b];
  free(part);
  const double nx = s->N[0], ny = s->N[1], nz = s->N[2];
  const double corr = sumInflow / (2 * (nx * ny + nx * nz + ny * nz));
  if (fabs(corr) < EPS) return;
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    for (int d = 0; d < 3; d++) {
      if (s->bcf[d] != ORC_BC_OPEN) continue;
      for (int side = 0; side < 2; side++) {
        if (bc[d] != (side ? s->bpd[d] - 1 : 0)) continue;
        for (int a = 0; a < BS; a++)
          for (int e = 0; e < BS; e++) {
            int id[3];
            id[d] = side ? BS - 1 : 0; id[(d + 1) % 3] = a; id[(d + 2) % 3] = e;
            FE* c = &s->blk[b].d[id[2]][id[1]][id[0]];
            double* un = d == 0 ? &c->u : (d == 1 ? &c->v : &c->w);
            *un += side ? -corr : corr;
          }
      }
    }
  }
}

void orc_advect_diffuse(orc_sim* s, double dt, double nu, const double uinf[3])
{
  const double h = s->h;
  const double facA = -dt / (2 * h);
  const double facD = (nu / h) * (dt / h);
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      lab_fade(s, b, &l, dt, uinf);
      Block* o = &s->blk[b];
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const FE* L = LAB(&l, ix, iy, iz);
            const FE *LW = LAB(&l, ix - 1, iy, iz), *LE = LAB(&l, ix + 1, iy, iz);
            const FE *LS = LAB(&l, ix, iy - 1, iz), *LN = LAB(&l, ix, iy + 1, iz);
            const FE *LF = LAB(&l, ix, iy, iz - 1), *LB = LAB(&l, ix, iy, iz + 1);
            const double dudx = LE->u - LW->u, dvdx = LE->v - LW->v, dwdx = LE->w - LW->w;
            const double dudy = LN->u - LS->u, dvdy = LN->v - LS->v, dwdy = LN->w - LS->w;
            const double dudz = LB->u - LF->u, dvdz = LB->v - LF->v, dwdz = LB->w - LF->w;
            const double u = L->u + uinf[0], v = L->v + uinf[1], w = L->w + uinf[2];
            const double duD = LN->u + LS->u + LE->u + LW->u + LF->u + LB->u - L->u * 6;
            const double dvD = LN->v + LS->v + LE->v + LW->v + LF->v + LB->v - L->v * 6;
            const double dwD = LN->w + LS->w + LE->w + LW->w + LF->w + LB->w - L->w * 6;
            const double duA = u * dudx + v * dudy + w * dudz;
            const double dvA = u * dvdx + v * dvdy + w * dvdz;
            const double dwA = u * dwdx + v * dwdy + w * dwdz;
            o->d[iz][iy][ix].tmpU = L->u + facA * duA + facD * duD;
            o->d[iz][iy][ix].tmpV = L->v + facA * dvA + facD * dvD;
            o->d[iz][iy][ix].tmpW = L->w + facA * dwA + facD * dwD;
          }
    }
    free(l.c);
  }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    FE* e = &s->blk[b].d[0][0][0];
    for (int i = 0; i < BS * BS * BS; i++) {
      e[i].u = e[i].tmpU;
      e[i].v = e[i].tmpV;
      e[i].w = e[i].tmpW;
    }
  }
  inflow_correction(s);
}

/* Third-order upwind-biased scheme with central diffusion (no reference counterpart: the
 * reference advects with the central scheme above). Per axis d, with the advecting velocity
 * a = (u + U)_d at the cell, the first derivative times 6h is taken on the upwind side:
 *   a > 0:  2 f(+1) + 3 f(0) - 6 f(-1) + f(-2)
 *   a <= 0: -f(+2) + 6 f(+1) - 3 f(0) - 2 f(-1)
 * (symbol i k h - (k h)^4 / 12 sign(a) + O((kh)^5): third order, dissipative). Diffusion is
 * the 7-point Laplacian. L(f) = -(f + U).grad_up3 f + nu lap f. */
static inline double up3(double a, double m2, double m1, double c, double p1, double p2)
{
  return a > 0 ? 2 * p1 + 3 * c - 6 * m1 + m2 : -p2 + 6 * p1 - 3 * c - 2 * m1;
}

static inline double fe_comp(const FE* e, int q, int tmp)
{
  if (tmp) return q == 0 ? e->tmpU : (q == 1 ? e->tmpV : e->tmpW);
  return q == 0 ? e->u : (q == 1 ? e->v : e->w);
}

/* out[q][cell] = u_q + fac * L(src)_q, src = (u,v,w) or (tmpU,tmpV,tmpW); out is
 * block-major, 3 x ncell. */
static void upwind3_update(orc_sim* s, int src_tmp, double fac, double nu, const double uinf[3], double* out)
{
  const double h = s->h;
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  #pragma omp parallel
  {
    Lab l = lab_alloc(2);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const FE* C = LAB(&l, ix, iy, iz);
            double a[3];
            for (int d = 0; d < 3; d++) a[d] = fe_comp(C, d, src_tmp) + uinf[d];
            const size_t cell = (size_t)b * BS * BS * BS + (iz * BS + iy) * BS + ix;
            for (int q = 0; q < 3; q++) {
              double f[3][5];
              for (int k = -2; k <= 2; k++) {
                f[0][k + 2] = fe_comp(LAB(&l, ix + k, iy, iz), q, src_tmp);
                f[1][k + 2] = fe_comp(LAB(&l, ix, iy + k, iz), q, src_tmp);
                f[2][k + 2] = fe_comp(LAB(&l, ix, iy, iz + k), q, src_tmp);
              }
              double adv = 0, lap = 0;
              for (int d = 0; d < 3; d++) {
                adv += a[d] * up3(a[d], f[d][0], f[d][1], f[d][2], f[d][3], f[d][4]);
                lap += f[d][1] + f[d][3] - 2 * f[d][2];
              }
              const double rhs = -adv / (6 * h) + nu * lap / (h * h);
              out[q * nc + cell] = fe_comp(&s->blk[b].d[iz][iy][ix], q, 0) + fac * rhs;
            }
          }
    }
    free(l.c);
  }
}

static void store_uvw_tmp(orc_sim* s, const double* in, int also_u)
{
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    FE* e = &s->blk[b].d[0][0][0];
    for (int i = 0; i < BS * BS * BS; i++) {
      const size_t c = (size_t)b * BS * BS * BS + i;
      e[i].tmpU = in[c];
      e[i].tmpV = in[nc + c];
      e[i].tmpW = in[2 * nc + c];
      if (also_u) {
        e[i].u = e[i].tmpU;
        e[i].v = e[i].tmpV;
        e[i].w = e[i].tmpW;
      }
    }
  }
}

void orc_advect_diffuse_scheme(orc_sim* s, double dt, double nu, const double uinf[3], int scheme)
{
  if (scheme == 0) {
    orc_advect_diffuse(s, dt, nu, uinf);
    return;
  }
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  double* out = (double*)malloc(3 * nc * sizeof(double));
  if (scheme == 1) {  /* Euler: u* = u + dt L(u) */
    upwind3_update(s, 0, dt, nu, uinf, out);
    store_uvw_tmp(s, out, 1);
  } else {            /* explicit midpoint: u1/2 = u + dt/2 L(u); u* = u + dt L(u1/2) */
    upwind3_update(s, 0, 0.5 * dt, nu, uinf, out);
    store_uvw_tmp(s, out, 0);
    upwind3_update(s, 1, dt, nu, uinf, out);
    store_uvw_tmp(s, out, 1);
  }
  free(out);
}

static inline size_t ridx(const orc_sim* s, int gx, int gy, int gz)
{
  return ((size_t)gx * s->N[1] + gy) * s->N[2] + gz;
}

/* KernelPressureRHS (PressureRHS.cpp:29-113) after the solver reset (:277, memset), then
 * with a sphere the perimeter correction (:322-343, KernelFinalizePerimeters :215-256). */
static void rhs_obstacle(orc_sim* s);
int orc_body_count(const orc_sim* s);

void orc_pressure_rhs(orc_sim* s, double dt)
{
  memset(s->rhs, 0, sizeof(double) * (size_t)s->N[0] * s->N[1] * s->N[2]);
  const double h = s->h, fac = 0.5 * h * h / dt;
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      int bc[3];
      bcoords(s, b, bc);
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const FE *LW = LAB(&l, ix - 1, iy, iz), *LE = LAB(&l, ix + 1, iy, iz);
            const FE *LS = LAB(&l, ix, iy - 1, iz), *LN = LAB(&l, ix, iy + 1, iz);
            const FE *LF = LAB(&l, ix, iy, iz - 1), *LB = LAB(&l, ix, iy, iz + 1);
            s->rhs[ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz)] =
                fac * (LE->u - LW->u + LN->v - LS->v + LB->w - LF->w);
          }
    }
    free(l.c);
  }
  /* the obstacles' masking and perimeter correction, body by body (their blocks are apart,
   * so this is the reference's all-masks-then-all-corrections order) */
  const int nbody = orc_body_count(s);
  for (int k = 0; k < nbody; k++) {
    orc_select_body(s, k);
    rhs_obstacle(s);
    orc_select_body(s, k);
  }
}

/* The active body's part of PressureRHS: rhs -= chi rhs where the grid chi does not exceed the
 * body's (PressureRHS.cpp:78-112), then the perimeter correction (:322-343, :215-256). */
static void rhs_obstacle(orc_sim* s)
{
  unsigned char* touch = (unsigned char*)calloc((size_t)s->N[0] * s->N[1] * s->N[2], 1);
  double sumRHS = 0, posRHS = 0, negRHS = 0;
  double* part = (double*)calloc(3 * (size_t)s->nb, sizeof(double));  /* per block, summed in block order */
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      if (s->ob[b] == NULL) continue;
      lab_load(s, b, &l);
      int bc[3];
      bcoords(s, b, bc);
      const OBlock* o = s->ob[b];
      double sumRHS = 0, posRHS = 0, negRHS = 0;
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            if (LAB(&l, ix, iy, iz)->chi > o->chi[iz][iy][ix]) continue;
            const size_t idx = ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz);
            const FE *LW = LAB(&l, ix - 1, iy, iz), *LE = LAB(&l, ix + 1, iy, iz);
            const FE *LS = LAB(&l, ix, iy - 1, iz), *LN = LAB(&l, ix, iy + 1, iz);
            const FE *LF = LAB(&l, ix, iy, iz - 1), *LB = LAB(&l, ix, iy, iz + 1);
            const double dXx = LE->chi - LW->chi, dXy = LN->chi - LS->chi, dXz = LB->chi - LF->chi;
            const int bSurface = dXx * dXx + dXy * dXy + dXz * dXz > 0;
            if (bSurface) touch[idx] = 1;
            const double srcBulk = -o->chi[iz][iy][ix] * s->rhs[idx];
            sumRHS += srcBulk;
            s->rhs[idx] += srcBulk;
            posRHS += bSurface * s->rhs[idx] * (s->rhs[idx] > 0);
            negRHS -= bSurface * s->rhs[idx] * (s->rhs[idx] < 0);
          }
      part[3 * b] = sumRHS;
      part[3 * b + 1] = posRHS;
      part[3 * b + 2] = negRHS;
    }
    free(l.c);
  }
  for (int b = 0; b < s->nb; b++) {
    sumRHS += part[3 * b];
    posRHS += part[3 * b + 1];
    negRHS += part[3 * b + 2];
  }
  free(part);
  const double corrDenom = sumRHS > 0 ? posRHS : negRHS;
  const double corr = (g_plant == ORC_PLANT_RHS_CORR_SIGN ? -1 : 1) * sumRHS / (corrDenom > EPS ? corrDenom : EPS);
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    if (s->ob[b] == NULL) continue;
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          const size_t idx = ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz);
          if (!touch[idx]) continue;
          if (s->rhs[idx] > 0 && corr > 0) s->rhs[idx] -= corr * s->rhs[idx];
          else if (s->rhs[idx] < 0 && corr < 0) s->rhs[idx] += corr * s->rhs[idx];
        }
  }
  free(touch);
}

void orc_get_rhs(const orc_sim* s, double* dst)
{
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    double* q = dst + (size_t)b * BS * BS * BS;
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++)
          q[(iz * BS + iy) * BS + ix] = s->rhs[ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz)];
  }
}

void orc_set_rhs(orc_sim* s, const double* src)
{
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    const double* q = src + (size_t)b * BS * BS * BS;
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++)
          s->rhs[ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz)] = q[(iz * BS + iy) * BS + ix];
  }
}

/* ---------------------------------------------------------------------- FFT */

/* In-place radix-2 decimation-in-time FFT of a contiguous line, unnormalized;
 * tw[k] = exp(sign*2*pi*i*k/n), k < n/2. */
static void fft_line(double complex* a, int n, const double complex* tw)
{
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      const double complex t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1, step = n / len;
    for (int i = 0; i < n; i += len)
      for (int k = 0; k < half; k++) {
        const double complex w = tw[k * step];
        const double complex x = a[i + k], y = a[i + k + half] * w;
        a[i + k] = x + y;
        a[i + k + half] = x - y;
      }
  }
}

/* Mixed-radix FFT for lengths that are not powers of two (FFTW plans any size,
 * PoissonSolverPeriodic.cpp:108-121): recursive decimation in time by the smallest prime
 * factor p of n. The p interleaved subsequences are transformed into out[r m .. r m + m),
 * m = n/p, then X[k + m q] = sum_r W_n^{r (k + m q)} Y_r[k]. twf[j] = exp(sign 2 pi i j / N)
 * for the top-level length N; at a level of length n the root W_n is twf[ts], ts = N/n. */
static void fft_mixed(const double complex* in, double complex* out, int n, int stride, const double complex* twf,
                      int ts)
{
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = 2;
  while (n % p) p++;
  const int m = n / p;
  for (int r = 0; r < p; r++) fft_mixed(in + (size_t)r * stride, out + (size_t)r * m, m, stride * p, twf, ts * p);
  double complex y[p], x[p];
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) y[r] = out[(size_t)r * m + k];
    for (int q = 0; q < p; q++) {
      double complex s = 0;
      for (int r = 0; r < p; r++) s += y[r] * twf[(size_t)(((long)r * (k + (long)m * q)) % n) * ts];
      x[q] = s;
    }
    for (int q = 0; q < p; q++) out[(size_t)q * m + k] = x[q];
  }
}

/* One line of any length n: the radix-2 loop above for powers of two, else fft_mixed. */
static void fft_any(double complex* a, int n, const double complex* tw, const double complex* twf,
                    double complex* scratch)
{
  if ((n & (n - 1)) == 0) {
    fft_line(a, n, tw);
    return;
  }
  fft_mixed(a, scratch, n, 1, twf, 1);
  memcpy(a, scratch, sizeof(double complex) * (size_t)n);
}

static double complex* make_tw_full(int n, int sign)
{
  double complex* tw = (double complex*)malloc(sizeof(double complex) * (size_t)n);
  for (int k = 0; k < n; k++) {
    const double ang = 2.0 * M_PI * (double)k / (double)n;
    tw[k] = cos(ang) + I * (sign * sin(ang));
  }
  return tw;
}

static double complex* make_tw(int n, int sign)
{
  double complex* tw = (double complex*)malloc(sizeof(double complex) * (n / 2 + 1));
  for (int k = 0; k < n / 2 + 1; k++) {
    const double ang = 2.0 * M_PI * (double)k / (double)n;
    tw[k] = cos(ang) + I * (sign * sin(ang));
  }
  return tw;
}

/* 3-D complex FFT of c[x][y][z] (z fastest), unnormalized. */
static void fft3d_c(double complex* c, int nx, int ny, int nz, int sign)
{
  double complex* twx = make_tw(nx, sign);
  double complex* twy = make_tw(ny, sign);
  double complex* twz = make_tw(nz, sign);
  double complex* tfx = make_tw_full(nx, sign);
  double complex* tfy = make_tw_full(ny, sign);
  double complex* tfz = make_tw_full(nz, sign);
  #pragma omp parallel
  {
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    double complex* line = (double complex*)malloc(sizeof(double complex) * nmax);
    double complex* scratch = (double complex*)malloc(sizeof(double complex) * nmax);
    #pragma omp for schedule(static)
    for (int xy = 0; xy < nx * ny; xy++) fft_any(c + (size_t)xy * nz, nz, twz, tfz, scratch);
    #pragma omp for schedule(static)
    for (int xz = 0; xz < nx * nz; xz++) {
      const int x = xz / nz, z = xz % nz;
      double complex* base = c + (size_t)x * ny * nz + z;
      for (int y = 0; y < ny; y++) line[y] = base[(size_t)y * nz];
      fft_any(line, ny, twy, tfy, scratch);
      for (int y = 0; y < ny; y++) base[(size_t)y * nz] = line[y];
    }
    #pragma omp for schedule(static)
    for (int yz = 0; yz < ny * nz; yz++) {
      double complex* base = c + yz;
      for (int x = 0; x < nx; x++) line[x] = base[(size_t)x * ny * nz];
      fft_any(line, nx, twx, tfx, scratch);
      for (int x = 0; x < nx; x++) base[(size_t)x * ny * nz] = line[x];
    }
    free(line);
    free(scratch);
  }
  free(twx); free(twy); free(twz);
  free(tfx); free(tfy); free(tfz);
}

void orc_fft3d(int nx, int ny, int nz, double* re, double* im, int sign)
{
  const size_t n = (size_t)nx * ny * nz;
  double complex* c = (double complex*)malloc(sizeof(double complex) * n);
  for (size_t i = 0; i < n; i++) c[i] = re[i] + I * im[i];
  fft3d_c(c, nx, ny, nz, sign < 0 ? -1 : +1);
  for (size_t i = 0; i < n; i++) { re[i] = creal(c[i]); im[i] = cimag(c[i]); }
  free(c);
}

/* PoissonSolverPeriodic::solve (PoissonSolverPeriodic.cpp:129-146) with _solve's
 * grid-consistent branch (:15-61): r2c, multiply by 4/(N0 h N1 N2) / (2D - 6) with
 * D = cos(4 pi i/N0) + cos(4 pi l/N1) + cos(4 pi k/N2), zero the 8 modes whose indices
 * are all in {0, N/2}, c2r; then _fftw2cub (PoissonSolver.cpp:116-130) copies into p. */
void orc_poisson_solve(orc_sim* s)
{
  const int N0 = s->N[0], N1 = s->N[1], N2 = s->N[2];
  const size_t n = (size_t)N0 * N1 * N2;
  double complex* c = (double complex*)malloc(sizeof(double complex) * n);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) c[i] = s->rhs[i];
  fft3d_c(c, N0, N1, N2, -1);
  const double h = s->h;
  const double norm_factor = 4 / ((double)N0 * h * N1 * N2);
  const double wFacX = 4 * M_PI / N0, wFacY = 4 * M_PI / N1, wFacZ = 4 * M_PI / N2;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < N0; i++)
    for (int l = 0; l < N1; l++)
      for (int k = 0; k < N2; k++) {
        const double D = cos(wFacX * i) + cos(wFacY * l) + cos(wFacZ * k);
        const double solutionFactor = norm_factor / (2 * D - 6);
        c[((size_t)i * N1 + l) * N2 + k] *= solutionFactor;
      }
  const int I0[2] = {0, N0 / 2}, L0[2] = {0, N1 / 2}, K0[2] = {0, N2 / 2};
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int d = 0; d < 2; d++) c[((size_t)I0[a] * N1 + L0[b]) * N2 + K0[d]] = 0;
  fft3d_c(c, N0, N1, N2, +1);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) s->rhs[i] = creal(c[i]);
  free(c);
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++)
          s->blk[b].d[iz][iy][ix].p = s->rhs[ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz)];
  }
}

/* KernelGradP (PressureProjection.cpp:27-51). */
/* PoissonSolverMixed::solve (PoissonSolverMixed.cpp:66-103, _solve PoissonSolverMixed.h:27-80)
 * on the RHS buffer: real-to-real transforms per axis as FFTW's r2r plans make them (R2HC /
 * HC2R on periodic axes, REDFT10 / REDFT01 on the others, PoissonSolverMixed.cpp:37-42), each
 * applied here as a dense N x N matrix; every coefficient times
 * norm / (denX + denY + denZ), den = (1 - tol)(cos(2 w i) - 1)/2 - tol ((k + s) w)^2 with
 * w = 2 pi/N, s = 0, k = min(i, N - i) (periodic) or w = pi/N, s = 1/2, k = i; tol = 0.01,
 * norm = (normX/h) normY normZ, norm = 1/N (periodic) or 1/(2N); the first coefficient zeroed.
 * The result goes to p (_fftw2cub, PoissonSolver.cpp:116-130). No FFT: N^4 work per axis, a
 * checker for small grids. */
static double* r2r_matrix(int n, int periodic, int forward)
{
  double* m = (double*)malloc(sizeof(double) * (size_t)n * n);  /* m[k][j]: out_k = sum_j m[k][j] in_j */
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) {
      double v;
      if (periodic && forward) {        /* R2HC, halfcomplex out: r_0..r_{n/2}, i_{(n+1)/2-1}..i_1 */
        v = k <= n / 2 ? cos(2 * M_PI * (double)j * k / n) : -sin(2 * M_PI * (double)j * (n - k) / n);
      } else if (periodic) {            /* HC2R: in_j is halfcomplex, out_k real */
        if (j == 0) v = 1;
        else if (2 * j == n) v = cos(M_PI * (double)k);
        else if (j < n / 2 || (j <= n / 2 && n % 2)) v = 2 * cos(2 * M_PI * (double)j * k / n);
        else v = -2 * sin(2 * M_PI * (double)(n - j) * k / n);
      } else if (forward) {             /* REDFT10 */
        v = 2 * cos(M_PI * (j + 0.5) * k / n);
      } else {                          /* REDFT01 */
        v = j == 0 ? 1 : 2 * cos(M_PI * (double)j * (k + 0.5) / n);
      }
      m[(size_t)k * n + j] = v;
    }
  return m;
}

/* a[x][y][z] (z fastest) along axis d by the matrix m. */
static void r2r_axis(double* a, const int N[3], int d, const double* m)
{
  const int n = N[d];
  const size_t sz = 1, sy = (size_t)N[2], sx = (size_t)N[1] * N[2];
  const size_t stride = d == 0 ? sx : (d == 1 ? sy : sz);
  const int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;  /* the other two axes */
  const size_t s1 = o1 == 0 ? sx : sy, s2 = o2 == 1 ? sy : sz;
  #pragma omp parallel
  {
    double* in = (double*)malloc(sizeof(double) * n);
    #pragma omp for schedule(static)
    for (int p = 0; p < N[o1] * N[o2]; p++) {
      double* base = a + (size_t)(p / N[o2]) * s1 + (size_t)(p % N[o2]) * s2;
      for (int j = 0; j < n; j++) in[j] = base[(size_t)j * stride];
      for (int k = 0; k < n; k++) {
        double acc = 0;
        for (int j = 0; j < n; j++) acc += m[(size_t)k * n + j] * in[j];
        base[(size_t)k * stride] = acc;
      }
    }
    free(in);
  }
}

void orc_poisson_mixed(orc_sim* s)
{
  const int* N = s->N;
  const double tol = 0.01, h = s->h;
  double* den[3];
  double norm = 1 / h;
  for (int d = 0; d < 3; d++) {
    const int per = s->bcf[d] == ORC_BC_PERIODIC, n = N[d];
    norm *= (per ? 1.0 : 0.5) / n;
    const double wf = (per ? 2 : 1) * M_PI / n;
    den[d] = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) {
      const int k = per ? (i <= n / 2 ? i : n - i) : i;
      const double rk2 = pow((k + (per ? 0 : 0.5)) * wf, 2);
      den[d][i] = (1 - tol) * (cos(2 * wf * i) - 1) / 2 - tol * rk2;
    }
  }
  for (int d = 0; d < 3; d++) {
    double* m = r2r_matrix(N[d], s->bcf[d] == ORC_BC_PERIODIC, 1);
    r2r_axis(s->rhs, N, d, m);
    free(m);
  }
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < N[0]; i++)
    for (int j = 0; j < N[1]; j++)
      for (int k = 0; k < N[2]; k++) {
        double* v = &s->rhs[((size_t)i * N[1] + j) * N[2] + k];
        *v *= norm / (den[0][i] + den[1][j] + den[2][k]);
      }
  s->rhs[0] = 0;
  for (int d = 0; d < 3; d++) {
    double* m = r2r_matrix(N[d], s->bcf[d] == ORC_BC_PERIODIC, 0);
    r2r_axis(s->rhs, N, d, m);
    free(m);
    free(den[d]);
  }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++)
          s->blk[b].d[iz][iy][ix].p = s->rhs[ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz)];
  }
}

/* PoissonSolverUnbounded::solve (PoissonSolverUnbounded.cpp:126-157, Green's function
 * :195-246): Eastwood and Brownrigg's cyclic convolution of the RHS, zero-padded to 2N-1
 * points per axis, with the free-space Green's function of the uniform mesh,
 * -h^2/(4 pi r) (r in cells; at r = 0 -0.1924173658 h^2, :216-218), normalized by
 * 1/(h^3 prod(2N-1)) (:50). On that period the cyclic convolution equals the linear one on
 * the physical points (distances stay below N), so the oracle evaluates the linear
 * convolution directly: p_i = sum_j g(|i - j|) rhs_j with g = -1/(4 pi h r), g(0) =
 * -0.1924173658/h. O(N^6): test grids only. */
void orc_poisson_unbounded(orc_sim* s)
{
  const int* N = s->N;
  const double h = s->h;
  const size_t nt = (size_t)N[0] * N[1] * N[2];
  double* g = (double*)malloc(sizeof(double) * nt);  /* g[dx][dy][dz], distances in cells */
  for (int i = 0; i < N[0]; i++)
    for (int j = 0; j < N[1]; j++)
      for (int k = 0; k < N[2]; k++) {
        const double r = sqrt((double)i * i + (double)j * j + (double)k * k);
        g[((size_t)i * N[1] + j) * N[2] + k] = r > 0 ? -1.0 / (4 * M_PI * h * r) : -0.1924173658 / h;
      }
  double* out = (double*)malloc(sizeof(double) * nt);
  #pragma omp parallel for schedule(dynamic, 16)
  for (size_t o = 0; o < nt; o++) {
    const int xi = (int)(o / ((size_t)N[1] * N[2])), yi = (int)((o / N[2]) % N[1]), zi = (int)(o % N[2]);
    double acc = 0;
    for (int x = 0; x < N[0]; x++) {
      const int dx = abs(xi - x);
      for (int y = 0; y < N[1]; y++) {
        const int dy = abs(yi - y);
        const double* gr = g + ((size_t)dx * N[1] + dy) * N[2];
        const double* rr = s->rhs + ((size_t)x * N[1] + y) * N[2];
        for (int z = 0; z < N[2]; z++) acc += gr[abs(zi - z)] * rr[z];
      }
    }
    out[o] = acc;
  }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++)
          s->blk[b].d[iz][iy][ix].p = out[ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz)];
  }
  free(out);
  free(g);
}

/* One r2r transform of a line (tests): kind 0 R2HC, 1 HC2R, 2 REDFT10, 3 REDFT01. */
void orc_r2r_line(double* a, int n, int kind)
{
  double* m = r2r_matrix(n, kind < 2, kind % 2 == 0);
  const int N[3] = {1, 1, n};
  r2r_axis(a, N, 2, m);
  free(m);
}

void orc_grad_p(orc_sim* s, double dt)
{
  const double fac = -0.5 * dt / s->h;
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      Block* o = &s->blk[b];
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            o->d[iz][iy][ix].u += fac * (LAB(&l, ix + 1, iy, iz)->p - LAB(&l, ix - 1, iy, iz)->p);
            o->d[iz][iy][ix].v += fac * (LAB(&l, ix, iy + 1, iz)->p - LAB(&l, ix, iy - 1, iz)->p);
            o->d[iz][iy][ix].w += fac * (LAB(&l, ix, iy, iz + 1)->p - LAB(&l, ix, iy, iz - 1)->p);
          }
    }
    free(l.c);
  }
}

/* KernelDissipation (ComputeDissipation.cpp:17-116); out[] in RDX order (:136-155). */
void orc_dissipation(const orc_sim* s, double nu, double out[20])
{
  double R[20] = {0};
  const double h = s->h, hCube = h * h * h, inv2h = .5 / h, invHh = 1 / (h * h);
  const double center[3] = {s->extent[0] / 2, s->extent[1] / 2, s->extent[2] / 2};
  double* part = (double*)calloc(20 * (size_t)s->nb, sizeof(double));  /* per block, summed in block order */
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      int bc[3];
      bcoords(s, b, bc);
      double R[20] = {0};
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const FE* L = LAB(&l, ix, iy, iz);
            const FE *LW = LAB(&l, ix - 1, iy, iz), *LE = LAB(&l, ix + 1, iy, iz);
            const FE *LS = LAB(&l, ix, iy - 1, iz), *LN = LAB(&l, ix, iy + 1, iz);
            const FE *LF = LAB(&l, ix, iy, iz - 1), *LB = LAB(&l, ix, iy, iz + 1);
            const double X = L->chi;
            double p[3];
            cell_pos(s, bc, ix, iy, iz, p);
            const double PX = p[0] - center[0], PY = p[1] - center[1], PZ = p[2] - center[2];
            const double WX = inv2h * ((LN->w - LS->w) - (LB->v - LF->v));
            const double WY = inv2h * ((LB->u - LF->u) - (LE->w - LW->w));
            const double WZ = inv2h * ((LE->v - LW->v) - (LN->u - LS->u));
            const double dPdx = inv2h * (LE->p - LW->p);
            const double dPdy = inv2h * (LN->p - LS->p);
            const double dPdz = inv2h * (LB->p - LF->p);
            R[15] -= (1 - X) * hCube * (dPdx * L->u + dPdy * L->v + dPdz * L->w);
            const double lapU = invHh * (LE->u + LW->u + LN->u + LS->u + LB->u + LF->u - 6 * L->u);
            const double lapV = invHh * (LE->v + LW->v + LN->v + LS->v + LB->v + LF->v - 6 * L->v);
            const double lapW = invHh * (LE->w + LW->w + LN->w + LS->w + LB->w + LF->w - 6 * L->w);
            const double V1 = lapU * L->u + lapV * L->v + lapW * L->w;
            const double D11 = inv2h * (LE->u - LW->u);
            const double D22 = inv2h * (LN->v - LS->v);
            const double D33 = inv2h * (LB->w - LF->w);
            const double D12 = inv2h * (LN->u - LS->u + LE->v - LW->v) / 2;
            const double D13 = inv2h * (LB->u - LF->u + LE->w - LW->w) / 2;
            const double D23 = inv2h * (LN->w - LS->w + LB->v - LF->v) / 2;
            const double V2 = D11 * D11 + D22 * D22 + D33 * D33 +
                              (g_plant == ORC_PLANT_DISS_SHEAR_WEIGHT ? 1 : 2) * (D12 * D12 + D13 * D13 + D23 * D23);
            R[16] += (1 - X) * hCube * nu * (V1 + 2 * V2);
            R[0] += hCube * WX;
            R[1] += hCube * WY;
            R[2] += hCube * WZ;
            R[3] += hCube / 2 * (PY * WZ - (g_plant == ORC_PLANT_DISS_IMPULSE_SIGN ? -1 : 1) * PZ * WY);
            R[4] += hCube / 2 * (PZ * WX - PX * WZ);
            R[5] += hCube / 2 * (PX * WY - PY * WX);
            R[6] += hCube * L->u;
            R[7] += hCube * L->v;
            R[8] += hCube * L->w;
            R[9] += hCube / 3 * (PX * (PY * WY + PZ * WZ) - WX * (PY * PY + (g_plant == ORC_PLANT_DISS_ANGIMP_TERM ? 0 : 1) * PZ * PZ));
            R[10] += hCube / 3 * (PY * (PX * WX + PZ * WZ) - WY * (PX * PX + PZ * PZ));
            R[11] += hCube / 3 * (PZ * (PX * WX + PY * WY) - WZ * (PX * PX + PY * PY));
            R[12] += hCube * (PY * L->w - PZ * L->v);
            R[13] += hCube * (PZ * L->u - PX * L->w);
            R[14] += hCube * (PX * L->v - PY * L->u);
            R[17] += hCube * (WX * L->u + WY * L->v + (g_plant == ORC_PLANT_DISS_HELICITY ? -1 : 1) * WZ * L->w);
            R[18] += hCube * (L->u * L->u + L->v * L->v + L->w * L->w) / 2;
            R[19] += sqrt(WX * WX + WY * WY + WZ * WZ);
          }
      for (int i = 0; i < 20; i++) part[20 * (size_t)b + i] = R[i];
    }
    free(l.c);
  }
  for (int b = 0; b < s->nb; b++)
    for (int i = 0; i < 20; i++) R[i] += part[20 * (size_t)b + i];
  free(part);
  for (int i = 0; i < 20; i++) out[i] = R[i];
}

/* ------------------------------------------------------------ spectral forcing */

static void field_to_grid(const orc_sim* s, int f, double complex* c)
{
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++)
          c[ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz)] =
              *fe_field((FE*)&s->blk[b].d[iz][iy][ix], f);
  }
}

static inline int signed_k(int i, int n) { return i <= n / 2 ? i : -(n - i); }

/* SpectralForcing::operator() (SpectralForcing.cpp:23-82) with
 * _compute_largeModesForcing (SpectralManipFFTW.cpp:25-108). The half-spectrum sums with
 * weight mult/2 equal full-spectrum sums with weight 1/2. */
/* The HIT branch of Analysis::operator() (Analysis.cpp:159-196): KernelAnalysis_gradStats
 * (:18-52) and SpectralAnalysis::run (SpectralAnalysis.cpp:99-106), i.e. the forward FFT of
 * u, v, w and SpectralManipFFTW::_compute_analysis (SpectralManipFFTW.cpp:110-207) with the
 * HITstatistics histogram (SpectralManip.h:51-110: nyquist = maxN/2, nBin = nyquist - 1).
 * Full complex spectra with weight 1 per mode, which by Hermitian symmetry equals the
 * reference's half spectrum with weights mult/2. out[9] = tke, eps, uprime, lambda, Re_lambda,
 * l_integral, tau_integral, grad_mean, grad_std; E (nBin values) the shell spectrum. Sums are
 * per x plane, added in plane order. */
void orc_hit_analysis(orc_sim* s, double nu, double out[9], double* E)
{
  const int N0 = s->N[0], N1 = s->N[1], N2 = s->N[2];
  const size_t n = (size_t)N0 * N1 * N2;
  int maxN = N0 > N1 ? N0 : N1;
  if (N2 > maxN) maxN = N2;
  const int nyq = maxN / 2, nbin = nyq - 1;
  const double nyq_scale = (nyq - 1) / (double)nyq;
  double complex* cu = (double complex*)malloc(sizeof(double complex) * n);
  double complex* cv = (double complex*)malloc(sizeof(double complex) * n);
  double complex* cw = (double complex*)malloc(sizeof(double complex) * n);
  field_to_grid(s, ORC_U, cu);
  field_to_grid(s, ORC_V, cv);
  field_to_grid(s, ORC_W, cw);
  fft3d_c(cu, N0, N1, N2, -1);
  fft3d_c(cv, N0, N1, N2, -1);
  fft3d_c(cw, N0, N1, N2, -1);
  const double wfx = 2.0 * M_PI / s->extent[0], wfy = 2.0 * M_PI / s->extent[1], wfz = 2.0 * M_PI / s->extent[2];
  const double dfx = 2 * M_PI / N0, dfy = 2 * M_PI / N1, dfz = 2 * M_PI / N2;
  double* part = (double*)calloc((3 + (size_t)nbin) * N0, sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < N0; i++) {
    double* P = part + (3 + (size_t)nbin) * i;
    for (int l = 0; l < N1; l++)
      for (int k = 0; k < N2; k++) {
        const size_t id = ((size_t)i * N1 + l) * N2 + k;
        const int ii = signed_k(i, N0), jj = signed_k(l, N1), kk = signed_k(k, N2);
        const double kx = ii * wfx, ky = jj * wfy, kz = kk * wfz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double dX = 2 * sin(dfx * ii), dY = 2 * sin(dfy * jj), dZ = 2 * sin(dfz * kk);
        const double complex U = cu[id], V = cv[id], W = cw[id];
        const double complex OX = I * dY * W - I * dZ * V;
        const double complex OY = I * dZ * U - I * dX * W;
        const double complex OZ = I * dX * V - I * dY * U;
        const double e = 0.5 * (creal(U * conj(U)) + creal(V * conj(V)) + creal(W * conj(W)));
        P[0] += e;
        P[1] += 0.5 * (creal(OX * conj(OX)) + creal(OY * conj(OY)) + creal(OZ * conj(OZ)));
        P[2] += k2 > 0 ? e / sqrt(k2) : 0;
        const long kind = (long)ii * ii + (long)jj * jj + (long)kk * kk;
        if (kind < (long)nyq * nyq) P[3 + (size_t)floor(sqrt((double)kind) * nyq_scale)] += e;
      }
  }
  double tke = 0, eps = 0, lint = 0;
  for (int b = 0; b < nbin; b++) E[b] = 0;
  for (int i = 0; i < N0; i++) {
    const double* P = part + (3 + (size_t)nbin) * i;
    tke += P[0]; eps += P[1]; lint += P[2];
    for (int b = 0; b < nbin; b++) E[b] += P[3 + b];
  }
  free(part); free(cu); free(cv); free(cw);
  const double normFFT = (double)n;
  for (int b = 0; b < nbin; b++) E[b] /= normFFT * normFFT;
  out[0] = tke / (normFFT * normFFT);
  out[1] = eps * 2 * nu / pow(normFFT * 2 * s->h, 2);
  out[2] = sqrt(2 * out[0] / 3);
  out[3] = sqrt(15 * nu / out[1]) * out[2];
  out[4] = out[2] * out[3] / nu;
  out[5] = lint / tke;
  out[6] = out[5] / out[2];
  /* gradient statistics: fac = 1/(4h^2), grad2 = fac sum of squared central differences */
  double* gp = (double*)calloc(2 * (size_t)s->nb, sizeof(double));
  const double fac = 1 / (4 * s->h * s->h);
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      double gm = 0, gs = 0;
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const FE *LW = LAB(&l, ix - 1, iy, iz), *LE = LAB(&l, ix + 1, iy, iz);
            const FE *LS = LAB(&l, ix, iy - 1, iz), *LN = LAB(&l, ix, iy + 1, iz);
            const FE *LF = LAB(&l, ix, iy, iz - 1), *LB = LAB(&l, ix, iy, iz + 1);
            const double ux = LE->u - LW->u, vx = LE->v - LW->v, wx = LE->w - LW->w;
            const double uy = LN->u - LS->u, vy = LN->v - LS->v, wy = LN->w - LS->w;
            const double uz = LB->u - LF->u, vz = LB->v - LF->v, wz = LB->w - LF->w;
            const double g2 = fac * (ux * ux + vx * vx + wx * wx + uy * uy + vy * vy + wy * wy +
                                     uz * uz + vz * vz + wz * wz);
            gm += sqrt(g2);
            gs += g2;
          }
      gp[2 * b] = gm;
      gp[2 * b + 1] = gs;
    }
    free(l.c);
  }
  double gm = 0, gs = 0;
  for (int b = 0; b < s->nb; b++) { gm += gp[2 * b]; gs += gp[2 * b + 1]; }
  free(gp);
  gm /= normFFT;
  gs /= normFFT;
  out[7] = gm;
  out[8] = sqrt(gs - gm * gm);
}

void orc_spectral_forcing(orc_sim* s, double dt, double sim_dt, double nu,
                          double tke_target, double injection_rate, double stats_out[5])
{
  const int N0 = s->N[0], N1 = s->N[1], N2 = s->N[2];
  const size_t n = (size_t)N0 * N1 * N2;
  double complex* cu = (double complex*)malloc(sizeof(double complex) * n);
  double complex* cv = (double complex*)malloc(sizeof(double complex) * n);
  double complex* cw = (double complex*)malloc(sizeof(double complex) * n);
  field_to_grid(s, ORC_U, cu);
  field_to_grid(s, ORC_V, cv);
  field_to_grid(s, ORC_W, cw);
  fft3d_c(cu, N0, N1, N2, -1);
  fft3d_c(cv, N0, N1, N2, -1);
  fft3d_c(cw, N0, N1, N2, -1);
  const double wfx = 2.0 * M_PI / s->extent[0], wfy = 2.0 * M_PI / s->extent[1], wfz = 2.0 * M_PI / s->extent[2];
  const double dfx = 2 * M_PI / N0, dfy = 2 * M_PI / N1, dfz = 2 * M_PI / N2;
  double eps = 0, tke = 0, tkeF = 0, lInt = 0;
  double* part = (double*)calloc(4 * (size_t)N0, sizeof(double));  /* per x plane, summed in plane order */
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < N0; i++) {
    double eps = 0, tke = 0, tkeF = 0, lInt = 0;
    for (int l = 0; l < N1; l++)
      for (int k = 0; k < N2; k++) {
        const size_t id = ((size_t)i * N1 + l) * N2 + k;
        const int ii = signed_k(i, N0), jj = signed_k(l, N1), kk = signed_k(k, N2);
        const double kx = ii * wfx, ky = jj * wfy, kz = kk * wfz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double dX = 2 * sin(dfx * ii), dY = 2 * sin(dfy * jj), dZ = 2 * sin(dfz * kk);
        const double complex U = cu[id], V = cv[id], W = cw[id];
        /* i*d applied: d(.)/dx -> i * dX * (.) */
        const double complex OX = I * dY * W - I * dZ * V;
        const double complex OY = I * dZ * U - I * dX * W;
        const double complex OZ = I * dX * V - I * dY * U;
        const double E = 0.5 * (creal(U * conj(U)) + creal(V * conj(V)) + creal(W * conj(W)));
        tke += E;
        eps += 0.5 * (creal(OX * conj(OX)) + creal(OY * conj(OY)) + creal(OZ * conj(OZ)));
        lInt += (k2 > 0) ? E / sqrt(k2) : 0;
        if (k2 > 0 && k2 <= 4) tkeF += E;
        else cu[id] = cv[id] = cw[id] = 0;
      }
    part[4 * i] = eps; part[4 * i + 1] = tke; part[4 * i + 2] = tkeF; part[4 * i + 3] = lInt;
  }
  for (int i = 0; i < N0; i++) {
    eps += part[4 * i]; tke += part[4 * i + 1]; tkeF += part[4 * i + 2]; lInt += part[4 * i + 3];
  }
  free(part);
  const double normFFT = (double)N0 * N1 * N2;
  const double st_tkeF = tkeF / (normFFT * normFFT);
  const double st_tke = tke / (normFFT * normFFT);
  const double st_eps = eps * 2 * nu / pow(normFFT * 2 * s->h, 2);
  const double st_lint = lInt / tke;
  double inj = 0;
  if (tke_target > 0) inj = (tke_target - st_tke) / dt;
  else if (injection_rate > 0) inj = injection_rate;
  const double fac = sim_dt * inj / (2 * st_tkeF);
  if (fac > 0) {
    fft3d_c(cu, N0, N1, N2, +1);
    fft3d_c(cv, N0, N1, N2, +1);
    fft3d_c(cw, N0, N1, N2, +1);
    const double f = fac / normFFT;
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      int bc[3];
      bcoords(s, b, bc);
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const size_t id = ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz);
            FE* e = &s->blk[b].d[iz][iy][ix];
            e->u += f * creal(cu[id]);
            e->v += f * creal(cv[id]);
            e->w += f * creal(cw[id]);
          }
    }
  }
  if (stats_out) {
    stats_out[0] = st_tke;
    stats_out[1] = st_eps;
    stats_out[2] = st_tkeF;
    stats_out[3] = st_lint;
    stats_out[4] = fac;
  }
  free(cu); free(cv); free(cw);
}

/* ---------------------------------------------------------- initial conditions */

/* KernelIC_taylorGreen (InitialConditions.cpp:32-52). */
void orc_ic_taylor_green(orc_sim* s, double uMax)
{
  const double a = 2 * M_PI / s->extent[0], b_ = 2 * M_PI / s->extent[1], c = 2 * M_PI / s->extent[2];
  const double A = uMax, B = -uMax * s->extent[1] / s->extent[0];
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          FE* e = &s->blk[b].d[iz][iy][ix];
          memset(e, 0, sizeof(FE));
          double p[3];
          cell_pos(s, bc, ix, iy, iz, p);
          e->u = A * cos(a * p[0]) * sin(b_ * p[1]) * sin(c * p[2]);
          e->v = B * sin(a * p[0]) * cos(b_ * p[1]) * sin(c * p[2]);
        }
  }
}

/* Philox-4x32-10 (Salmon et al., SC'11), the counter-based generator of the seeded
 * isotropic-turbulence initial condition. */
static void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox4x32_10(ctr, key, out); }

/* Seeded isotropic-turbulence IC. Construction of each mode follows
 * SpectralManipFFTW::_compute_IC (SpectralManipFFTW.cpp:217-288) -- Rogallo's solenoidal
 * combination of two random complex amplitudes -- with the "art" spectrum
 * E(k) = 32/(3 k0) sqrt(2/pi) (k/k0)^4 exp(-2 (k/k0)^2) (SpectralIcGenerator.cpp:63-81)
 * evaluated at |k|, with these fixed differences: phases from Philox-4x32-10 keyed by the
 * seed and counted by the canonical mode index, the imaginary part of u built from
 * noise_a[1] (the reference reads noise_a[0] there, :277), exact Hermitian symmetry
 * (every mode with a negative partner takes the conjugate of its partner), DC and every
 * Nyquist index zeroed, and the field scaled so that sqrt(<u^2+v^2+w^2>/3) = u_rms. */
/* EnergySpectrum::interpE (SpectralManip.cpp:96-116): with i the first index of K above k,
 * the line through (K[i-1], E[i-1]) and (K[i], E[i]); below K[0] the reference's form
 * E[0] - (E[1] - E[0]) / (K[1] - K[0]) (K[1] - k); beyond the table the last segment's line;
 * negative values give 0. */
double orc_spectrum_interp(const double* K, const double* E, int n, double k)
{
  int i = 0;
  while (i < n && !(k < K[i])) i++;
  double e;
  if (i > 0 && i < n) e = E[i] + (E[i - 1] - E[i]) / (K[i] - K[i - 1]) * (K[i] - k);
  else if (i == 0) e = E[0] - (E[1] - E[0]) / (K[1] - K[0]) * (K[1] - k);
  else e = E[n - 1] + (E[n - 2] - E[n - 1]) / (K[n - 1] - K[n - 2]) * (K[n - 1] - k);
  return e < 0 ? 0.0 : e;
}

static void ic_hit_core(orc_sim* s, double k0, const double* tK, const double* tE, int tn, double u_rms,
                        uint64_t seed);

void orc_ic_hit(orc_sim* s, double k0, double u_rms, uint64_t seed) { ic_hit_core(s, k0, NULL, NULL, 0, u_rms, seed); }

/* The same construction with E(|k|) from a target table (SpectralIcGenerator::_generateTarget,
 * SpectralIcGenerator.cpp:21-98, read by interpE in _compute_IC); u_rms <= 0 keeps the
 * table's energy (the reference does not rescale). */
void orc_ic_hit_spectrum(orc_sim* s, const double* K, const double* E, int n, double u_rms, uint64_t seed)
{
  ic_hit_core(s, 0.0, K, E, n, u_rms, seed);
}

static void ic_hit_core(orc_sim* s, double k0, const double* tK, const double* tE, int tn, double u_rms,
                        uint64_t seed)
{
  const int N0 = s->N[0], N1 = s->N[1], N2 = s->N[2];
  const size_t n = (size_t)N0 * N1 * N2;
  double complex* cu = (double complex*)calloc(n, sizeof(double complex));
  double complex* cv = (double complex*)calloc(n, sizeof(double complex));
  double complex* cw = (double complex*)calloc(n, sizeof(double complex));
  const double wfx = 2.0 * M_PI / s->extent[0], wfy = 2.0 * M_PI / s->extent[1], wfz = 2.0 * M_PI / s->extent[2];
  const double coef = tn ? 0.0 : 32.0 / (3 * k0) * sqrt(2.0 / M_PI);
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < N0; i++)
    for (int l = 0; l < N1; l++)
      for (int k = 0; k < N2; k++) {
        const size_t id = ((size_t)i * N1 + l) * N2 + k;
        int ii = signed_k(i, N0), jj = signed_k(l, N1), kk = signed_k(k, N2);
        if (2 * abs(ii) == N0 || 2 * abs(jj) == N1 || 2 * abs(kk) == N2) continue;
        if (ii == 0 && jj == 0 && kk == 0) continue;
        const int canonical = kk > 0 || (kk == 0 && (jj > 0 || (jj == 0 && ii > 0)));
        if (!canonical) { ii = -ii; jj = -jj; kk = -kk; }
        const size_t cid = ((size_t)wrap(ii, N0) * N1 + wrap(jj, N1)) * N2 + wrap(kk, N2);
        const double kx = ii * wfx, ky = jj * wfy, kz = kk * wfz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double k_xy = sqrt(kx * kx + ky * ky), k_norm = sqrt(k2);
        const double r = tn ? 0.0 : k_norm / k0;
        const double E_k = tn ? orc_spectrum_interp(tK, tE, tn, k_norm) : coef * r * r * r * r * exp(-2 * r * r);
        const double amp = sqrt(E_k / (2 * M_PI * pow(k_norm / wfx, 2)));
        const uint32_t ctr[4] = {(uint32_t)cid, (uint32_t)((uint64_t)cid >> 32), 0, 0};
        uint32_t o[4];
        philox4x32_10(ctr, key, o);
        const double theta1 = (o[0] + 0.5) * (1.0 / 4294967296.0) * 2 * M_PI;
        const double theta2 = (o[1] + 0.5) * (1.0 / 4294967296.0) * 2 * M_PI;
        const double phi = (o[2] + 0.5) * (1.0 / 4294967296.0) * 2 * M_PI;
        const double complex na = amp * cos(phi) * (cos(theta1) + I * sin(theta1));
        const double complex nb = amp * sin(phi) * (cos(theta2) + I * sin(theta2));
        const double fac = k_norm * k_xy, invFac = fac <= 0 ? 0 : 1 / fac;
        double complex U = invFac * (na * k_norm * ky + nb * kx * kz);
        double complex V = invFac * (nb * ky * kz - na * k_norm * kx);
        double complex W = -nb * k_xy / k_norm;
        if (!canonical) { U = conj(U); V = conj(V); W = conj(W); }
        cu[id] = U; cv[id] = V; cw[id] = W;
      }
  fft3d_c(cu, N0, N1, N2, +1);
  fft3d_c(cv, N0, N1, N2, +1);
  fft3d_c(cw, N0, N1, N2, +1);
  double ssq = 0;
  double* part = (double*)calloc((size_t)N0, sizeof(double));  /* per x plane, summed in plane order */
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < N0; i++) {
    double a2 = 0;
    for (size_t j = (size_t)i * N1 * N2; j < (size_t)(i + 1) * N1 * N2; j++) {
      const double a = creal(cu[j]), b = creal(cv[j]), c = creal(cw[j]);
      a2 += a * a + b * b + c * c;
    }
    part[i] = a2;
  }
  for (int i = 0; i < N0; i++) ssq += part[i];
  free(part);
  const double scale = u_rms > 0 ? u_rms / sqrt(ssq / (3.0 * n)) : 1.0;
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          const size_t id = ridx(s, bc[0] * BS + ix, bc[1] * BS + iy, bc[2] * BS + iz);
          FE* e = &s->blk[b].d[iz][iy][ix];
          memset(e, 0, sizeof(FE));
          e->u = scale * creal(cu[id]);
          e->v = scale * creal(cv[id]);
          e->w = scale * creal(cw[id]);
        }
  }
  free(cu); free(cv); free(cw);
}

/* ------------------------------------------------------------------ sphere */

static void ob_write(OBlock* o, int ix, int iy, int iz, double delta, double gx, double gy, double gz)
{
  if (o->npts == o->cap) {
    o->cap = o->cap ? 2 * o->cap : 64;
    o->pix = (int*)realloc(o->pix, sizeof(int) * o->cap);
    o->dchidx = (double*)realloc(o->dchidx, sizeof(double) * o->cap);
    o->dchidy = (double*)realloc(o->dchidy, sizeof(double) * o->cap);
    o->dchidz = (double*)realloc(o->dchidz, sizeof(double) * o->cap);
    o->delta = (double*)realloc(o->delta, sizeof(double) * o->cap);
  }
  const int k = o->npts++;
  o->pix[k] = (iz * BS + iy) * BS + ix;
  o->dchidx[k] = -delta * gx; /* ObstacleBlock::write (ObstacleBlock.h:157-166) */
  o->dchidy[k] = -delta * gy;
  o->dchidz[k] = -delta * gz;
  o->delta[k] = delta;
}

/* CreateObstacles for one sphere (ObstaclesCreate.cpp:444-540): reset chi=0 and
 * tmpU=-1, SphereObstacle::FillBlocks over the blocks its box touches (Sphere.cpp:17-50,
 * ObstacleLibrary.h:41-69, Obstacle.h:186-207), KernelCharacteristicFunction
 * (ObstaclesCreate.cpp:22-137), KernelComputeGridCoM (:230-256). The sphere's udef is
 * zero, so the udef-momentum steps (:258-440) leave everything unchanged. */
/* Mollified Heaviside of the signed distance d with half-width e (no reference
 * counterpart): 0 for d <= -e, 1 for d >= e, (1 + d/e + sin(pi d/e)/pi)/2 between. It is
 * C1: dH/dd = (1 + cos(pi d/e))/(2e) vanishes at d = +-e. */
static double moll_H(double d, double e)
{
  if (d <= -e) return 0.0;
  if (d >= e) return 1.0;
  return 0.5 * (1.0 + d / e + sin(M_PI * d / e) / M_PI);
}

static double moll_dH(double d, double e)
{
  if (d <= -e || d >= e) return 0.0;
  return (1.0 + cos(M_PI * d / e)) / (2.0 * e);
}

/* d chi / d centre at a cell of position p for the mollified chi: dH(R - |p - c|) (p - c)/|p - c|. */
static void moll_dchi_dc(const orc_sim* s, const double p[3], double g[3])
{
  const double r[3] = {p[0] - s->spos[0], p[1] - s->spos[1], p[2] - s->spos[2]};
  const double dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  const double dH = dist > 0 ? moll_dH(s->radius - dist, s->moll * s->h) / dist : 0.0;
  for (int k = 0; k < 3; k++) g[k] = dH * r[k];
}

static void sphere_build(orc_sim* s, double radius, const double pos[3]);
static void sphere_build_blocks(orc_sim* s, double radius, const double pos[3], int reset);

void orc_sphere_create(orc_sim* s, double radius, const double pos[3])
{
  free_other_bodies(s);
  s->moll = 0;
  s->hemi = 0;
  sphere_build(s, radius, pos);
}

/* ObstacleVector with several spheres (ObstaclesCreate.cpp:444-480 for each): the grid chi
 * and SDF field are reset once and every body writes its blocks (which must be apart, as the
 * device requires); the new body takes index nmore + 1 (0 when it is the first). */
static void bodies_rebuild(orc_sim* s)
{
  for (int b = 0; b < s->nb; b++) {
    FE* e = &s->blk[b].d[0][0][0];
    for (int i = 0; i < BS * BS * BS; i++) { e[i].chi = 0; e[i].tmpU = -1; }
  }
  for (int k = 0; k <= s->nmore; k++) {
    orc_select_body(s, k);
    const double r = s->radius, p[3] = {s->spos[0], s->spos[1], s->spos[2]};
    sphere_build_blocks(s, r, p, 0);
    orc_select_body(s, k);
  }
}

int orc_add_sphere(orc_sim* s, double radius, const double pos[3], int hemi)
{
  if (!s->has_sphere) {
    s->moll = 0;
    s->hemi = hemi;
    sphere_build(s, radius, pos);
    return 0;
  }
  s->more = (OBody*)realloc(s->more, (size_t)(s->nmore + 1) * sizeof(OBody));
  OBody nb = {0, radius, {pos[0], pos[1], pos[2]}, 0.0, hemi, NULL, {0, 0, 0}, 0};
  s->more[s->nmore++] = nb;
  bodies_rebuild(s);
  return s->nmore;
}

/* Body k at a new centre: CreateObstacles resets chi and the SDF field and every body fills
 * its blocks again (ObstaclesCreate.cpp:449-466). Returns -1 for no such body. */
int orc_move_body(orc_sim* s, int k, const double pos[3])
{
  if (k < 0 || k >= orc_body_count(s)) return -1;
  double* c = k == 0 ? s->spos : s->more[k - 1].spos;
  for (int d = 0; d < 3; d++) c[d] = pos[d];
  bodies_rebuild(s);
  return 0;
}

int orc_body_count(const orc_sim* s) { return s->has_sphere ? 1 + s->nmore : 0; }

/* HemiSphereObstacle::FillBlocks (Sphere.cpp:53-84): SDF min(-(x - c_x), R - |x - c|) and the
 * block box [c_x - R - safety, c_x + safety] in x; chi by Towers' formula. */
void orc_hemisphere_create(orc_sim* s, double radius, const double pos[3])
{
  free_other_bodies(s);
  s->moll = 0;
  s->hemi = 1;
  sphere_build(s, radius, pos);
}

/* The sphere with chi = moll_H(sdf, width h) instead of Towers' formula (width in (0, 2]
 * cells keeps the chi band inside the obstacle blocks); SDF, surface points and the rest
 * as orc_sphere_create. */
void orc_sphere_create_mollified(orc_sim* s, double radius, const double pos[3], double width)
{
  free_other_bodies(s);
  s->moll = width;
  s->hemi = 0;
  sphere_build(s, radius, pos);
}

static void sphere_build(orc_sim* s, double radius, const double pos[3]) { sphere_build_blocks(s, radius, pos, 1); }

static void sphere_build_blocks(orc_sim* s, double radius, const double pos[3], int reset)
{
  free_sphere(s);
  s->has_sphere = 1;
  s->radius = radius;
  for (int d = 0; d < 3; d++) s->spos[d] = pos[d];
  s->ob = (OBlock**)calloc((size_t)s->nb, sizeof(OBlock*));
  const double h = s->h, safety = (2 + SURFDH) * h;
  double box[3][2];
  for (int d = 0; d < 3; d++) {
    box[d][0] = pos[d] - (radius + safety);
    box[d][1] = pos[d] + (radius + safety);
  }
  if (s->hemi) box[0][1] = pos[0] + safety;
  if (reset) {
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      FE* e = &s->blk[b].d[0][0][0];
      for (int i = 0; i < BS * BS * BS; i++) { e[i].chi = 0; e[i].tmpU = -1; }
    }
  }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    int bc[3];
    bcoords(s, b, bc);
    double mn[3], mx[3];
    cell_pos(s, bc, 0, 0, 0, mn);
    cell_pos(s, bc, BS - 1, BS - 1, BS - 1, mx);
    int touching = 1;
    for (int d = 0; d < 3; d++) {
      const double lo = mn[d] > box[d][0] ? mn[d] : box[d][0];
      const double hi = mx[d] < box[d][1] ? mx[d] : box[d][1];
      if (!(hi - lo > 0)) touching = 0;
    }
    if (!touching) continue;
    OBlock* o = (OBlock*)calloc(1, sizeof(OBlock));
    s->ob[b] = o;
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          double p[3];
          cell_pos(s, bc, ix, iy, iz, p);
          const double dx = p[0] - pos[0], dy = p[1] - pos[1], dz = p[2] - pos[2];
          const double ball = radius - sqrt(dx * dx + dy * dy + dz * dz);
          const double dist = s->hemi && -dx < ball ? -dx : ball;
          o->sdf[iz][iy][ix] = dist;
          FE* e = &s->blk[b].d[iz][iy][ix];
          e->tmpU = dist > e->tmpU ? dist : e->tmpU;
        }
  }
  const double inv2h = .5 / h, fac1 = .5 * h * h, vol = h * h * h;
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      OBlock* o = s->ob[b];
      if (!o) continue;
      lab_load(s, b, &l);
      int bc[3];
      bcoords(s, b, bc);
      o->CoM[0] = o->CoM[1] = o->CoM[2] = o->mass = 0;
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const double sdf0 = o->sdf[iz][iy][ix];
            if (s->moll > 0) {
              o->chi[iz][iy][ix] = moll_H(sdf0, s->moll * h);
            } else if (sdf0 > +2 * h || sdf0 < -2 * h) {
              o->chi[iz][iy][ix] = sdf0 > 0 ? 1 : 0;
            } else {
              const double distPx = LAB(&l, ix + 1, iy, iz)->tmpU, distMx = LAB(&l, ix - 1, iy, iz)->tmpU;
              const double distPy = LAB(&l, ix, iy + 1, iz)->tmpU, distMy = LAB(&l, ix, iy - 1, iz)->tmpU;
              const double distPz = LAB(&l, ix, iy, iz + 1)->tmpU, distMz = LAB(&l, ix, iy, iz - 1)->tmpU;
              const double gradUX = inv2h * (distPx - distMx);
              const double gradUY = inv2h * (distPy - distMy);
              const double gradUZ = inv2h * (distPz - distMz);
              const double gradUSq = gradUX * gradUX + gradUY * gradUY + gradUZ * gradUZ + EPS;
              const double IplusX = distPx < 0 ? 0 : distPx, IminuX = distMx < 0 ? 0 : distMx;
              const double IplusY = distPy < 0 ? 0 : distPy, IminuY = distMy < 0 ? 0 : distMy;
              const double IplusZ = distPz < 0 ? 0 : distPz, IminuZ = distMz < 0 ? 0 : distMz;
              const double gradIX = inv2h * (IplusX - IminuX);
              const double gradIY = inv2h * (IplusY - IminuY);
              const double gradIZ = inv2h * (IplusZ - IminuZ);
              const double numH = gradIX * gradUX + gradIY * gradUY + gradIZ * gradUZ;
              o->chi[iz][iy][ix] = numH / gradUSq;
            }
            double p[3];
            cell_pos(s, bc, ix, iy, iz, p);
            FE* e = &s->blk[b].d[iz][iy][ix];
            const double X = o->chi[iz][iy][ix];
            e->chi = X > e->chi ? X : e->chi;
            o->CoM[0] += X * vol * p[0];
            o->CoM[1] += X * vol * p[1];
            o->CoM[2] += X * vol * p[2];
            o->mass += X * vol;
            const double sdf = sdf0 + h * SURFDH;
            if (sdf > +2 * h || sdf < -2 * h) continue;
            const double distPx = LAB(&l, ix + 1, iy, iz)->tmpU + h * SURFDH;
            const double distMx = LAB(&l, ix - 1, iy, iz)->tmpU + h * SURFDH;
            const double distPy = LAB(&l, ix, iy + 1, iz)->tmpU + h * SURFDH;
            const double distMy = LAB(&l, ix, iy - 1, iz)->tmpU + h * SURFDH;
            const double distPz = LAB(&l, ix, iy, iz + 1)->tmpU + h * SURFDH;
            const double distMz = LAB(&l, ix, iy, iz - 1)->tmpU + h * SURFDH;
            const double gradUX = inv2h * (distPx - distMx);
            const double gradUY = inv2h * (distPy - distMy);
            const double gradUZ = inv2h * (distPz - distMz);
            const double gradUSq = gradUX * gradUX + gradUY * gradUY + gradUZ * gradUZ + EPS;
            const double HplusX = fabs(distPx) < EPS ? 0.5 : (distPx < 0 ? 0 : 1);
            const double HminuX = fabs(distMx) < EPS ? 0.5 : (distMx < 0 ? 0 : 1);
            const double HplusY = fabs(distPy) < EPS ? 0.5 : (distPy < 0 ? 0 : 1);
            const double HminuY = fabs(distMy) < EPS ? 0.5 : (distMy < 0 ? 0 : 1);
            const double HplusZ = fabs(distPz) < EPS ? 0.5 : (distPz < 0 ? 0 : 1);
            const double HminuZ = fabs(distMz) < EPS ? 0.5 : (distMz < 0 ? 0 : 1);
            const double gradHX = HplusX - HminuX, gradHY = HplusY - HminuY, gradHZ = HplusZ - HminuZ;
            const double numD = gradHX * gradUX + gradHY * gradUY + gradHZ * gradUZ;
            const double Delta = fac1 * numD / gradUSq;
            if (Delta > EPS) ob_write(o, ix, iy, iz, Delta, gradUX, gradUY, gradUZ);
          }
    }
    free(l.c);
  }
  double com[4] = {0, 0, 0, 0};
  for (int b = 0; b < s->nb; b++) {
    if (!s->ob[b]) continue;
    com[0] += s->ob[b]->mass;
    com[1] += s->ob[b]->CoM[0];
    com[2] += s->ob[b]->CoM[1];
    com[3] += s->ob[b]->CoM[2];
  }
  s->mass = com[0];
  for (int d = 0; d < 3; d++) s->CoM[d] = com[d + 1] / com[0];
}

/* The active body's udef from three block-major fields (3 x ncell, component-major); cells
 * outside its obstacle blocks are ignored. */
void orc_set_udef(orc_sim* s, const double* udef)
{
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  for (int b = 0; s->ob && b < s->nb; b++) {
    OBlock* o = s->ob[b];
    if (!o) continue;
    for (int q = 0; q < 3; q++) memcpy(o->udef[q], udef + q * nc + (size_t)b * BS * BS * BS, sizeof(o->udef[q]));
  }
}

/* The active body's udef as three block-major fields, 0 outside its obstacle blocks. */
void orc_get_udef(const orc_sim* s, double* udef)
{
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  memset(udef, 0, 3 * nc * sizeof(double));
  for (int b = 0; s->ob && b < s->nb; b++) {
    const OBlock* o = s->ob[b];
    if (!o) continue;
    for (int q = 0; q < 3; q++) memcpy(udef + q * nc + (size_t)b * BS * BS * BS, o->udef[q], sizeof(o->udef[q]));
  }
}

/* The momentum removal of the deformation velocity, for the active body:
 * KernelIntegrateUdefMomenta (ObstaclesCreate.cpp:258-332): per obstacle block, over cells with
 * chi > 0 and p = x - CoM, V = sum chi dv, F = sum chi udef dv, T = sum chi p x (udef - c_old) dv
 * (c_old the previous removal's linear correction) and the inertia J; KernelAccumulate
 * UdefMomenta (:334-384): transVel_correction = F / V, angVel_correction = J^-1 T;
 * KernelRemoveUdefMomenta (:386-440): udef -= transVel_correction + angVel_correction x p on
 * every cell of the obstacle blocks. out: mass, the two corrections, J0..J5. The block sums
 * are added in block order. */
void orc_remove_udef_momenta(orc_sim* s, double out[13])
{
  double M[13] = {0};
  const double dv = s->h * s->h * s->h;
  for (int b = 0; s->ob && b < s->nb; b++) {
    const OBlock* o = s->ob[b];
    if (!o) continue;
    int bc[3];
    bcoords(s, b, bc);
    double m[13] = {0};
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          const double X = o->chi[iz][iy][ix];
          if (X <= 0) continue;
          double p[3];
          cell_pos(s, bc, ix, iy, iz, p);
          for (int d = 0; d < 3; d++) p[d] -= s->CoM[d];
          const double U = o->udef[0][iz][iy][ix], V = o->udef[1][iz][iy][ix], W = o->udef[2][iz][iy][ix];
          const double dU = U - s->corr[0], dV = V - s->corr[1], dW = W - s->corr[2];
          m[0] += X * dv;
          m[1] += X * U * dv;
          m[2] += X * V * dv;
          m[3] += X * W * dv;
          m[4] += X * (p[1] * dW - p[2] * dV) * dv;
          m[5] += X * (p[2] * dU - p[0] * dW) * dv;
          m[6] += X * (p[0] * dV - p[1] * dU) * dv;
          m[7] += X * (p[1] * p[1] + p[2] * p[2]) * dv;
          m[8] += X * (p[0] * p[0] + p[2] * p[2]) * dv;
          m[9] += X * (p[0] * p[0] + p[1] * p[1]) * dv;
          m[10] -= X * p[0] * p[1] * dv;
          m[11] -= X * p[0] * p[2] * dv;
          m[12] -= X * p[1] * p[2] * dv;
        }
    for (int q = 0; q < 13; q++) M[q] += m[q];
  }
  /* inverse of the symmetric J (diagonal J0..J2, off-diagonal J3 = xy, J4 = xz, J5 = yz) by cofactors */
  const double* J = M + 7;
  const double c0 = J[1] * J[2] - J[5] * J[5], c1 = J[0] * J[2] - J[4] * J[4], c2 = J[0] * J[1] - J[3] * J[3];
  const double c3 = J[4] * J[5] - J[3] * J[2], c4 = J[3] * J[5] - J[4] * J[1], c5 = J[3] * J[4] - J[0] * J[5];
  const double det = J[0] * c0 + J[3] * c3 + J[4] * c4;
  /* invertSym (MatArrayMath.h:18-38): a zero inverse when |det| <= DBL_MIN */
  double iJ[6] = {0, 0, 0, 0, 0, 0};
  if (!(fabs(det) <= DBL_MIN)) {
    iJ[0] = c0 / det; iJ[1] = c1 / det; iJ[2] = c2 / det;
    iJ[3] = c3 / det; iJ[4] = c4 / det; iJ[5] = c5 / det;
  }
  const double T[3] = {M[4], M[5], M[6]};
  double ct[3], cw[3];
  for (int d = 0; d < 3; d++) ct[d] = M[1 + d] / M[0];
  cw[0] = iJ[0] * T[0] + iJ[3] * T[1] + iJ[4] * T[2];
  cw[1] = iJ[3] * T[0] + iJ[1] * T[1] + iJ[5] * T[2];
  cw[2] = iJ[4] * T[0] + iJ[5] * T[1] + iJ[2] * T[2];
  for (int b = 0; s->ob && b < s->nb; b++) {
    OBlock* o = s->ob[b];
    if (!o) continue;
    int bc[3];
    bcoords(s, b, bc);
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          double p[3];
          cell_pos(s, bc, ix, iy, iz, p);
          for (int d = 0; d < 3; d++) p[d] -= s->CoM[d];
          const double r[3] = {cw[1] * p[2] - cw[2] * p[1], cw[2] * p[0] - cw[0] * p[2], cw[0] * p[1] - cw[1] * p[0]};
          for (int d = 0; d < 3; d++) o->udef[d][iz][iy][ix] -= ct[d] + r[d];
        }
  }
  for (int d = 0; d < 3; d++) { s->corr[d] = ct[d]; s->corr[3 + d] = cw[d]; }
  out[0] = M[0];
  for (int d = 0; d < 3; d++) { out[1 + d] = ct[d]; out[4 + d] = cw[d]; }
  for (int q = 0; q < 6; q++) out[7 + q] = J[q];
}

void orc_sphere_info(const orc_sim* s, double out[8])
{
  int nbk = 0, npts = 0;
  for (int b = 0; s->ob && b < s->nb; b++)
    if (s->ob[b]) { nbk++; npts += s->ob[b]->npts; }
  out[0] = s->CoM[0]; out[1] = s->CoM[1]; out[2] = s->CoM[2];
  out[3] = s->mass; out[4] = nbk; out[5] = npts; out[6] = s->radius; out[7] = 0;
}

/* KernelPenalization<implicit> + KernelFinalizePenalizationForce
 * (Penalization.cpp:20-124), one body with udef = 0 and lambda_factor = 1. */
void orc_penalize(orc_sim* s, double dt, double lambda, int implicit,
                  const double vel[3], const double omega[3], double ft_out[6])
{
  double M[6] = {0};
  if (!s->has_sphere) { for (int i = 0; i < 6; i++) ft_out[i] = 0; return; }
  const double dv = pow(s->h, 3);
  const double lambdaFac = implicit ? lambda : 1.0 / dt;
  double* part = (double*)calloc(6 * (size_t)s->nb, sizeof(double));  /* per block, summed in block order */
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    const OBlock* o = s->ob[b];
    if (!o) continue;
    int bc[3];
    bcoords(s, b, bc);
    double M[6] = {0};
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          FE* e = &s->blk[b].d[iz][iy][ix];
          const double X = o->chi[iz][iy][ix];
          if (e->chi > X) continue;
          if (X <= 0) continue;
          double p[3];
          cell_pos(s, bc, ix, iy, iz, p);
          p[0] -= s->CoM[0]; p[1] -= s->CoM[1]; p[2] -= s->CoM[2];
          const double U_TOT[3] = {vel[0] + omega[1] * p[2] - omega[2] * p[1] + o->udef[0][iz][iy][ix],
                                   vel[1] + omega[2] * p[0] - omega[0] * p[2] + o->udef[1][iz][iy][ix],
                                   vel[2] + omega[0] * p[1] - omega[1] * p[0] + o->udef[2][iz][iy][ix]};
          const double penalFac = implicit ? X * lambdaFac / (1 + lambdaFac * dt * X) : X * lambdaFac;
          const double FPX = penalFac * (U_TOT[0] - e->u);
          const double FPY = penalFac * (U_TOT[1] - e->v);
          const double FPZ = penalFac * (U_TOT[2] - e->w);
          e->u = e->u + dt * FPX;
          e->v = e->v + dt * FPY;
          e->w = e->w + dt * FPZ;
          M[0] += dv * FPX; M[1] += dv * FPY; M[2] += dv * FPZ;
          M[3] += dv * (p[1] * FPZ - p[2] * FPY);
          M[4] += dv * (p[2] * FPX - p[0] * FPZ);
          M[5] += dv * (p[0] * FPY - p[1] * FPX);
        }
    for (int i = 0; i < 6; i++) part[6 * (size_t)b + i] = M[i];
  }
  for (int b = 0; b < s->nb; b++)
    for (int i = 0; i < 6; i++) M[i] += part[6 * (size_t)b + i];
  free(part);
  for (int i = 0; i < 6; i++) ft_out[i] = M[i];
}

/* KernelIntegrateFluidMomenta (ObstaclesUpdate.cpp:15-131) for the sphere (udef = 0), each
 * obstacle block summed over its cells in (iz, iy, ix) order and the blocks in block order
 * (KernelFinalizeObstacleVel, :133-170). out[0] V = sum chi dv; [1..3] sum chi u dv;
 * [4..6] sum chi (p x u) dv; [7..12] J0..J5; with implicit penalization also [13] GfX,
 * [14..16] Gp, [17..22] Gj, [23..25] Gu, [26..28] Ga weighted by
 * penalFac = dv lambda dt chi / (1 + chi lambda dt), else zero. p = x - CoM; the obstacle's
 * chi, cells with chi <= 0 skipped (:82). */
void orc_obstacle_momenta(const orc_sim* s, double dt, double lambda, int implicit, double out[29])
{
  for (int i = 0; i < 29; i++) out[i] = 0;
  if (!s->has_sphere) return;
  const double dv = pow(s->h, 3), lambdt = lambda * dt;
  double* part = (double*)calloc(29 * (size_t)s->nb, sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    const OBlock* o = s->ob[b];
    if (!o) continue;
    int bc[3];
    bcoords(s, b, bc);
    double* M = part + 29 * (size_t)b;
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          const double X = o->chi[iz][iy][ix];
          if (X <= 0) continue;
          const FE* e = &s->blk[b].d[iz][iy][ix];
          double p[3];
          cell_pos(s, bc, ix, iy, iz, p);
          p[0] -= s->CoM[0]; p[1] -= s->CoM[1]; p[2] -= s->CoM[2];
          const double Xd = X * dv;
          M[0] += Xd;
          M[1] += Xd * e->u; M[2] += Xd * e->v; M[3] += Xd * e->w;
          M[4] += Xd * (p[1] * e->w - p[2] * e->v);
          M[5] += Xd * (p[2] * e->u - p[0] * e->w);
          M[6] += Xd * (p[0] * e->v - p[1] * e->u);
          M[7] += Xd * (p[1] * p[1] + p[2] * p[2]);
          M[8] += Xd * (p[0] * p[0] + p[2] * p[2]);
          M[9] += Xd * (p[0] * p[0] + p[1] * p[1]);
          M[10] -= Xd * p[0] * p[1];
          M[11] -= Xd * p[0] * p[2];
          M[12] -= Xd * p[1] * p[2];
          if (implicit) {
            const double pf = dv * lambdt * X / (1 + X * lambdt);
            M[13] += pf;
            M[14] += pf * p[0]; M[15] += pf * p[1]; M[16] += pf * p[2];
            M[17] += pf * (p[1] * p[1] + p[2] * p[2]);
            M[18] += pf * (p[0] * p[0] + p[2] * p[2]);
            M[19] += pf * (p[0] * p[0] + p[1] * p[1]);
            M[20] -= pf * p[0] * p[1];
            M[21] -= pf * p[0] * p[2];
            M[22] -= pf * p[1] * p[2];
            const double DU = e->u - o->udef[0][iz][iy][ix], DV = e->v - o->udef[1][iz][iy][ix],
                         DW = e->w - o->udef[2][iz][iy][ix];  /* ObstaclesUpdate.cpp:117-126 */
            M[23] += pf * DU; M[24] += pf * DV; M[25] += pf * DW;
            M[26] += pf * (p[1] * DW - p[2] * DV);
            M[27] += pf * (p[2] * DU - p[0] * DW);
            M[28] += pf * (p[0] * DV - p[1] * DU);
          }
        }
  }
  for (int b = 0; b < s->nb; b++)
    for (int i = 0; i < 29; i++) out[i] += part[29 * (size_t)b + i];
  free(part);
}

/* KernelComputeForces (FluidSolidForces.cpp:17-158) summed as in Obstacle::computeForces
 * (Obstacle.cpp:266-286); out[] in ObstacleBlock::sumQoI order (ObstacleBlock.h:81-93). */
void orc_surface_forces(const orc_sim* s, double nu, const double uTrans[3], const double omega[3],
                        double out[22])
{
  double Q[22] = {0};
  if (!s->has_sphere) { for (int i = 0; i < 22; i++) out[i] = 0; return; }
  double velUnit[3] = {0, 0, 0};
  const double vel_norm = sqrt(uTrans[0] * uTrans[0] + uTrans[1] * uTrans[1] + uTrans[2] * uTrans[2]);
  if (vel_norm > 1e-9)
    for (int d = 0; d < 3; d++) velUnit[d] = uTrans[d] / vel_norm;
  const double _1oH = nu / s->h;
  const double* CM = s->CoM;
  double* part = (double*)calloc(22 * (size_t)s->nb, sizeof(double));  /* per block, summed in block order */
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      const OBlock* o = s->ob[b];
      if (!o || o->npts == 0) continue;
      lab_load(s, b, &l);
      int bc[3];
      bcoords(s, b, bc);
      double Q[22] = {0};
      for (int i = 0; i < o->npts; i++) {
        const int ix = o->pix[i] % BS, iy = (o->pix[i] / BS) % BS, iz = o->pix[i] / (BS * BS);
        double p[3];
        cell_pos(s, bc, ix, iy, iz, p);
        const double D11 = _1oH * (LAB(&l, ix + 1, iy, iz)->u - LAB(&l, ix - 1, iy, iz)->u);
        const double D22 = _1oH * (LAB(&l, ix, iy + 1, iz)->v - LAB(&l, ix, iy - 1, iz)->v);
        const double D33 = _1oH * (LAB(&l, ix, iy, iz + 1)->w - LAB(&l, ix, iy, iz - 1)->w);
        const double half = g_plant == ORC_PLANT_FORCES_SHEAR_HALF ? 1.0 : .5;
        const double D12 = half * _1oH * (LAB(&l, ix, iy + 1, iz)->u - LAB(&l, ix, iy - 1, iz)->u +
                                          LAB(&l, ix + 1, iy, iz)->v - LAB(&l, ix - 1, iy, iz)->v);
        const double D13 = half * _1oH * (LAB(&l, ix, iy, iz + 1)->u - LAB(&l, ix, iy, iz - 1)->u +
                                          LAB(&l, ix + 1, iy, iz)->w - LAB(&l, ix - 1, iy, iz)->w);
        const double D23 = half * _1oH * (LAB(&l, ix, iy + 1, iz)->w - LAB(&l, ix, iy - 1, iz)->w +
                                          LAB(&l, ix, iy, iz + 1)->v - LAB(&l, ix, iy, iz - 1)->v);
        const FE* C = LAB(&l, ix, iy, iz);
        const double P = C->p;
        const double normX = o->dchidx[i], normY = o->dchidy[i], normZ = o->dchidz[i];
        const double fXV = D11 * normX + D12 * normY + D13 * normZ;
        const double fYV = D12 * normX + D22 * normY + D23 * normZ;
        const double fZV = D13 * normX + D23 * normY + D33 * normZ;
        const double fXP = -P * normX, fYP = -P * normY, fZP = -P * normZ;
        const double fXT = fXV + fXP, fYT = fYV + fYP, fZT = fZV + fZP;
        const double vX = C->u, vY = C->v, vZ = C->w;
        Q[0] += fXT; Q[1] += fYT; Q[2] += fZT;
        Q[3] += fXP; Q[4] += fYP; Q[5] += fZP;
        Q[6] += fXV; Q[7] += fYV; Q[8] += fZV;
        Q[9] += (p[1] - CM[1]) * fZT - (p[2] - CM[2]) * fYT;
        Q[10] += (p[2] - CM[2]) * fXT - (p[0] - CM[0]) * fZT;
        Q[11] += (p[0] - CM[0]) * fYT - (p[1] - CM[1]) * fXT;
        Q[12] += normY * vZ - normZ * vY;
        Q[13] += normZ * vX - normX * vZ;
        Q[14] += normX * vY - normY * vX;
        const double forcePar = fXT * velUnit[0] + fYT * velUnit[1] + fZT * velUnit[2];
        const int swap = g_plant == ORC_PLANT_FORCES_THRUST_DRAG;
        Q[swap ? 15 : 16] += .5 * (forcePar + fabs(forcePar));
        Q[swap ? 16 : 15] -= .5 * (forcePar - fabs(forcePar));
        const double powOut = fXT * vX + fYT * vY + fZT * vZ;
        const double powDef = fXT * o->udef[0][iz][iy][ix] + fYT * o->udef[1][iz][iy][ix] +
                              fZT * o->udef[2][iz][iy][ix]; /* FluidSolidForces.cpp:110-112,145 */
        Q[17] += powOut;
        Q[18] += powOut < 0 ? powOut : 0;
        Q[19] += powDef;
        Q[20] += powDef < 0 ? powDef : 0;
        const double rV[3] = {p[0] - CM[0], p[1] - CM[1], p[2] - CM[2]};
        const double uS[3] = {uTrans[0] + omega[1] * rV[2] - rV[1] * omega[2],
                              uTrans[1] + omega[2] * rV[0] - rV[2] * omega[0],
                              uTrans[2] + omega[0] * rV[1] - rV[0] * omega[1]};
        Q[21] += fXT * uS[0] + fYT * uS[1] + fZT * uS[2];
      }
      for (int q = 0; q < 22; q++) part[22 * (size_t)b + q] = Q[q];
    }
    free(l.c);
  }
  for (int b = 0; b < s->nb; b++)
    for (int q = 0; q < 22; q++) Q[q] += part[22 * (size_t)b + q];
  free(part);
  for (int i = 0; i < 22; i++) out[i] = Q[i];
}

/* ------------------------------------------------------------------ adjoint */

/* Adjoint (vector-Jacobian product) of one step u_{n+1} = P(A(u_n)) without forcing or
 * body, central advection with Euler (AdvectionDiffusion.cpp:106-145, then PressureRHS,
 * PressureProjection). No reference counterpart. The forward state u_n is the sim's
 * (u,v,w); lam (3 x ncell, block-major) holds dJ/du_{n+1} on entry and dJ/du_n on exit.
 *  - P = I - G S D with central D and G = -D^T and a real even symbol S: P^T = P, so
 *    lam* = P lam is the forward projection applied to lam (dt cancels).
 *  - A(u)_q = u_q + facA sum_d a_d (u_q(+e_d) - u_q(-e_d)) + facD lap u_q, a = u + U:
 *    A'(u)^T lam*_r = lam*_r + facD lap lam*_r - facA sum_d [(a_d lam*_r)(+e_d) - (a_d lam*_r)(-e_d)]
 *                     + facA sum_q lam*_q (u_q(+e_r) - u_q(-e_r)). */
/* lam <- A'(u_n)^T lam with u_n the sim's velocity (overwrites tmp). */
static void adj_advect(orc_sim* s, double dt, double nu, const double uinf[3], double* lam)
{
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  /* lam* into tmp (so that a lab carries both u and lam*), then the transposed operator */
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    FE* e = &s->blk[b].d[0][0][0];
    for (int i = 0; i < BS * BS * BS; i++) {
      const size_t c = (size_t)b * BS * BS * BS + i;
      e[i].tmpU = lam[c];
      e[i].tmpV = lam[nc + c];
      e[i].tmpW = lam[2 * nc + c];
    }
  }
  const double h = s->h;
  const double facA = -dt / (2 * h);
  const double facD = (nu / h) * (dt / h);
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      lab_load(s, b, &l);
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const FE* C = LAB(&l, ix, iy, iz);
            const FE* nbr[3][2] = {{LAB(&l, ix - 1, iy, iz), LAB(&l, ix + 1, iy, iz)},
                                   {LAB(&l, ix, iy - 1, iz), LAB(&l, ix, iy + 1, iz)},
                                   {LAB(&l, ix, iy, iz - 1), LAB(&l, ix, iy, iz + 1)}};
            const size_t cell = (size_t)b * BS * BS * BS + (iz * BS + iy) * BS + ix;
            for (int r = 0; r < 3; r++) {
              const double lc = fe_comp(C, r, 1);
              double lap = 0, adv = 0, src = 0;
              for (int d = 0; d < 3; d++) {
                const FE *m = nbr[d][0], *p = nbr[d][1];
                lap += fe_comp(m, r, 1) + fe_comp(p, r, 1) - 2 * lc;
                adv += (fe_comp(p, d, 0) + uinf[d]) * fe_comp(p, r, 1) - (fe_comp(m, d, 0) + uinf[d]) * fe_comp(m, r, 1);
              }
              for (int q = 0; q < 3; q++) src += fe_comp(C, q, 1) * (fe_comp(nbr[r][1], q, 0) - fe_comp(nbr[r][0], q, 0));
              lam[r * nc + cell] = lc + facD * lap - facA * adv + facA * src;
            }
          }
    }
    free(l.c);
  }
}

/* lam <- P lam: the forward projection (without the sphere) on the adjoint, P symmetric. */
static void adj_project(orc_sim* s, double dt, double* lam)
{
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  double* save = (double*)malloc(4 * nc * sizeof(double));
  const int keep[4] = {ORC_U, ORC_V, ORC_W, ORC_P};
  for (int q = 0; q < 4; q++) orc_get_field(s, keep[q], save + q * nc);
  for (int q = 0; q < 3; q++) orc_set_field(s, ORC_U + q, lam + q * nc);
  const int sphere = s->has_sphere;
  s->has_sphere = 0;
  orc_pressure_rhs(s, dt);
  orc_poisson_solve(s);
  orc_grad_p(s, dt);
  s->has_sphere = sphere;
  for (int q = 0; q < 3; q++) orc_get_field(s, ORC_U + q, lam + q * nc);
  for (int q = 0; q < 4; q++) orc_set_field(s, keep[q], save + q * nc);
  free(save);
}

/* Block-major index of the periodic cell (gx, gy, gz). */
static inline size_t bm_index(const orc_sim* s, int gx, int gy, int gz)
{
  gx = (gx % s->N[0] + s->N[0]) % s->N[0];
  gy = (gy % s->N[1] + s->N[1]) % s->N[1];
  gz = (gz % s->N[2] + s->N[2]) % s->N[2];
  const int b = gx / BS + s->bpd[0] * (gy / BS + s->bpd[1] * (gz / BS));
  return (size_t)b * BS * BS * BS + ((gz % BS) * BS + gy % BS) * BS + gx % BS;
}

/* Transposed linearisation of the upwind operator L (upwind3_update) at phi, applied to lam
 * (no reference counterpart): with a = phi + U and S^{+/-} f = sum_k c^{+/-}_k f(+k) along an
 * axis (c^+ = (1, -6, 3, 2, 0), c^- = (0, -2, -3, 6, -1) for k = -2..2),
 *   out_p(y) = -(1/6h) [ sum_q lam_q(y) S^{sign a_p(y)}_p phi_q (y)
 *                        + sum_d sum_k lam_p(y - k e_d) a_d(y - k e_d) c^{sign a_d(y - k e_d)}_k ]
 *              + nu/h^2 lap lam_p(y).
 * The first sum differentiates the advecting velocity, the second the advected field; the
 * upwind side is held where it is (a.e. differentiable). Fields block-major, 3 x ncell. */
static void up3_jt(const orc_sim* s, const double* phi, const double* lam, double nu, const double uinf[3],
                   double* out)
{
  static const double cp[5] = {1, -6, 3, 2, 0}, cm[5] = {0, -2, -3, 6, -1};
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  const double h = s->h;
  #pragma omp parallel for collapse(2) schedule(static)
  for (int gz = 0; gz < s->N[2]; gz++)
    for (int gy = 0; gy < s->N[1]; gy++)
      for (int gx = 0; gx < s->N[0]; gx++) {
        const int g0[3] = {gx, gy, gz};
        const size_t y = bm_index(s, gx, gy, gz);
        for (int p = 0; p < 3; p++) {
          double t1 = 0, t2 = 0, lap = 0;
          /* sum_q lam_q(y) S_p phi_q (y), upwind side of a_p(y) */
          const double ap = phi[p * nc + y] + uinf[p];
          const double* c = ap > 0 ? cp : cm;
          for (int q = 0; q < 3; q++) {
            double S = 0;
            for (int k = -2; k <= 2; k++) {
              int o[3] = {gx, gy, gz};
              o[p] += k;
              S += c[k + 2] * phi[q * nc + bm_index(s, o[0], o[1], o[2])];
            }
            t1 += lam[q * nc + y] * S;
          }
          for (int d = 0; d < 3; d++) {
            for (int k = -2; k <= 2; k++) {
              int o[3] = {gx, gy, gz};
              o[d] -= k;
              const size_t x = bm_index(s, o[0], o[1], o[2]);
              const double a = phi[d * nc + x] + uinf[d];
              t2 += lam[p * nc + x] * a * (a > 0 ? cp[k + 2] : cm[k + 2]);
            }
            int e[3] = {gx, gy, gz}, f[3] = {gx, gy, gz};
            e[d] += 1;
            f[d] -= 1;
            lap += lam[p * nc + bm_index(s, e[0], e[1], e[2])] + lam[p * nc + bm_index(s, f[0], f[1], f[2])] -
                   2 * lam[p * nc + y];
          }
          (void)g0;
          out[p * nc + y] = -(t1 + t2) / (6 * h) + nu * lap / (h * h);
        }
      }
}

/* Adjoint of one step of the upwind scheme (no reference counterpart): lam holds
 * dJ/du_{n+1} on entry, dJ/du_n on exit, with u_n the sim's velocity. lam* = P lam, then
 * Euler:    lam_n = lam* + dt L'(u_n)^T lam*;
 * midpoint: u_h = u_n + dt/2 L(u_n), mu = dt L'(u_h)^T lam*, lam_n = lam* + mu + dt/2 L'(u_n)^T mu.
 * scheme 0 is the central Euler step (orc_adjoint_step). */
static void adj_advection(orc_sim* s, double dt, double nu, const double uinf[3], int scheme, double* lam);
static void adj_forcing(orc_sim* s, double dt, double nu, const double uinf[3], int scheme, double tke_target,
                        double injection_rate, double* lam);

void orc_adjoint_step_scheme(orc_sim* s, double dt, double nu, const double uinf[3], int scheme, double* lam)
{
  if (scheme == 0) {
    orc_adjoint_step(s, dt, nu, uinf, lam);
    return;
  }
  adj_project(s, dt, lam);
  adj_advection(s, dt, nu, uinf, scheme, lam);
}

/* out = P in for three block-major components: the band filter of SpectralForcing, keeping
 * the modes 0 < |k|^2 <= 4 in physical wavenumbers (the forward's filter), an orthogonal
 * projector on the grid. */
static void low_project(const orc_sim* s, const double* in, double* out)
{
  const int N0 = s->N[0], N1 = s->N[1], N2 = s->N[2];
  const size_t n = (size_t)N0 * N1 * N2, nc = (size_t)s->nb * BS * BS * BS;
  const double wf[3] = {2.0 * M_PI / s->extent[0], 2.0 * M_PI / s->extent[1], 2.0 * M_PI / s->extent[2]};
  double complex* g = (double complex*)malloc(sizeof(double complex) * n);
  for (int q = 0; q < 3; q++) {
    for (int gz = 0; gz < N2; gz++)
      for (int gy = 0; gy < N1; gy++)
        for (int gx = 0; gx < N0; gx++) g[ridx(s, gx, gy, gz)] = in[q * nc + bm_index(s, gx, gy, gz)];
    fft3d_c(g, N0, N1, N2, -1);
    for (int i = 0; i < N0; i++)
      for (int l = 0; l < N1; l++)
        for (int k = 0; k < N2; k++) {
          const double kx = signed_k(i, N0) * wf[0], ky = signed_k(l, N1) * wf[1], kz = signed_k(k, N2) * wf[2];
          const double k2 = kx * kx + ky * ky + kz * kz;
          if (!(k2 > 0 && k2 <= 4)) g[((size_t)i * N1 + l) * N2 + k] = 0;
        }
    fft3d_c(g, N0, N1, N2, +1);
    for (int gz = 0; gz < N2; gz++)
      for (int gy = 0; gy < N1; gy++)
        for (int gx = 0; gx < N0; gx++) out[q * nc + bm_index(s, gx, gy, gz)] = creal(g[ridx(s, gx, gy, gz)]) / n;
  }
  free(g);
}

/* Adjoint of one step with the linear forcing (no reference counterpart). The forward step
 * is u* = A(u_n) (the scheme's advection), u** = u* + fac(u*) P u* when fac > 0
 * (SpectralForcing.cpp:37-77: P the band filter, tke = <|u*|^2>/2, tke_F = <u*, P u*>/2
 * (Parseval), fac = (tke_target - tke)/(2 tke_F), or dt r/(2 tke_F) for the constant
 * injection rate r), then the projection. With lam* = P_proj lam and s = <P u*, lam*> summed
 * over cells and components (n cells):
 *   lam** = lam* + fac P lam* + s grad fac,
 *   grad fac = -u* / (2 n tke_F) [target only] - fac P u* / (n tke_F),
 * then lam_n = A'(u_n)^T lam**. The switch fac > 0 is held at the forward state. */
void orc_adjoint_step_forcing(orc_sim* s, double dt, double nu, const double uinf[3], int scheme,
                              double tke_target, double injection_rate, double* lam)
{
  if (!(tke_target > 0) && !(injection_rate > 0)) {
    orc_adjoint_step_scheme(s, dt, nu, uinf, scheme, lam);
    return;
  }
  adj_project(s, dt, lam);
  adj_forcing(s, dt, nu, uinf, scheme, tke_target, injection_rate, lam);
  adj_advection(s, dt, nu, uinf, scheme, lam);
}

/* lam <- F'(u*)^T lam for the linear forcing at u* = A(u_n) (the scheme's advection of the
 * sim's velocity u_n, which is kept): lam += fac P lam + <P u*, lam> grad fac. */
static void adj_forcing(orc_sim* s, double dt, double nu, const double uinf[3], int scheme, double tke_target,
                        double injection_rate, double* lam)
{
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  const double n = (double)s->N[0] * s->N[1] * s->N[2];
  double* un = (double*)malloc(3 * nc * sizeof(double));
  double* us = (double*)malloc(3 * nc * sizeof(double));
  double* Pu = (double*)malloc(3 * nc * sizeof(double));
  double* Pl = (double*)malloc(3 * nc * sizeof(double));
  for (int q = 0; q < 3; q++) orc_get_field(s, ORC_U + q, un + q * nc);
  orc_advect_diffuse_scheme(s, dt, nu, uinf, scheme);
  for (int q = 0; q < 3; q++) orc_get_field(s, ORC_U + q, us + q * nc);
  for (int q = 0; q < 3; q++) orc_set_field(s, ORC_U + q, un + q * nc);
  low_project(s, us, Pu);
  low_project(s, lam, Pl);
  double tke = 0, tkeF = 0, sl = 0;
  for (size_t i = 0; i < 3 * nc; i++) {
    tke += us[i] * us[i];
    tkeF += us[i] * Pu[i];
    sl += Pu[i] * lam[i];
  }
  tke *= 0.5 / n;
  tkeF *= 0.5 / n;
  const double inj = tke_target > 0 ? (tke_target - tke) / dt : injection_rate;
  const double fac = dt * inj / (2 * tkeF);
  if (fac > 0) {
    const double a_u = tke_target > 0 ? -sl / (2 * n * tkeF) : 0.0;
    const double a_p = -sl * fac / (n * tkeF);
    for (size_t i = 0; i < 3 * nc; i++) lam[i] += fac * Pl[i] + a_u * us[i] + a_p * Pu[i];
  }
  free(un);
  free(us);
  free(Pu);
  free(Pl);
}

/* lam <- A'(u_n)^T lam for the scheme's advection, u_n the sim's velocity. */
static void adj_advection(orc_sim* s, double dt, double nu, const double uinf[3], int scheme, double* lam)
{
  if (scheme == 0) {
    adj_advect(s, dt, nu, uinf, lam);
    return;
  }
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  double* phi = (double*)malloc(3 * nc * sizeof(double));
  double* t = (double*)malloc(3 * nc * sizeof(double));
  for (int q = 0; q < 3; q++) orc_get_field(s, ORC_U + q, phi + q * nc);
  if (scheme == 1) {
    up3_jt(s, phi, lam, nu, uinf, t);
    for (size_t i = 0; i < 3 * nc; i++) lam[i] += dt * t[i];
  } else {
    double* uh = (double*)malloc(3 * nc * sizeof(double));
    double* mu = (double*)malloc(3 * nc * sizeof(double));
    upwind3_update(s, 0, 0.5 * dt, nu, uinf, uh);
    up3_jt(s, uh, lam, nu, uinf, mu);
    for (size_t i = 0; i < 3 * nc; i++) mu[i] *= dt;
    up3_jt(s, phi, mu, nu, uinf, t);
    for (size_t i = 0; i < 3 * nc; i++) lam[i] += mu[i] + 0.5 * dt * t[i];
    free(uh);
    free(mu);
  }
  free(phi);
  free(t);
}

void orc_adjoint_step(orc_sim* s, double dt, double nu, const double uinf[3], double* lam)
{
  adj_project(s, dt, lam);  /* lam* */
  adj_advect(s, dt, nu, uinf, lam);
}

/* Adjoint of one step with the sphere (forced, velocity V, no rotation), central advection
 * with Euler, no forcing. No reference counterpart. The step is
 *   u* = A(u_n); r = M(R u*) (PressureRHS.cpp:78-112 masking and :322-343 perimeter
 *   correction); u' = u* - (dt/2h) grad_c L^{-1} r; u_{n+1} = u' + dt pf (V - u')
 *   (Penalization.cpp:40-98), with the step's drag F_x = sum h^3 pf (V_x - u'_x) (:93-96).
 * With J = sum_n w_n F_x^n + g(u_N): lam holds dJ/du_{n+1} on entry and dJ/du_n on exit, and
 * dJdV accumulates dJ/dV of this step (wdrag = w_n). The masking's sign pattern (which
 * surface cells the correction scales) is held at the forward state; the transposed
 * projection is the forward one with M'^T applied to the pressure instead of M to the RHS. */
void orc_adjoint_step_body(orc_sim* s, double dt, double nu, const double uinf[3], double lambda, int implicit,
                           const double V[3], double wdrag, double* lam, double dJdV[3], double dJdc[3])
{
  const double W[3] = {0, 0, 0};
  orc_adjoint_step_body_full(s, dt, nu, uinf, 0, 0.0, 0.0, lambda, implicit, V, W, wdrag, lam, dJdV, NULL, dJdc);
}

/* The active body's masking linearised at u* (the sim's velocity): per cell of its blocks the
 * flags (bit 0 masked, bit 1 surface), rhs_new and, for the centre gradient, r0; its totals S,
 * D and the correction, as rhs_obstacle forms them. */
typedef struct {
  double S, D, denom, corr, sg;
} MaskTot;

static MaskTot mask_linearise(orc_sim* s, double dt, double* rnew, double* r0c, unsigned char* flag)
{
  const double h = s->h, fac = 0.5 * h * h / dt;
  double* part = (double*)calloc(3 * (size_t)s->nb, sizeof(double));
  #pragma omp parallel
  {
    Lab l = lab_alloc(1);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; b++) {
      const OBlock* o = s->ob[b];
      if (!o) continue;
      lab_load(s, b, &l);
      double S = 0, pos = 0, neg = 0;
      for (int iz = 0; iz < BS; iz++)
        for (int iy = 0; iy < BS; iy++)
          for (int ix = 0; ix < BS; ix++) {
            const size_t c = (size_t)b * BS * BS * BS + (iz * BS + iy) * BS + ix;
            const FE *LW = LAB(&l, ix - 1, iy, iz), *LE = LAB(&l, ix + 1, iy, iz);
            const FE *LS = LAB(&l, ix, iy - 1, iz), *LN = LAB(&l, ix, iy + 1, iz);
            const FE *LF = LAB(&l, ix, iy, iz - 1), *LB = LAB(&l, ix, iy, iz + 1);
            if (LAB(&l, ix, iy, iz)->chi > o->chi[iz][iy][ix]) continue;
            const double r0 = fac * (LE->u - LW->u + LN->v - LS->v + LB->w - LF->w);
            const double dXx = LE->chi - LW->chi, dXy = LN->chi - LS->chi, dXz = LB->chi - LF->chi;
            const int surf = dXx * dXx + dXy * dXy + dXz * dXz > 0;
            const double srcBulk = -o->chi[iz][iy][ix] * r0;
            S += srcBulk;
            const double r1 = r0 + srcBulk;
            rnew[c] = r1;
            if (r0c) r0c[c] = r0;
            flag[c] = 1 | (surf ? 2 : 0);
            pos += surf * r1 * (r1 > 0);
            neg -= surf * r1 * (r1 < 0);
          }
      part[3 * b] = S;
      part[3 * b + 1] = pos;
      part[3 * b + 2] = neg;
    }
    free(l.c);
  }
  double S = 0, pos = 0, neg = 0;
  for (int b = 0; b < s->nb; b++) {
    S += part[3 * b];
    pos += part[3 * b + 1];
    neg += part[3 * b + 2];
  }
  free(part);
  MaskTot m;
  m.S = S;
  m.denom = S > 0 ? pos : neg;
  m.D = m.denom > EPS ? m.denom : EPS;
  m.corr = S / m.D;
  m.sg = m.corr > 0 ? 1.0 : (m.corr < 0 ? -1.0 : 0.0);
  return m;
}

/* The active body's penalization and drag transposed: lam' = (1 - dt pf) lam - wdrag h^3 pf e_x
 * on its penalized cells. The target velocity V + W x r (r = x - CoM) enters as
 * g . (V + W x r) with g_d = dt pf lam_d + wdrag h^3 pf [d = x]: acc[0..2] += dJ/dV = sum g,
 * acc[6..8] += dJ/dW = sum r x g. With uin (the penalization's input P_M(u*)), acc[3..5] +=
 * the chi part of dJ/d(centre): dpf/dchi (dt sum_d lam_d (V_d - u_in,d) + wdrag h^3 (V_x -
 * u_in,x)) dchi/dc. */
static void pen_transpose(orc_sim* s, double dt, double lambda, int implicit, const double V[3], double wdrag,
                          const double* uin, double* lam, double acc[9])
{
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  const double h3 = s->h * s->h * s->h;
  const double lf = implicit ? lambda : 1.0 / dt;
  double* pv = (double*)calloc(9 * (size_t)s->nb, sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    const OBlock* o = s->ob[b];
    if (!o) continue;
    int bc[3];
    bcoords(s, b, bc);
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int iz = 0; iz < BS; iz++)
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          const double X = o->chi[iz][iy][ix];
          if (s->blk[b].d[iz][iy][ix].chi > X || X <= 0) continue;
          const double pf = implicit ? X * lf / (1 + lf * dt * X) : X * lf;
          const size_t c = (size_t)b * BS * BS * BS + (iz * BS + iy) * BS + ix;
          if (uin) {
            const double den = 1 + lf * dt * X;
            const double dpf = implicit ? lf / (den * den) : lf;
            double gx = wdrag * h3 * (V[0] - uin[c]);
            for (int d = 0; d < 3; d++) gx += dt * lam[d * nc + c] * (V[d] - uin[d * nc + c]);
            double p[3], dX[3];
            cell_pos(s, bc, ix, iy, iz, p);
            moll_dchi_dc(s, p, dX);
            for (int k = 0; k < 3; k++) a[3 + k] += dpf * gx * dX[k];
          }
          double gv[3];
          for (int d = 0; d < 3; d++) {
            gv[d] = dt * pf * lam[d * nc + c] + (d == 0 ? wdrag * h3 * pf : 0.0);
            a[d] += gv[d];
            lam[d * nc + c] = (1 - dt * pf) * lam[d * nc + c] - (d == 0 ? wdrag * h3 * pf : 0.0);
          }
          double r[3];
          cell_pos(s, bc, ix, iy, iz, r);
          for (int d = 0; d < 3; d++) r[d] -= s->CoM[d];
          a[6] += r[1] * gv[2] - r[2] * gv[1];
          a[7] += r[2] * gv[0] - r[0] * gv[2];
          a[8] += r[0] * gv[1] - r[1] * gv[0];
        }
    for (int d = 0; d < 9; d++) pv[9 * b + d] = a[d];
  }
  for (int b = 0; b < s->nb; b++)
    for (int d = 0; d < 9; d++) acc[d] += pv[9 * b + d];
  free(pv);
}

/* y <- M'^T y on the active body's blocks (the masking's transpose at its linearisation m):
 * with g = sum over its corrected cells Q of y (-sg rhs_new), a_S = g/D and a_D = -g corr/D,
 * adj(rhs_new) = y (1 - sg corr) + a_D sg on Q and y elsewhere, and the unmasked RHS's adjoint is
 * (1 - X) adj(rhs_new) - X a_S. With dc (the centre gradient): rhs_new = (1 - X) r0 and S =
 * -sum X r0 give dJ/dX = -r0 (adj(rhs_new) + a_S), scaled by yscale and contracted with dX/dc. */
static void mask_transpose(orc_sim* s, const MaskTot* m, const double* rnew, const double* r0c,
                           const unsigned char* flag, double yscale, double* y, double dc[3])
{
  const int n3 = BS * BS * BS;
  double g = 0;
  for (int b = 0; b < s->nb; b++) {
    if (!s->ob[b]) continue;
    for (int i = 0; i < n3; i++) {
      const size_t c = (size_t)b * n3 + i;
      if ((flag[c] & 2) && ((rnew[c] > 0 && m->corr > 0) || (rnew[c] < 0 && m->corr < 0)))
        g += y[c] * (-m->sg * rnew[c]);
    }
  }
  const double aS = g / m->D, aD = m->denom > EPS ? -g * m->corr / m->D : 0.0;
  double* pm = (double*)calloc(3 * (size_t)s->nb, sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; b++) {
    const OBlock* o = s->ob[b];
    if (!o) continue;
    int bc[3];
    bcoords(s, b, bc);
    double a[3] = {0, 0, 0};
    for (int i = 0; i < n3; i++) {
      const size_t c = (size_t)b * n3 + i;
      if (!(flag[c] & 1)) continue;
      const double X = (&o->chi[0][0][0])[i];
      const int inQ = (flag[c] & 2) && ((rnew[c] > 0 && m->corr > 0) || (rnew[c] < 0 && m->corr < 0));
      const double an = inQ ? y[c] * (1 - m->sg * m->corr) + aD * m->sg : y[c];
      if (dc) {
        double p[3], dX[3];
        cell_pos(s, bc, i % BS, (i / BS) % BS, i / (BS * BS), p);
        moll_dchi_dc(s, p, dX);
        for (int k = 0; k < 3; k++) a[k] += -yscale * r0c[c] * (an + aS) * dX[k];
      }
      y[c] = (1 - X) * an - X * aS;
    }
    for (int k = 0; k < 3; k++) pm[3 * b + k] = a[k];
  }
  if (dc)
    for (int b = 0; b < s->nb; b++)
      for (int k = 0; k < 3; k++) dc[k] += pm[3 * b + k];
  free(pm);
}

/* The adjoint step with nbody bodies: nbody == 1 takes the active body, nbody > 1 bodies
 * 0..nbody-1 (the sim's bodies, whose obstacle blocks are apart, so each body's transposes act
 * on its own blocks). Body k has velocity V[3k..], rotation W[3k..] and drag weight wdrag[k];
 * dJdV[3k..] and dJdW[3k..] (if non-NULL) accumulate. dJdc: the active body's centre (nbody ==
 * 1, mollified chi, no rotation), else NaN. */
static void adjoint_bodies(orc_sim* s, double dt, double nu, const double uinf[3], int scheme, double tke_target,
                           double injection_rate, double lambda, int implicit, int nbody, const double* V,
                           const double* W, const double* wdrag, double* lam, double* dJdV, double* dJdW,
                           double dJdc[3])
{
  const int forced = tke_target > 0 || injection_rate > 0;
  const int rot = W[0] != 0 || W[1] != 0 || W[2] != 0;
  /* Towers' chi, or a rotating body (its lever arms move with the centre): no centre gradient */
  const int want_c = dJdc != NULL && nbody == 1 && s->moll > 0 && !rot;
  if (dJdc && !want_c) dJdc[0] = dJdc[1] = dJdc[2] = NAN;
  const size_t nc = (size_t)s->nb * BS * BS * BS;
  const double h = s->h;
#define SEL(k) do { if (nbody > 1) orc_select_body(s, (k)); } while (0)
  /* forward state u_n kept aside; u* = F(A(u_n)), the scheme's advection and the forcing, is
   * the masking's linearisation point */
  double* save = (double*)malloc(4 * nc * sizeof(double));
  const int keep[4] = {ORC_U, ORC_V, ORC_W, ORC_P};
  for (int q = 0; q < 4; q++) orc_get_field(s, keep[q], save + q * nc);
  orc_advect_diffuse_scheme(s, dt, nu, uinf, scheme);
  if (forced) {
    double st[5];
    orc_spectral_forcing(s, dt, dt, nu, tke_target, injection_rate, st);
  }
  double* rnew = (double*)calloc(nc, sizeof(double));
  double* r0c = want_c ? (double*)calloc(nc, sizeof(double)) : NULL;
  unsigned char* flag = (unsigned char*)calloc(nc, 1);
  MaskTot* mt = (MaskTot*)malloc((size_t)nbody * sizeof(MaskTot));
  for (int k = 0; k < nbody; k++) {
    SEL(k);
    mt[k] = mask_linearise(s, dt, rnew, r0c, flag);
    SEL(k);
  }
  /* the penalization's input u_in = P_M(u*), for its chi sensitivity */
  double* uin = NULL;
  if (want_c) {
    orc_pressure_rhs(s, dt);
    orc_poisson_solve(s);
    orc_grad_p(s, dt);
    uin = (double*)malloc(3 * nc * sizeof(double));
    for (int q = 0; q < 3; q++) orc_get_field(s, ORC_U + q, uin + q * nc);
  }
  for (int q = 0; q < 4; q++) orc_set_field(s, keep[q], save + q * nc);  /* u_n back */
  double dc[3] = {0, 0, 0};
  for (int k = 0; k < nbody; k++) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    SEL(k);
    pen_transpose(s, dt, lambda, implicit, V + 3 * k, wdrag[k], uin, lam, acc);
    SEL(k);
    for (int d = 0; d < 3; d++) {
      dJdV[3 * k + d] += acc[d];
      dc[d] += acc[3 + d];
      if (dJdW) dJdW[3 * k + d] += acc[6 + d];
    }
  }
  free(uin);
  /* transposed projection: y = L^{-1} R lam' (plain RHS), y <- M'^T y, lam* = lam' - (dt/2h) grad y.
   * y is the adjoint of the masked RHS times h^3/dt^2: the transposed projection applies the RHS
   * operator (h^2/2dt) div where the transpose of the gradient step is (dt/2h) div, and the
   * gradient step after it applies the inverse factor, so lam* is exact; the chi term is scaled. */
  for (int q = 0; q < 3; q++) orc_set_field(s, ORC_U + q, lam + q * nc);
  const int sphere = s->has_sphere;
  s->has_sphere = 0;
  orc_pressure_rhs(s, dt);
  orc_poisson_solve(s);
  double* y = (double*)malloc(nc * sizeof(double));
  orc_get_field(s, ORC_P, y);
  s->has_sphere = sphere;
  const double yscale = dt * dt / (h * h * h);
  for (int k = 0; k < nbody; k++) {
    SEL(k);
    mask_transpose(s, &mt[k], rnew, r0c, flag, yscale, y, want_c ? dc : NULL);
    SEL(k);
  }
#undef SEL
  if (want_c)
    for (int k = 0; k < 3; k++) dJdc[k] += dc[k];
  orc_set_field(s, ORC_P, y);
  s->has_sphere = 0;
  orc_grad_p(s, dt);
  s->has_sphere = sphere;
  for (int q = 0; q < 3; q++) orc_get_field(s, ORC_U + q, lam + q * nc);  /* lam* */
  for (int q = 0; q < 4; q++) orc_set_field(s, keep[q], save + q * nc);
  free(save); free(y); free(rnew); free(flag); free(r0c); free(mt);
  if (forced) adj_forcing(s, dt, nu, uinf, scheme, tke_target, injection_rate, lam);
  adj_advection(s, dt, nu, uinf, scheme, lam);
}

void orc_adjoint_step_body_full(orc_sim* s, double dt, double nu, const double uinf[3], int scheme,
                                double tke_target, double injection_rate, double lambda, int implicit,
                                const double V[3], const double W[3], double wdrag, double* lam, double dJdV[3],
                                double dJdW[3], double dJdc[3])
{
  if (!s->has_sphere) { orc_adjoint_step_forcing(s, dt, nu, uinf, scheme, tke_target, injection_rate, lam); return; }
  adjoint_bodies(s, dt, nu, uinf, scheme, tke_target, injection_rate, lambda, implicit, 1, V, W, &wdrag, lam, dJdV,
                 dJdW, dJdc);
}

void orc_adjoint_step_bodies(orc_sim* s, double dt, double nu, const double uinf[3], int scheme, double tke_target,
                             double injection_rate, double lambda, int implicit, int nbody, const double* V,
                             const double* W, const double* wdrag, double* lam, double* dJdV, double* dJdW)
{
  if (nbody != orc_body_count(s) || nbody < 1) {
    for (int i = 0; i < 3 * (nbody > 0 ? nbody : 0); i++) dJdV[i] = dJdW[i] = NAN;
    return;
  }
  adjoint_bodies(s, dt, nu, uinf, scheme, tke_target, injection_rate, lambda, implicit, nbody, V, W, wdrag, lam,
                 dJdV, dJdW, NULL);
}

/* ---------------------------------------------------------------- timestep */

/* Simulation::timestep (Simulation.cpp:593-623) over the pipeline of setupOperators
 * (:307-401) for the periodic, non-iterative configuration: [CreateObstacles: a static
 * sphere is created once by orc_sphere_create], AdvectionDiffusion, [SpectralForcing],
 * PressureRHS, PressureProjection, [UpdateObstacles: forced body, velocities unchanged],
 * [Penalization], [ComputeForces]. ComputeDissipation is called separately. */
void orc_timestep(orc_sim* s, double dt, const orc_params* prm, double forcing_stats[5],
                  double penal_ft[6], double surf_qoi[22])
{
  orc_advect_diffuse_scheme(s, dt, prm->nu, prm->uinf, prm->scheme);
  if (prm->spectral_forcing)
    orc_spectral_forcing(s, dt, dt, prm->nu, prm->tke_target, prm->injection_rate, forcing_stats);
  orc_pressure_rhs(s, dt);
  /* PressureProjection's solver (PressureProjection.cpp:86-105): FFT with every face periodic
   * (bUseFourierBC), else PoissonSolverMixed */
  if (s->bcf[0] == ORC_BC_PERIODIC && s->bcf[1] == ORC_BC_PERIODIC && s->bcf[2] == ORC_BC_PERIODIC) orc_poisson_solve(s);
  else if (s->unbounded) orc_poisson_unbounded(s);
  else orc_poisson_mixed(s);
  orc_grad_p(s, dt);
  if (prm->has_sphere && s->has_sphere) {
    double ft[6];
    orc_penalize(s, dt, prm->lambda, prm->implicit_penalization, prm->trans_vel, prm->ang_vel, ft);
    if (penal_ft) for (int i = 0; i < 6; i++) penal_ft[i] = ft[i];
    double q[22];
    orc_surface_forces(s, prm->nu, prm->trans_vel, prm->ang_vel, q);
    if (surf_qoi) for (int i = 0; i < 22; i++) surf_qoi[i] = q[i];
  }
}
