// cup_internal.h -- device context, shared device helpers and the launch functions each
// .hip file provides. Internal to libcup.so; the public surface is include/cup.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cup.h"

namespace cup {

constexpr int BS = 8;     // cells per block edge (the configs' CUP_BLOCK_SIZE)
constexpr int BS3 = 512;  // cells per block
constexpr int NT = 256;   // threads per workgroup of the block-store kernels

template <typename T>
struct C2 {
  T x, y;
};

// Two consecutive values stored with the non-temporal hint (a vector store streamed past the
// caches, for outputs the next kernel reads in another order), and loaded with it (for inputs
// read for the last time).
template <typename T>
__device__ __forceinline__ void store2_nt(T* p, T a, T b)
{
  typedef T v2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(v2{a, b}, reinterpret_cast<v2*>(p));
}
template <typename T>
__device__ __forceinline__ void load2_nt(const T* p, T& a, T& b)
{
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  a = v.x;
  b = v.y;
}

// Four-wide vector of the context precision (16- or 32-byte loads of 4 consecutive cells).
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

// Step scalars resident on the device. The fused driver never copies them to the host
// inside a step; the host reads them on request.
struct DevScalars {
  double dt;         // dt of the step being taken
  double last_dt;    // dt of the last completed step
  double umax;       // max_i |u_i + uinf_i| measured on the current velocity
  double time;
  double ke;         // sum over cells of (u^2+v^2+w^2)/2 * h^3 on the current velocity
  long long step;
  double forcing_fac;
  double tke, tke_low, eps, lint;
  double red[32];    // generic finalized reduction results
  unsigned long long nan_count[4];  // cup_check_nan: NaN cells of u, v, w, p
  // max |u+uinf| of the fused step, folded per workgroup by atomic max on the bit pattern
  // of a non-negative double, spread over UMAX_SLOTS addresses to avoid contention. Two
  // banks: a step's gradient writes one while the next step's advection reads the other.
  unsigned long long umax_slots[2][256];
};
constexpr int UMAX_SLOTS = 256;

// Radix plan of a transform length that is not a power of two (FFT passes with the length
// known at run time): n = r[0] r[1] ... r[ns-1], radices 8, 4, 2, 3, 5.
constexpr int FFT_MAX_N = 1024;
struct FftPlan {
  int n = 0, ns = 0;
  int r[16] = {};
};

// Canonical low modes of the band-limited forcing (0 < |k|^2 <= 4, one member of each
// conjugate pair) and their positions in the packed half spectrum; passed by value.
constexpr int LOW_MAX = 32;  // modes supported by the fused forcing
constexpr int LOW_KX = 2;    // largest |kx| among them
struct LowModes {
  int n;                  // canonical modes
  int k[LOW_MAX][3];      // (kx, ky, kz) integer wavenumbers
  int npos;               // packed positions: 1 per mode with kx != 0, 2 with kx == 0
  int pos[2 * LOW_MAX][4];  // slot (|kx|), ky index, kz index, (mode << 1) | conjugate
  double sk[LOW_MAX][3];    // sin(2 pi k_d / N_d): the central-difference symbol of each mode
};

// Geometry passed by value to kernels.
struct GridArgs {
  int nbx, nby, nbz;  // blocks per dimension
  int Nx, Ny, Nz;     // cells per dimension
  int nb;             // blocks
  double h;
  int bc[3];          // per axis: CUP_BC_PERIODIC, CUP_BC_OPEN, CUP_BC_WALL
  int bcany;          // some axis is not periodic
};

// How a halo value is formed at a wall face (BlockLabBC::applyBCfaceWall,
// Definitions.h:302-334): velocities (and the SDF the obstacle build keeps in tmpU) are the
// negated mirror image, chi is zero, everything else (p) copies. Open faces copy every field.
enum HaloKind { HK_COPY = 0, HK_VEL = 1, HK_CHI = 2 };

// A sphere turned into chi and surface data (static body: built once).
struct Body {
  bool present = false;
  double radius = 0, center[3] = {0, 0, 0};
  double moll = 0;              // half-width in cells of the mollified chi; 0: Towers' chi
  bool hemi = false;            // the hemisphere (HemiSphereObstacle): the ball's part at x < c_x
  double com[3] = {0, 0, 0}, mass = 0;
  int nob = 0;                  // obstacle blocks
  int* ob_ids = nullptr;        // device: block id of each obstacle block
  int* ob_slot = nullptr;       // device: per block, its obstacle slot or -1
  double* ob_chi = nullptr;     // device: nob*512 obstacle chi (FP64)
  double* ob_sdf = nullptr;     // device: nob*512 SDF
  double* ob_surf = nullptr;    // device: nob*512*4 (delta, dchidx, dchidy, dchidz)
  void* ob_rhs = nullptr;       // device: nob*512 corrected RHS (context precision)
  unsigned char* ob_touch = nullptr;  // device: nob*512 surface flag of the RHS correction
  void* ob_r0 = nullptr;        // device: nob*512 unmasked RHS (adjoint with the centre; lazy)
  std::vector<int> ids_host;    // the obstacle blocks' ids, to reuse the buffers when a moved body keeps them
  double* ob_udef = nullptr;    // device: nob*3*512 deformation velocity [slot][component][cell] (lazy; 0 when absent)
  double corr[6] = {0, 0, 0, 0, 0, 0};  // transVel_correction, angVel_correction of the last udef removal
  double vel[6] = {0, 0, 0, 0, 0, 0};   // velocity and rotation the device-resident loop penalizes with (bodies >= 1)
};

struct Ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  int precision = 8;  // 4 or 8
  int decomposition = CUP_FUSED;
  int bpd[3] = {0, 0, 0}, N[3] = {0, 0, 0}, nb = 0;
  size_t ncell = 0;
  double h = 0, maxextent = 0, extent[3] = {0, 0, 0};
  size_t esz = 8;

  void* f[14] = {};       // block-layout fields chi,u,v,w,p,tmpU,tmpV,tmpW; adjoint u,v,w (lazy);
                          // F_ADJ2..+2: the upwind midpoint adjoint's scratch (lazy)
  int ustar = CUP_TMPU;   // first of the three fields holding the step's u* after advection
  int umax_bank = 0;      // the bank of DevScalars::umax_slots that the max |u| writers fill
  int bc[3] = {CUP_BC_PERIODIC, CUP_BC_PERIODIC, CUP_BC_PERIODIC};
  double umax_measured = 0;  // SimulationData::uMax_measured of the last cup_calc_max_timestep
  double host_dt = 0;        // dt of the last host-driven operator call (the fades' fac)
  bool periodic() const { return bc[0] == CUP_BC_PERIODIC && bc[1] == CUP_BC_PERIODIC && bc[2] == CUP_BC_PERIODIC; }
  bool plain_rhs = false; // the fused x pass ignores the obstacle RHS (the adjoint's projection)
  void* spec = nullptr;   // packed half spectrum, complex, [z][y][kx], kx < Nx/2
  double* ana = nullptr;  // HIT analysis: per-line sums [3 + nbin][Ny*Nz], then the row totals (lazy)
  void* spec2 = nullptr;  // extra spectra for the plain forcing path
  void* spec3 = nullptr;
  void* rbuf = nullptr;   // block-layout real scratch (plain RHS / solution)
  // p's x faces as the fused step's x inverse pass writes them beside p: per block, side 0
  // (ix = 0) and side 1 (ix = 7), 64 values each at iz*8 + iy, read by the gradient that follows
  void* xface = nullptr;
  bool xface_ok = false;  // xface holds the faces of the current p
  void* tw[3] = {};       // FFT twiddles per axis, e^{-2 pi i t / N}, context precision
  double* cos4[3] = {};   // cos(4 pi k / N) per axis (FP64), the Poisson symbol
  FftPlan plan[3];        // per axis, for the lengths that are not powers of two
  void* r2r[3][2] = {};   // PoissonSolverMixed: per axis the forward and backward r2r matrix,
                          // transposed (m[j][k] = M[k][j]), context precision (lazy)
  double* mden[3] = {};   // PoissonSolverMixed's per-axis symbol terms (lazy)
  bool unbounded = false; // freespace faces on every axis: PoissonSolverUnbounded (bc[] holds them as open)
  Ctx* unb = nullptr;     // PoissonSolverUnbounded's padded periodic grid (lazy)
  void* ghat = nullptr;   // (padded grid) the Green's function's spectrum, FP64 complex
  double* partials = nullptr;
  size_t partials_cap = 0;  // doubles
  double* ke_part = nullptr;  // [nb][4] partial sums |u|^2 of the fused step's velocity (per wave)
  DevScalars* sc = nullptr;
  DevScalars* sc_host = nullptr;  // pinned
  double* host_red = nullptr;     // pinned scratch for reductions
  // forcing: canonical low modes and the fused path's device tables
  int nlow = 0;
  bool low_fused_ok = false;   // modes fit LowModes (count, |kx|)
  LowModes low{};
  LowModes* low_dev = nullptr;  // device copy (kernels index it per lane)
  double* low_part = nullptr;  // x-pass row values [Nz][Ny][16]: per component the row's x-DFT at kx 0..2, |u|^2
  double* low_red = nullptr;   // plane DFTs in y [Nz][16][5], then the W mode sums (6 per mode: re/im of u,v,w; |u|^2)
  double* low_coef = nullptr;  // per mode: RHS addition (re,im), synthesis of u,v,w (re,im) x3
  unsigned* low_counter = nullptr;
  void* low_rows = nullptr;    // per grid row (y, z): 16 synthesis coefficients (context precision)
  double* body_ft = nullptr;   // device: 6 per body, each body's last penalization force/torque in the loop
  int body_ft_n = 0;
  void* adjf = nullptr;        // the forcing adjoint's 9 fields: n P lam, n P u*, u_n kept aside (lazy)
  Body body;                // the active body: body 0, or body k inside a body-indexed call
  std::vector<Body> more;   // the other bodies (more[k-1] holds body k while body 0 is active)
  int* ob_slot_all = nullptr;     // device [nb]: every body's obstacle blocks, slot into ob_rhs_all
  void* ob_rhs_all = nullptr;     // device: the corrected RHS of every body's blocks (ob_rhs slices)
  size_t ob_rhs_all_cap = 0;      // bytes
  std::string err;
  // measurement hook: HIP events around one stage of the fused step
  int timing_stage = -1, timing_cap = 0, timing_count = 0;
  bool timing_skip = false;  // the first step of a cup_run_steps call is not timed
  std::vector<hipEvent_t> tev;
  GridArgs grid() const {
    GridArgs g;
    g.nbx = bpd[0]; g.nby = bpd[1]; g.nbz = bpd[2];
    g.Nx = N[0]; g.Ny = N[1]; g.Nz = N[2];
    g.nb = nb; g.h = h;
    g.bcany = 0;
    for (int a = 0; a < 3; ++a) {
      g.bc[a] = bc[a];
      g.bcany |= bc[a] != CUP_BC_PERIODIC;
    }
    return g;
  }
};

// ----------------------------------------------------------------- device helpers

// Workgroup i is dispatched to XCD i % 8. Remap so that each XCD walks a contiguous
// range of logical ids (neighbouring blocks share that XCD's L2).
__device__ __forceinline__ int xcd_remap(int i, int n)
{
  if ((n & 7) != 0) return i;
  return (i & 7) * (n >> 3) + (i >> 3);
}

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ size_t cell_index(const GridArgs& g, int gx, int gy, int gz)
{
  const int b = (gx >> 3) + g.nbx * ((gy >> 3) + g.nby * (gz >> 3));
  return (size_t)b * BS3 + (((gz & 7) * BS + (gy & 7)) * BS + (gx & 7));
}

// The boundary condition of a block face (0..5 = x-, x+, y-, y+, z-, z+) when it lies on a
// non-periodic domain face, else 0 (the halo comes from the periodic neighbour block).
__device__ __forceinline__ int domain_face(const GridArgs& g, int bx, int by, int bz, int face)
{
  const int a = face >> 1;
  if (g.bc[a] == CUP_BC_PERIODIC) return 0;
  const int bi = a == 0 ? bx : (a == 1 ? by : bz), nbk = a == 0 ? g.nbx : (a == 1 ? g.nby : g.nbz);
  return ((face & 1) ? bi == nbk - 1 : bi == 0) ? g.bc[a] : 0;
}

// A halo value from its source cell at a domain face of type bcf (0: not a domain face).
template <typename T>
__device__ __forceinline__ T halo_value(T v, int kind, int bcf)
{
  if (bcf != CUP_BC_WALL) return v;
  return kind == HK_VEL ? -v : (kind == HK_CHI ? (T)0 : v);
}

// The neighbour of global cell (gx, gy, gz) one cell along axis a in direction dir = +-1
// through the grid's boundaries (BlockLabBC, Definitions.h:261-378): the periodic image, or
// at a non-periodic domain face the boundary cell itself (open faces copy it; a wall mirrors
// across the face, and one cell out the mirror image is the boundary cell). *bcf is the face
// type crossed (0 inside the domain or across a periodic face).
__device__ __forceinline__ size_t nb_cell(const GridArgs& g, int gx, int gy, int gz, int a, int dir, int* bcf)
{
  int c[3] = {gx, gy, gz};
  const int n = a == 0 ? g.Nx : (a == 1 ? g.Ny : g.Nz);
  int v = c[a] + dir, f = 0;
  if (v < 0 || v >= n) {
    if (g.bc[a] == CUP_BC_PERIODIC) v = v < 0 ? v + n : v - n;
    else { f = g.bc[a]; v = c[a]; }
  }
  c[a] = v;
  *bcf = f;
  return cell_index(g, c[0], c[1], c[2]);
}

template <typename T>
__device__ __forceinline__ T nb_value(const T* __restrict__ f, const GridArgs& g, int gx, int gy, int gz, int a,
                                      int dir, int kind)
{
  int bcf;
  const size_t i = nb_cell(g, gx, gy, gz, a, dir, &bcf);
  return halo_value(f[i], kind, bcf);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// ----------------------------------------------------------------- launchers
// cup_stencil.hip
int launch_lab_shift(Ctx& c, int g, int dx, int dy, int dz);
int launch_max_u(Ctx& c, const double uinf[3], bool prime = false);  // partials[nb][2]; prime: atomic max + ke_part
// The previous step's bookkeeping folded into the advection (central scheme only): every
// workgroup forms dt from the max |u| slots of bank `slots` by the CFL rule for step `step`
// (0-based) and workgroup 0 advances step, time, last_dt, umax and dt as k_finalize_step does.
struct StepFold {
  const unsigned long long* slots;
  long long step;
  double CFL;
  int rampup;
  double fixed_dt;
};
int launch_advdiff(Ctx& c, double nu, const double uinf[3], int scheme, const StepFold* fold = nullptr);  // -> tmp
int launch_copy_tmp_to_u(Ctx& c);
int launch_copy_u_to_tmp(Ctx& c);
int launch_adjoint_advdiff(Ctx& c, double nu, const double uinf[3]);
constexpr int F_ADJ2 = 11;
int launch_adjoint_up3(Ctx& c, int s0, int l0, int b10, int b20, int out0, double nu, const double uinf[3],
                       double mult);
int launch_up3_stage(Ctx& c, int src0, int base0, int out0, double nu, const double uinf[3], double mult);
int launch_gradp_inplace(Ctx& c, int field0);
// The adjoint step with the sphere, in phases (cup_adjoint_step_body orders them).
enum AdjBodyPhase {
  ADJB_FWD_MASK,     // forward masking of u*: r_new in ob_rhs, flags, sums (and r0 with the centre)
  ADJB_FWD_CORRECT,  // the forward perimeter correction of ob_rhs (for the forward re-solve)
  ADJB_CHI_PEN,      // chi sensitivity of the penalization and its drag, contracted with dchi/dc
  ADJB_PENALIZE,     // penalization and drag transpose on the adjoint; dJ/dV
  ADJB_MASK          // transposed masking on the adjoint pressure (and its chi sensitivity)
};
struct AdjBodyArgs {
  double dt = 0, lambda = 0, wdrag = 0, vel[3] = {0, 0, 0};
  int implicit = 1;
  bool centre = false;
};
int launch_adjoint_body(Ctx& c, AdjBodyPhase phase, const AdjBodyArgs& a);
int launch_rhs_plain(Ctx& c);                                   // -> rbuf
int launch_gradp_plain(Ctx& c);                                 // u += -dt/2h grad p
// xface: take p's x faces from c.xface when the fused Poisson solve just wrote them
int launch_gradp_fused(Ctx& c, const double uinf[3], const int* skip_slot, bool force = false,
                       bool xface = false);  // u = tmp - dt/2h grad p; max, ke_part
int launch_diagnostics(Ctx& c, double nu, int* nparts);  // partials[nparts][20]
int launch_ic_taylor_green(Ctx& c, double uMax);
int launch_fill(Ctx& c, void* ptr, size_t n, double value);
int launch_copy(Ctx& c, void* dst, const void* src, size_t n);
int launch_set_dt(Ctx& c, double dt);
int launch_reduce(Ctx& c, int n, int width, unsigned maxmask, double* out_dev, const double* part = nullptr);
int launch_count_nan(Ctx& c, long long out[4]);  // NaN cells of u, v, w, p (synchronizes)
// Records the message of a failed call that has no context (read by cup_last_error(NULL)).
int free_error(int code, const std::string& msg);
int launch_reset_umax(Ctx& c);  // clears both banks
int launch_inflow_correction(Ctx& c);  // UpdateAndCorrectInflow on the open faces of u
int launch_poisson_mixed(Ctx& c);      // PoissonSolverMixed: rbuf (RHS) -> rbuf (solution)
int launch_poisson_unbounded(Ctx& c);  // PoissonSolverUnbounded: rbuf (RHS) -> p
void unbounded_free(Ctx& c);
// Device address of bank k of the max |u| slots (host pointer arithmetic, no access).
inline unsigned long long* umax_slots_dev(const Ctx& c, int k) { return &c.sc->umax_slots[k][0]; }
int launch_finalize_step(Ctx& c, double CFL, double nu, int rampup, double fixed_dt, int advance);
// cup_fft.hip
int fft_setup(Ctx& c);
// xface: the x inverse pass also writes p's x faces to c.xface (for launch_gradp_fused)
int launch_poisson_fused(Ctx& c, bool force = false, double tke_target = 0, double inj_rate = 0,
                         bool xface = false);  // div(tmp) -> p
int launch_poisson_plain(Ctx& c);   // rhs in rbuf -> solution in rbuf (separate divide pass)
int launch_fft_forward_field(Ctx& c, const void* real_blocks, void* spec);  // 3 passes
// HIT analysis: out[0..2] = raw sums tke, eps, l_integral numerator; E[nbin] raw shell sums;
// grad[2] = sum sqrt(grad2), sum grad2 (see cup_hit_analysis)
int launch_hit_analysis(Ctx& c, double* tke_eps_lint, double* E, int nbin, double grad[2]);
int launch_grad_stats(Ctx& c, double out[2]);  // sums of sqrt(grad2) and grad2 (synchronous)
int launch_fft_inverse_field(Ctx& c, void* spec, void* real_blocks);        // 3 passes
// cup_spectral.hip
int spectral_setup(Ctx& c);
int launch_ic_hit(Ctx& c, double k0, const double* K, const double* E, int tn, double u_rms, uint64_t seed);
// field: the first of the three velocity fields forced in place (-1: u* under FUSED, u under PLAIN)
int forcing_plain(Ctx& c, double dt, double sim_dt, double nu, double tke_target, double inj,
                  double stats[5], int field = -1);
// planes = false: the plane stage ran beside the y pass (forcing_planes_extra)
int launch_forcing_reduce(Ctx& c, double tke_target, double inj_rate, bool planes = true);
struct ColExtra;
ColExtra forcing_planes_extra(Ctx& c);  // the plane and row stages' buffers, for the column passes
int launch_forcing_rows(Ctx& c);
// adjoint fields += the forcing's transpose at u* (fields us0..us0+2); synchronous
int launch_adjoint_forcing(Ctx& c, int us0, double dt, double tke_target, double inj_rate);
// cup_obstacle.hip
int body_create(Ctx& c, const double center[3], double radius, double moll, bool hemi = false);
// Body k made the active body for the scope (bodies are swapped, not copied).
struct BodyScope {
  Ctx& c;
  int k;
  BodyScope(Ctx& ctx, int id) : c(ctx), k(id)
  {
    if (k > 0) std::swap(c.body, c.more[k - 1]);
  }
  ~BodyScope()
  {
    if (k > 0) std::swap(c.body, c.more[k - 1]);
  }
};
inline int body_count(const Ctx& c) { return c.body.present ? 1 + (int)c.more.size() : 0; }
void bodies_free(Ctx& c);
void body_free(Ctx& c);
int launch_penalize(Ctx& c, double lambda, int implicit, const double vel[3], const double omega[3],
                    double* ft_dev = nullptr);
int launch_penalize_step(Ctx& c, double lambda, int implicit, const double vel[3], const double omega[3],
                         const double uinf[3], double* ft_dev = nullptr);  // also refreshes the fused step's max/energy partials
int launch_surface_forces(Ctx& c, double nu, const double vel[3], const double omega[3]);
int launch_obstacle_momenta(Ctx& c, double dt, double lambda, int implicit);
int launch_udef_remove(Ctx& c, double out[13]);  // the active body's udef momentum removal (synchronous)
int launch_add_uniform(Ctx& c, const double du[3]);
int launch_rhs_obstacle(Ctx& c, void* rhs_blocks);

// Stages of the fused step, for the measurement hook.
enum Stage { ST_ADVDIFF = 0, ST_XFWD, ST_FREDUCE, ST_YFWD, ST_ZPASS, ST_YINV, ST_XINV, ST_GRADP, ST_PENAL, ST_FINAL };
void stage_mark(Ctx& c, int stage, bool end);

// helpers (cup_capi.cpp)
int ensure_partials(Ctx& c, size_t ndoubles);
int set_error(Ctx& c, int code, const std::string& msg);
int check_hip(Ctx& c, hipError_t e, const char* what);

}  // namespace cup

struct cup_ctx {
  cup::Ctx c;
};
