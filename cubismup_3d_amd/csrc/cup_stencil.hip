// cup_stencil.hip -- block-store kernels: LDS-staged periodic halo gather, advection-
// diffusion, pressure RHS, pressure gradient, reductions, diagnostics, initial conditions.
//
// Layout: every field is block-major, block b = bx + nbx*(by + nby*bz) holds 512 cells
// [iz][iy][ix] contiguously. One workgroup of 256 threads owns one 8^3 block: it stages
// the block plus the halo its stencil needs in LDS (the device form of Cubism's
// BlockLab::load with periodic wrap), then computes 2 cells per thread.
#include "cup_internal.h"

#include <cmath>
#include <mutex>

namespace cup {

// 10^3 lab (halo 1), faces only: the 7-point stencils never read edges or corners.
constexpr int L1 = 10;
constexpr int L1C = L1 * L1 * L1;

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};


// kinds[f] (HaloKind) says how field f's halo forms at a wall face; periodic grids ignore it.
template <typename T, int NF>
__device__ __forceinline__ void gather_g1(const T* const* src, T (*lab)[L1C], const GridArgs& g, int b,
                                          int bx, int by, int bz, const int (&kinds)[NF])
{
  const int tid = threadIdx.x;
  static_assert(2 * NT == BS3, "two cells per thread");
  {
    const int c = 2 * tid, ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const int l = ((iz + 1) * L1 + iy + 1) * L1 + ix + 1;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const T* q = src[f] + (size_t)b * BS3 + c;
      const T a0 = q[0], a1 = q[1];
      lab[f][l] = a0;
      lab[f][l + 1] = a1;
    }
  }
  for (int c = tid; c < 6 * 64; c += NT) {
    const int face = c >> 6, a = (c >> 3) & 7, e = c & 7;
    int qx = bx, qy = by, qz = bz, sx, sy, sz, lx, ly, lz;
    switch (face) {
      case 0: qx = wrapi(bx - 1, g.nbx); sx = 7; sy = e; sz = a; lx = 0; ly = e + 1; lz = a + 1; break;
      case 1: qx = wrapi(bx + 1, g.nbx); sx = 0; sy = e; sz = a; lx = 9; ly = e + 1; lz = a + 1; break;
      case 2: qy = wrapi(by - 1, g.nby); sx = e; sy = 7; sz = a; lx = e + 1; ly = 0; lz = a + 1; break;
      case 3: qy = wrapi(by + 1, g.nby); sx = e; sy = 0; sz = a; lx = e + 1; ly = 9; lz = a + 1; break;
      case 4: qz = wrapi(bz - 1, g.nbz); sx = e; sy = a; sz = 7; lx = e + 1; ly = a + 1; lz = 0; break;
      default: qz = wrapi(bz + 1, g.nbz); sx = e; sy = a; sz = 0; lx = e + 1; ly = a + 1; lz = 9; break;
    }
    const int bf = g.bcany ? domain_face(g, bx, by, bz, face) : 0;
    if (bf) {  // a non-periodic domain face: the block's own boundary cell (BlockLabBC)
      qx = bx; qy = by; qz = bz;
      const int edge = (face & 1) ? 7 : 0;
      if (face < 2) sx = edge;
      else if (face < 4) sy = edge;
      else sz = edge;
    }
    const size_t s = (size_t)(qx + g.nbx * (qy + g.nby * qz)) * BS3 + (sz * BS + sy) * BS + sx;
    const int l = (lz * L1 + ly) * L1 + lx;
#pragma unroll
    for (int f = 0; f < NF; ++f) lab[f][l] = halo_value(src[f][s], kinds[f], bf);
  }
}

__device__ __forceinline__ void block_coords(const GridArgs& g, int b, int& bx, int& by, int& bz)
{
  bx = b % g.nbx;
  by = (b / g.nbx) % g.nby;
  bz = b / (g.nbx * g.nby);
}

// ------------------------------------------------------- pipelined lab staging
// The fused step's stencil kernels walk several blocks per workgroup (the gradient and the
// adjoint stacked in z, the advection along x) and fetch the next blocks' cells into
// registers while the current one is computed from LDS, which keeps about twice the bytes in
// flight per CU of a load-sync-compute workgroup.
// A thread's cells: interior c = 2t, 2t+1; halo (faces only, 6*64 cells) c = t (faces
// x-, x+, y-, y+: one face per wave) and, for t < 128, c = NT + t (faces z-, z+).
// The pipelined lab pads x to a row stride PXV of at least 12 (interior at x = 2..9, halo at
// 1 and 10): a thread's x-adjacent cell pair, and its y and z neighbour pairs, are 8-byte
// aligned (one 64-bit LDS access each). A wave's half covers the 8 rows of one z plane; with
// rows 24 apart their 8-float interiors fall on distinct LDS banks (24 r mod 64 is 8 times a
// permutation of 0..7), rows 14 apart share fewer banks than 12; the stride is chosen per
// kernel by measurement (the advection's lab of three fields at 24, the gradient's at 14).
template <int PXV>
struct LabGeom {
  static constexpr int PY = PXV;            // y stride
  static constexpr int PZ = PXV * L1;       // z stride
  static constexpr int PLC = PXV * L1 * L1;  // cells per field
};
constexpr int PX_ADV = 24;
constexpr int PX_GRAD = 14;
constexpr int ADV_PD = 3;  // blocks in flight in the advection's walk along x
struct LabSlots {
  int li;        // LDS index of interior cell 2t
  int l[2];      // LDS index of the halo cells
  int off[2];    // their offsets inside the neighbour block
  int off_bc[2]; // at a non-periodic domain face: the offset of the block's own boundary cell
  int face[2];   // 0..5 = x-, x+, y-, y+, z-, z+
};

__device__ __forceinline__ bool has_halo2() { return threadIdx.x < 6 * 64 - NT; }

template <int PXV>
__device__ __forceinline__ LabSlots lab_slots()
{
  constexpr int PX = PXV;
  LabSlots s;
  const int t = threadIdx.x, c0 = 2 * t;
  s.li = (((c0 >> 6) + 1) * L1 + ((c0 >> 3) & 7) + 1) * PX + (c0 & 7) + 2;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int c = k == 0 ? t : (has_halo2() ? NT + t : 4 * 64);
    const int face = c >> 6, a = (c >> 3) & 7, e = c & 7;
    int sx, sy, sz, lx, ly, lz;
    switch (face) {
      case 0: sx = 7; sy = e; sz = a; lx = 0; ly = e + 1; lz = a + 1; break;
      case 1: sx = 0; sy = e; sz = a; lx = 9; ly = e + 1; lz = a + 1; break;
      case 2: sx = e; sy = 7; sz = a; lx = e + 1; ly = 0; lz = a + 1; break;
      case 3: sx = e; sy = 0; sz = a; lx = e + 1; ly = 9; lz = a + 1; break;
      case 4: sx = e; sy = a; sz = 7; lx = e + 1; ly = a + 1; lz = 0; break;
      default: sx = e; sy = a; sz = 0; lx = e + 1; ly = a + 1; lz = 9; break;
    }
    s.face[k] = face;
    s.off[k] = (sz * BS + sy) * BS + sx;
    // the boundary cell sits at the other end of the normal axis: 7 - (source index)
    s.off_bc[k] = face < 2 ? (sz * BS + sy) * BS + (7 - sx) : (face < 4 ? (sz * BS + (7 - sy)) * BS + sx
                                                                        : ((7 - sz) * BS + sy) * BS + sx);
    s.l[k] = (lz * L1 + ly) * PX + lx + 1;
  }
  return s;
}

// Periodic neighbour of block (bx, by, bz) across face 0..5.
__device__ __forceinline__ int face_block(const GridArgs& g, int bx, int by, int bz, int face)
{
  const int d = (face & 1) ? 1 : -1;
  if (face < 2) bx = wrapi(bx + d, g.nbx);
  else if (face < 4) by = wrapi(by + d, g.nby);
  else bz = wrapi(bz + d, g.nbz);
  return bx + g.nbx * (by + g.nby * bz);
}

template <typename T, int NF>
struct LabRegs {
  T in[NF][2];
  T hv[NF][2];
};

// BC: the grid has a non-periodic axis (compiled separately, so periodic grids run none of
// the domain-face tests).
template <bool BC, typename T, int NF>
__device__ __forceinline__ void lab_fetch(LabRegs<T, NF>& r, const T* const* src, const LabSlots& s,
                                          const GridArgs& g, int b, const int (&kinds)[NF])
{
  using V2 = typename Vec2<T>::type;
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const V2 a = *reinterpret_cast<const V2*>(src[f] + (size_t)b * BS3 + 2 * threadIdx.x);
    r.in[f][0] = a.x;
    r.in[f][1] = a.y;
  }
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k == 0 || has_halo2()) {
      // a halo slot's face is the same across its wave: the neighbour block is scalar work
      const int face = __builtin_amdgcn_readfirstlane(s.face[k]);
      if (!BC) {
        const size_t q = (size_t)face_block(g, bx, by, bz, face) * BS3 + s.off[k];
#pragma unroll
        for (int f = 0; f < NF; ++f) r.hv[f][k] = src[f][q];
        continue;
      }
      const int bf = domain_face(g, bx, by, bz, face);
      const size_t q = bf ? (size_t)b * BS3 + s.off_bc[k] : (size_t)face_block(g, bx, by, bz, face) * BS3 + s.off[k];
#pragma unroll
      for (int f = 0; f < NF; ++f) r.hv[f][k] = halo_value(src[f][q], kinds[f], bf);
    }
}

template <int PXV, typename T, int NF>
__device__ __forceinline__ void lab_put(const LabRegs<T, NF>& r, T (*lab)[LabGeom<PXV>::PLC], const LabSlots& s)
{
  using V2 = typename Vec2<T>::type;
#pragma unroll
  for (int f = 0; f < NF; ++f) *reinterpret_cast<V2*>(&lab[f][s.li]) = V2{r.in[f][0], r.in[f][1]};
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k == 0 || has_halo2())
#pragma unroll
      for (int f = 0; f < NF; ++f) lab[f][s.l[k]] = r.hv[f][k];
}

// The pipelined kernels' walk: workgroup w takes nbw blocks stacked in z,
// b_j = bxy + nxy*(zg*nbw + j) with bxy = w % nxy, zg = w / nxy (nxy = nbx*nby), so that each
// block after the first finds its lower z-face in the block just loaded (still in L2).
// nbw = 4 when the blocks along z allow it.
inline int blocks_per_wg(int nbz) { return nbz % 4 == 0 ? 4 : (nbz % 2 == 0 ? 2 : 1); }

__device__ __forceinline__ int walk_first(const GridArgs& g, int w, int nbw)
{
  const int nxy = g.nbx * g.nby;
  return w % nxy + nxy * (w / nxy) * nbw;
}

// Block-wide reduction of NV doubles per thread into out[0..NV) (thread 0 holds them).
// Fixed shuffle tree and fixed LDS order: deterministic.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const unsigned maxmask, double* out)
{
  __shared__ double red[NT / 64][NV > 0 ? NV : 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = ((maxmask >> i) & 1) ? wave_max(v[i]) : wave_sum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[wv][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double a = red[0][i];
      for (int w = 1; w < NT / 64; ++w) a = ((maxmask >> i) & 1) ? fmax(a, red[w][i]) : a + red[w][i];
      out[i] = a;
    }
  }
}

// The CFL rule of Simulation::calcMaxTimestep (Simulation.cpp:403-424), ramp included:
// the dt of step `step` (0-based) from the velocity's max |u + uinf|.
__device__ __forceinline__ double cfl_dt(double umax, long long step, double CFL, double nu, double h, int rampup,
                                         double fixed_dt)
{
  if (fixed_dt > 0) return fixed_dt;
  const double dtDif = nu > 0 ? h * h / nu : INFINITY;
  const double dtAdv = h / (umax + 1e-8);
  const double mm = fmin(dtDif, dtAdv);
  double dt = CFL * mm;
  if (step < rampup) {
    const double x = (step + 1.0) / rampup;
    const double rampCFL = exp(log(1e-3) * (1 - x) + log(CFL) * x);
    dt = rampCFL * mm;
  }
  return dt;
}

// ------------------------------------------------------------------- lab shift test

template <typename T, int G>
__global__ __launch_bounds__(NT) void k_lab_shift(const T* __restrict__ u, T* __restrict__ out, GridArgs g, int dx,
                                                  int dy, int dz)
{
  constexpr int L = BS + 2 * G;
  __shared__ T lab[L * L * L];
  const int b = blockIdx.x;
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
  for (int c = threadIdx.x; c < L * L * L; c += NT) {
    const int lx = c % L, ly = (c / L) % L, lz = c / (L * L);
    const int gx = wrapi(bx * BS + lx - G, g.Nx), gy = wrapi(by * BS + ly - G, g.Ny), gz = wrapi(bz * BS + lz - G, g.Nz);
    lab[c] = u[cell_index(g, gx, gy, gz)];
  }
  __syncthreads();
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    out[(size_t)b * BS3 + c] = lab[((iz + G - dz) * L + (iy + G - dy)) * L + (ix + G - dx)];
  }
}

template <typename T>
static int lab_shift_t(Ctx& c, int G, int dx, int dy, int dz)
{
  const T* u = (const T*)c.f[CUP_U];
  T* o = (T*)c.f[CUP_TMPU];
  if (G == 1) hipLaunchKernelGGL((k_lab_shift<T, 1>), dim3(c.nb), dim3(NT), 0, c.stream, u, o, c.grid(), dx, dy, dz);
  else hipLaunchKernelGGL((k_lab_shift<T, 2>), dim3(c.nb), dim3(NT), 0, c.stream, u, o, c.grid(), dx, dy, dz);
  return check_hip(c, hipGetLastError(), "k_lab_shift");
}

int launch_lab_shift(Ctx& c, int g, int dx, int dy, int dz)
{
  if (g < 1 || g > 2 || abs(dx) > g || abs(dy) > g || abs(dz) > g)
    return set_error(c, CUP_E_INVALID, "lab_shift: need 1 <= g <= 2 and |shift| <= g");
  return c.precision == 4 ? lab_shift_t<float>(c, g, dx, dy, dz) : lab_shift_t<double>(c, g, dx, dy, dz);
}

// ------------------------------------------------------------------- max |u + uinf|
// findMaxU (ProcessHelpers.h:74-99) per block, with the sum of |u|^2 beside it:
// part[2b] = max_i |u_i + uinf_i|, part[2b+1] = sum (u^2+v^2+w^2).

template <typename T>
__global__ __launch_bounds__(NT) void k_max_u(const T* __restrict__ u, const T* __restrict__ v,
                                              const T* __restrict__ w, int nb, double ux, double uy, double uz,
                                              double* __restrict__ part, unsigned long long* __restrict__ slots,
                                              double* __restrict__ ke_part)
{
  const int b = blockIdx.x;
  double m[2] = {0.0, 0.0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const size_t i = (size_t)b * BS3 + c;
    const double a = fabs((double)u[i] + ux), bb = fabs((double)v[i] + uy), cc = fabs((double)w[i] + uz);
    m[0] = fmax(m[0], fmax(a, fmax(bb, cc)));
    m[1] += (double)u[i] * u[i] + (double)v[i] * v[i] + (double)w[i] * w[i];
  }
  __shared__ double out[2];
  block_reduce<2>(m, 1u, out);
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * (size_t)b] = out[0];
    part[2 * (size_t)b + 1] = out[1];
    if (slots) {
      atomicMax(&slots[b & (UMAX_SLOTS - 1)], (unsigned long long)__double_as_longlong(out[0]));
      ke_part[4 * (size_t)b] = out[1];
      ke_part[4 * (size_t)b + 1] = ke_part[4 * (size_t)b + 2] = ke_part[4 * (size_t)b + 3] = 0.0;
    }
  }
}

template <typename T>
static int max_u_t(Ctx& c, const double uinf[3], bool prime)
{
  hipLaunchKernelGGL((k_max_u<T>), dim3(c.nb), dim3(NT), 0, c.stream, (const T*)c.f[CUP_U], (const T*)c.f[CUP_V],
                     (const T*)c.f[CUP_W], c.nb, uinf[0], uinf[1], uinf[2], c.partials,
                     prime ? umax_slots_dev(c, c.umax_bank) : nullptr, c.ke_part);
  return check_hip(c, hipGetLastError(), "k_max_u");
}

int launch_max_u(Ctx& c, const double uinf[3], bool prime)
{
  if (int e = ensure_partials(c, 2 * (size_t)c.nb)) return e;
  return c.precision == 4 ? max_u_t<float>(c, uinf, prime) : max_u_t<double>(c, uinf, prime);
}

__global__ void k_reset_umax(DevScalars* sc) { sc->umax_slots[0][threadIdx.x] = sc->umax_slots[1][threadIdx.x] = 0ull; }

int launch_reset_umax(Ctx& c)
{
  hipLaunchKernelGGL(k_reset_umax, dim3(1), dim3(UMAX_SLOTS), 0, c.stream, c.sc);
  return check_hip(c, hipGetLastError(), "k_reset_umax");
}

// ------------------------------------------------------------ advection-diffusion

// KernelAdvectDiffuse (AdvectionDiffusion.cpp:106-145), same operation order:
// tmp = u + facA*((u+U).grad_c u) + facD*(sum of 6 neighbours - 6u), facA = -dt/(2h),
// facD = (nu/h)*(dt/h).
// A workgroup walks NBW blocks along x: block j's x- halo is block j-1's
// ix = 7 plane and its x+ halo block j+1's ix = 0 plane, both taken from the registers the
// pipeline holds (threads t % 4 == 3 hold ix = 6, 7; t % 4 == 0 hold ix = 0, 1), so only
// the walk's two end faces are gathered. PD blocks are in flight (register sets rotate
// through PD + 1 slots; the loop unrolls fully, so the slot indices are constants).
// fade[face]: the advection's inflow fade of an open domain face (1: none), applied to the
// velocities of that face's halo (AdvectionDiffusion.cpp:39-104).
template <bool BC, typename T, int NF>
__device__ __forceinline__ void lab_fetch_xw(LabRegs<T, NF>& r, const T* const* src, const LabSlots& s,
                                             const GridArgs& g, int b, bool x_lo, bool x_hi, const T (&fade)[6])
{
  using V2 = typename Vec2<T>::type;
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const V2 a = *reinterpret_cast<const V2*>(src[f] + (size_t)b * BS3 + 2 * threadIdx.x);
    r.in[f][0] = a.x;
    r.in[f][1] = a.y;
  }
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k == 0 || has_halo2()) {
      const int face = __builtin_amdgcn_readfirstlane(s.face[k]);
      if ((face == 0 && !x_lo) || (face == 1 && !x_hi)) continue;
      if (!BC) {
        const size_t q = (size_t)face_block(g, bx, by, bz, face) * BS3 + s.off[k];
#pragma unroll
        for (int f = 0; f < NF; ++f) r.hv[f][k] = src[f][q];
        continue;
      }
      const int bf = domain_face(g, bx, by, bz, face);
      const size_t q = bf ? (size_t)b * BS3 + s.off_bc[k] : (size_t)face_block(g, bx, by, bz, face) * BS3 + s.off[k];
      const T fd = bf == CUP_BC_OPEN ? fade[face] : (T)1;
#pragma unroll
      for (int f = 0; f < NF; ++f) r.hv[f][k] = bf ? halo_value(src[f][q], HK_VEL, bf) * fd : src[f][q];
    }
}

template <int PXV, typename T, int NF>
__device__ __forceinline__ void lab_put_xw(const LabRegs<T, NF>& r, const LabRegs<T, NF>& prev,
                                           const LabRegs<T, NF>& next, T (*lab)[LabGeom<PXV>::PLC], const LabSlots& s,
                                           bool x_lo,
                                           bool x_hi)
{
  using V2 = typename Vec2<T>::type;
#pragma unroll
  for (int f = 0; f < NF; ++f) *reinterpret_cast<V2*>(&lab[f][s.li]) = V2{r.in[f][0], r.in[f][1]};
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k == 0 || has_halo2()) {
      const int face = __builtin_amdgcn_readfirstlane(s.face[k]);
      if ((face == 0 && !x_lo) || (face == 1 && !x_hi)) continue;
#pragma unroll
      for (int f = 0; f < NF; ++f) lab[f][s.l[k]] = r.hv[f][k];
    }
  const int t4 = threadIdx.x & 3;
  if (!x_lo && t4 == 3)
#pragma unroll
    for (int f = 0; f < NF; ++f) lab[f][s.li - 7] = prev.in[f][1];
  if (!x_hi && t4 == 0)
#pragma unroll
    for (int f = 0; f < NF; ++f) lab[f][s.li + 8] = next.in[f][0];
}

// First block of logical walk L along x (NBW blocks of a row of blocks). Rows of blocks
// (by, bz) are ordered in tiles of ZT planes of blocks with bz fastest, so that a walk's z
// neighbours run a few walks earlier or later (on the same XCD, whose consecutive logical ids
// xcd_remap keeps together) and its y and z halo faces are still in L2. The rows run from the
// last by down to 0: the gradient writes u from by = 0 up, so the advection of the next step
// starts on the rows written last, which the last-level cache still holds.
template <int NBW>
__device__ __forceinline__ int xwalk_first(const GridArgs& g, int L)
{
  const int nq = g.nbx / NBW;  // walks per row of blocks
  const int xq = L % nq, row = L / nq;
  const int zt = g.nbz % 4 == 0 ? 4 : (g.nbz % 2 == 0 ? 2 : 1);
  const int tile = row / (g.nby * zt), within = row - tile * (g.nby * zt);
  const int by = g.nby - 1 - within / zt, bz = tile * zt + (within - (within / zt) * zt);
  return xq * NBW + g.nbx * (by + g.nby * bz);
}

template <typename T>
struct Fade6 {
  T f[6];
};

template <typename T, int NBW, int PD, bool BC>
__global__ __launch_bounds__(NT) void k_advdiff_xwalk(const T* __restrict__ u, const T* __restrict__ v,
                                                      const T* __restrict__ w, T* __restrict__ ou,
                                                      T* __restrict__ ov, T* __restrict__ ow, GridArgs g,
                                                      DevScalars* __restrict__ sc, T Ux, T Uy, T Uz, double nu,
                                                      StepFold fold, Fade6<T> fd)
{
  using LG = LabGeom<PX_ADV>;
  __shared__ __align__(16) T lab[3][LG::PLC];
  const LabSlots s = lab_slots<PX_ADV>();
  const T* src[3] = {u, v, w};
  const int b0 = xwalk_first<NBW>(g, xcd_remap(blockIdx.x, gridDim.x));
  static_assert(PD >= 2, "a block's x+ halo comes from the next block, which must be in registers");
  constexpr int NS = PD + 1;
  LabRegs<T, 3> r[NS];
#pragma unroll
  for (int p = 0; p < PD && p < NBW; ++p) lab_fetch_xw<BC>(r[p], src, s, g, b0 + p, p == 0, p == NBW - 1, fd.f);
  double dt;
  if (fold.slots) {
    // the previous step's finalize: no workgroup reads the scalars workgroup 0 updates
    double m[1] = {__longlong_as_double((long long)fold.slots[threadIdx.x])};
    __shared__ double fold_umax[1];
    block_reduce<1>(m, 1u, fold_umax);
    __syncthreads();
    const double umax = fold_umax[0];
    dt = cfl_dt(umax, fold.step, fold.CFL, nu, g.h, fold.rampup, fold.fixed_dt);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const double dt_prev = sc->dt;
      sc->umax = umax;
      sc->step += 1;
      sc->time += dt_prev;
      sc->last_dt = dt_prev;
      sc->dt = dt;
    }
  } else {
    dt = sc->dt;
  }
  const T facA = (T)(-dt / (2 * g.h));
  const T facD = (T)((nu / g.h) * (dt / g.h));
  const int c0 = 2 * threadIdx.x;
  using V2 = typename Vec2<T>::type;
#pragma unroll
  for (int j = 0; j < NBW; ++j) {
    if (j) __syncthreads();  // the previous block's compute is done with the lab
    lab_put_xw<PX_ADV>(r[j % NS], r[(j + NS - 1) % NS], r[(j + 1) % NS], lab, s, j == 0, j == NBW - 1);
    __syncthreads();
    if (j + PD < NBW) lab_fetch_xw<BC>(r[(j + PD) % NS], src, s, g, b0 + j + PD, false, j + PD == NBW - 1, fd.f);
    T av[3][2];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const V2 C = *reinterpret_cast<const V2*>(lab[q] + s.li);
      av[q][0] = C.x;
      av[q][1] = C.y;
    }
    const T uu[2] = {av[0][0] + Ux, av[0][1] + Ux}, vv[2] = {av[1][0] + Uy, av[1][1] + Uy},
            ww[2] = {av[2][0] + Uz, av[2][1] + Uz};
    T o[3][2];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T* F = lab[q];
      const int l = s.li;
      const V2 Nn = *reinterpret_cast<const V2*>(F + l + LG::PY), Sn = *reinterpret_cast<const V2*>(F + l - LG::PY);
      const V2 Bn = *reinterpret_cast<const V2*>(F + l + LG::PZ), Fn = *reinterpret_cast<const V2*>(F + l - LG::PZ);
      const T Wx = F[l - 1], Ex = F[l + 2];
      const T fc[2] = {av[q][0], av[q][1]};
      const T fe[2] = {av[q][1], Ex}, fw[2] = {Wx, av[q][0]};
      const T fn[2] = {Nn.x, Nn.y}, fs[2] = {Sn.x, Sn.y}, fb[2] = {Bn.x, Bn.y}, ff[2] = {Fn.x, Fn.y};
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const T ddx = fe[k] - fw[k], ddy = fn[k] - fs[k], ddz = fb[k] - ff[k];
        const T dD = fn[k] + fs[k] + fe[k] + fw[k] + ff[k] + fb[k] - fc[k] * 6;
        const T dA = uu[k] * ddx + vv[k] * ddy + ww[k] * ddz;
        o[q][k] = fc[k] + facA * dA + facD * dD;
      }
    }
    const size_t q = (size_t)(b0 + j) * BS3 + c0;
    store2_nt(ou + q, o[0][0], o[0][1]);
    store2_nt(ov + q, o[1][0], o[1][1]);
    store2_nt(ow + q, o[2][0], o[2][1]);
  }
}

template <typename T, int NBW>
static void launch_advdiff_xwalk(Ctx& c, double nu, const double uinf[3], const StepFold& fold, const Fade6<T>& fd)
{
  if (c.periodic())
    hipLaunchKernelGGL((k_advdiff_xwalk<T, NBW, ADV_PD, false>), dim3(c.nb / NBW), dim3(NT), 0, c.stream,
                       (const T*)c.f[CUP_U], (const T*)c.f[CUP_V], (const T*)c.f[CUP_W], (T*)c.f[CUP_TMPU],
                       (T*)c.f[CUP_TMPV], (T*)c.f[CUP_TMPW], c.grid(), c.sc, (T)uinf[0], (T)uinf[1], (T)uinf[2], nu,
                       fold, fd);
  else
    hipLaunchKernelGGL((k_advdiff_xwalk<T, NBW, ADV_PD, true>), dim3(c.nb / NBW), dim3(NT), 0, c.stream,
                       (const T*)c.f[CUP_U], (const T*)c.f[CUP_V], (const T*)c.f[CUP_W], (T*)c.f[CUP_TMPU],
                       (T*)c.f[CUP_TMPV], (T*)c.f[CUP_TMPW], c.grid(), c.sc, (T)uinf[0], (T)uinf[1], (T)uinf[2], nu,
                       fold, fd);
}

template <typename T>
static int advdiff_t(Ctx& c, double nu, const double uinf[3], const StepFold* fold_in)
{
  const StepFold fold = fold_in ? *fold_in : StepFold{nullptr, 0, 0.0, 0, 0.0};
  // KernelAdvectDiffuseBase's fades (AdvectionDiffusion.cpp:25-33): per open domain face
  // 1 - fac (max or min of uinf_d over max |uinf|)^2, fac = min(1, uMax_measured dt / h); the
  // kernel applies them on open faces only.
  Fade6<T> fd;
  {
    double fac = fmin(1.0, c.umax_measured * c.host_dt / c.h);
    double nor = 2.220446049250313e-16;
    for (int a = 0; a < 3; ++a) nor = fmax(nor, fabs(uinf[a]));
    for (int a = 0; a < 3; ++a) {
      const double lo = fmax(uinf[a], 0.0) / nor, hi = fmin(uinf[a], 0.0) / nor;
      fd.f[2 * a] = (T)(1 - fac * lo * lo);
      fd.f[2 * a + 1] = (T)(1 - fac * hi * hi);
    }
  }
  const int nbx = c.bpd[0];
  if (nbx % 8 == 0) launch_advdiff_xwalk<T, 8>(c, nu, uinf, fold, fd);
  else if (nbx % 4 == 0) launch_advdiff_xwalk<T, 4>(c, nu, uinf, fold, fd);
  else if (nbx % 2 == 0) launch_advdiff_xwalk<T, 2>(c, nu, uinf, fold, fd);
  else launch_advdiff_xwalk<T, 1>(c, nu, uinf, fold, fd);
  return check_hip(c, hipGetLastError(), "k_advdiff_xwalk");
}

// ------------------------------------------------ third-order upwind advection
// The upwind scheme (no reference counterpart; the reference's advection is the central
// kernel above): per axis d with advecting velocity a = (s + U)_d at the cell, 6h times the
// first derivative is 2f(+1) + 3f(0) - 6f(-1) + f(-2) for a > 0 and
// -f(+2) + 6f(+1) - 3f(0) - 2f(-1) otherwise; diffusion is the 7-point Laplacian.
// out = base + mult*dt * (-(s + U).grad_up3 s + nu lap s), with s read through a lab of
// halo 2 (faces only, two layers deep: 12^3 in LDS) and base read at the cell only, so
// out may alias base (the midpoint stage u* = u + dt L(u_half) updates u in place).
constexpr int L2 = 12;
constexpr int L2C = L2 * L2 * L2;

template <typename T>
__device__ __forceinline__ void gather_g2(const T* const* src, T (*lab)[L2C], const GridArgs& g, int b, int bx,
                                          int by, int bz)
{
  const int tid = threadIdx.x;
  {
    const int c = 2 * tid, ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const int l = ((iz + 2) * L2 + iy + 2) * L2 + ix + 2;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const T* q = src[f] + (size_t)b * BS3 + c;
      lab[f][l] = q[0];
      lab[f][l + 1] = q[1];
    }
  }
  for (int c = tid; c < 6 * 128; c += NT) {
    const int face = c >> 7, layer = (c >> 6) & 1, a = (c >> 3) & 7, e = c & 7;
    const bool plus = face & 1;
    const int sd = plus ? layer : 6 + layer;      // source index along the face normal
    const int ld = plus ? 10 + layer : layer;     // lab index along the normal
    int qx = bx, qy = by, qz = bz, sx, sy, sz, lx, ly, lz;
    if (face < 2) {
      qx = wrapi(bx + (plus ? 1 : -1), g.nbx); sx = sd; sy = e; sz = a; lx = ld; ly = e + 2; lz = a + 2;
    } else if (face < 4) {
      qy = wrapi(by + (plus ? 1 : -1), g.nby); sx = e; sy = sd; sz = a; lx = e + 2; ly = ld; lz = a + 2;
    } else {
      qz = wrapi(bz + (plus ? 1 : -1), g.nbz); sx = e; sy = a; sz = sd; lx = e + 2; ly = a + 2; lz = ld;
    }
    const size_t sidx = (size_t)(qx + g.nbx * (qy + g.nby * qz)) * BS3 + (sz * BS + sy) * BS + sx;
    const int l = (lz * L2 + ly) * L2 + lx;
#pragma unroll
    for (int f = 0; f < 3; ++f) lab[f][l] = src[f][sidx];
  }
}

template <typename T>
__device__ __forceinline__ T up3(T a, T m2, T m1, T c, T p1, T p2)
{
  return a > 0 ? 2 * p1 + 3 * c - 6 * m1 + m2 : -p2 + 6 * p1 - 3 * c - 2 * m1;
}

// The stage runs as a walk of NBW blocks along x with PD blocks in flight, as the central
// kernel walks: the lab pads x to U3X = 14 (interior at x = 2..9, two halo layers each side),
// y and z to L2 = 12. A block's x halos are the neighbouring blocks' ix = 6, 7 and 0, 1 planes,
// held in registers by the threads t % 4 == 3 and t % 4 == 0; the walk's two end faces (two
// layers, 256 cells: one per thread) and the four y/z faces (two layers, 512 cells: two per
// thread, one face and layer per wave) are gathered. Each thread's cell pair and all its
// stencil neighbours are 8-byte aligned pairs in LDS (one 64-bit access each). SEP: base is
// not src (read at the cells one block ahead); else base is the lab's own centre.
constexpr int U3X = 14;            // y stride of the walk's lab (x padded to 14)
constexpr int U3Z = U3X * L2;      // z stride
constexpr int U3C = U3X * L2 * L2; // cells per field
struct Up3Slots {
  int li;       // LDS index of interior cell 2t
  int l[2];     // the thread's two y/z face cells: LDS index,
  int off[2];   // offset inside the neighbour block,
  int face[2];  // face 2..5
  int lx, offx, facex;  // its walk-end x face cell (face 0 or 1)
};

__device__ __forceinline__ Up3Slots up3_slots()
{
  Up3Slots s;
  const int t = threadIdx.x, c0 = 2 * t;
  s.li = (((c0 >> 6) + 2) * L2 + ((c0 >> 3) & 7) + 2) * U3X + (c0 & 7) + 2;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int c = t + NT * k;  // 0..511: face 2 + (c >> 7), layer (c >> 6) & 1
    const int face = 2 + (c >> 7), layer = (c >> 6) & 1, a = (c >> 3) & 7, e = c & 7;
    const bool plus = face & 1;
    const int sd = plus ? layer : 6 + layer, ld = plus ? 10 + layer : layer;
    int sx = e, sy, sz, ly, lz;
    if (face < 4) { sy = sd; sz = a; ly = ld; lz = a + 2; }
    else { sy = a; sz = sd; ly = a + 2; lz = ld; }
    s.face[k] = face;
    s.off[k] = (sz * BS + sy) * BS + sx;
    s.l[k] = (lz * L2 + ly) * U3X + e + 2;
  }
  {
    const int face = t >> 7, layer = (t >> 6) & 1, a = (t >> 3) & 7, e = t & 7;
    const bool plus = face & 1;
    const int sd = plus ? layer : 6 + layer, ld = plus ? 10 + layer : layer;
    s.facex = face;
    s.offx = (a * BS + e) * BS + sd;
    s.lx = ((a + 2) * L2 + e + 2) * U3X + ld;
  }
  return s;
}

template <typename T>
struct Up3Regs {
  T in[3][2];
  T hv[3][2];
  T hx[3];
};

template <typename T>
__device__ __forceinline__ void up3_fetch(Up3Regs<T>& r, const T* const* src, const Up3Slots& s, const GridArgs& g,
                                          int b, bool x_lo, bool x_hi)
{
  using V2 = typename Vec2<T>::type;
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const V2 a = *reinterpret_cast<const V2*>(src[f] + (size_t)b * BS3 + 2 * threadIdx.x);
    r.in[f][0] = a.x;
    r.in[f][1] = a.y;
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int face = __builtin_amdgcn_readfirstlane(s.face[k]);  // one face per wave
    const size_t q = (size_t)face_block(g, bx, by, bz, face) * BS3 + s.off[k];
#pragma unroll
    for (int f = 0; f < 3; ++f) r.hv[f][k] = src[f][q];
  }
  const int fx = __builtin_amdgcn_readfirstlane(s.facex);
  if ((fx == 0 && x_lo) || (fx == 1 && x_hi)) {
    const size_t q = (size_t)face_block(g, bx, by, bz, fx) * BS3 + s.offx;
#pragma unroll
    for (int f = 0; f < 3; ++f) r.hx[f] = src[f][q];
  }
}

template <typename T>
__device__ __forceinline__ void up3_put(const Up3Regs<T>& r, const Up3Regs<T>& prev, const Up3Regs<T>& next,
                                        T (*lab)[U3C], const Up3Slots& s, bool x_lo, bool x_hi)
{
  using V2 = typename Vec2<T>::type;
#pragma unroll
  for (int f = 0; f < 3; ++f) *reinterpret_cast<V2*>(&lab[f][s.li]) = V2{r.in[f][0], r.in[f][1]};
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int f = 0; f < 3; ++f) lab[f][s.l[k]] = r.hv[f][k];
  const int fx = __builtin_amdgcn_readfirstlane(s.facex);
  if ((fx == 0 && x_lo) || (fx == 1 && x_hi))
#pragma unroll
    for (int f = 0; f < 3; ++f) lab[f][s.lx] = r.hx[f];
  const int t4 = threadIdx.x & 3;
  if (!x_lo && t4 == 3)  // the previous block's ix = 6, 7 at x = 0, 1
#pragma unroll
    for (int f = 0; f < 3; ++f) *reinterpret_cast<V2*>(&lab[f][s.li - 8]) = V2{prev.in[f][0], prev.in[f][1]};
  if (!x_hi && t4 == 0)  // the next block's ix = 0, 1 at x = 10, 11
#pragma unroll
    for (int f = 0; f < 3; ++f) *reinterpret_cast<V2*>(&lab[f][s.li + 8]) = V2{next.in[f][0], next.in[f][1]};
}

// (Without SEP the walk asks for five waves per SIMD, which costs it a few bytes of spill.)
template <typename T, int NBW, int PD, bool SEP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(SEP || sizeof(T) > 4 ? 1 : 5))) void k_advdiff_up3_xwalk(const T* __restrict__ su, const T* __restrict__ sv,
                                                          const T* __restrict__ sw, const T* bu, const T* bv,
                                                          const T* bw, T* ou, T* ov, T* ow, GridArgs g,
                                                          const DevScalars* __restrict__ sc, T Ux, T Uy, T Uz,
                                                          double nu, double mult)
{
  using V2 = typename Vec2<T>::type;
  __shared__ __align__(16) T lab[3][U3C];
  __shared__ __align__(16) T bl[SEP ? 3 : 1][BS3];  // the block's base (SEP)
  const Up3Slots s = up3_slots();
  const T* src[3] = {su, sv, sw};
  const T* base[3] = {bu, bv, bw};
  const int b0 = xwalk_first<NBW>(g, xcd_remap(blockIdx.x, gridDim.x));
  static_assert(PD >= 2, "a block's x+ halo comes from the next block, which must be in registers");
  constexpr int NS = PD + 1;
  Up3Regs<T> r[NS];
#pragma unroll
  for (int p = 0; p < PD && p < NBW; ++p) up3_fetch(r[p], src, s, g, b0 + p, p == 0, p == NBW - 1);
  const double fac = mult * sc->dt;
  const T facA = (T)(-fac / (6 * g.h));
  const T facD = (T)(fac * nu / (g.h * g.h));
  const int l = s.li;
  // base at the cells (SEP), loaded one block ahead and staged in LDS with the lab
  V2 bnext[SEP ? 3 : 1];
  if (SEP)
#pragma unroll
    for (int q = 0; q < 3; ++q) load2_nt(base[q] + (size_t)b0 * BS3 + 2 * threadIdx.x, bnext[q].x, bnext[q].y);
#pragma unroll
  for (int j = 0; j < NBW; ++j) {
    if (j) __syncthreads();  // the previous block's compute is done with the lab
    up3_put(r[j % NS], r[(j + NS - 1) % NS], r[(j + 1) % NS], lab, s, j == 0, j == NBW - 1);
    if (SEP)
#pragma unroll
      for (int q = 0; q < 3; ++q) *reinterpret_cast<V2*>(&bl[SEP ? q : 0][2 * threadIdx.x]) = bnext[q];
    __syncthreads();
    if (j + PD < NBW) up3_fetch(r[(j + PD) % NS], src, s, g, b0 + j + PD, false, j + PD == NBW - 1);
    const size_t o = (size_t)(b0 + j) * BS3 + 2 * threadIdx.x;
    if (SEP && j + 1 < NBW)
#pragma unroll
      for (int q = 0; q < 3; ++q) load2_nt(base[q] + o + BS3, bnext[q].x, bnext[q].y);
    T a0[2], a1[2], a2[2];
    {
      const V2 A = *reinterpret_cast<const V2*>(&lab[0][l]), B = *reinterpret_cast<const V2*>(&lab[1][l]),
               C = *reinterpret_cast<const V2*>(&lab[2][l]);
      a0[0] = A.x + Ux; a0[1] = A.y + Ux;
      a1[0] = B.x + Uy; a1[1] = B.y + Uy;
      a2[0] = C.x + Uz; a2[1] = C.y + Uz;
    }
    // one field at a time (not unrolled: the stencil's 22 values per field stay in registers
    // once, which keeps the walk's in-flight blocks within the register budget)
#pragma unroll 1
    for (int q = 0; q < 3; ++q) {
      const T* F = lab[q];
      const V2 xm = *reinterpret_cast<const V2*>(F + l - 2), xc = *reinterpret_cast<const V2*>(F + l),
               xp = *reinterpret_cast<const V2*>(F + l + 2);
      const V2 ym2 = *reinterpret_cast<const V2*>(F + l - 2 * U3X), ym1 = *reinterpret_cast<const V2*>(F + l - U3X),
               yp1 = *reinterpret_cast<const V2*>(F + l + U3X), yp2 = *reinterpret_cast<const V2*>(F + l + 2 * U3X);
      const V2 zm2 = *reinterpret_cast<const V2*>(F + l - 2 * U3Z), zm1 = *reinterpret_cast<const V2*>(F + l - U3Z),
               zp1 = *reinterpret_cast<const V2*>(F + l + U3Z), zp2 = *reinterpret_cast<const V2*>(F + l + 2 * U3Z);
      // x neighbours of cell 0: xm.x xm.y | xc.x | xc.y xp.x; of cell 1: xm.y xc.x | xc.y | xp.x xp.y
      const T X[6] = {xm.x, xm.y, xc.x, xc.y, xp.x, xp.y};
      const T Ym2[2] = {ym2.x, ym2.y}, Ym1[2] = {ym1.x, ym1.y}, Yp1[2] = {yp1.x, yp1.y}, Yp2[2] = {yp2.x, yp2.y};
      const T Zm2[2] = {zm2.x, zm2.y}, Zm1[2] = {zm1.x, zm1.y}, Zp1[2] = {zp1.x, zp1.y}, Zp2[2] = {zp2.x, zp2.y};
      T res[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const T fc = X[2 + k];
        // a D(a) per axis without a branch: with S = D+ + D- and T = D+ - D- (up3's two sides),
        // a D(a) = (a S + |a| T) / 2 for either sign; m1 + p1 is shared with the Laplacian
        T pm[3], adv2 = 0;
        {
          const T m2[3] = {X[k], Ym2[k], Zm2[k]}, m1[3] = {X[k + 1], Ym1[k], Zm1[k]},
                  p1[3] = {X[k + 3], Yp1[k], Zp1[k]}, p2[3] = {X[k + 4], Yp2[k], Zp2[k]};
          const T a[3] = {a0[k], a1[k], a2[k]};
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            pm[d] = m1[d] + p1[d];
            const T S = 8 * (p1[d] - m1[d]) + (m2[d] - p2[d]);
            const T Tq = (m2[d] + p2[d]) - 4 * pm[d] + 6 * fc;
            adv2 += a[d] * S + fabs(a[d]) * Tq;
          }
        }
        const T lap = pm[0] + pm[1] + pm[2] - 6 * fc;
        res[k] = (facA * (T)0.5) * adv2 + facD * lap;
      }
      const V2 bb = SEP ? *reinterpret_cast<const V2*>(&bl[SEP ? q : 0][2 * threadIdx.x]) : xc;
      T* dst = q == 0 ? ou : (q == 1 ? ov : ow);
      store2_nt(dst + o, bb.x + res[0], bb.y + res[1]);
    }
  }
}

template <typename T, int NBW>
static void up3_walk_launch(Ctx& c, int src0, int base0, int out0, double nu, const double uinf[3], double mult)
{
  constexpr int PD = 2;
  const bool sep = base0 != src0;
  auto args = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(c.nb / NBW), dim3(NT), 0, c.stream, (const T*)c.f[src0], (const T*)c.f[src0 + 1],
                       (const T*)c.f[src0 + 2], (const T*)c.f[base0], (const T*)c.f[base0 + 1],
                       (const T*)c.f[base0 + 2], (T*)c.f[out0], (T*)c.f[out0 + 1], (T*)c.f[out0 + 2], c.grid(), c.sc,
                       (T)uinf[0], (T)uinf[1], (T)uinf[2], nu, mult);
  };
  if (sep) args(k_advdiff_up3_xwalk<T, NBW, PD, true>);
  else args(k_advdiff_up3_xwalk<T, NBW, PD, false>);
}

template <typename T>
static int up3_stage(Ctx& c, int src0, int base0, int out0, double nu, const double uinf[3], double mult)
{
  const int nbx = c.bpd[0];
  if (nbx % 8 == 0) up3_walk_launch<T, 8>(c, src0, base0, out0, nu, uinf, mult);
  else if (nbx % 4 == 0) up3_walk_launch<T, 4>(c, src0, base0, out0, nu, uinf, mult);
  else if (nbx % 2 == 0) up3_walk_launch<T, 2>(c, src0, base0, out0, nu, uinf, mult);
  else up3_walk_launch<T, 1>(c, src0, base0, out0, nu, uinf, mult);
  return check_hip(c, hipGetLastError(), "k_advdiff_up3_xwalk");
}

// Advection-diffusion of the scheme into the step's intermediate velocity u*: the central
// and upwind Euler forms write u* to tmp; the midpoint form writes u_half to tmp and then u*
// over u. c.ustar names the field set holding u* afterwards (the fused step's consumers read
// it; the plain decomposition copies it so that u = tmp = u* as after the reference's copy).
int launch_advdiff(Ctx& c, double nu, const double uinf[3], int scheme, const StepFold* fold)
{
  c.ustar = CUP_TMPU;
  if (scheme == CUP_CENTRAL_EULER)
    return c.precision == 4 ? advdiff_t<float>(c, nu, uinf, fold) : advdiff_t<double>(c, nu, uinf, fold);
  if (scheme != CUP_UPWIND3_EULER && scheme != CUP_UPWIND3_RK2)
    return set_error(c, CUP_E_INVALID, "advect_diffuse: unknown scheme");
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "advect_diffuse: the upwind schemes need a periodic box");
  if (fold) return set_error(c, CUP_E_STATE, "advect_diffuse: the folded step bookkeeping takes the central scheme");
  auto stage = [&](int src0, int base0, int out0, double mult) {
    return c.precision == 4 ? up3_stage<float>(c, src0, base0, out0, nu, uinf, mult)
                            : up3_stage<double>(c, src0, base0, out0, nu, uinf, mult);
  };
  if (scheme == CUP_UPWIND3_EULER) return stage(CUP_U, CUP_U, CUP_TMPU, 1.0);
  if (int rc = stage(CUP_U, CUP_U, CUP_TMPU, 0.5)) return rc;
  if (int rc = stage(CUP_TMPU, CUP_U, CUP_U, 1.0)) return rc;
  c.ustar = CUP_U;
  return CUP_OK;
}

// ------------------------------------------------------------- inflow correction
// UpdateAndCorrectInflow::operate (AdvectionDiffusion.cpp:209-296) after the copy tmp -> u:
// the net outflow sum_faces (+-) u_n over the boundary cells of the open domain faces (neither
// periodic nor wall), then corr = sum / (2 (NxNy + NxNz + NyNz)) is added to the inflow
// faces' and removed from the outflow faces' normal velocity, unless |corr| < DBL_EPSILON.
// One workgroup per block; blocks off the open faces return at once.
__device__ __forceinline__ bool open_face(const GridArgs& g, int bx, int by, int bz, int d, int side)
{
  if (g.bc[d] != CUP_BC_OPEN) return false;
  const int bi = d == 0 ? bx : (d == 1 ? by : bz), nbk = d == 0 ? g.nbx : (d == 1 ? g.nby : g.nbz);
  return side ? bi == nbk - 1 : bi == 0;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_inflow_sum(const T* __restrict__ u, const T* __restrict__ v,
                                                   const T* __restrict__ w, GridArgs g, double* __restrict__ part)
{
  const int b = blockIdx.x;
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
  const T* f[3] = {u, v, w};
  double acc[1] = {0.0};
  for (int d = 0; d < 3; ++d)
    for (int side = 0; side < 2; ++side) {
      if (!open_face(g, bx, by, bz, d, side)) continue;
      if (threadIdx.x < BS * BS) {
        int id[3];
        id[d] = side ? BS - 1 : 0;
        id[(d + 1) % 3] = threadIdx.x & 7;
        id[(d + 2) % 3] = threadIdx.x >> 3;
        const double un = (double)f[d][(size_t)b * BS3 + (id[2] * BS + id[1]) * BS + id[0]];
        acc[0] += side ? un : -un;
      }
    }
  block_reduce<1>(acc, 0u, part + b);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_inflow_apply(T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                                                     GridArgs g, const DevScalars* __restrict__ sc)
{
  const int b = blockIdx.x;
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
  const double nx = g.Nx, ny = g.Ny, nz = g.Nz;
  const double corr = sc->red[0] / (2 * (nx * ny + nx * nz + ny * nz));
  if (fabs(corr) < 2.220446049250313e-16) return;
  T* f[3] = {u, v, w};
  for (int d = 0; d < 3; ++d)
    for (int side = 0; side < 2; ++side) {
      if (!open_face(g, bx, by, bz, d, side) || threadIdx.x >= BS * BS) continue;
      int id[3];
      id[d] = side ? BS - 1 : 0;
      id[(d + 1) % 3] = threadIdx.x & 7;
      id[(d + 2) % 3] = threadIdx.x >> 3;
      T& un = f[d][(size_t)b * BS3 + (id[2] * BS + id[1]) * BS + id[0]];
      un = (T)((double)un + (side ? -corr : corr));
    }
}

template <typename T>
static int inflow_t(Ctx& c)
{
  hipLaunchKernelGGL((k_inflow_sum<T>), dim3(c.nb), dim3(NT), 0, c.stream, (const T*)c.f[CUP_U],
                     (const T*)c.f[CUP_V], (const T*)c.f[CUP_W], c.grid(), c.partials);
  if (int rc = launch_reduce(c, c.nb, 1, 0u, c.sc->red)) return rc;
  hipLaunchKernelGGL((k_inflow_apply<T>), dim3(c.nb), dim3(NT), 0, c.stream, (T*)c.f[CUP_U], (T*)c.f[CUP_V],
                     (T*)c.f[CUP_W], c.grid(), c.sc);
  return check_hip(c, hipGetLastError(), "inflow correction");
}

int launch_inflow_correction(Ctx& c)
{
  if (c.bc[0] != CUP_BC_OPEN && c.bc[1] != CUP_BC_OPEN && c.bc[2] != CUP_BC_OPEN) return CUP_OK;
  if (int e = ensure_partials(c, (size_t)c.nb)) return e;
  return c.precision == 4 ? inflow_t<float>(c) : inflow_t<double>(c);
}

// -------------------------------------------------------------------- copies, fills

template <typename T>
__global__ void k_copy3(T* __restrict__ a, T* __restrict__ b, T* __restrict__ cc, const T* __restrict__ sa,
                        const T* __restrict__ sb, const T* __restrict__ scc, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    a[i] = sa[i];
    b[i] = sb[i];
    cc[i] = scc[i];
  }
}

template <typename T>
__global__ void k_copy1(T* __restrict__ a, const T* __restrict__ sa, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = sa[i];
}

template <typename T>
__global__ void k_fill(T* __restrict__ a, size_t n, T value)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = value;
}

static dim3 ew_grid(size_t n)
{
  size_t g = (n + NT - 1) / NT;
  if (g > 65536) g = 65536;
  return dim3((unsigned)g);
}

template <typename T>
__global__ void k_add_const(T* __restrict__ a, size_t n, double value)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = (T)((double)a[i] + value);
}

// KernelExternalForcing (ExternalForcing.cpp:17-31): du added to every cell of u, and of u* when
// the advection left u* apart from u.
int launch_add_uniform(Ctx& c, const double du[3])
{
  for (int q = 0; q < 3; ++q) {
    if (du[q] == 0) continue;
    const int fields[2] = {CUP_U + q, c.ustar + q};
    for (int k = 0; k < (c.ustar == CUP_U ? 1 : 2); ++k) {
      if (c.precision == 4)
        hipLaunchKernelGGL((k_add_const<float>), ew_grid(c.ncell), dim3(NT), 0, c.stream, (float*)c.f[fields[k]],
                           c.ncell, du[q]);
      else
        hipLaunchKernelGGL((k_add_const<double>), ew_grid(c.ncell), dim3(NT), 0, c.stream, (double*)c.f[fields[k]],
                           c.ncell, du[q]);
    }
  }
  return check_hip(c, hipGetLastError(), "k_add_const");
}

// u -> tmp (after the midpoint scheme, which leaves u* in u).
int launch_copy_u_to_tmp(Ctx& c)
{
  if (c.precision == 4)
    hipLaunchKernelGGL((k_copy3<float>), ew_grid(c.ncell), dim3(NT), 0, c.stream, (float*)c.f[CUP_TMPU],
                       (float*)c.f[CUP_TMPV], (float*)c.f[CUP_TMPW], (const float*)c.f[CUP_U],
                       (const float*)c.f[CUP_V], (const float*)c.f[CUP_W], c.ncell);
  else
    hipLaunchKernelGGL((k_copy3<double>), ew_grid(c.ncell), dim3(NT), 0, c.stream, (double*)c.f[CUP_TMPU],
                       (double*)c.f[CUP_TMPV], (double*)c.f[CUP_TMPW], (const double*)c.f[CUP_U],
                       (const double*)c.f[CUP_V], (const double*)c.f[CUP_W], c.ncell);
  return check_hip(c, hipGetLastError(), "k_copy3");
}

int launch_copy_tmp_to_u(Ctx& c)
{
  if (c.precision == 4)
    hipLaunchKernelGGL((k_copy3<float>), ew_grid(c.ncell), dim3(NT), 0, c.stream, (float*)c.f[CUP_U],
                       (float*)c.f[CUP_V], (float*)c.f[CUP_W], (const float*)c.f[CUP_TMPU],
                       (const float*)c.f[CUP_TMPV], (const float*)c.f[CUP_TMPW], c.ncell);
  else
    hipLaunchKernelGGL((k_copy3<double>), ew_grid(c.ncell), dim3(NT), 0, c.stream, (double*)c.f[CUP_U],
                       (double*)c.f[CUP_V], (double*)c.f[CUP_W], (const double*)c.f[CUP_TMPU],
                       (const double*)c.f[CUP_TMPV], (const double*)c.f[CUP_TMPW], c.ncell);
  return check_hip(c, hipGetLastError(), "k_copy3");
}

int launch_copy(Ctx& c, void* dst, const void* src, size_t n)
{
  if (c.precision == 4)
    hipLaunchKernelGGL((k_copy1<float>), ew_grid(n), dim3(NT), 0, c.stream, (float*)dst, (const float*)src, n);
  else
    hipLaunchKernelGGL((k_copy1<double>), ew_grid(n), dim3(NT), 0, c.stream, (double*)dst, (const double*)src, n);
  return check_hip(c, hipGetLastError(), "k_copy1");
}

int launch_fill(Ctx& c, void* ptr, size_t n, double value)
{
  if (c.precision == 4)
    hipLaunchKernelGGL((k_fill<float>), ew_grid(n), dim3(NT), 0, c.stream, (float*)ptr, n, (float)value);
  else
    hipLaunchKernelGGL((k_fill<double>), ew_grid(n), dim3(NT), 0, c.stream, (double*)ptr, n, value);
  return check_hip(c, hipGetLastError(), "k_fill");
}

__global__ void k_set_dt(DevScalars* sc, double dt) { sc->dt = dt; }

int launch_set_dt(Ctx& c, double dt)
{
  c.host_dt = dt;
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c.stream, c.sc, dt);
  return check_hip(c, hipGetLastError(), "k_set_dt");
}

// ----------------------------------------------------------------- pressure RHS

// KernelPressureRHS (PressureRHS.cpp:51-65): h^2/(2dt) * (u_E-u_W + v_N-v_S + w_B-w_F),
// written to a block-layout buffer (the solver's input).
template <typename T>
__global__ __launch_bounds__(NT) void k_rhs(const T* __restrict__ u, const T* __restrict__ v,
                                            const T* __restrict__ w, T* __restrict__ rhs, GridArgs g,
                                            const DevScalars* __restrict__ sc)
{
  __shared__ T lab[3][L1C];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
  const T* src[3] = {u, v, w};
  gather_g1<T, 3>(src, lab, g, b, bx, by, bz, {HK_VEL, HK_VEL, HK_VEL});
  __syncthreads();
  const T fac = (T)(0.5 * g.h * g.h / sc->dt);
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const int l = ((iz + 1) * L1 + iy + 1) * L1 + ix + 1;
    rhs[(size_t)b * BS3 + c] = fac * (lab[0][l + 1] - lab[0][l - 1] + lab[1][l + L1] - lab[1][l - L1] +
                                      lab[2][l + L1 * L1] - lab[2][l - L1 * L1]);
  }
}

int launch_rhs_plain(Ctx& c)
{
  const int src0 = c.decomposition == CUP_FUSED ? c.ustar : CUP_U;
  if (c.precision == 4)
    hipLaunchKernelGGL((k_rhs<float>), dim3(c.nb), dim3(NT), 0, c.stream, (const float*)c.f[src0],
                       (const float*)c.f[src0 + 1], (const float*)c.f[src0 + 2], (float*)c.rbuf, c.grid(), c.sc);
  else
    hipLaunchKernelGGL((k_rhs<double>), dim3(c.nb), dim3(NT), 0, c.stream, (const double*)c.f[src0],
                       (const double*)c.f[src0 + 1], (const double*)c.f[src0 + 2], (double*)c.rbuf, c.grid(), c.sc);
  return check_hip(c, hipGetLastError(), "k_rhs");
}

// --------------------------------------------------------------- pressure gradient

// lab_fetch of p on a periodic grid with the x faces taken from the x inverse pass's face rows
// (xf[(b*2 + side)*64 + iz*8 + iy], side 1 = ix 7): a wave reads its face as 64 contiguous values.
// The x faces are the halo slots k = 0 of waves 0 (x-) and 1 (x+), where the slot's (sz, sy) is
// (c >> 3 & 7, c & 7) with c = t, i.e. iz*8 + iy = t & 63.
template <typename T>
__device__ __forceinline__ void lab_fetch_xf(LabRegs<T, 1>& r, const T* __restrict__ p, const T* __restrict__ xf,
                                             const LabSlots& s, const GridArgs& g, int b)
{
  using V2 = typename Vec2<T>::type;
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
  const V2 a = *reinterpret_cast<const V2*>(p + (size_t)b * BS3 + 2 * threadIdx.x);
  r.in[0][0] = a.x;
  r.in[0][1] = a.y;
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k == 0 || has_halo2()) {
      const int face = __builtin_amdgcn_readfirstlane(s.face[k]);
      const size_t nb = (size_t)face_block(g, bx, by, bz, face);
      if (face < 2) r.hv[0][k] = xf[(nb * 2 + (face == 0)) * 64 + (threadIdx.x & 63)];
      else r.hv[0][k] = p[nb * BS3 + s.off[k]];
    }
}

// KernelGradP (PressureProjection.cpp:38-50): u += -dt/(2h) (p_E - p_W), etc.
// PLAIN: in place on u (which holds u* after the advection copy).
// FUSED: u = tmp + fac*grad p, plus the step's reductions: max |u+uinf| over the
// workgroup's blocks folded into the device scalar by one atomic max (order independent, so
// deterministic), and each wave's sum |u|^2 of each block into ke_part[4b + wave]. Blocks a
// penalized body covers (skip_slot >= 0) leave the max to the penalization kernel, which
// changes their velocity afterwards.
// Each thread owns two x-adjacent cells c = 2t, 2t+1 of the block (vector loads/stores);
// a workgroup walks nbw blocks stacked in z, fetching the next block's p lab, tmp velocity
// and forcing rows while the current one is computed.
// FORCE adds the fused forcing's velocity modes, fac*u_low = R0 + 2 Re(R1 e^{i a x}) +
// 2 Re(R2 e^{2 i a x}) per component with the row coefficients R of k_forcing_rows; the
// block's 64 rows x 16 coefficients are staged in LDS by one coalesced 4-value load per
// thread. A thread's x is the same in every block of its walk, so its e^{-i a x} and
// e^{-2 i a x} are formed once.
template <typename T, bool FUSED, bool FORCE, bool BC, bool XF = false>
__global__ __launch_bounds__(NT) void k_gradp(const T* __restrict__ p, const T* __restrict__ su,
                                              const T* __restrict__ sv, const T* __restrict__ sw, T* __restrict__ u,
                                              T* __restrict__ v, T* __restrict__ w, GridArgs g,
                                              DevScalars* __restrict__ sc, double Ux, double Uy, double Uz,
                                              double* __restrict__ ke_part, const int* __restrict__ skip_slot,
                                              const T* __restrict__ rows, const C2<T>* __restrict__ twx, int nbw,
                                              unsigned long long* __restrict__ slots,
                                              unsigned long long* __restrict__ clear_slots, const T* __restrict__ xf)
{
  using V2 = typename Vec2<T>::type;
  using V4 = typename Vec4<T>::type;
  using LG = LabGeom<PX_GRAD>;
  __shared__ __align__(16) T lab[1][LG::PLC];
  __shared__ __align__(16) T rl[FORCE ? 64 * 16 : 4];
  const int t = threadIdx.x;
  const LabSlots s = lab_slots<PX_GRAD>();
  const T* src[1] = {p};
  const int nxy = g.nbx * g.nby;
  const int b0 = walk_first(g, xcd_remap(blockIdx.x, gridDim.x), nbw);
  const int c0 = 2 * t, ix0 = c0 & 7, iy = (c0 >> 3) & 7, iz = c0 >> 6;
  LabRegs<T, 1> r;
  T us[3][2];
  T rr[4];
  auto fetch = [&](int b) {
    if (XF) lab_fetch_xf(r, p, xf, s, g, b);
    else lab_fetch<BC>(r, src, s, g, b, {HK_COPY});
    const size_t o = (size_t)b * BS3 + c0;
    load2_nt(su + o, us[0][0], us[0][1]);
    load2_nt(sv + o, us[1][0], us[1][1]);
    load2_nt(sw + o, us[2][0], us[2][1]);
    if (FORCE) {
      // rows (y = by*8 + iy', z = bz*8 + t/32): 8 runs of 128 contiguous values
      int bx, by, bz;
      block_coords(g, b, bx, by, bz);
      const T* q = rows + ((size_t)(bz * BS + (t >> 5)) * g.Ny + by * BS) * 16 + (t & 31) * 4;
      const V2 q0 = *reinterpret_cast<const V2*>(q), q1 = *reinterpret_cast<const V2*>(q + 2);
      rr[0] = q0.x; rr[1] = q0.y; rr[2] = q1.x; rr[3] = q1.y;
    }
  };
  fetch(b0);
  C2<T> e1[2], e2[2];
  if (FORCE) {
    const int x0 = (b0 % g.nbx) * BS + ix0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      e1[k] = twx[x0 + k];                                                     // e^{-i a x}
      e2[k] = C2<T>{e1[k].x * e1[k].x - e1[k].y * e1[k].y, 2 * e1[k].x * e1[k].y};  // e^{-2 i a x}
    }
  }
  const T fac = (T)(-0.5 * sc->dt / g.h);
  double mx = 0.0;
  for (int j = 0; j < nbw; ++j) {
    const int b = b0 + j * nxy;
    if (j) __syncthreads();  // the previous block is done with the lab
    lab_put<PX_GRAD>(r, lab, s);
    if (FORCE) *reinterpret_cast<V4*>(rl + 4 * t) = V4{rr[0], rr[1], rr[2], rr[3]};
    T nu_[2] = {us[0][0], us[0][1]}, nv_[2] = {us[1][0], us[1][1]}, nw_[2] = {us[2][0], us[2][1]};
    __syncthreads();
    if (j + 1 < nbw) fetch(b + nxy);
    {
      const T* P = lab[0];
      const int l = s.li;
      const V2 C = *reinterpret_cast<const V2*>(P + l);
      const V2 Nn = *reinterpret_cast<const V2*>(P + l + LG::PY), Sn = *reinterpret_cast<const V2*>(P + l - LG::PY);
      const V2 Bn = *reinterpret_cast<const V2*>(P + l + LG::PZ), Fn = *reinterpret_cast<const V2*>(P + l - LG::PZ);
      const T Wx = P[l - 1], Ex = P[l + 2];
      nu_[0] += fac * (C.y - Wx);
      nu_[1] += fac * (Ex - C.x);
      nv_[0] += fac * (Nn.x - Sn.x);
      nv_[1] += fac * (Nn.y - Sn.y);
      nw_[0] += fac * (Bn.x - Fn.x);
      nw_[1] += fac * (Bn.y - Fn.y);
    }
    if (FORCE) {
      // the row's 15 coefficients (3 components x {R0, Re R1, Im R1, Re R2, Im R2}) as four
      // 16-byte LDS reads
      const V4* R4 = reinterpret_cast<const V4*>(rl + (iz * BS + iy) * 16);
      const V4 A = R4[0], B = R4[1], Cc = R4[2], D = R4[3];
      const T Q[16] = {A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w, Cc.x, Cc.y, Cc.z, Cc.w, D.x, D.y, D.z, D.w};
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        // Re(R e^{+i a x}) with e^{+i a x} = conj(e1)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const T* R = Q + 5 * q;
          const T add = R[0] + 2 * (R[1] * e1[k].x + R[2] * e1[k].y) + 2 * (R[3] * e2[k].x + R[4] * e2[k].y);
          if (q == 0) nu_[k] += add;
          else if (q == 1) nv_[k] += add;
          else nw_[k] += add;
        }
      }
    }
    const size_t o = (size_t)b * BS3 + c0;
    *reinterpret_cast<V2*>(u + o) = V2{nu_[0], nu_[1]};
    *reinterpret_cast<V2*>(v + o) = V2{nv_[0], nv_[1]};
    *reinterpret_cast<V2*>(w + o) = V2{nw_[0], nw_[1]};
    if (FUSED) {
      double ke = 0.0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double aa = fabs((double)nu_[k] + Ux), b2 = fabs((double)nv_[k] + Uy), c2 = fabs((double)nw_[k] + Uz);
        if (!skip_slot || skip_slot[b] < 0) mx = fmax(mx, fmax(aa, fmax(b2, c2)));
        ke += (double)nu_[k] * nu_[k] + (double)nv_[k] * nv_[k] + (double)nw_[k] * nw_[k];
      }
      ke = wave_sum(ke);
      if ((t & 63) == 0) ke_part[4 * (size_t)b + (t >> 6)] = ke;
    }
  }
  if (FUSED) {
    double m[1] = {mx};
    __shared__ double out[1];
    block_reduce<1>(m, 1u, out);
    if (t == 0) atomicMax(&slots[b0 & (UMAX_SLOTS - 1)], (unsigned long long)__double_as_longlong(out[0]));
    // the other bank, read by this step's advection, is cleared for the next step's gradient
    if (blockIdx.x == 0) clear_slots[t] = 0ull;
  }
}

template <typename T>
static int gradp_t(Ctx& c, bool fused, const double uinf[3], const int* skip, bool force, bool xface = false)
{
  const T* p = (const T*)c.f[CUP_P];
  const int s = fused ? c.ustar : CUP_U;
  const C2<T>* tx = (const C2<T>*)c.tw[0];
  const int nbw = blocks_per_wg(c.bpd[2]);
  const dim3 grid(c.nb / nbw);
  const T* xf = (const T*)c.xface;
  if (fused && force && xface)
    hipLaunchKernelGGL((k_gradp<T, true, true, false, true>), grid, dim3(NT), 0, c.stream, p, (const T*)c.f[s],
                       (const T*)c.f[s + 1], (const T*)c.f[s + 2], (T*)c.f[CUP_U], (T*)c.f[CUP_V], (T*)c.f[CUP_W],
                       c.grid(), c.sc, uinf[0], uinf[1], uinf[2], c.ke_part, skip, (const T*)c.low_rows, tx, nbw,
                       umax_slots_dev(c, c.umax_bank), umax_slots_dev(c, c.umax_bank ^ 1), xf);
  else if (fused && xface)
    hipLaunchKernelGGL((k_gradp<T, true, false, false, true>), grid, dim3(NT), 0, c.stream, p, (const T*)c.f[s],
                       (const T*)c.f[s + 1], (const T*)c.f[s + 2], (T*)c.f[CUP_U], (T*)c.f[CUP_V], (T*)c.f[CUP_W],
                       c.grid(), c.sc, uinf[0], uinf[1], uinf[2], c.ke_part, skip, (const T*)nullptr, tx, nbw,
                       umax_slots_dev(c, c.umax_bank), umax_slots_dev(c, c.umax_bank ^ 1), xf);
  else if (fused && force)
    hipLaunchKernelGGL((k_gradp<T, true, true, false>), grid, dim3(NT), 0, c.stream, p, (const T*)c.f[s],
                       (const T*)c.f[s + 1], (const T*)c.f[s + 2], (T*)c.f[CUP_U], (T*)c.f[CUP_V], (T*)c.f[CUP_W],
                       c.grid(), c.sc, uinf[0], uinf[1], uinf[2], c.ke_part, skip, (const T*)c.low_rows, tx, nbw,
                       umax_slots_dev(c, c.umax_bank), umax_slots_dev(c, c.umax_bank ^ 1), nullptr);
  else if (fused)
    hipLaunchKernelGGL((k_gradp<T, true, false, false>), grid, dim3(NT), 0, c.stream, p, (const T*)c.f[s],
                       (const T*)c.f[s + 1], (const T*)c.f[s + 2], (T*)c.f[CUP_U], (T*)c.f[CUP_V], (T*)c.f[CUP_W],
                       c.grid(), c.sc, uinf[0], uinf[1], uinf[2], c.ke_part, skip, (const T*)nullptr, tx, nbw,
                       umax_slots_dev(c, c.umax_bank), umax_slots_dev(c, c.umax_bank ^ 1), nullptr);
  else if (c.periodic())
    hipLaunchKernelGGL((k_gradp<T, false, false, false>), grid, dim3(NT), 0, c.stream, p, (const T*)c.f[s],
                       (const T*)c.f[s + 1], (const T*)c.f[s + 2], (T*)c.f[CUP_U], (T*)c.f[CUP_V], (T*)c.f[CUP_W],
                       c.grid(), c.sc, 0.0, 0.0, 0.0, (double*)nullptr, (const int*)nullptr, (const T*)nullptr, tx,
                       nbw, (unsigned long long*)nullptr, (unsigned long long*)nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_gradp<T, false, false, true>), grid, dim3(NT), 0, c.stream, p, (const T*)c.f[s],
                       (const T*)c.f[s + 1], (const T*)c.f[s + 2], (T*)c.f[CUP_U], (T*)c.f[CUP_V], (T*)c.f[CUP_W],
                       c.grid(), c.sc, 0.0, 0.0, 0.0, (double*)nullptr, (const int*)nullptr, (const T*)nullptr, tx,
                       nbw, (unsigned long long*)nullptr, (unsigned long long*)nullptr, nullptr);
  return check_hip(c, hipGetLastError(), "k_gradp");
}

// Gradient step in place on the three fields from field0 (no reductions): the adjoint
// projection's last pass.
int launch_gradp_inplace(Ctx& c, int field0)
{
  const int nbw = blocks_per_wg(c.bpd[2]);
  const dim3 grid(c.nb / nbw);
  if (c.precision == 4) {
    float* f[3] = {(float*)c.f[field0], (float*)c.f[field0 + 1], (float*)c.f[field0 + 2]};
    hipLaunchKernelGGL((k_gradp<float, false, false, false>), grid, dim3(NT), 0, c.stream, (const float*)c.f[CUP_P], f[0],
                       f[1], f[2], f[0], f[1], f[2], c.grid(), c.sc, 0.0, 0.0, 0.0, (double*)nullptr,
                       (const int*)nullptr, (const float*)nullptr, (const C2<float>*)c.tw[0], nbw,
                       (unsigned long long*)nullptr, (unsigned long long*)nullptr, nullptr);
  } else {
    double* f[3] = {(double*)c.f[field0], (double*)c.f[field0 + 1], (double*)c.f[field0 + 2]};
    hipLaunchKernelGGL((k_gradp<double, false, false, false>), grid, dim3(NT), 0, c.stream, (const double*)c.f[CUP_P],
                       f[0], f[1], f[2], f[0], f[1], f[2], c.grid(), c.sc, 0.0, 0.0, 0.0, (double*)nullptr,
                       (const int*)nullptr, (const double*)nullptr, (const C2<double>*)c.tw[0], nbw,
                       (unsigned long long*)nullptr, (unsigned long long*)nullptr, nullptr);
  }
  return check_hip(c, hipGetLastError(), "k_gradp");
}

// ------------------------------------------------------------------------ adjoint
// Transposed linearisation of the central advection-diffusion step
// A(u)_q = u_q + facA sum_d a_d (u_q(+e_d) - u_q(-e_d)) + facD lap u_q, a = u + U
// (AdvectionDiffusion.cpp:106-145), applied to the adjoint lam* (no reference counterpart):
// out_r = lam*_r + facD lap lam*_r - facA sum_d [(a_d lam*_r)(+e_d) - (a_d lam*_r)(-e_d)]
//         + facA sum_q lam*_q (u_q(+e_r) - u_q(-e_r)).
// One lab of six fields (u, v, w of the forward state and the three adjoint components).
template <typename T>
__global__ __launch_bounds__(NT) void k_adj_advdiff(const T* __restrict__ u, const T* __restrict__ v,
                                                    const T* __restrict__ w, const T* __restrict__ lu,
                                                    const T* __restrict__ lv, const T* __restrict__ lw,
                                                    T* __restrict__ ou, T* __restrict__ ov, T* __restrict__ ow,
                                                    GridArgs g, const DevScalars* __restrict__ sc, T Ux, T Uy, T Uz,
                                                    double nu)
{
  __shared__ T lab[6][L1C];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
  const T* src[6] = {u, v, w, lu, lv, lw};
  gather_g1<T, 6>(src, lab, g, b, bx, by, bz, {HK_VEL, HK_VEL, HK_VEL, HK_VEL, HK_VEL, HK_VEL});
  __syncthreads();
  const double dt = sc->dt;
  const T facA = (T)(-dt / (2 * g.h));
  const T facD = (T)((nu / g.h) * (dt / g.h));
  const T Uc[3] = {Ux, Uy, Uz};
  constexpr int st[3] = {1, L1, L1 * L1};
  const int c0 = 2 * threadIdx.x, ix0 = c0 & 7, iy = (c0 >> 3) & 7, iz = c0 >> 6;
  T r[3][2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int l = ((iz + 1) * L1 + iy + 1) * L1 + ix0 + k + 1;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T* Lq = lab[3 + q];
      const T lc = Lq[l];
      T lap = 0, adv = 0, srcs = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T* A = lab[d];
        lap += Lq[l - st[d]] + Lq[l + st[d]] - 2 * lc;
        adv += (A[l + st[d]] + Uc[d]) * Lq[l + st[d]] - (A[l - st[d]] + Uc[d]) * Lq[l - st[d]];
      }
#pragma unroll
      for (int p = 0; p < 3; ++p) srcs += lab[3 + p][l] * (lab[p][l + st[q]] - lab[p][l - st[q]]);
      r[q][k] = lc + facD * lap - facA * adv + facA * srcs;
    }
  }
  using V2 = typename Vec2<T>::type;
  const size_t o = (size_t)b * BS3 + c0;
  *reinterpret_cast<V2*>(ou + o) = V2{r[0][0], r[0][1]};
  *reinterpret_cast<V2*>(ov + o) = V2{r[1][0], r[1][1]};
  *reinterpret_cast<V2*>(ow + o) = V2{r[2][0], r[2][1]};
}

template <typename T>
static int adj_advdiff_t(Ctx& c, double nu, const double uinf[3])
{
  hipLaunchKernelGGL((k_adj_advdiff<T>), dim3(c.nb), dim3(NT), 0, c.stream, (const T*)c.f[CUP_U],
                     (const T*)c.f[CUP_V], (const T*)c.f[CUP_W], (const T*)c.f[CUP_ADJ_U], (const T*)c.f[CUP_ADJ_V],
                     (const T*)c.f[CUP_ADJ_W], (T*)c.f[CUP_TMPU], (T*)c.f[CUP_TMPV], (T*)c.f[CUP_TMPW], c.grid(),
                     c.sc, (T)uinf[0], (T)uinf[1], (T)uinf[2], nu);
  return check_hip(c, hipGetLastError(), "k_adj_advdiff");
}

int launch_adjoint_advdiff(Ctx& c, double nu, const double uinf[3])
{
  return c.precision == 4 ? adj_advdiff_t<float>(c, nu, uinf) : adj_advdiff_t<double>(c, nu, uinf);
}

// Transposed linearisation of the upwind operator L(s) = -(s + U).grad_up3 s + nu lap s at s
// (k_advdiff_up3), applied to lam (no reference counterpart):
// out = b1 + b2 + mult*dt * L'(s)^T lam, b1 and b2 optional (null: zero), with a = s + U,
// S^{+/-} f = sum_k c^{+/-}_k f(+k) along an axis, c^+ = (1, -6, 3, 2, 0), c^- = (0, -2, -3, 6, -1):
// (L'^T lam)_p(y) = -(1/6h) [sum_q lam_q(y) S_p^{sign a_p(y)} s_q (y)
//                   + sum_d sum_k lam_p(y - k e_d) a_d(y - k e_d) c^{sign a_d(y - k e_d)}_k]
//                   + nu/h^2 lap lam_p(y).
// Two labs of halo 2 (s and lam); every term moves along one axis, so faces suffice.
template <typename T>
__global__ __launch_bounds__(NT) void k_adj_up3(const T* __restrict__ su, const T* __restrict__ sv,
                                                const T* __restrict__ sw, const T* __restrict__ lu,
                                                const T* __restrict__ lv, const T* __restrict__ lw, const T* b1u,
                                                const T* b1v, const T* b1w, const T* b2u, const T* b2v,
                                                const T* b2w, T* __restrict__ ou, T* __restrict__ ov,
                                                T* __restrict__ ow, GridArgs g, const DevScalars* __restrict__ sc,
                                                T Ux, T Uy, T Uz, double nu, double mult)
{
  __shared__ T labs[3][L2C];
  __shared__ T labl[3][L2C];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int bx, by, bz;
  block_coords(g, b, bx, by, bz);
  const T* src[3] = {su, sv, sw};
  const T* lsrc[3] = {lu, lv, lw};
  gather_g2<T>(src, labs, g, b, bx, by, bz);
  gather_g2<T>(lsrc, labl, g, b, bx, by, bz);
  __syncthreads();
  const double fac = mult * sc->dt;
  const T fA = (T)(-fac / (6 * g.h));
  const T fD = (T)(fac * nu / (g.h * g.h));
  const T Uc[3] = {Ux, Uy, Uz};
  constexpr int st[3] = {1, L2, L2 * L2};
  constexpr T cp[5] = {1, -6, 3, 2, 0}, cm[5] = {0, -2, -3, 6, -1};
  const int c0 = 2 * threadIdx.x, ix0 = c0 & 7, iy = (c0 >> 3) & 7, iz = c0 >> 6;
  T r[3][2];
#pragma unroll
  for (int k2 = 0; k2 < 2; ++k2) {
    const int l = ((iz + 2) * L2 + iy + 2) * L2 + ix0 + k2 + 2;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const T ap = labs[p][l] + Uc[p];
      T t1 = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const T* F = labs[q];
        t1 += labl[q][l] * up3(ap, F[l - 2 * st[p]], F[l - st[p]], F[l], F[l + st[p]], F[l + 2 * st[p]]);
      }
      T t2 = 0, lap = 0;
      const T* Lp = labl[p];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int k = -2; k <= 2; ++k) {
          const int x = l - k * st[d];
          const T a = labs[d][x] + Uc[d];
          t2 += Lp[x] * a * (a > 0 ? cp[k + 2] : cm[k + 2]);
        }
        lap += Lp[l - st[d]] + Lp[l + st[d]] - 2 * Lp[l];
      }
      r[p][k2] = fA * (t1 + t2) + fD * lap;
    }
  }
  using V2 = typename Vec2<T>::type;
  const size_t o = (size_t)b * BS3 + c0;
  const T* b1[3] = {b1u, b1v, b1w};
  const T* b2[3] = {b2u, b2v, b2w};
  T* out[3] = {ou, ov, ow};
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    T v0 = r[p][0], v1 = r[p][1];
    if (b1[p]) {
      const V2 a = *reinterpret_cast<const V2*>(b1[p] + o);
      v0 += a.x;
      v1 += a.y;
    }
    if (b2[p]) {
      const V2 a = *reinterpret_cast<const V2*>(b2[p] + o);
      v0 += a.x;
      v1 += a.y;
    }
    *reinterpret_cast<V2*>(out[p] + o) = V2{v0, v1};
  }
}

// out0 = b1 + b2 + mult*dt * L'(s)^T lam on field triples (b1, b2 may be -1: zero).
template <typename T>
static int adj_up3_t(Ctx& c, int s0, int l0, int b10, int b20, int out0, double nu, const double uinf[3], double mult)
{
  auto f = [&](int base, int q) { return base < 0 ? (const T*)nullptr : (const T*)c.f[base + q]; };
  hipLaunchKernelGGL((k_adj_up3<T>), dim3(c.nb), dim3(NT), 0, c.stream, f(s0, 0), f(s0, 1), f(s0, 2), f(l0, 0),
                     f(l0, 1), f(l0, 2), f(b10, 0), f(b10, 1), f(b10, 2), f(b20, 0), f(b20, 1), f(b20, 2),
                     (T*)c.f[out0], (T*)c.f[out0 + 1], (T*)c.f[out0 + 2], c.grid(), c.sc, (T)uinf[0], (T)uinf[1],
                     (T)uinf[2], nu, mult);
  return check_hip(c, hipGetLastError(), "k_adj_up3");
}

int launch_adjoint_up3(Ctx& c, int s0, int l0, int b10, int b20, int out0, double nu, const double uinf[3],
                       double mult)
{
  return c.precision == 4 ? adj_up3_t<float>(c, s0, l0, b10, b20, out0, nu, uinf, mult)
                          : adj_up3_t<double>(c, s0, l0, b10, b20, out0, nu, uinf, mult);
}

int launch_up3_stage(Ctx& c, int src0, int base0, int out0, double nu, const double uinf[3], double mult)
{
  return c.precision == 4 ? up3_stage<float>(c, src0, base0, out0, nu, uinf, mult)
                          : up3_stage<double>(c, src0, base0, out0, nu, uinf, mult);
}

int launch_gradp_plain(Ctx& c)
{
  return c.precision == 4 ? gradp_t<float>(c, false, nullptr, nullptr, false)
                          : gradp_t<double>(c, false, nullptr, nullptr, false);
}

int launch_gradp_fused(Ctx& c, const double uinf[3], const int* skip_slot, bool force, bool xface)
{
  // the faces are those of the p the last fused Poisson solve wrote, used once
  const bool xf = xface && c.xface_ok && c.periodic();
  c.xface_ok = false;
  return c.precision == 4 ? gradp_t<float>(c, true, uinf, skip_slot, force, xf)
                          : gradp_t<double>(c, true, uinf, skip_slot, force, xf);
}

// ------------------------------------------------------------------- reductions

// One workgroup per component; component i of partial j at part[j*width + i].
__global__ __launch_bounds__(NT) void k_reduce(const double* __restrict__ part, int n, int width, unsigned maxmask,
                                               double* __restrict__ out)
{
  const int i = blockIdx.x;
  const bool ismax = (maxmask >> i) & 1;
  double a[1] = {ismax ? -INFINITY : 0.0};
  for (int j = threadIdx.x; j < n; j += NT) {
    const double x = part[(size_t)j * width + i];
    a[0] = ismax ? fmax(a[0], x) : a[0] + x;
  }
  block_reduce<1>(a, ismax ? 1u : 0u, out + i);
}

int launch_reduce(Ctx& c, int n, int width, unsigned maxmask, double* out_dev, const double* part)
{
  hipLaunchKernelGGL(k_reduce, dim3(width), dim3(NT), 0, c.stream, part ? part : c.partials, n, width, maxmask,
                     out_dev);
  return check_hip(c, hipGetLastError(), "k_reduce");
}

// ----------------------------------------------------------------- NaN check

// NaN by its bit pattern (all-ones exponent, non-zero mantissa), which no floating-point
// flag of the build can fold away.
__device__ inline unsigned is_nan_bits(float x) { return (__float_as_uint(x) & 0x7fffffffu) > 0x7f800000u; }
__device__ inline unsigned is_nan_bits(double x)
{
  return ((unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull) > 0x7ff0000000000000ull;
}

// Operator::check's scan (Operator.h:23-50) as counts: NaN cells of u, v, w and p, summed per
// lane over a grid-stride loop, per wave by shuffles, then one atomic per wave and field.
template <typename T>
__global__ __launch_bounds__(NT) void k_count_nan(const T* __restrict__ u, const T* __restrict__ v,
                                                  const T* __restrict__ w, const T* __restrict__ p, size_t n,
                                                  unsigned long long* __restrict__ out)
{
  unsigned cnt[4] = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
    cnt[0] += is_nan_bits(u[i]);
    cnt[1] += is_nan_bits(v[i]);
    cnt[2] += is_nan_bits(w[i]);
    cnt[3] += is_nan_bits(p[i]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    unsigned x = cnt[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(out + k, (unsigned long long)x);
  }
}

__global__ void k_clear_u64(unsigned long long* p, int n)
{
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

int launch_count_nan(Ctx& c, long long out[4])
{
  unsigned long long* cnt = c.sc->nan_count;
  hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(64), 0, c.stream, cnt, 4);
  const int grid = (int)std::min<size_t>(4096, (c.ncell + NT - 1) / NT);
  if (c.precision == 4)
    hipLaunchKernelGGL(k_count_nan<float>, dim3(grid), dim3(NT), 0, c.stream, (const float*)c.f[CUP_U],
                       (const float*)c.f[CUP_V], (const float*)c.f[CUP_W], (const float*)c.f[CUP_P], c.ncell, cnt);
  else
    hipLaunchKernelGGL(k_count_nan<double>, dim3(grid), dim3(NT), 0, c.stream, (const double*)c.f[CUP_U],
                       (const double*)c.f[CUP_V], (const double*)c.f[CUP_W], (const double*)c.f[CUP_P], c.ncell, cnt);
  if (int rc = check_hip(c, hipGetLastError(), "k_count_nan")) return rc;
  unsigned long long h[4];
  if (int rc = check_hip(c, hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, c.stream), "nan counts"))
    return rc;
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "synchronize")) return rc;
  for (int k = 0; k < 4; ++k) out[k] = (long long)h[k];
  return CUP_OK;
}

// Step bookkeeping after the fused gradient: umax of the new velocity from the atomic
// max over one bank of slots (then cleared), step += 1 and time += dt when advancing, then
// the next dt by the CFL rule.
__global__ __launch_bounds__(UMAX_SLOTS) void k_finalize_step(DevScalars* sc, unsigned long long* slots, double CFL,
                                                               double nu, double h, int rampup, double fixed_dt,
                                                               int advance)
{
  double m[1] = {__longlong_as_double((long long)slots[threadIdx.x])};
  slots[threadIdx.x] = 0ull;
  __shared__ double out[1];
  block_reduce<1>(m, 1u, out);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double umax = out[0];
  sc->umax = umax;
  if (advance) {
    sc->step += 1;
    sc->time += sc->dt;
    sc->last_dt = sc->dt;
  }
  sc->dt = cfl_dt(umax, sc->step, CFL, nu, h, rampup, fixed_dt);
}

int launch_finalize_step(Ctx& c, double CFL, double nu, int rampup, double fixed_dt, int advance)
{
  hipLaunchKernelGGL(k_finalize_step, dim3(1), dim3(UMAX_SLOTS), 0, c.stream, c.sc, umax_slots_dev(c, c.umax_bank), CFL,
                     nu, c.h, rampup, fixed_dt, advance);
  return check_hip(c, hipGetLastError(), "k_finalize_step");
}

// ------------------------------------------------------------------ diagnostics

// KernelDissipation (ComputeDissipation.cpp:46-115): 20 FP64 integrals. A workgroup sums
// nbw consecutive blocks (one block reduction per workgroup) into part[20*wg].
template <typename T>
__global__ __launch_bounds__(NT) void k_diagnostics(const T* __restrict__ u, const T* __restrict__ v,
                                                    const T* __restrict__ w, const T* __restrict__ p,
                                                    const T* __restrict__ chi, GridArgs g, double nu, double cx,
                                                    double cy, double cz, double* __restrict__ part, int nbw)
{
  __shared__ T lab[4][L1C];
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const T* src[4] = {u, v, w, p};
  const double h = g.h, hCube = h * h * h, inv2h = .5 / h, invHh = 1 / (h * h);
  double R[20];
#pragma unroll
  for (int i = 0; i < 20; ++i) R[i] = 0;
  for (int j = 0; j < nbw; ++j) {
    const int b = wg * nbw + j;
    int bx, by, bz;
    block_coords(g, b, bx, by, bz);
    if (j) __syncthreads();  // the previous block is done with the lab
    gather_g1<T, 4>(src, lab, g, b, bx, by, bz, {HK_VEL, HK_VEL, HK_VEL, HK_COPY});
    __syncthreads();
    for (int c = threadIdx.x; c < BS3; c += NT) {
      const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
      const int l = ((iz + 1) * L1 + iy + 1) * L1 + ix + 1;
      const int E = l + 1, W = l - 1, N = l + L1, S = l - L1, B = l + L1 * L1, F = l - L1 * L1;
      const T* U = lab[0];
      const T* V = lab[1];
      const T* Wv = lab[2];
      const T* P = lab[3];
      const double X = (double)chi[(size_t)b * BS3 + c];
      const double PX = (bx * BS + ix + 0.5) * h - cx, PY = (by * BS + iy + 0.5) * h - cy, PZ = (bz * BS + iz + 0.5) * h - cz;
      const double Lu = U[l], Lv = V[l], Lw = Wv[l];
      const double WX = inv2h * (((double)Wv[N] - Wv[S]) - ((double)V[B] - V[F]));
      const double WY = inv2h * (((double)U[B] - U[F]) - ((double)Wv[E] - Wv[W]));
      const double WZ = inv2h * (((double)V[E] - V[W]) - ((double)U[N] - U[S]));
      const double dPdx = inv2h * ((double)P[E] - P[W]);
      const double dPdy = inv2h * ((double)P[N] - P[S]);
      const double dPdz = inv2h * ((double)P[B] - P[F]);
      R[15] -= (1 - X) * hCube * (dPdx * Lu + dPdy * Lv + dPdz * Lw);
      const double lapU = invHh * ((double)U[E] + U[W] + U[N] + U[S] + U[B] + U[F] - 6 * Lu);
      const double lapV = invHh * ((double)V[E] + V[W] + V[N] + V[S] + V[B] + V[F] - 6 * Lv);
      const double lapW = invHh * ((double)Wv[E] + Wv[W] + Wv[N] + Wv[S] + Wv[B] + Wv[F] - 6 * Lw);
      const double V1 = lapU * Lu + lapV * Lv + lapW * Lw;
      const double D11 = inv2h * ((double)U[E] - U[W]);
      const double D22 = inv2h * ((double)V[N] - V[S]);
      const double D33 = inv2h * ((double)Wv[B] - Wv[F]);
      const double D12 = inv2h * ((double)U[N] - U[S] + V[E] - V[W]) / 2;
      const double D13 = inv2h * ((double)U[B] - U[F] + Wv[E] - Wv[W]) / 2;
      const double D23 = inv2h * ((double)Wv[N] - Wv[S] + V[B] - V[F]) / 2;
      const double V2 = D11 * D11 + D22 * D22 + D33 * D33 + 2 * (D12 * D12 + D13 * D13 + D23 * D23);
      R[16] += (1 - X) * hCube * nu * (V1 + 2 * V2);
      R[0] += hCube * WX;
      R[1] += hCube * WY;
      R[2] += hCube * WZ;
      R[3] += hCube / 2 * (PY * WZ - PZ * WY);
      R[4] += hCube / 2 * (PZ * WX - PX * WZ);
      R[5] += hCube / 2 * (PX * WY - PY * WX);
      R[6] += hCube * Lu;
      R[7] += hCube * Lv;
      R[8] += hCube * Lw;
      R[9] += hCube / 3 * (PX * (PY * WY + PZ * WZ) - WX * (PY * PY + PZ * PZ));
      R[10] += hCube / 3 * (PY * (PX * WX + PZ * WZ) - WY * (PX * PX + PZ * PZ));
      R[11] += hCube / 3 * (PZ * (PX * WX + PY * WY) - WZ * (PX * PX + PY * PY));
      R[12] += hCube * (PY * Lw - PZ * Lv);
      R[13] += hCube * (PZ * Lu - PX * Lw);
      R[14] += hCube * (PX * Lv - PY * Lu);
      R[17] += hCube * (WX * Lu + WY * Lv + WZ * Lw);
      R[18] += hCube * (Lu * Lu + Lv * Lv + Lw * Lw) / 2;
      R[19] += sqrt(WX * WX + WY * WY + WZ * WZ);
    }
  }
  block_reduce<20>(R, 0u, part + 20 * (size_t)wg);
}

// Blocks per diagnostics workgroup: 8 when the x rows allow it.
inline int diag_nbw_x(int bpdx) { return bpdx % 8 == 0 ? 8 : (bpdx % 4 == 0 ? 4 : (bpdx % 2 == 0 ? 2 : 1)); }
inline int diag_nbw(const Ctx& c) { return diag_nbw_x(c.bpd[0]); }

template <typename T>
static int diagnostics_t(Ctx& c, double nu)
{
  const int nbw = diag_nbw(c);
  hipLaunchKernelGGL((k_diagnostics<T>), dim3(c.nb / nbw), dim3(NT), 0, c.stream, (const T*)c.f[CUP_U],
                     (const T*)c.f[CUP_V], (const T*)c.f[CUP_W], (const T*)c.f[CUP_P], (const T*)c.f[CUP_CHI],
                     c.grid(), nu, c.extent[0] / 2, c.extent[1] / 2, c.extent[2] / 2, c.partials, nbw);
  return check_hip(c, hipGetLastError(), "k_diagnostics");
}

// KernelAnalysis_gradStats (Analysis.cpp:18-52): per cell grad2 = sum of the nine squared
// central differences of u, v, w over 4h^2; sums of sqrt(grad2) and grad2, nbw blocks per
// workgroup, into part[2*wg].
template <typename T>
__global__ __launch_bounds__(NT) void k_grad_stats(const T* __restrict__ u, const T* __restrict__ v,
                                                   const T* __restrict__ w, GridArgs g, double* __restrict__ part,
                                                   int nbw)
{
  __shared__ T lab[3][L1C];
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const T* src[3] = {u, v, w};
  const double fac = 1 / (4 * g.h * g.h);
  double R[2] = {0, 0};
  for (int j = 0; j < nbw; ++j) {
    const int b = wg * nbw + j;
    int bx, by, bz;
    block_coords(g, b, bx, by, bz);
    if (j) __syncthreads();  // the previous block is done with the lab
    gather_g1<T, 3>(src, lab, g, b, bx, by, bz, {HK_VEL, HK_VEL, HK_VEL});
    __syncthreads();
    for (int c = threadIdx.x; c < BS3; c += NT) {
      const int l = (((c >> 6) + 1) * L1 + ((c >> 3) & 7) + 1) * L1 + (c & 7) + 1;
      double g2 = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const T* F = lab[q];
        const double dx = (double)F[l + 1] - F[l - 1], dy = (double)F[l + L1] - F[l - L1],
                     dz = (double)F[l + L1 * L1] - F[l - L1 * L1];
        g2 += dx * dx + dy * dy + dz * dz;
      }
      g2 *= fac;
      R[0] += sqrt(g2);
      R[1] += g2;
    }
  }
  block_reduce<2>(R, 0u, part + 2 * (size_t)wg);
}

int launch_grad_stats(Ctx& c, double out[2])
{
  if (int e = ensure_partials(c, 2 * (size_t)c.nb)) return e;
  const int nbw = diag_nbw(c);
  if (c.precision == 4)
    hipLaunchKernelGGL((k_grad_stats<float>), dim3(c.nb / nbw), dim3(NT), 0, c.stream, (const float*)c.f[CUP_U],
                       (const float*)c.f[CUP_V], (const float*)c.f[CUP_W], c.grid(), c.partials, nbw);
  else
    hipLaunchKernelGGL((k_grad_stats<double>), dim3(c.nb / nbw), dim3(NT), 0, c.stream, (const double*)c.f[CUP_U],
                       (const double*)c.f[CUP_V], (const double*)c.f[CUP_W], c.grid(), c.partials, nbw);
  if (int rc = launch_reduce(c, c.nb / nbw, 2, 0u, c.sc->red)) return rc;
  if (int rc = check_hip(c, hipMemcpyAsync(out, c.sc->red, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                         "hipMemcpy grad stats"))
    return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "synchronize");
}

int launch_diagnostics(Ctx& c, double nu, int* nparts)
{
  if (int e = ensure_partials(c, 20 * (size_t)c.nb)) return e;
  *nparts = c.nb / diag_nbw(c);
  return c.precision == 4 ? diagnostics_t<float>(c, nu) : diagnostics_t<double>(c, nu);
}

// ------------------------------------------------- diagnostics without allocation

// The per-workgroup partial sums of cup_diagnostics_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends. Host code never passes this name to a kernel:
// each device's address comes from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) double scratch[(size_t)CUP_NOALLOC_MAX_PARTS * 20];

constexpr int SCRATCH_MAX_DEVICES = 64;

// The device address of `scratch` on the current device `dev`, looked up once per device.
static int scratch_ptr(int dev, double** out)
{
  static std::mutex mu;
  static double* ptr[SCRATCH_MAX_DEVICES] = {};
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return free_error(CUP_E_DEVICE, std::string("diagnostics_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    ptr[dev] = (double*)p;
  }
  *out = ptr[dev];
  return CUP_OK;
}

template <typename T>
static int diagnostics_noalloc_t(const void* const fields[5], const GridArgs& g, const double half_extent[3],
                                 double nu, int nbw, double* part, hipStream_t s, double* out_dev)
{
  hipLaunchKernelGGL((k_diagnostics<T>), dim3(g.nb / nbw), dim3(NT), 0, s, (const T*)fields[CUP_U],
                     (const T*)fields[CUP_V], (const T*)fields[CUP_W], (const T*)fields[CUP_P],
                     (const T*)fields[CUP_CHI], g, nu, half_extent[0], half_extent[1], half_extent[2], part, nbw);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_reduce, dim3(20), dim3(NT), 0, s, (const double*)part, g.nb / nbw, 20, 0u, out_dev);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return free_error(CUP_E_DEVICE, std::string("diagnostics_noalloc: launch: ") + hipGetErrorString(e));
  return CUP_OK;
}

// ----------------------------------------------------------- initial conditions

// KernelIC_taylorGreen (InitialConditions.cpp:32-52): every field cleared, then
// u = A cos(ax) sin(by) sin(cz), v = B sin(ax) cos(by) sin(cz), B = -A Ly/Lx.
template <typename T>
__global__ void k_ic_tgv(T* const* f, GridArgs g, double a, double bb, double cc, double A, double B)
{
  const size_t n = (size_t)g.nb * BS3;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(i / BS3), c = (int)(i % BS3);
    int bx, by, bz;
    block_coords(g, b, bx, by, bz);
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const double x = (bx * BS + ix + 0.5) * g.h, y = (by * BS + iy + 0.5) * g.h, z = (bz * BS + iz + 0.5) * g.h;
    for (int k = 0; k < 8; ++k) f[k][i] = 0;
    f[CUP_U][i] = (T)(A * cos(a * x) * sin(bb * y) * sin(cc * z));
    f[CUP_V][i] = (T)(B * sin(a * x) * cos(bb * y) * sin(cc * z));
  }
}

template <typename T>
static int ic_tgv_t(Ctx& c, double uMax)
{
  T** dptrs = nullptr;
  if (hipMalloc(&dptrs, 8 * sizeof(T*)) != hipSuccess) return set_error(c, CUP_E_NOMEM, "ic: hipMalloc");
  T* hp[8];
  for (int k = 0; k < 8; ++k) hp[k] = (T*)c.f[k];
  hipMemcpyAsync(dptrs, hp, sizeof(hp), hipMemcpyHostToDevice, c.stream);
  const double a = 2 * M_PI / c.extent[0], bb = 2 * M_PI / c.extent[1], cc = 2 * M_PI / c.extent[2];
  hipLaunchKernelGGL((k_ic_tgv<T>), ew_grid(c.ncell), dim3(NT), 0, c.stream, dptrs, c.grid(), a, bb, cc, uMax,
                     -uMax * c.extent[1] / c.extent[0]);
  int e = check_hip(c, hipGetLastError(), "k_ic_tgv");
  hipStreamSynchronize(c.stream);
  hipFree(dptrs);
  return e;
}

int launch_ic_taylor_green(Ctx& c, double uMax)
{
  return c.precision == 4 ? ic_tgv_t<float>(c, uMax) : ic_tgv_t<double>(c, uMax);
}

}  // namespace cup

extern "C" int cup_diagnostics_noalloc(const cup_grid_desc* d, const void* const fields[5], double nu,
                                       void* hip_stream, double* out_dev)
{
  using namespace cup;
  if (!d || !fields || !out_dev) return free_error(CUP_E_INVALID, "diagnostics_noalloc: null argument");
  for (int k = 0; k < 5; ++k)
    if (!fields[k]) return free_error(CUP_E_INVALID, "diagnostics_noalloc: null field pointer");
  for (int a = 0; a < 3; ++a)
    if (d->bpd[a] < 1) return free_error(CUP_E_INVALID, "diagnostics_noalloc: blocks per dimension must be >= 1");
  if (d->precision != 4 && d->precision != 8)
    return free_error(CUP_E_INVALID, "diagnostics_noalloc: precision must be 4 (FP32) or 8 (FP64)");
  if (!(d->maxextent > 0)) return free_error(CUP_E_INVALID, "diagnostics_noalloc: maxextent must be > 0");
  if (d->device < 0 || d->device >= SCRATCH_MAX_DEVICES)
    return free_error(CUP_E_INVALID, "diagnostics_noalloc: device ordinal out of range");
  const long long nb = (long long)d->bpd[0] * d->bpd[1] * d->bpd[2];
  const int nbw = diag_nbw_x(d->bpd[0]);
  if (nb / nbw > CUP_NOALLOC_MAX_PARTS)
    return free_error(CUP_E_INVALID, "diagnostics_noalloc: the grid needs " + std::to_string(nb / nbw) +
                                         " partial sums, more than the static scratch holds (" +
                                         std::to_string(CUP_NOALLOC_MAX_PARTS) + ")");
  GridArgs g;
  g.nbx = d->bpd[0]; g.nby = d->bpd[1]; g.nbz = d->bpd[2];
  g.Nx = g.nbx * BS; g.Ny = g.nby * BS; g.Nz = g.nbz * BS;
  g.nb = (int)nb;
  g.bcany = 0;
  for (int a = 0; a < 3; ++a) {
    if (d->bc[a] != CUP_BC_PERIODIC && d->bc[a] != CUP_BC_OPEN && d->bc[a] != CUP_BC_WALL)
      return free_error(CUP_E_INVALID, "diagnostics_noalloc: unknown boundary condition");
    g.bc[a] = d->bc[a];
    g.bcany |= d->bc[a] != CUP_BC_PERIODIC;
  }
  const int maxN = std::max(g.Nx, std::max(g.Ny, g.Nz));
  g.h = d->maxextent / maxN;
  const double half_extent[3] = {0.5 * ((double)g.Nx / maxN) * d->maxextent, 0.5 * ((double)g.Ny / maxN) * d->maxextent,
                                 0.5 * ((double)g.Nz / maxN) * d->maxextent};
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) return free_error(CUP_E_DEVICE, "diagnostics_noalloc: hipGetDevice");
  if (hipSetDevice(d->device) != hipSuccess) return free_error(CUP_E_DEVICE, "diagnostics_noalloc: hipSetDevice");
  double* part = nullptr;
  int rc = scratch_ptr(d->device, &part);
  if (!rc)
    rc = d->precision == 4 ? diagnostics_noalloc_t<float>(fields, g, half_extent, nu, nbw, part, (hipStream_t)hip_stream, out_dev)
                           : diagnostics_noalloc_t<double>(fields, g, half_extent, nu, nbw, part, (hipStream_t)hip_stream, out_dev);
  hipSetDevice(prev);
  return rc;
}
