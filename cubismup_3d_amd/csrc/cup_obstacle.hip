// cup_obstacle.hip -- a sphere on the block grid: SDF and chi (Towers' smoothed
// Heaviside), surface points, the pressure-RHS masking with its perimeter correction,
// Brinkman penalization and the surface force integrals. All body arithmetic is FP64;
// only the velocity it updates and the grid chi are stored in the context precision.
// The body touches few blocks (a few hundred of tens of thousands), so every kernel here
// runs one workgroup per obstacle block.
#include "cup_internal.h"

#include <cfloat>
#include <cmath>
#include <limits>
#include <vector>

namespace cup {

static constexpr double DEPS = DBL_EPSILON;
static constexpr int SURFDH = 1;  // ObstacleBlock.h:16

__device__ __forceinline__ void bcoords(const GridArgs& g, int b, int& bx, int& by, int& bz)
{
  bx = b % g.nbx;
  by = (b / g.nbx) % g.nby;
  bz = b / (g.nbx * g.nby);
}

// FP64 block-wide reduction, thread 0 writes out[0..NV).
template <int NV>
__device__ __forceinline__ void breduce(double (&v)[NV], double* out)
{
  __shared__ double red[NT / 64][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[wv][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double a = red[0][i];
      for (int w = 1; w < NT / 64; ++w) a += red[w][i];
      out[i] = a;
    }
  __syncthreads();
}

// SDF of the sphere on obstacle blocks: R - |x - c| (Sphere.cpp:44-48); of the hemisphere
// min(-(x - c_x), R - |x - c|) (Sphere.cpp:78-83).
__global__ void k_sphere_sdf(const int* __restrict__ ob_ids, double* __restrict__ sdf, GridArgs g, double cx,
                             double cy, double cz, double R, int hemi)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const double dx = (bx * BS + ix + 0.5) * g.h - cx, dy = (by * BS + iy + 0.5) * g.h - cy,
                 dz = (bz * BS + iz + 0.5) * g.h - cz;
    const double ball = R - sqrt(dx * dx + dy * dy + dz * dz);
    sdf[(size_t)s * BS3 + c] = hemi && -dx < ball ? -dx : ball;
  }
}

// SDF as the reference's grid tmpU after CreateObstacles' reset (-1) and the block fill
// (max with the body's SDF): the body's SDF on obstacle blocks, -1 elsewhere.
// Outside a non-periodic domain face the lab holds the boundary cell's value, negated at a
// wall (the reference keeps the SDF in tmpU, which BlockLabBC mirrors with its sign flipped).
__device__ __forceinline__ double sdf_at(const GridArgs& g, const int* __restrict__ slot, const double* __restrict__ sdf,
                                         int gx, int gy, int gz)
{
  double sign = 1.0;
  if (g.bcany) {
    int c[3] = {gx, gy, gz};
    const int n[3] = {g.Nx, g.Ny, g.Nz};
    for (int a = 0; a < 3; ++a)
      if (g.bc[a] != CUP_BC_PERIODIC && (c[a] < 0 || c[a] >= n[a])) {
        c[a] = c[a] < 0 ? 0 : n[a] - 1;
        if (g.bc[a] == CUP_BC_WALL) sign = -sign;
      }
    gx = c[0]; gy = c[1]; gz = c[2];
  }
  gx = wrapi(gx, g.Nx); gy = wrapi(gy, g.Ny); gz = wrapi(gz, g.Nz);
  const int b = (gx >> 3) + g.nbx * ((gy >> 3) + g.nby * (gz >> 3));
  const int s = slot[b];
  if (s < 0) return -sign;
  return sign * sdf[(size_t)s * BS3 + (((gz & 7) * BS + (gy & 7)) * BS + (gx & 7))];
}

// Mollified Heaviside of the signed distance d with half-width e (no reference counterpart):
// 0 for d <= -e, 1 for d >= e, (1 + d/e + sin(pi d/e)/pi)/2 between. It is C1, which makes chi
// differentiable in the sphere centre; dH/dd = (1 + cos(pi d/e))/(2e) vanishes at d = +-e.
__device__ __forceinline__ double moll_H(double d, double e)
{
  if (d <= -e) return 0.0;
  if (d >= e) return 1.0;
  return 0.5 * (1.0 + d / e + sin(M_PI * d / e) / M_PI);
}

// d chi / d centre at the cell (gx, gy, gz) for the mollified chi H(R - |x - c|):
// dH (x - c)/|x - c|.
__device__ __forceinline__ void moll_dchi_dc(const GridArgs& g, int gx, int gy, int gz, double cx, double cy,
                                             double cz, double R, double e, double out[3])
{
  const double rx = (gx + 0.5) * g.h - cx, ry = (gy + 0.5) * g.h - cy, rz = (gz + 0.5) * g.h - cz;
  const double dist = sqrt(rx * rx + ry * ry + rz * rz);
  const double d = R - dist;
  const double dH = (dist > 0 && d > -e && d < e) ? (1.0 + cos(M_PI * d / e)) / (2.0 * e) / dist : 0.0;
  out[0] = dH * rx;
  out[1] = dH * ry;
  out[2] = dH * rz;
}

// KernelCharacteristicFunction (ObstaclesCreate.cpp:51-137) on one obstacle block:
// chi by Towers' formula inside the +-2h band, grid chi = max, CoM and mass partials,
// surface points (delta, dchi) where Delta > eps (ObstacleBlock::write, ObstacleBlock.h:157-166).
// With moll_eps > 0 chi is instead moll_H(sdf, moll_eps); the surface points are unchanged.
template <typename T>
__global__ __launch_bounds__(NT) void k_sphere_chi(const int* __restrict__ ob_ids, const int* __restrict__ slot,
                                                   const double* __restrict__ sdf, double* __restrict__ ob_chi,
                                                   double* __restrict__ surf, T* __restrict__ chi, GridArgs g,
                                                   double* __restrict__ part, double moll_eps)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  const double h = g.h, inv2h = .5 / h, fac1 = .5 * h * h, vol = h * h * h;
  double acc[4] = {0, 0, 0, 0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const int gx = bx * BS + ix, gy = by * BS + iy, gz = bz * BS + iz;
    const double sdf0 = sdf[(size_t)s * BS3 + c];
    const double dPx = sdf_at(g, slot, sdf, gx + 1, gy, gz), dMx = sdf_at(g, slot, sdf, gx - 1, gy, gz);
    const double dPy = sdf_at(g, slot, sdf, gx, gy + 1, gz), dMy = sdf_at(g, slot, sdf, gx, gy - 1, gz);
    const double dPz = sdf_at(g, slot, sdf, gx, gy, gz + 1), dMz = sdf_at(g, slot, sdf, gx, gy, gz - 1);
    double X;
    if (moll_eps > 0) {
      X = moll_H(sdf0, moll_eps);
    } else if (sdf0 > +2 * h || sdf0 < -2 * h) {
      X = sdf0 > 0 ? 1 : 0;
    } else {
      const double gUX = inv2h * (dPx - dMx), gUY = inv2h * (dPy - dMy), gUZ = inv2h * (dPz - dMz);
      const double gUSq = gUX * gUX + gUY * gUY + gUZ * gUZ + DEPS;
      const double gIX = inv2h * (fmax(dPx, 0.0) - fmax(dMx, 0.0));
      const double gIY = inv2h * (fmax(dPy, 0.0) - fmax(dMy, 0.0));
      const double gIZ = inv2h * (fmax(dPz, 0.0) - fmax(dMz, 0.0));
      X = (gIX * gUX + gIY * gUY + gIZ * gUZ) / gUSq;
    }
    ob_chi[(size_t)s * BS3 + c] = X;
    const size_t o = (size_t)b * BS3 + c;
    chi[o] = (T)fmax((double)chi[o], X);
    const double px = (gx + 0.5) * h, py = (gy + 0.5) * h, pz = (gz + 0.5) * h;
    acc[0] += X * vol * px;
    acc[1] += X * vol * py;
    acc[2] += X * vol * pz;
    acc[3] += X * vol;
    double* sd = surf + ((size_t)s * BS3 + c) * 4;
    sd[0] = sd[1] = sd[2] = sd[3] = 0;
    const double sdfs = sdf0 + h * SURFDH;
    if (sdfs > +2 * h || sdfs < -2 * h) continue;
    const double aPx = dPx + h * SURFDH, aMx = dMx + h * SURFDH, aPy = dPy + h * SURFDH, aMy = dMy + h * SURFDH,
                 aPz = dPz + h * SURFDH, aMz = dMz + h * SURFDH;
    const double gUX = inv2h * (aPx - aMx), gUY = inv2h * (aPy - aMy), gUZ = inv2h * (aPz - aMz);
    const double gUSq = gUX * gUX + gUY * gUY + gUZ * gUZ + DEPS;
    auto H = [](double d) { return fabs(d) < DEPS ? 0.5 : (d < 0 ? 0.0 : 1.0); };
    const double gHX = H(aPx) - H(aMx), gHY = H(aPy) - H(aMy), gHZ = H(aPz) - H(aMz);
    const double Delta = fac1 * (gHX * gUX + gHY * gUY + gHZ * gUZ) / gUSq;
    if (Delta > DEPS) {
      sd[0] = Delta;
      sd[1] = -Delta * gUX;
      sd[2] = -Delta * gUY;
      sd[3] = -Delta * gUZ;
    }
  }
  breduce<4>(acc, part + 4 * (size_t)s);
}

template <typename T>
__global__ void k_fill_t(T* p, size_t n, T v)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

void body_free(Ctx& c)
{
  Body& B = c.body;
  if (B.ob_ids) hipFree(B.ob_ids);
  if (B.ob_slot) hipFree(B.ob_slot);
  if (B.ob_chi) hipFree(B.ob_chi);
  if (B.ob_sdf) hipFree(B.ob_sdf);
  if (B.ob_surf) hipFree(B.ob_surf);
  if (B.ob_touch) hipFree(B.ob_touch);
  if (B.ob_r0) hipFree(B.ob_r0);
  if (B.ob_udef) hipFree(B.ob_udef);
  B = Body();
}

// Every body's buffers and the merged tables.
void bodies_free(Ctx& c)
{
  for (size_t j = 0; j < c.more.size(); ++j) {
    BodyScope scope(c, (int)j + 1);
    body_free(c);
  }
  c.more.clear();
  body_free(c);
  if (c.ob_slot_all) hipFree(c.ob_slot_all);
  if (c.ob_rhs_all) hipFree(c.ob_rhs_all);
  c.ob_slot_all = nullptr;
  c.ob_rhs_all = nullptr;
  c.ob_rhs_all_cap = 0;
}

// The merged obstacle-slot table and RHS buffer of all bodies (their blocks are disjoint):
// body j's blocks take slots [offset_j, offset_j + nob_j), and its ob_rhs is that slice.
static int bodies_merge(Ctx& c)
{
  std::vector<Body*> all;
  if (c.body.present) all.push_back(&c.body);
  for (auto& b : c.more)
    if (b.present) all.push_back(&b);
  size_t total = 0;
  for (Body* b : all) total += (size_t)b->nob;
  const size_t bytes = total * BS3 * c.esz;
  if (bytes > c.ob_rhs_all_cap) {
    if (c.ob_rhs_all) hipFree(c.ob_rhs_all);
    c.ob_rhs_all = nullptr;
    c.ob_rhs_all_cap = 0;
    if (check_hip(c, hipMalloc(&c.ob_rhs_all, bytes), "hipMalloc obstacle RHS")) return CUP_E_NOMEM;
    c.ob_rhs_all_cap = bytes;
  }
  if (!c.ob_slot_all && check_hip(c, hipMalloc(&c.ob_slot_all, c.nb * sizeof(int)), "hipMalloc obstacle slots"))
    return CUP_E_NOMEM;
  std::vector<int> slot(c.nb, -1);
  size_t off = 0;
  for (Body* b : all) {
    b->ob_rhs = (char*)c.ob_rhs_all + off * BS3 * c.esz;
    for (int s = 0; s < b->nob; ++s) slot[b->ids_host[s]] = (int)(off + s);
    off += b->nob;
  }
  return check_hip(c, hipMemcpy(c.ob_slot_all, slot.data(), c.nb * sizeof(int), hipMemcpyHostToDevice),
                   "obstacle slots");
}

template <typename T>
static int body_chi_t(Ctx& c, bool fill = true)
{
  Body& B = c.body;
  const size_t n = c.ncell;
  size_t gsz = (n + NT - 1) / NT;
  if (gsz > 65536) gsz = 65536;
  if (fill) hipLaunchKernelGGL((k_fill_t<T>), dim3((unsigned)gsz), dim3(NT), 0, c.stream, (T*)c.f[CUP_CHI], n, (T)0);
  hipLaunchKernelGGL(k_sphere_sdf, dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_sdf, c.grid(), B.center[0],
                     B.center[1], B.center[2], B.radius, B.hemi ? 1 : 0);
  hipLaunchKernelGGL((k_sphere_chi<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_slot, B.ob_sdf, B.ob_chi,
                     B.ob_surf, (T*)c.f[CUP_CHI], c.grid(), c.partials, B.moll * c.h);
  return check_hip(c, hipGetLastError(), "k_sphere_chi");
}

// CreateObstacles for one sphere (ObstaclesCreate.cpp:444-540); moll > 0 selects the
// mollified chi of that half-width in cells. Called again for a body that moved: the
// per-block buffers are kept when the set of obstacle blocks is unchanged.
int body_create(Ctx& c, const double center[3], double radius, double moll, bool hemi)
{
  hipStreamSynchronize(c.stream);
  Body& B = c.body;
  // Blocks whose cell-centre box strictly intersects the sphere's box of half-width
  // R + (2+SURFDH) h (Sphere.cpp:21-42, Obstacle.h:186-207).
  const double safety = (2 + SURFDH) * c.h;
  std::vector<int> ids, slot(c.nb, -1);
  for (int b = 0; b < c.nb; ++b) {
    const int bc[3] = {b % c.bpd[0], (b / c.bpd[0]) % c.bpd[1], b / (c.bpd[0] * c.bpd[1])};
    bool touching = true;
    for (int a = 0; a < 3; ++a) {
      const double mn = (bc[a] * BS + 0.5) * c.h, mx = (bc[a] * BS + BS - 1 + 0.5) * c.h;
      // the hemisphere's box ends at c_x + safety in x (Sphere.cpp:59-63)
      const double top = hemi && a == 0 ? center[a] + safety : center[a] + (radius + safety);
      const double lo = std::fmax(mn, center[a] - (radius + safety)), hi = std::fmin(mx, top);
      if (!(hi - lo > 0)) touching = false;
    }
    if (touching) {
      slot[b] = (int)ids.size();
      ids.push_back(b);
    }
  }
  // Both checks come before B is touched: a rejected placement (a new body, or a move of an
  // existing one) leaves the body, its buffers and the merged tables as they were.
  if (ids.empty()) return set_error(c, CUP_E_INVALID, "chi_from_sphere: the sphere touches no block");
  // the other bodies' obstacle blocks must be apart from this one's
  for (const Body& o : c.more)
    if (o.present)
      for (int b : ids)
        if (std::find(o.ids_host.begin(), o.ids_host.end(), b) != o.ids_host.end())
          return set_error(c, CUP_E_INVALID, "add_sphere: the bodies' obstacle blocks overlap (bodies too close)");
  const bool reuse = B.present && B.radius == radius && B.moll == moll && B.hemi == hemi && B.ids_host == ids;
  // a rebuilt body has no deformation velocity until it is set again; the last udef
  // corrections stay (they are the body's, as Obstacle::transVel_correction)
  double corr[6];
  std::copy(B.corr, B.corr + 6, corr);
  if (B.ob_udef) hipFree(B.ob_udef);
  B.ob_udef = nullptr;
  if (!reuse) body_free(c);
  std::copy(corr, corr + 6, B.corr);
  B.radius = radius;
  B.moll = moll;
  B.hemi = hemi;
  for (int a = 0; a < 3; ++a) B.center[a] = center[a];
  B.nob = (int)ids.size();
  const size_t nc = (size_t)B.nob * BS3;
  if (!reuse && (check_hip(c, hipMalloc(&B.ob_ids, B.nob * sizeof(int)), "hipMalloc") ||
      check_hip(c, hipMalloc(&B.ob_slot, c.nb * sizeof(int)), "hipMalloc") ||
      check_hip(c, hipMalloc(&B.ob_chi, nc * sizeof(double)), "hipMalloc") ||
      check_hip(c, hipMalloc(&B.ob_sdf, nc * sizeof(double)), "hipMalloc") ||
      check_hip(c, hipMalloc(&B.ob_surf, nc * 4 * sizeof(double)), "hipMalloc") ||
      check_hip(c, hipMalloc(&B.ob_touch, nc), "hipMalloc"))) {
    body_free(c);
    return CUP_E_NOMEM;
  }
  if (!reuse) {
    hipMemcpy(B.ob_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice);
    hipMemcpy(B.ob_slot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice);
    B.ids_host = ids;
  }
  B.present = true;
  if (int rc = bodies_merge(c)) return rc;
  if (int rc = ensure_partials(c, 32 * (size_t)B.nob + 2 * (size_t)c.nb)) return rc;
  int rc = c.precision == 4 ? body_chi_t<float>(c) : body_chi_t<double>(c);
  if (rc) return rc;
  std::vector<double> part(4 * (size_t)B.nob);
  if (int rc2 = check_hip(c, hipMemcpyAsync(part.data(), c.partials, part.size() * sizeof(double),
                                            hipMemcpyDeviceToHost, c.stream), "chi_from_sphere"))
    return rc2;
  if (int rc2 = check_hip(c, hipStreamSynchronize(c.stream), "chi_from_sphere")) return rc2;
  double com[4] = {0, 0, 0, 0};
  for (int s = 0; s < B.nob; ++s)
    for (int i = 0; i < 4; ++i) com[i] += part[4 * (size_t)s + i];
  B.mass = com[3];
  for (int a = 0; a < 3; ++a) B.com[a] = com[a] / com[3];
  // the grid chi was cleared: the other bodies write theirs again (after this body's centre of
  // mass was read, since they reuse the partials)
  for (size_t j = 0; j < c.more.size() && !rc; ++j) {
    if (!c.more[j].present) continue;
    std::swap(c.body, c.more[j]);
    rc = c.precision == 4 ? body_chi_t<float>(c, false) : body_chi_t<double>(c, false);
    std::swap(c.body, c.more[j]);
  }
  if (rc) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "chi_from_sphere");
}

// ------------------------------------------------------- pressure RHS with a body

// KernelPressureRHS's visitor (PressureRHS.cpp:78-112) on one obstacle block: where the
// grid chi does not exceed the body's chi, rhs *= (1 - chi) with the bulk source summed,
// and the positive/negative RHS of surface cells (nonzero central grad chi) summed.
// FUSED (rhs == nullptr): the unmasked RHS of the block is first formed from tmp
// (h^2/(2dt) times the central divergence) into ob_rhs; PLAIN: rhs is the block buffer.
template <typename T>
__global__ __launch_bounds__(NT) void k_rhs_obs_mask(const int* __restrict__ ob_ids, const double* __restrict__ ob_chi,
                                                     const T* __restrict__ chi, const T* __restrict__ u,
                                                     const T* __restrict__ v, const T* __restrict__ w,
                                                     T* __restrict__ rhs, T* __restrict__ ob_rhs,
                                                     unsigned char* __restrict__ touch, GridArgs g,
                                                     const DevScalars* __restrict__ sc, double* __restrict__ part,
                                                     T* __restrict__ r0_out)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  const T fac = (T)(0.5 * g.h * g.h / sc->dt);
  double acc[3] = {0, 0, 0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const int gx = bx * BS + ix, gy = by * BS + iy, gz = bz * BS + iz;
    const size_t o = (size_t)b * BS3 + c;
    T r;
    if (rhs) {
      r = rhs[o];
    } else {
      r = fac * (nb_value(u, g, gx, gy, gz, 0, 1, HK_VEL) - nb_value(u, g, gx, gy, gz, 0, -1, HK_VEL) +
                 nb_value(v, g, gx, gy, gz, 1, 1, HK_VEL) - nb_value(v, g, gx, gy, gz, 1, -1, HK_VEL) +
                 nb_value(w, g, gx, gy, gz, 2, 1, HK_VEL) - nb_value(w, g, gx, gy, gz, 2, -1, HK_VEL));
    }
    const double X = ob_chi[(size_t)s * BS3 + c];
    if (r0_out) r0_out[(size_t)s * BS3 + c] = r;
    unsigned char t = 0;
    if (!(chi[o] > (T)X)) {  // in the grid's precision: the grid chi holds the body's chi rounded to T
      const double dXx = (double)nb_value(chi, g, gx, gy, gz, 0, 1, HK_CHI) - nb_value(chi, g, gx, gy, gz, 0, -1, HK_CHI);
      const double dXy = (double)nb_value(chi, g, gx, gy, gz, 1, 1, HK_CHI) - nb_value(chi, g, gx, gy, gz, 1, -1, HK_CHI);
      const double dXz = (double)nb_value(chi, g, gx, gy, gz, 2, 1, HK_CHI) - nb_value(chi, g, gx, gy, gz, 2, -1, HK_CHI);
      const bool surf = dXx * dXx + dXy * dXy + dXz * dXz > 0;
      t = surf ? 1 : 0;
      const double srcBulk = -X * (double)r;
      acc[0] += srcBulk;
      const T nr = (T)((double)r + srcBulk);
      r = nr;
      if (surf) {
        if (nr > 0) acc[1] += (double)nr;
        if (nr < 0) acc[2] -= (double)nr;
      }
    }
    touch[(size_t)s * BS3 + c] = t;
    if (rhs) rhs[o] = r;
    else ob_rhs[(size_t)s * BS3 + c] = r;
  }
  breduce<3>(acc, part + 3 * (size_t)s);
}

// KernelFinalizePerimeters (PressureRHS.cpp:215-256) after the factor of :331-334.
// Every workgroup reduces the per-block sums itself (same order everywhere).
template <typename T>
__global__ __launch_bounds__(NT) void k_rhs_obs_correct(const int* __restrict__ ob_ids, int nob,
                                                        const unsigned char* __restrict__ touch, T* __restrict__ rhs,
                                                        T* __restrict__ ob_rhs, const double* __restrict__ part)
{
  __shared__ double corr_s;
  double acc[3] = {0, 0, 0};
  for (int j = threadIdx.x; j < nob; j += NT) {
    acc[0] += part[3 * (size_t)j];
    acc[1] += part[3 * (size_t)j + 1];
    acc[2] += part[3 * (size_t)j + 2];
  }
  __shared__ double tot[3];
  breduce<3>(acc, tot);
  if (threadIdx.x == 0) {
    const double sumRHS = tot[0], posRHS = tot[1], negRHS = tot[2];
    const double denom = sumRHS > 0 ? posRHS : negRHS;
    corr_s = sumRHS / fmax(denom, DEPS);
  }
  __syncthreads();
  const double corr = corr_s;
  const int s = blockIdx.x, b = ob_ids[s];
  for (int c = threadIdx.x; c < BS3; c += NT) {
    if (!touch[(size_t)s * BS3 + c]) continue;
    T* p = rhs ? &rhs[(size_t)b * BS3 + c] : &ob_rhs[(size_t)s * BS3 + c];
    const double r = (double)*p;
    if (r > 0 && corr > 0) *p = (T)(r - corr * r);
    else if (r < 0 && corr < 0) *p = (T)(r + corr * r);
  }
}

template <typename T>
static int rhs_obs_t(Ctx& c, void* rhs_blocks)
{
  Body& B = c.body;
  const int s0 = c.decomposition == CUP_FUSED ? c.ustar : CUP_U;
  hipLaunchKernelGGL((k_rhs_obs_mask<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                     (const T*)c.f[CUP_CHI], (const T*)c.f[s0], (const T*)c.f[s0 + 1], (const T*)c.f[s0 + 2],
                     (T*)rhs_blocks, (T*)B.ob_rhs, B.ob_touch, c.grid(), c.sc, c.partials + 2 * (size_t)c.nb,
                     (T*)nullptr);
  hipLaunchKernelGGL((k_rhs_obs_correct<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.nob, B.ob_touch,
                     (T*)rhs_blocks, (T*)B.ob_rhs, (const double*)(c.partials + 2 * (size_t)c.nb));
  return check_hip(c, hipGetLastError(), "rhs_obstacle");
}

// PressureRHS's obstacle visitor (PressureRHS.cpp:78-112, 322-343) for every body in turn.
int launch_rhs_obstacle(Ctx& c, void* rhs_blocks)
{
  const int n = body_count(c);
  for (int k = 0; k < n; ++k) {
    BodyScope scope(c, k);
    if (int rc = c.precision == 4 ? rhs_obs_t<float>(c, rhs_blocks) : rhs_obs_t<double>(c, rhs_blocks)) return rc;
  }
  return CUP_OK;
}

// ------------------------------------------------------------------ penalization

// KernelPenalization<implicit> (Penalization.cpp:40-98), one body, udef = 0. In the fused
// step it also folds the penalized blocks into the step's max |u+uinf| and energy.
template <typename T>
__global__ __launch_bounds__(NT) void k_penalize(const int* __restrict__ ob_ids, const double* __restrict__ ob_chi,
                                                 const double* __restrict__ ob_udef, const T* __restrict__ chi, T* __restrict__ u, T* __restrict__ v,
                                                 T* __restrict__ w, GridArgs g, DevScalars* __restrict__ sc,
                                                 double lambda, int implicit, double vx, double vy, double vz,
                                                 double ox, double oy, double oz, double cmx, double cmy, double cmz,
                                                 double Ux, double Uy, double Uz, double* __restrict__ part,
                                                 double* __restrict__ ke_part, unsigned long long* __restrict__ slots)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  const double dt = sc->dt, dv = g.h * g.h * g.h;
  const double lambdaFac = implicit ? lambda : 1.0 / dt;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  double mk[2] = {0, 0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const size_t o = (size_t)b * BS3 + c;
    const double X = ob_chi[(size_t)s * BS3 + c];
    double uu = u[o], vv = v[o], ww = w[o];
    if (!(chi[o] > (T)X) && X > 0) {  // compared in the grid's precision, as the grid chi is stored
      const double px = (bx * BS + ix + 0.5) * g.h - cmx, py = (by * BS + iy + 0.5) * g.h - cmy,
                   pz = (bz * BS + iz + 0.5) * g.h - cmz;
      // U_tot = V + Omega x p + udef (Penalization.cpp:74-80); no udef: + 0, the same values
      double d0 = 0, d1 = 0, d2 = 0;
      if (ob_udef) {
        const double* D = ob_udef + (size_t)s * 3 * BS3;
        d0 = D[c];
        d1 = D[BS3 + c];
        d2 = D[2 * BS3 + c];
      }
      const double UT0 = vx + oy * pz - oz * py + d0, UT1 = vy + oz * px - ox * pz + d1,
                   UT2 = vz + ox * py - oy * px + d2;
      const double pf = implicit ? X * lambdaFac / (1 + lambdaFac * dt * X) : X * lambdaFac;
      const double FX = pf * (UT0 - uu), FY = pf * (UT1 - vv), FZ = pf * (UT2 - ww);
      uu = uu + dt * FX;
      vv = vv + dt * FY;
      ww = ww + dt * FZ;
      u[o] = (T)uu;
      v[o] = (T)vv;
      w[o] = (T)ww;
      uu = u[o]; vv = v[o]; ww = w[o];
      acc[0] += dv * FX; acc[1] += dv * FY; acc[2] += dv * FZ;
      acc[3] += dv * (py * FZ - pz * FY);
      acc[4] += dv * (pz * FX - px * FZ);
      acc[5] += dv * (px * FY - py * FX);
    }
    mk[0] = fmax(mk[0], fmax(fabs(uu + Ux), fmax(fabs(vv + Uy), fabs(ww + Uz))));
    mk[1] += uu * uu + vv * vv + ww * ww;
  }
  breduce<6>(acc, part + 6 * (size_t)s);
  if (ke_part) {
    __shared__ double red[NT / 64][2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    mk[0] = wave_max(mk[0]);
    mk[1] = wave_sum(mk[1]);
    if (lane == 0) { red[wv][0] = mk[0]; red[wv][1] = mk[1]; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = red[0][0], bsum = red[0][1];
      for (int q = 1; q < NT / 64; ++q) { a = fmax(a, red[q][0]); bsum += red[q][1]; }
      atomicMax(&slots[b & (UMAX_SLOTS - 1)], (unsigned long long)__double_as_longlong(a));
      ke_part[4 * (size_t)b] = bsum;  // the block's partials: [4b] its sum, [4b+1..3] zero
      ke_part[4 * (size_t)b + 1] = ke_part[4 * (size_t)b + 2] = ke_part[4 * (size_t)b + 3] = 0.0;
    }
  }
}

__global__ void k_sum_rows(const double* __restrict__ part, int n, int width, double* __restrict__ out);

// ------------------------------------------------------------ adjoint with a body
// (no reference counterpart) The transposes of the penalization, of its drag F_x and of the
// RHS masking with its perimeter correction, for the adjoint step with a forced sphere
// (velocity V and rotation W about its centre of mass). Partials per obstacle block, summed
// by every workgroup in the same order where a total is needed (deterministic).

// lam' = (1 - dt pf) lam - wdrag h^3 pf e_x on penalized cells. The target velocity
// V + W x r (r = x - CoM) enters as g . (V + W x r) with g_d = dt pf lam_d + wdrag h^3 pf [d = x]:
// part[6s+d] += this block's dJ/dV_d = sum g_d, part[6s+3+d] += its dJ/dW_d = sum (r x g)_d.
template <typename T>
__global__ __launch_bounds__(NT) void k_adj_penalize(const int* __restrict__ ob_ids, const double* __restrict__ ob_chi,
                                                     const T* __restrict__ chi, T* __restrict__ lu, T* __restrict__ lv,
                                                     T* __restrict__ lw, GridArgs g, const DevScalars* __restrict__ sc,
                                                     double lambda, int implicit, double wdrag, double cmx,
                                                     double cmy, double cmz, double* __restrict__ part)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  const double dt = sc->dt, h3 = g.h * g.h * g.h;
  const double lf = implicit ? lambda : 1.0 / dt;
  T* L[3] = {lu, lv, lw};
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const size_t o = (size_t)b * BS3 + c;
    const double X = ob_chi[(size_t)s * BS3 + c];
    if (chi[o] > (T)X || !(X > 0)) continue;
    const double pf = implicit ? X * lf / (1 + lf * dt * X) : X * lf;
    double gv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double l = L[d][o];
      const double dr = d == 0 ? wdrag * h3 * pf : 0.0;
      gv[d] = dt * pf * l + dr;
      acc[d] += gv[d];
      L[d][o] = (T)((1 - dt * pf) * l - dr);
    }
    const double px = (bx * BS + (c & 7) + 0.5) * g.h - cmx, py = (by * BS + ((c >> 3) & 7) + 0.5) * g.h - cmy,
                 pz = (bz * BS + (c >> 6) + 0.5) * g.h - cmz;
    acc[3] += py * gv[2] - pz * gv[1];
    acc[4] += pz * gv[0] - px * gv[2];
    acc[5] += px * gv[1] - py * gv[0];
  }
  breduce<6>(acc, part + 6 * (size_t)s);
}

// Totals of the forward masking sums (S, pos, neg per block) and the correction factor,
// as k_rhs_obs_correct forms them.
__device__ __forceinline__ void mask_totals(const double* part, int nob, double& S, double& D, double& corr,
                                            bool& live)
{
  __shared__ double tot[3];
  double acc[3] = {0, 0, 0};
  for (int j = threadIdx.x; j < nob; j += NT) {
    acc[0] += part[3 * (size_t)j];
    acc[1] += part[3 * (size_t)j + 1];
    acc[2] += part[3 * (size_t)j + 2];
  }
  breduce<3>(acc, tot);
  __syncthreads();
  S = tot[0];
  const double denom = S > 0 ? tot[1] : tot[2];
  D = fmax(denom, DEPS);
  live = denom > DEPS;  // the denominator varies with the RHS (else it is the floor)
  corr = S / D;
}

// g = sum over corrected cells Q of y * dr_final/dcorr = y * (-sign(corr) r_new), per block.
template <typename T>
__global__ __launch_bounds__(NT) void k_adj_mask_g(const int* __restrict__ ob_ids, const double* __restrict__ ob_chi,
                                                   const T* __restrict__ chi, const T* __restrict__ rnew,
                                                   const unsigned char* __restrict__ touch, const T* __restrict__ y,
                                                   const double* __restrict__ msum, int nob, double* __restrict__ gpart)
{
  double S, D, corr;
  bool live;
  mask_totals(msum, nob, S, D, corr, live);
  const double sg = corr > 0 ? 1.0 : (corr < 0 ? -1.0 : 0.0);
  const int s = blockIdx.x, b = ob_ids[s];
  double acc[1] = {0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const size_t o = (size_t)b * BS3 + c, q = (size_t)s * BS3 + c;
    if (chi[o] > (T)ob_chi[q] || !touch[q]) continue;
    const double r = rnew[q];
    if ((r > 0 && corr > 0) || (r < 0 && corr < 0)) acc[0] += (double)y[o] * (-sg * r);
  }
  breduce<1>(acc, gpart + s);
}

// y <- M'^T y on the masked cells of the obstacle blocks: with a_S = g/D, a_D = -g corr/D,
// the adjoint of r_new is y (1 - sg corr) + a_D sg on corrected cells and y elsewhere, and
// that of the unmasked RHS is (1 - X) adj(r_new) - X a_S.
// With cpart (the mollified chi): r_new = (1 - X) r0 and S = -sum X r0 give the chi
// sensitivity -r0 (adj(r_new) + a_S), times yscale = dt^2/h^3 because y is the RHS adjoint
// scaled by the plain RHS operator's h^2/(2dt) in place of the gradient step's dt/(2h);
// contracted with dchi/dc into cpart[3s..3s+2].
template <typename T>
__global__ __launch_bounds__(NT) void k_adj_mask_apply(const int* __restrict__ ob_ids,
                                                       const double* __restrict__ ob_chi, const T* __restrict__ chi,
                                                       const T* __restrict__ rnew,
                                                       const unsigned char* __restrict__ touch, T* __restrict__ y,
                                                       const double* __restrict__ msum, int nob,
                                                       const double* __restrict__ gpart, GridArgs g,
                                                       const T* __restrict__ r0, double yscale, double cx, double cy,
                                                       double cz, double R, double e, double* __restrict__ cpart)
{
  double S, D, corr;
  bool live;
  mask_totals(msum, nob, S, D, corr, live);
  __shared__ double gtot;
  {
    double acc[1] = {0};
    for (int j = threadIdx.x; j < nob; j += NT) acc[0] += gpart[j];
    __shared__ double t1[1];
    breduce<1>(acc, t1);
    __syncthreads();
    if (threadIdx.x == 0) gtot = t1[0];
    __syncthreads();
  }
  const double sg = corr > 0 ? 1.0 : (corr < 0 ? -1.0 : 0.0);
  const double aS = gtot / D, aD = live ? -gtot * corr / D : 0.0;
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  double acc[3] = {0, 0, 0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const size_t o = (size_t)b * BS3 + c, q = (size_t)s * BS3 + c;
    const double X = ob_chi[q];
    if (chi[o] > (T)X) continue;
    const double r = rnew[q], yy = y[o];
    const bool inQ = touch[q] && ((r > 0 && corr > 0) || (r < 0 && corr < 0));
    const double an = inQ ? yy * (1 - sg * corr) + aD * sg : yy;
    y[o] = (T)((1 - X) * an - X * aS);
    if (cpart) {
      double dX[3];
      moll_dchi_dc(g, bx * BS + (c & 7), by * BS + ((c >> 3) & 7), bz * BS + (c >> 6), cx, cy, cz, R, e, dX);
      const double gX = -yscale * (double)r0[q] * (an + aS);
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] += gX * dX[k];
    }
  }
  if (cpart) breduce<3>(acc, cpart + 3 * (size_t)s);
}

// The chi sensitivity of the penalization and its drag, contracted with dchi/dc (the
// mollified chi): per block sum dpf/dX (dt sum_d lam_d (V_d - u_in,d) + wdrag h^3 (V_x - u_in,x))
// dchi/dc over penalized cells, lam the adjoint before the penalization's transpose and
// u_in = u* + fac (p(+e_d) - p(-e_d)), fac = -dt/(2h), formed as the gradient kernel forms it.
template <typename T>
__global__ __launch_bounds__(NT) void k_adj_chi_pen(const int* __restrict__ ob_ids, const double* __restrict__ ob_chi,
                                                    const T* __restrict__ chi, const T* __restrict__ us,
                                                    const T* __restrict__ vs, const T* __restrict__ ws,
                                                    const T* __restrict__ p, const T* __restrict__ lu,
                                                    const T* __restrict__ lv, const T* __restrict__ lw, GridArgs g,
                                                    const DevScalars* __restrict__ sc, double lambda, int implicit,
                                                    double wdrag, double vx, double vy, double vz, double cx,
                                                    double cy, double cz, double R, double e,
                                                    double* __restrict__ part)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  const double dt = sc->dt, h3 = g.h * g.h * g.h;
  const double lf = implicit ? lambda : 1.0 / dt;
  const T fac = (T)(-0.5 * dt / g.h);
  double acc[3] = {0, 0, 0};
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int gx = bx * BS + (c & 7), gy = by * BS + ((c >> 3) & 7), gz = bz * BS + (c >> 6);
    const size_t o = (size_t)b * BS3 + c;
    const double X = ob_chi[(size_t)s * BS3 + c];
    if (chi[o] > (T)X || !(X > 0)) continue;
    const T ui = us[o] + fac * (p[cell_index(g, wrapi(gx + 1, g.Nx), gy, gz)] - p[cell_index(g, wrapi(gx - 1, g.Nx), gy, gz)]);
    const T vi = vs[o] + fac * (p[cell_index(g, gx, wrapi(gy + 1, g.Ny), gz)] - p[cell_index(g, gx, wrapi(gy - 1, g.Ny), gz)]);
    const T wi = ws[o] + fac * (p[cell_index(g, gx, gy, wrapi(gz + 1, g.Nz))] - p[cell_index(g, gx, gy, wrapi(gz - 1, g.Nz))]);
    const double den = 1 + lf * dt * X;
    const double dpf = implicit ? lf / (den * den) : lf;
    const double gpen = wdrag * h3 * (vx - (double)ui) +
                        dt * ((double)lu[o] * (vx - (double)ui) + (double)lv[o] * (vy - (double)vi) +
                              (double)lw[o] * (vz - (double)wi));
    double dX[3];
    moll_dchi_dc(g, gx, gy, gz, cx, cy, cz, R, e, dX);
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] += dpf * gpen * dX[k];
  }
  breduce<3>(acc, part + 3 * (size_t)s);
}

template <typename T>
static int adj_body_t(Ctx& c, AdjBodyPhase phase, const AdjBodyArgs& a)
{
  Body& B = c.body;
  double* msum = c.partials + 2 * (size_t)c.nb;  // written by k_rhs_obs_mask
  double* gpart = msum + 3 * (size_t)B.nob;
  double* vpart = gpart + B.nob;
  double* cpart = vpart + 6 * (size_t)B.nob;
  const double e = B.moll * c.h;
  switch (phase) {
    case ADJB_FWD_MASK: {  // the forward masking of u*: r_new in ob_rhs, touch flags, sums
      if (a.centre && !B.ob_r0)
        if (int rc = check_hip(c, hipMalloc(&B.ob_r0, (size_t)B.nob * BS3 * c.esz), "hipMalloc")) return rc;
      const int s0 = c.ustar;
      hipLaunchKernelGGL((k_rhs_obs_mask<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                         (const T*)c.f[CUP_CHI], (const T*)c.f[s0], (const T*)c.f[s0 + 1], (const T*)c.f[s0 + 2],
                         (T*)nullptr, (T*)B.ob_rhs, B.ob_touch, c.grid(), c.sc, msum,
                         a.centre ? (T*)B.ob_r0 : (T*)nullptr);
      break;
    }
    case ADJB_FWD_CORRECT:  // r_new -> the corrected RHS the forward Poisson solve reads
      hipLaunchKernelGGL((k_rhs_obs_correct<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.nob, B.ob_touch,
                         (T*)nullptr, (T*)B.ob_rhs, (const double*)msum);
      break;
    case ADJB_CHI_PEN:  // u* in c.ustar, the forward pressure in P, lam in the adjoint fields
      hipLaunchKernelGGL((k_adj_chi_pen<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                         (const T*)c.f[CUP_CHI], (const T*)c.f[c.ustar], (const T*)c.f[c.ustar + 1],
                         (const T*)c.f[c.ustar + 2], (const T*)c.f[CUP_P], (const T*)c.f[CUP_ADJ_U],
                         (const T*)c.f[CUP_ADJ_V], (const T*)c.f[CUP_ADJ_W], c.grid(), c.sc, a.lambda, a.implicit,
                         a.wdrag, a.vel[0], a.vel[1], a.vel[2], B.center[0], B.center[1], B.center[2], B.radius, e,
                         cpart);
      hipLaunchKernelGGL(k_sum_rows, dim3(3), dim3(NT), 0, c.stream, (const double*)cpart, B.nob, 3, c.sc->red + 11);
      break;
    case ADJB_PENALIZE:  // penalization and drag, dJ/dV into sc->red[8..10], dJ/dW into red[17..19]
      hipLaunchKernelGGL((k_adj_penalize<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                         (const T*)c.f[CUP_CHI], (T*)c.f[CUP_ADJ_U], (T*)c.f[CUP_ADJ_V], (T*)c.f[CUP_ADJ_W],
                         c.grid(), c.sc, a.lambda, a.implicit, a.wdrag, B.com[0], B.com[1], B.com[2], vpart);
      hipLaunchKernelGGL(k_sum_rows, dim3(3), dim3(NT), 0, c.stream, (const double*)vpart, B.nob, 6, c.sc->red + 8);
      hipLaunchKernelGGL(k_sum_rows, dim3(3), dim3(NT), 0, c.stream, (const double*)(vpart + 3), B.nob, 6,
                         c.sc->red + 17);
      break;
    case ADJB_MASK: {  // the masking's transpose on the adjoint pressure; with the centre, red[14..16]
      const double yscale = a.dt * a.dt / (c.h * c.h * c.h);
      hipLaunchKernelGGL((k_adj_mask_g<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                         (const T*)c.f[CUP_CHI], (const T*)B.ob_rhs, B.ob_touch, (const T*)c.f[CUP_P], msum, B.nob,
                         gpart);
      hipLaunchKernelGGL((k_adj_mask_apply<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                         (const T*)c.f[CUP_CHI], (const T*)B.ob_rhs, B.ob_touch, (T*)c.f[CUP_P], msum, B.nob, gpart,
                         c.grid(), (const T*)B.ob_r0, yscale, B.center[0], B.center[1], B.center[2], B.radius, e,
                         a.centre ? cpart : (double*)nullptr);
      if (a.centre)
        hipLaunchKernelGGL(k_sum_rows, dim3(3), dim3(NT), 0, c.stream, (const double*)cpart, B.nob, 3,
                           c.sc->red + 14);
      break;
    }
  }
  return check_hip(c, hipGetLastError(), "adjoint body");
}

int launch_adjoint_body(Ctx& c, AdjBodyPhase phase, const AdjBodyArgs& a)
{
  if (!c.body.present) return set_error(c, CUP_E_STATE, "adjoint: no obstacle");
  if (a.centre && !(c.body.moll > 0))
    return set_error(c, CUP_E_STATE, "adjoint: the centre gradient needs the mollified chi");
  return c.precision == 4 ? adj_body_t<float>(c, phase, a) : adj_body_t<double>(c, phase, a);
}

__global__ __launch_bounds__(NT) void k_sum_rows(const double* __restrict__ part, int n, int width,
                                                 double* __restrict__ out)
{
  const int i = blockIdx.x;
  double a[1] = {0};
  for (int j = threadIdx.x; j < n; j += NT) a[0] += part[(size_t)j * width + i];
  breduce<1>(a, out + i);
}

template <typename T>
static int penalize_t(Ctx& c, double lambda, int implicit, const double vel[3], const double om[3],
                      const double uinf[3], bool refresh, double* ft_dev = nullptr)
{
  Body& B = c.body;
  double* bp = c.partials + 2 * (size_t)c.nb;
  hipLaunchKernelGGL((k_penalize<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                     (const double*)B.ob_udef, (const T*)c.f[CUP_CHI],
                     (T*)c.f[CUP_U], (T*)c.f[CUP_V], (T*)c.f[CUP_W], c.grid(), c.sc, lambda, implicit, vel[0], vel[1],
                     vel[2], om[0], om[1], om[2], B.com[0], B.com[1], B.com[2], uinf[0], uinf[1], uinf[2], bp,
                     refresh ? c.ke_part : (double*)nullptr, umax_slots_dev(c, c.umax_bank));
  hipLaunchKernelGGL(k_sum_rows, dim3(6), dim3(NT), 0, c.stream, (const double*)bp, B.nob, 6,
                     ft_dev ? ft_dev : c.sc->red);
  return check_hip(c, hipGetLastError(), "k_penalize");
}

int launch_penalize(Ctx& c, double lambda, int implicit, const double vel[3], const double omega[3], double* ft_dev)
{
  const double zero[3] = {0, 0, 0};
  return c.precision == 4 ? penalize_t<float>(c, lambda, implicit, vel, omega, zero, false, ft_dev)
                          : penalize_t<double>(c, lambda, implicit, vel, omega, zero, false, ft_dev);
}

int launch_penalize_step(Ctx& c, double lambda, int implicit, const double vel[3], const double omega[3],
                         const double uinf[3], double* ft_dev)
{
  return c.precision == 4 ? penalize_t<float>(c, lambda, implicit, vel, omega, uinf, true, ft_dev)
                          : penalize_t<double>(c, lambda, implicit, vel, omega, uinf, true, ft_dev);
}

// ------------------------------------------------------------ body momenta

// KernelIntegrateFluidMomenta (ObstaclesUpdate.cpp:15-131) on one obstacle block, udef = 0:
// part[29 s + q] in KernelFinalizeObstacleVel's order (:139-158): V; sum chi u dv (3);
// sum chi (p x u) dv (3); J0..J5; and with implicit penalization GfX, Gp (3), Gj (6),
// Gu (3), Ga (3) weighted by penalFac = dv lambda dt chi / (1 + chi lambda dt), zero
// otherwise. p = x - CoM; the obstacle's chi, cells with chi <= 0 skipped.
template <typename T>
__global__ __launch_bounds__(NT) void k_obstacle_momenta(const int* __restrict__ ob_ids,
                                                         const double* __restrict__ ob_chi,
                                                         const double* __restrict__ ob_udef, const T* __restrict__ u,
                                                         const T* __restrict__ v, const T* __restrict__ w, GridArgs g,
                                                         double lambdt, int implicit, double cmx, double cmy,
                                                         double cmz, double* __restrict__ part)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  const double dv = g.h * g.h * g.h;
  double M[29];
#pragma unroll
  for (int q = 0; q < 29; ++q) M[q] = 0;
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const double X = ob_chi[(size_t)s * BS3 + c];
    if (!(X > 0)) continue;
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const double px = (bx * BS + ix + 0.5) * g.h - cmx, py = (by * BS + iy + 0.5) * g.h - cmy,
                 pz = (bz * BS + iz + 0.5) * g.h - cmz;
    const size_t o = (size_t)b * BS3 + c;
    const double U = u[o], V = v[o], W = w[o];
    const double Xd = X * dv;
    M[0] += Xd;
    M[1] += Xd * U;
    M[2] += Xd * V;
    M[3] += Xd * W;
    M[4] += Xd * (py * W - pz * V);
    M[5] += Xd * (pz * U - px * W);
    M[6] += Xd * (px * V - py * U);
    M[7] += Xd * (py * py + pz * pz);
    M[8] += Xd * (px * px + pz * pz);
    M[9] += Xd * (px * px + py * py);
    M[10] -= Xd * px * py;
    M[11] -= Xd * px * pz;
    M[12] -= Xd * py * pz;
    if (implicit) {
      const double pf = dv * lambdt * X / (1 + X * lambdt);
      M[13] += pf;
      M[14] += pf * px;
      M[15] += pf * py;
      M[16] += pf * pz;
      M[17] += pf * (py * py + pz * pz);
      M[18] += pf * (px * px + pz * pz);
      M[19] += pf * (px * px + py * py);
      M[20] -= pf * px * py;
      M[21] -= pf * px * pz;
      M[22] -= pf * py * pz;
      // u - udef (ObstaclesUpdate.cpp:117-126); no udef: - 0, the same values
      double d0 = 0, d1 = 0, d2 = 0;
      if (ob_udef) {
        const double* D = ob_udef + (size_t)s * 3 * BS3;
        d0 = D[c];
        d1 = D[BS3 + c];
        d2 = D[2 * BS3 + c];
      }
      const double DU = U - d0, DV = V - d1, DW = W - d2;
      M[23] += pf * DU;
      M[24] += pf * DV;
      M[25] += pf * DW;
      M[26] += pf * (py * DW - pz * DV);
      M[27] += pf * (pz * DU - px * DW);
      M[28] += pf * (px * DV - py * DU);
    }
  }
  breduce<29>(M, part + 29 * (size_t)s);
}

template <typename T>
static int momenta_t(Ctx& c, double lambdt, int implicit)
{
  Body& B = c.body;
  double* bp = c.partials + 2 * (size_t)c.nb;
  hipLaunchKernelGGL((k_obstacle_momenta<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi,
                     (const double*)B.ob_udef, (const T*)c.f[CUP_U], (const T*)c.f[CUP_V], (const T*)c.f[CUP_W], c.grid(), lambdt, implicit,
                     B.com[0], B.com[1], B.com[2], bp);
  hipLaunchKernelGGL(k_sum_rows, dim3(29), dim3(NT), 0, c.stream, (const double*)bp, B.nob, 29, c.sc->red);
  return check_hip(c, hipGetLastError(), "k_obstacle_momenta");
}

int launch_obstacle_momenta(Ctx& c, double dt, double lambda, int implicit)
{
  return c.precision == 4 ? momenta_t<float>(c, lambda * dt, implicit) : momenta_t<double>(c, lambda * dt, implicit);
}

// ---------------------------------------------------------------- udef momenta

// KernelIntegrateUdefMomenta::visit (ObstaclesCreate.cpp:280-331) on one obstacle block:
// over cells with chi > 0, p = x - CoM and the previous linear correction c_old,
// part[13 s + q] = V, F (3), T = p x (udef - c_old) (3), J0..J5, each weighted by chi dv.
__global__ __launch_bounds__(NT) void k_udef_momenta(const int* __restrict__ ob_ids, const double* __restrict__ ob_chi,
                                                     const double* __restrict__ ob_udef, GridArgs g, double cmx,
                                                     double cmy, double cmz, double ocx, double ocy, double ocz,
                                                     double* __restrict__ part)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  const double dv = g.h * g.h * g.h;
  const double* U = ob_udef + (size_t)s * 3 * BS3;
  double M[13];
#pragma unroll
  for (int q = 0; q < 13; ++q) M[q] = 0;
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const double X = ob_chi[(size_t)s * BS3 + c];
    if (X <= 0) continue;
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const double p0 = (bx * BS + ix + 0.5) * g.h - cmx, p1 = (by * BS + iy + 0.5) * g.h - cmy,
                 p2 = (bz * BS + iz + 0.5) * g.h - cmz;
    const double ux = U[c], uy = U[BS3 + c], uz = U[2 * BS3 + c];
    const double dU = ux - ocx, dV = uy - ocy, dW = uz - ocz;
    M[0] += X * dv;
    M[1] += X * ux * dv;
    M[2] += X * uy * dv;
    M[3] += X * uz * dv;
    M[4] += X * (p1 * dW - p2 * dV) * dv;
    M[5] += X * (p2 * dU - p0 * dW) * dv;
    M[6] += X * (p0 * dV - p1 * dU) * dv;
    M[7] += X * (p1 * p1 + p2 * p2) * dv;
    M[8] += X * (p0 * p0 + p2 * p2) * dv;
    M[9] += X * (p0 * p0 + p1 * p1) * dv;
    M[10] -= X * p0 * p1 * dv;
    M[11] -= X * p0 * p2 * dv;
    M[12] -= X * p1 * p2 * dv;
  }
  breduce<13>(M, part + 13 * (size_t)s);
}

// KernelRemoveUdefMomenta::visit (ObstaclesCreate.cpp:392-439): udef -= ct + cw x (x - CoM) on
// every cell of the obstacle blocks.
__global__ __launch_bounds__(NT) void k_udef_remove(const int* __restrict__ ob_ids, double* __restrict__ ob_udef,
                                                    GridArgs g, double cmx, double cmy, double cmz, double ctx,
                                                    double cty, double ctz, double cwx, double cwy, double cwz)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  double* U = ob_udef + (size_t)s * 3 * BS3;
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const double p0 = (bx * BS + ix + 0.5) * g.h - cmx, p1 = (by * BS + iy + 0.5) * g.h - cmy,
                 p2 = (bz * BS + iz + 0.5) * g.h - cmz;
    U[c] -= ctx + (cwy * p2 - cwz * p1);
    U[BS3 + c] -= cty + (cwz * p0 - cwx * p2);
    U[2 * BS3 + c] -= ctz + (cwx * p1 - cwy * p0);
  }
}

// KernelIntegrateUdefMomenta, KernelAccumulateUdefMomenta (ObstaclesCreate.cpp:334-384: the
// corrections F/V and J^-1 T, J inverted by cofactors on the host) and KernelRemoveUdefMomenta
// for the active body. out: mass, linear correction (3), angular correction (3), J0..J5.
int launch_udef_remove(Ctx& c, double out[13])
{
  Body& B = c.body;
  const size_t n = (size_t)B.nob * 3 * BS3;
  if (!B.ob_udef) {  // no deformation velocity set: zero
    if (check_hip(c, hipMalloc(&B.ob_udef, n * sizeof(double)), "hipMalloc udef")) return CUP_E_NOMEM;
    const std::vector<double> zero(n, 0.0);
    if (int rc = check_hip(c, hipMemcpy(B.ob_udef, zero.data(), n * sizeof(double), hipMemcpyHostToDevice), "udef"))
      return rc;
  }
  if (int rc = ensure_partials(c, 32 * (size_t)B.nob + 2 * (size_t)c.nb)) return rc;
  double* bp = c.partials + 2 * (size_t)c.nb;
  hipLaunchKernelGGL(k_udef_momenta, dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_chi, B.ob_udef, c.grid(),
                     B.com[0], B.com[1], B.com[2], B.corr[0], B.corr[1], B.corr[2], bp);
  hipLaunchKernelGGL(k_sum_rows, dim3(13), dim3(NT), 0, c.stream, (const double*)bp, B.nob, 13, c.sc->red);
  if (int rc = check_hip(c, hipGetLastError(), "k_udef_momenta")) return rc;
  double M[13];
  if (int rc = check_hip(c, hipMemcpyAsync(M, c.sc->red, sizeof(M), hipMemcpyDeviceToHost, c.stream), "udef"))
    return rc;
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "udef")) return rc;
  if (!(M[0] > 0)) return set_error(c, CUP_E_STATE, "remove_udef_momenta: the body has no volume on the grid");
  const double* J = M + 7;  // J0..J2 diagonal, J3 = xy, J4 = xz, J5 = yz
  const double c0 = J[1] * J[2] - J[5] * J[5], c1 = J[0] * J[2] - J[4] * J[4], c2 = J[0] * J[1] - J[3] * J[3];
  const double c3 = J[4] * J[5] - J[3] * J[2], c4 = J[3] * J[5] - J[4] * J[1], c5 = J[3] * J[4] - J[0] * J[5];
  const double det = J[0] * c0 + J[3] * c3 + J[4] * c4;
  // a singular J (a body of one cell, or of collinear cells) has no angular correction, as
  // invertSym returns a zero inverse for |det| <= DBL_MIN (MatArrayMath.h:18-38)
  const bool singular = std::fabs(det) <= std::numeric_limits<double>::min();
  double iJ[6] = {0, 0, 0, 0, 0, 0};
  if (!singular) {
    iJ[0] = c0 / det; iJ[1] = c1 / det; iJ[2] = c2 / det;
    iJ[3] = c3 / det; iJ[4] = c4 / det; iJ[5] = c5 / det;
  }
  const double ct[3] = {M[1] / M[0], M[2] / M[0], M[3] / M[0]};
  const double cw[3] = {iJ[0] * M[4] + iJ[3] * M[5] + iJ[4] * M[6], iJ[3] * M[4] + iJ[1] * M[5] + iJ[5] * M[6],
                        iJ[4] * M[4] + iJ[5] * M[5] + iJ[2] * M[6]};
  hipLaunchKernelGGL(k_udef_remove, dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_udef, c.grid(), B.com[0],
                     B.com[1], B.com[2], ct[0], ct[1], ct[2], cw[0], cw[1], cw[2]);
  if (int rc = check_hip(c, hipGetLastError(), "k_udef_remove")) return rc;
  for (int d = 0; d < 3; ++d) {
    B.corr[d] = ct[d];
    B.corr[3 + d] = cw[d];
  }
  out[0] = M[0];
  for (int d = 0; d < 3; ++d) {
    out[1 + d] = ct[d];
    out[4 + d] = cw[d];
  }
  for (int q = 0; q < 6; ++q) out[7 + q] = J[q];
  return check_hip(c, hipStreamSynchronize(c.stream), "udef");
}

// ---------------------------------------------------------------- surface forces

// KernelComputeForces::visit (FluidSolidForces.cpp:42-157) on one obstacle block:
// 22 QoIs in ObstacleBlock::sumQoI order (ObstacleBlock.h:81-93); the deformation powers
// (defPower, defPowerBnd, FluidSolidForces.cpp:110-112,145-147) from the body's udef when set.
template <typename T>
__global__ __launch_bounds__(NT) void k_surface_forces(const int* __restrict__ ob_ids, const double* __restrict__ surf,
                                                       const double* __restrict__ ob_udef, const T* __restrict__ u, const T* __restrict__ v,
                                                       const T* __restrict__ w, const T* __restrict__ p, GridArgs g,
                                                       double nu, double vx, double vy, double vz, double ox,
                                                       double oy, double oz, double cmx, double cmy, double cmz,
                                                       double* __restrict__ part)
{
  const int s = blockIdx.x, b = ob_ids[s];
  int bx, by, bz;
  bcoords(g, b, bx, by, bz);
  double Q[22];
#pragma unroll
  for (int i = 0; i < 22; ++i) Q[i] = 0;
  const double vn = sqrt(vx * vx + vy * vy + vz * vz);
  double vu[3] = {0, 0, 0};
  if (vn > 1e-9) { vu[0] = vx / vn; vu[1] = vy / vn; vu[2] = vz / vn; }
  const double _1oH = nu / g.h;
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const double* sd = surf + ((size_t)s * BS3 + c) * 4;
    if (!(sd[0] > DEPS)) continue;
    const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
    const int gx = bx * BS + ix, gy = by * BS + iy, gz = bz * BS + iz;
    // neighbours through the grid's boundaries (a wall face negates the velocities)
    int fE, fW, fN, fS, fB, fF;
    const size_t E = nb_cell(g, gx, gy, gz, 0, 1, &fE), W = nb_cell(g, gx, gy, gz, 0, -1, &fW);
    const size_t N = nb_cell(g, gx, gy, gz, 1, 1, &fN), S = nb_cell(g, gx, gy, gz, 1, -1, &fS);
    const size_t Bk = nb_cell(g, gx, gy, gz, 2, 1, &fB), F = nb_cell(g, gx, gy, gz, 2, -1, &fF);
    const size_t C = (size_t)b * BS3 + c;
    auto vel = [](const T* f, size_t i, int bcf) { return (double)halo_value(f[i], HK_VEL, bcf); };
    const double D11 = _1oH * (vel(u, E, fE) - vel(u, W, fW));
    const double D22 = _1oH * (vel(v, N, fN) - vel(v, S, fS));
    const double D33 = _1oH * (vel(w, Bk, fB) - vel(w, F, fF));
    const double D12 = .5 * _1oH * (vel(u, N, fN) - vel(u, S, fS) + vel(v, E, fE) - vel(v, W, fW));
    const double D13 = .5 * _1oH * (vel(u, Bk, fB) - vel(u, F, fF) + vel(w, E, fE) - vel(w, W, fW));
    const double D23 = .5 * _1oH * (vel(w, N, fN) - vel(w, S, fS) + vel(v, Bk, fB) - vel(v, F, fF));
    const double P = p[C];
    const double nX = sd[1], nY = sd[2], nZ = sd[3];
    const double fXV = D11 * nX + D12 * nY + D13 * nZ;
    const double fYV = D12 * nX + D22 * nY + D23 * nZ;
    const double fZV = D13 * nX + D23 * nY + D33 * nZ;
    const double fXP = -P * nX, fYP = -P * nY, fZP = -P * nZ;
    const double fXT = fXV + fXP, fYT = fYV + fYP, fZT = fZV + fZP;
    const double uX = u[C], uY = v[C], uZ = w[C];
    const double px = (gx + 0.5) * g.h, py = (gy + 0.5) * g.h, pz = (gz + 0.5) * g.h;
    Q[0] += fXT; Q[1] += fYT; Q[2] += fZT;
    Q[3] += fXP; Q[4] += fYP; Q[5] += fZP;
    Q[6] += fXV; Q[7] += fYV; Q[8] += fZV;
    Q[9] += (py - cmy) * fZT - (pz - cmz) * fYT;
    Q[10] += (pz - cmz) * fXT - (px - cmx) * fZT;
    Q[11] += (px - cmx) * fYT - (py - cmy) * fXT;
    Q[12] += nY * uZ - nZ * uY;
    Q[13] += nZ * uX - nX * uZ;
    Q[14] += nX * uY - nY * uX;
    const double fpar = fXT * vu[0] + fYT * vu[1] + fZT * vu[2];
    Q[16] += .5 * (fpar + fabs(fpar));
    Q[15] -= .5 * (fpar - fabs(fpar));
    const double powOut = fXT * uX + fYT * uY + fZT * uZ;
    Q[17] += powOut;
    Q[18] += fmin(0.0, powOut);
    if (ob_udef) {
      const double* D = ob_udef + (size_t)s * 3 * BS3;
      const double powDef = fXT * D[c] + fYT * D[BS3 + c] + fZT * D[2 * BS3 + c];
      Q[19] += powDef;
      Q[20] += fmin(0.0, powDef);
    }
    const double rx = px - cmx, ry = py - cmy, rz = pz - cmz;
    Q[21] += fXT * (vx + oy * rz - ry * oz) + fYT * (vy + oz * rx - rz * ox) + fZT * (vz + ox * ry - rx * oy);
  }
  breduce<22>(Q, part + 22 * (size_t)s);
}

template <typename T>
static int surface_forces_t(Ctx& c, double nu, const double vel[3], const double om[3])
{
  Body& B = c.body;
  double* bp = c.partials + 2 * (size_t)c.nb;
  hipLaunchKernelGGL((k_surface_forces<T>), dim3(B.nob), dim3(NT), 0, c.stream, B.ob_ids, B.ob_surf,
                     (const double*)B.ob_udef, (const T*)c.f[CUP_U], (const T*)c.f[CUP_V], (const T*)c.f[CUP_W], (const T*)c.f[CUP_P], c.grid(),
                     nu, vel[0], vel[1], vel[2], om[0], om[1], om[2], B.com[0], B.com[1], B.com[2], bp);
  hipLaunchKernelGGL(k_sum_rows, dim3(22), dim3(NT), 0, c.stream, (const double*)bp, B.nob, 22, c.sc->red);
  return check_hip(c, hipGetLastError(), "k_surface_forces");
}

int launch_surface_forces(Ctx& c, double nu, const double vel[3], const double omega[3])
{
  return c.precision == 4 ? surface_forces_t<float>(c, nu, vel, omega) : surface_forces_t<double>(c, nu, vel, omega);
}

}  // namespace cup
