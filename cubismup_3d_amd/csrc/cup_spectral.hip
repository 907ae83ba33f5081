// cup_spectral.hip -- spectral pieces of isotropic turbulence: the seeded initial
// condition and band-limited linear forcing in the reference's decomposition.
#include "cup_internal.h"
#include "cup_forcing_dev.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace cup {

// The forcing's plane stage splits each z plane into up to FR_PARTS_MAX equal row ranges of
// at least 64 rows, one workgroup each.
constexpr int FR_PARTS_MAX = 4;
inline int forcing_plane_parts(int ny)
{
  int p = ny / 64;
  p = p < 1 ? 1 : (p > FR_PARTS_MAX ? FR_PARTS_MAX : p);
  while (ny % p) --p;
  return p;
}

__device__ __forceinline__ int signed_k(int i, int n) { return i <= n / 2 ? i : i - n; }

// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as
// 1, 2, 3", SC'11): 10 rounds of the 4x32 multiply-xor bijection with a Weyl key schedule.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4])
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Packed half spectrum of the seeded isotropic-turbulence velocity, in FP64. Entry
// (kz, ky, s): s >= 1 is mode (kx=s, ky, kz); s = 0 is mode (0, ky, kz) (+ i * the
// kx = Nx/2 plane, which is zero). Construction per mode: SpectralManipFFTW::_compute_IC
// (SpectralManipFFTW.cpp:254-285) with noise_a[1] for Im(u) and exact Hermitian symmetry.
// EnergySpectrum::interpE (SpectralManip.cpp:96-116) on a table with K increasing: with i the
// first index of K above k (a binary search), the line through (K[i-1], E[i-1]) and
// (K[i], E[i]); below K[0] the reference's E[0] - (E[1] - E[0]) / (K[1] - K[0]) (K[1] - k);
// beyond the table the last segment's line; negative values give 0.
__device__ double spectrum_interp(const double* __restrict__ K, const double* __restrict__ E, int n, double k)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (k < K[mid]) hi = mid;
    else lo = mid + 1;
  }
  const int i = lo;
  double e;
  if (i > 0 && i < n) e = E[i] + (E[i - 1] - E[i]) / (K[i] - K[i - 1]) * (K[i] - k);
  else if (i == 0) e = E[0] - (E[1] - E[0]) / (K[1] - K[0]) * (K[1] - k);
  else e = E[n - 1] + (E[n - 2] - E[n - 1]) / (K[n - 1] - K[n - 2]) * (K[n - 1] - k);
  return e < 0 ? 0.0 : e;
}

__global__ void k_hit_spectrum(C2<double>* __restrict__ su, C2<double>* __restrict__ sv, C2<double>* __restrict__ sw,
                               int N0, int N1, int N2, double wfx, double wfy, double wfz, double k0, uint32_t key0,
                               uint32_t key1, const double* __restrict__ tK, const double* __restrict__ tE, int tn)
{
  const int ncols = N0 / 2;
  const size_t n = (size_t)ncols * N1 * N2;
  const double coef = tn ? 0.0 : 32.0 / (3 * k0) * sqrt(2.0 / M_PI);
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(id % ncols);
    const int ky = (int)((id / ncols) % N1);
    const int kz = (int)(id / ((size_t)ncols * N1));
    int ii = s, jj = signed_k(ky, N1), kk = signed_k(kz, N2);
    C2<double> U{0, 0}, V{0, 0}, W{0, 0};
    const bool nyq = 2 * abs(jj) == N1 || 2 * abs(kk) == N2;
    if (!nyq && !(ii == 0 && jj == 0 && kk == 0)) {
      const bool canonical = kk > 0 || (kk == 0 && (jj > 0 || (jj == 0 && ii > 0)));
      if (!canonical) { ii = -ii; jj = -jj; kk = -kk; }
      const int wi = ii < 0 ? ii + N0 : ii, wj = jj < 0 ? jj + N1 : jj, wk = kk < 0 ? kk + N2 : kk;
      const uint64_t cid = ((uint64_t)wi * N1 + wj) * N2 + wk;
      const double kx = ii * wfx, kyv = jj * wfy, kz_ = kk * wfz;
      const double k2 = kx * kx + kyv * kyv + kz_ * kz_;
      const double k_xy = sqrt(kx * kx + kyv * kyv), k_norm = sqrt(k2);
      const double r = tn ? 0.0 : k_norm / k0;
      const double E_k = tn ? spectrum_interp(tK, tE, tn, k_norm) : coef * r * r * r * r * exp(-2 * r * r);
      const double kn = k_norm / wfx;
      const double amp = sqrt(E_k / (2 * M_PI * kn * kn));
      uint32_t o[4];
      philox4x32_10((uint32_t)cid, (uint32_t)(cid >> 32), 0u, 0u, key0, key1, o);
      const double inv32 = 1.0 / 4294967296.0;
      const double theta1 = (o[0] + 0.5) * inv32 * 2 * M_PI;
      const double theta2 = (o[1] + 0.5) * inv32 * 2 * M_PI;
      const double phi = (o[2] + 0.5) * inv32 * 2 * M_PI;
      const double ca = amp * cos(phi), cb = amp * sin(phi);
      const C2<double> na{ca * cos(theta1), ca * sin(theta1)};
      const C2<double> nb{cb * cos(theta2), cb * sin(theta2)};
      const double fac = k_norm * k_xy, invFac = fac <= 0 ? 0 : 1 / fac;
      U = C2<double>{invFac * (na.x * k_norm * kyv + nb.x * kx * kz_), invFac * (na.y * k_norm * kyv + nb.y * kx * kz_)};
      V = C2<double>{invFac * (nb.x * kyv * kz_ - na.x * k_norm * kx), invFac * (nb.y * kyv * kz_ - na.y * k_norm * kx)};
      W = C2<double>{-nb.x * k_xy / k_norm, -nb.y * k_xy / k_norm};
      if (!canonical) { U.y = -U.y; V.y = -V.y; W.y = -W.y; }
    }
    su[id] = U;
    sv[id] = V;
    sw[id] = W;
  }
}

template <typename T>
__global__ void k_scale3(T* __restrict__ a, T* __restrict__ b, T* __restrict__ c, size_t n, double s)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    a[i] = (T)(a[i] * s);
    b[i] = (T)(b[i] * s);
    c[i] = (T)(c[i] * s);
  }
}

static dim3 grid_for(size_t n)
{
  size_t g = (n + NT - 1) / NT;
  if (g > 65536) g = 65536;
  return dim3((unsigned)g);
}

static int alloc_extra_spectra(Ctx& c)
{
  if (!c.spec2 && check_hip(c, hipMalloc(&c.spec2, c.ncell * c.esz), "hipMalloc spectrum 2")) return CUP_E_NOMEM;
  if (!c.spec3 && check_hip(c, hipMalloc(&c.spec3, c.ncell * c.esz), "hipMalloc spectrum 3")) return CUP_E_NOMEM;
  return CUP_OK;
}

int launch_ic_hit(Ctx& c, double k0, const double* K, const double* E, int tn, double u_rms, uint64_t seed)
{
  if (c.precision == 4) {
    // Generate in an FP64 context of the same grid, then store rounded to FP32.
    cup_grid_desc d;
    for (int a = 0; a < 3; ++a) d.bpd[a] = c.bpd[a];
    d.maxextent = c.maxextent;
    d.precision = 8;
    d.decomposition = CUP_FUSED;
    d.device = c.dev;
    d.bc[0] = d.bc[1] = d.bc[2] = CUP_BC_PERIODIC;
    cup_ctx* t = nullptr;
    if (int rc = cup_create(&d, &t)) return set_error(c, rc, "ic_hit: FP64 work context");
    int rc = tn ? cup_ic_hit_spectrum(t, K, E, tn, u_rms, seed) : cup_ic_hit(t, k0, u_rms, seed);
    std::vector<double> host(c.ncell);
    for (int f = 0; f < 8 && !rc; ++f) {
      rc = cup_get_field(t, f, host.data());
      if (!rc) rc = cup_set_field(reinterpret_cast<cup_ctx*>(&c), f, host.data());
    }
    if (rc) c.err = std::string("ic_hit: ") + cup_last_error(t);
    cup_destroy(t);
    return rc;
  }
  if (int rc = alloc_extra_spectra(c)) return rc;
  for (int k = 0; k < 8; ++k)
    if (int rc = launch_fill(c, c.f[k], c.ncell, 0.0)) return rc;
  const size_t n = (size_t)(c.N[0] / 2) * c.N[1] * c.N[2];
  const double wf[3] = {2 * M_PI / c.extent[0], 2 * M_PI / c.extent[1], 2 * M_PI / c.extent[2]};
  double* table = nullptr;  // the target table on the device, freed once the spectrum is built
  if (tn) {
    if (check_hip(c, hipMalloc(&table, 2 * (size_t)tn * sizeof(double)), "hipMalloc spectrum table")) return CUP_E_NOMEM;
    hipMemcpy(table, K, tn * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(table + tn, E, tn * sizeof(double), hipMemcpyHostToDevice);
  }
  hipLaunchKernelGGL(k_hit_spectrum, grid_for(n), dim3(NT), 0, c.stream, (C2<double>*)c.spec, (C2<double>*)c.spec2,
                     (C2<double>*)c.spec3, c.N[0], c.N[1], c.N[2], wf[0], wf[1], wf[2], k0, (uint32_t)seed,
                     (uint32_t)(seed >> 32), (const double*)table, (const double*)(table ? table + tn : nullptr), tn);
  const int krc = check_hip(c, hipGetLastError(), "k_hit_spectrum");
  if (table) {
    hipStreamSynchronize(c.stream);
    hipFree(table);
  }
  if (krc) return krc;
  if (int rc = launch_fft_inverse_field(c, c.spec, c.f[CUP_U])) return rc;
  if (int rc = launch_fft_inverse_field(c, c.spec2, c.f[CUP_V])) return rc;
  if (int rc = launch_fft_inverse_field(c, c.spec3, c.f[CUP_W])) return rc;
  const double zero[3] = {0, 0, 0};
  if (int rc = launch_max_u(c, zero)) return rc;
  if (int rc = launch_reduce(c, c.nb, 2, 1u, c.sc->red)) return rc;
  double red[2];
  if (int rc = check_hip(c, hipMemcpyAsync(red, c.sc->red, sizeof(red), hipMemcpyDeviceToHost, c.stream), "ic_hit"))
    return rc;
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "ic_hit")) return rc;
  if (!(u_rms > 0)) return CUP_OK;  // a target table's energy is kept
  const double scale = u_rms / std::sqrt(red[1] / (3.0 * (double)c.ncell));
  hipLaunchKernelGGL((k_scale3<double>), grid_for(c.ncell), dim3(NT), 0, c.stream, (double*)c.f[CUP_U],
                     (double*)c.f[CUP_V], (double*)c.f[CUP_W], c.ncell, scale);
  return check_hip(c, hipGetLastError(), "k_scale3");
}

// ------------------------------------------------------------- forcing (plain)

// SpectralManipFFTW::_compute_largeModesForcing (SpectralManipFFTW.cpp:25-108) over the
// packed half spectra of u, v, w: per-mode energy, central-difference enstrophy, E/|k|,
// energy of 0 < k^2 <= 4, and every other mode zeroed. The x axis is the halved one here
// (weight 2 for 0 < kx < Nx/2); slot 0 holds the kx = 0 and kx = Nx/2 planes as A + iB and
// is handled in (k, -k) pairs. part[4*wg + {0..3}] = tke, eps, tke_filtered, lIntegral sums.
struct ModeStat {
  double E, eps, lint;
  bool keep;
};

__device__ __forceinline__ ModeStat mode_stat(int ii, int jj, int kk, int N0, int N1, int N2, double wfx, double wfy,
                                              double wfz, C2<double> U, C2<double> V, C2<double> W)
{
  ModeStat m;
  const double kx = ii * wfx, ky = jj * wfy, kz = kk * wfz;
  const double k2 = kx * kx + ky * ky + kz * kz;
  const double dX = 2 * sin(2 * M_PI / N0 * ii), dY = 2 * sin(2 * M_PI / N1 * jj), dZ = 2 * sin(2 * M_PI / N2 * kk);
  // d/dx -> i*dX: (a+ib)*i*d = (-b d, a d)
  const double OXr = -W.y * dY + V.y * dZ, OXi = W.x * dY - V.x * dZ;
  const double OYr = -U.y * dZ + W.y * dX, OYi = U.x * dZ - W.x * dX;
  const double OZr = -V.y * dX + U.y * dY, OZi = V.x * dX - U.x * dY;
  m.E = 0.5 * (U.x * U.x + U.y * U.y + V.x * V.x + V.y * V.y + W.x * W.x + W.y * W.y);
  m.eps = 0.5 * (OXr * OXr + OXi * OXi + OYr * OYr + OYi * OYi + OZr * OZr + OZi * OZi);
  m.lint = k2 > 0 ? m.E / sqrt(k2) : 0;
  m.keep = k2 > 0 && k2 <= 4;
  return m;
}

template <typename T>
__device__ __forceinline__ C2<double> ld(const C2<T>* p, size_t i)
{
  const C2<T> v = p[i];
  return C2<double>{(double)v.x, (double)v.y};
}

template <typename T>
__global__ __launch_bounds__(NT) void k_forcing_filter(C2<T>* __restrict__ su, C2<T>* __restrict__ sv,
                                                       C2<T>* __restrict__ sw, int N0, int N1, int N2, double wfx,
                                                       double wfy, double wfz, double* __restrict__ part)
{
  const int ncols = N0 / 2;
  const size_t n = (size_t)ncols * N1 * N2;
  double acc[4] = {0, 0, 0, 0};
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(id % ncols);
    const int ky = (int)((id / ncols) % N1);
    const int kz = (int)(id / ((size_t)ncols * N1));
    const int jj = signed_k(ky, N1), kk = signed_k(kz, N2);
    if (s > 0) {
      const ModeStat m = mode_stat(s, jj, kk, N0, N1, N2, wfx, wfy, wfz, ld(su, id), ld(sv, id), ld(sw, id));
      acc[0] += 2 * m.E; acc[1] += 2 * m.eps; acc[3] += 2 * m.lint;
      if (m.keep) acc[2] += 2 * m.E;
      else { su[id] = C2<T>{0, 0}; sv[id] = C2<T>{0, 0}; sw[id] = C2<T>{0, 0}; }
      continue;
    }
    // slot 0: pair (ky,kz) with its mirror (-ky,-kz); the canonical member does both.
    // Exactly one member of each pair is canonical: order pairs by raw (kz, ky) index
    // (signed indices are ambiguous on the Nyquist planes, where k and -k share a sign).
    const int my = ky == 0 ? 0 : N1 - ky, mz = kz == 0 ? 0 : N2 - kz;
    const bool self = my == ky && mz == kz;
    const bool canon = kz < mz || (kz == mz && ky <= my);
    if (!canon) continue;
    const size_t mid = ((size_t)mz * N1 + my) * ncols;
    C2<double> Z[3] = {ld(su, id), ld(sv, id), ld(sw, id)};
    C2<double> Zm[3] = {ld(su, mid), ld(sv, mid), ld(sw, mid)};
    C2<double> A[3], B[3];
    for (int q = 0; q < 3; ++q) {
      A[q] = C2<double>{0.5 * (Z[q].x + Zm[q].x), 0.5 * (Z[q].y - Zm[q].y)};
      B[q] = C2<double>{0.5 * (Z[q].y + Zm[q].y), 0.5 * (Zm[q].x - Z[q].x)};
    }
    const double w = self ? 1.0 : 2.0;
    const ModeStat ma = mode_stat(0, jj, kk, N0, N1, N2, wfx, wfy, wfz, A[0], A[1], A[2]);
    const ModeStat mb = mode_stat(N0 / 2, jj, kk, N0, N1, N2, wfx, wfy, wfz, B[0], B[1], B[2]);
    acc[0] += w * (ma.E + mb.E);
    acc[1] += w * (ma.eps + mb.eps);
    acc[3] += w * (ma.lint + mb.lint);
    if (ma.keep) acc[2] += w * ma.E;
    if (mb.keep) acc[2] += w * mb.E;
    C2<T>* sp[3] = {su, sv, sw};
    for (int q = 0; q < 3; ++q) {
      const C2<double> a = ma.keep ? A[q] : C2<double>{0, 0};
      const C2<double> bq = mb.keep ? B[q] : C2<double>{0, 0};
      // Z = a + i b ; Z(-k) = conj(a) + i conj(b)
      sp[q][id] = C2<T>{(T)(a.x - bq.y), (T)(a.y + bq.x)};
      if (!self) sp[q][mid] = C2<T>{(T)(a.x + bq.y), (T)(bq.x - a.y)};
    }
  }
  // block reduction
  __shared__ double red[NT / 64][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = 0; i < 4; ++i) acc[i] = wave_sum(acc[i]);
  if (lane == 0)
    for (int i = 0; i < 4; ++i) red[wv][i] = acc[i];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < 4; ++i) {
      double a = 0;
      for (int q = 0; q < NT / 64; ++q) a += red[q][i];
      part[4 * (size_t)blockIdx.x + i] = a;
    }
}

template <typename T>
__global__ void k_axpy_blocks(T* __restrict__ y, const T* __restrict__ x, size_t n, double a)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (T)((double)y[i] + a * (double)x[i]);
}

template <typename T>
static int forcing_plain_t(Ctx& c, double dt, double sim_dt, double nu, double tke_target, double inj_rate,
                           double stats[5], int field)
{
  if (int rc = alloc_extra_spectra(c)) return rc;
  const int v0 = field >= 0 ? field : (c.decomposition == CUP_FUSED ? c.ustar : CUP_U);
  // SpectralForcing::_cub2fftw + runFwd: three forward transforms.
  if (int rc = launch_fft_forward_field(c, c.f[v0], c.spec)) return rc;
  if (int rc = launch_fft_forward_field(c, c.f[v0 + 1], c.spec2)) return rc;
  if (int rc = launch_fft_forward_field(c, c.f[v0 + 2], c.spec3)) return rc;
  const size_t n = (size_t)(c.N[0] / 2) * c.N[1] * c.N[2];
  dim3 g = grid_for(n);
  if (g.x > 4096) g.x = 4096;
  if (int rc = ensure_partials(c, 4 * (size_t)g.x)) return rc;
  const double wf[3] = {2 * M_PI / c.extent[0], 2 * M_PI / c.extent[1], 2 * M_PI / c.extent[2]};
  hipLaunchKernelGGL((k_forcing_filter<T>), g, dim3(NT), 0, c.stream, (C2<T>*)c.spec, (C2<T>*)c.spec2,
                     (C2<T>*)c.spec3, c.N[0], c.N[1], c.N[2], wf[0], wf[1], wf[2], c.partials);
  if (int rc = check_hip(c, hipGetLastError(), "k_forcing_filter")) return rc;
  if (int rc = launch_reduce(c, (int)g.x, 4, 0u, c.sc->red)) return rc;
  if (int rc = check_hip(c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream),
                         "forcing"))
    return rc;
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "forcing")) return rc;
  if (dt < 0) dt = c.sc_host->dt;
  if (sim_dt < 0) sim_dt = c.sc_host->dt;
  const double normFFT = (double)c.N[0] * c.N[1] * c.N[2];
  const double* r = c.sc_host->red;
  const double tke = r[0] / (normFFT * normFFT);
  const double eps = r[1] * 2 * nu / std::pow(normFFT * 2 * c.h, 2);
  const double tkeF = r[2] / (normFFT * normFFT);
  const double lint = r[3] / r[0];
  double inj = 0;
  if (tke_target > 0) inj = (tke_target - tke) / dt;
  else if (inj_rate > 0) inj = inj_rate;
  const double fac = sim_dt * inj / (2 * tkeF);
  if (fac > 0) {
    void* specs[3] = {c.spec, c.spec2, c.spec3};
    for (int q = 0; q < 3; ++q) {
      if (int rc = launch_fft_inverse_field(c, specs[q], c.rbuf)) return rc;
      hipLaunchKernelGGL((k_axpy_blocks<T>), grid_for(c.ncell), dim3(NT), 0, c.stream, (T*)c.f[v0 + q],
                         (const T*)c.rbuf, c.ncell, fac / normFFT);
      if (int rc = check_hip(c, hipGetLastError(), "k_axpy_blocks")) return rc;
    }
  }
  stats[0] = tke; stats[1] = eps; stats[2] = tkeF; stats[3] = lint; stats[4] = fac;
  return CUP_OK;
}

int forcing_plain(Ctx& c, double dt, double sim_dt, double nu, double tke_target, double inj, double stats[5],
                  int field)
{
  return c.precision == 4 ? forcing_plain_t<float>(c, dt, sim_dt, nu, tke_target, inj, stats, field)
                          : forcing_plain_t<double>(c, dt, sim_dt, nu, tke_target, inj, stats, field);
}

// sum over cells of a0 b0 + a1 b1 + a2 b2 into part[blockIdx.x] (fixed grid: deterministic)
template <typename T>
__global__ __launch_bounds__(NT) void k_dot3(const T* __restrict__ a0, const T* __restrict__ a1,
                                             const T* __restrict__ a2, const T* __restrict__ b0,
                                             const T* __restrict__ b1, const T* __restrict__ b2, size_t n,
                                             double* __restrict__ part)
{
  double acc = 0;
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
    acc += (double)a0[i] * b0[i] + (double)a1[i] * b1[i] + (double)a2[i] * b2[i];
  acc = wave_sum(acc);
  __shared__ double red[NT / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int q = 0; q < NT / 64; ++q) s += red[q];
    part[blockIdx.x] = s;
  }
}

// Transpose of the linear forcing at u* (no reference counterpart). The forward
// (SpectralForcing.cpp:37-77) is u** = u* + fac P u* for fac > 0, with P the band filter
// (k_forcing_filter's mask, an orthogonal projector), tke = <|u*|^2>/2, tke_F = <u*, P u*>/2
// and fac = (tke_target - tke)/(2 tke_F), or dt r/(2 tke_F) at the constant injection rate r.
// With s = <P u*, lam> over the n cells and three components:
//   lam += fac P lam + s grad fac,  grad fac = -u*/(2 n tke_F) [target only] - fac P u*/(n tke_F).
// P u* and P lam come from forward transforms, the filter and inverse transforms (n times P,
// unnormalized); the switch fac > 0 is held at the forward state.
template <typename T>
static int adjoint_forcing_t(Ctx& c, int us0, double dt, double tke_target, double inj_rate)
{
  if (int rc = alloc_extra_spectra(c)) return rc;
  const size_t fb = c.ncell * c.esz;
  if (!c.adjf && check_hip(c, hipMalloc(&c.adjf, 9 * fb), "hipMalloc forcing adjoint")) return CUP_E_NOMEM;
  T* X = (T*)c.adjf;                 // n P lam, three fields
  T* Y = (T*)((char*)c.adjf + 3 * fb);  // n P u*
  const size_t nh = (size_t)(c.N[0] / 2) * c.N[1] * c.N[2];
  dim3 g = grid_for(nh), gd = grid_for(c.ncell);
  if (g.x > 4096) g.x = 4096;
  if (gd.x > 4096) gd.x = 4096;
  if (int rc = ensure_partials(c, std::max(4 * (size_t)g.x, (size_t)gd.x))) return rc;
  const double wf[3] = {2 * M_PI / c.extent[0], 2 * M_PI / c.extent[1], 2 * M_PI / c.extent[2]};
  void* specs[3] = {c.spec, c.spec2, c.spec3};
  auto filtered = [&](int f0) {  // spectra of fields f0.. filtered, band sums in c.partials
    for (int q = 0; q < 3; ++q)
      if (int rc = launch_fft_forward_field(c, c.f[f0 + q], specs[q])) return rc;
    hipLaunchKernelGGL((k_forcing_filter<T>), g, dim3(NT), 0, c.stream, (C2<T>*)c.spec, (C2<T>*)c.spec2,
                       (C2<T>*)c.spec3, c.N[0], c.N[1], c.N[2], wf[0], wf[1], wf[2], c.partials);
    return check_hip(c, hipGetLastError(), "k_forcing_filter");
  };
  auto inverse = [&](T* dst) {
    for (int q = 0; q < 3; ++q)
      if (int rc = launch_fft_inverse_field(c, specs[q], dst + q * c.ncell)) return rc;
    return CUP_OK;
  };
  auto scalars = [&]() {
    if (int rc = check_hip(c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream),
                           "forcing adjoint"))
      return rc;
    return check_hip(c, hipStreamSynchronize(c.stream), "forcing adjoint");
  };
  // the forward's statistics at u*
  if (int rc = filtered(us0)) return rc;
  if (int rc = launch_reduce(c, (int)g.x, 4, 0u, c.sc->red)) return rc;
  if (int rc = scalars()) return rc;
  const double n = (double)c.N[0] * c.N[1] * c.N[2];
  const double tke = c.sc_host->red[0] / (n * n), tkeF = c.sc_host->red[2] / (n * n);
  const double inj = tke_target > 0 ? (tke_target - tke) / dt : inj_rate;
  const double fac = dt * inj / (2 * tkeF);
  if (!(fac > 0)) return CUP_OK;
  if (int rc = inverse(Y)) return rc;
  if (int rc = filtered(CUP_ADJ_U)) return rc;
  if (int rc = inverse(X)) return rc;
  // s = <P u*, lam> = <Y, lam>/n
  hipLaunchKernelGGL((k_dot3<T>), gd, dim3(NT), 0, c.stream, Y, Y + c.ncell, Y + 2 * c.ncell,
                     (const T*)c.f[CUP_ADJ_U], (const T*)c.f[CUP_ADJ_V], (const T*)c.f[CUP_ADJ_W], c.ncell,
                     c.partials);
  if (int rc = check_hip(c, hipGetLastError(), "k_dot3")) return rc;
  if (int rc = launch_reduce(c, (int)gd.x, 1, 0u, c.sc->red)) return rc;
  if (int rc = scalars()) return rc;
  const double s = c.sc_host->red[0] / n;
  const double a_x = fac / n;                                            // on n P lam
  const double a_u = tke_target > 0 ? -s / (2 * n * tkeF) : 0.0;          // on u*
  const double a_y = -s * fac / (n * tkeF) / n;                          // on n P u*
  for (int q = 0; q < 3; ++q) {
    T* lam = (T*)c.f[CUP_ADJ_U + q];
    hipLaunchKernelGGL((k_axpy_blocks<T>), grid_for(c.ncell), dim3(NT), 0, c.stream, lam,
                       (const T*)(X + q * c.ncell), c.ncell, a_x);
    if (a_u != 0)
      hipLaunchKernelGGL((k_axpy_blocks<T>), grid_for(c.ncell), dim3(NT), 0, c.stream, lam,
                         (const T*)c.f[us0 + q], c.ncell, a_u);
    hipLaunchKernelGGL((k_axpy_blocks<T>), grid_for(c.ncell), dim3(NT), 0, c.stream, lam,
                       (const T*)(Y + q * c.ncell), c.ncell, a_y);
    if (int rc = check_hip(c, hipGetLastError(), "k_axpy_blocks")) return rc;
  }
  return CUP_OK;
}

int launch_adjoint_forcing(Ctx& c, int us0, double dt, double tke_target, double inj_rate)
{
  return c.precision == 4 ? adjoint_forcing_t<float>(c, us0, dt, tke_target, inj_rate)
                          : adjoint_forcing_t<double>(c, us0, dt, tke_target, inj_rate);
}

// Low modes of the linear forcing, 0 < |k|^2 <= 4 in physical wavenumbers, canonical
// member of each conjugate pair (kz > 0, or kz = 0 and ky > 0, or kz = ky = 0 and kx > 0),
// and their positions in the packed half spectrum (x halved): kx > 0 at (kx, ky, kz);
// kx < 0 through the partner (-kx, -ky, -kz) conjugated; kx = 0 in slot 0 at (ky, kz) and,
// conjugated, at (-ky, -kz).
int spectral_setup(Ctx& c)
{
  std::vector<int> ks;
  const double wf[3] = {2 * M_PI / c.extent[0], 2 * M_PI / c.extent[1], 2 * M_PI / c.extent[2]};
  int lim[3];
  for (int a = 0; a < 3; ++a) {
    lim[a] = (int)std::floor(2.0 / wf[a] + 1e-9);
    if (lim[a] > c.N[a] / 2 - 1) lim[a] = c.N[a] / 2 - 1;
  }
  for (int kz = 0; kz <= lim[2]; ++kz)
    for (int ky = -lim[1]; ky <= lim[1]; ++ky)
      for (int kx = -lim[0]; kx <= lim[0]; ++kx) {
        const bool canonical = kz > 0 || (kz == 0 && (ky > 0 || (ky == 0 && kx > 0)));
        if (!canonical) continue;
        const double k2 = std::pow(kx * wf[0], 2) + std::pow(ky * wf[1], 2) + std::pow(kz * wf[2], 2);
        if (k2 > 0 && k2 <= 4) { ks.push_back(kx); ks.push_back(ky); ks.push_back(kz); }
      }
  c.nlow = (int)ks.size() / 3;
  c.low = LowModes{};
  c.low_fused_ok = c.nlow > 0 && c.nlow <= LOW_MAX;
  for (int m = 0; m < c.nlow && c.low_fused_ok; ++m)
    if (std::abs(ks[3 * m]) > LOW_KX) c.low_fused_ok = false;
  if (!c.low_fused_ok) return CUP_OK;
  LowModes& L = c.low;
  L.n = c.nlow;
  auto wrapn = [](int k, int n) { return ((k % n) + n) % n; };
  for (int m = 0; m < c.nlow; ++m) {
    const int kx = ks[3 * m], ky = ks[3 * m + 1], kz = ks[3 * m + 2];
    L.k[m][0] = kx; L.k[m][1] = ky; L.k[m][2] = kz;
    for (int d = 0; d < 3; ++d) L.sk[m][d] = std::sin(2.0 * M_PI * ks[3 * m + d] / c.N[d]);
    auto add = [&](int slot, int yy, int zz, int conj) {
      int* q = L.pos[L.npos++];
      q[0] = slot; q[1] = wrapn(yy, c.N[1]); q[2] = wrapn(zz, c.N[2]); q[3] = (m << 1) | conj;
    };
    if (kx > 0) add(kx, ky, kz, 0);
    else if (kx < 0) add(-kx, -ky, -kz, 1);
    else { add(0, ky, kz, 0); add(0, -ky, -kz, 1); }
  }
  const int W = 6 * c.nlow + 1;
  const size_t nrow = (size_t)c.N[1] * c.N[2];
  if (check_hip(c, hipMalloc(&c.low_part, nrow * 16 * sizeof(double)), "hipMalloc low row values") ||
      check_hip(c, hipMalloc(&c.low_red, ((size_t)c.N[2] * FR_PARTS_MAX * 16 * 5 + W) * sizeof(double)),
                "hipMalloc low red") ||
      check_hip(c, hipMalloc(&c.low_coef, (size_t)c.nlow * 8 * sizeof(double)), "hipMalloc low coef") ||
      check_hip(c, hipMalloc(&c.low_counter, sizeof(unsigned)), "hipMalloc counter") ||
      check_hip(c, hipMalloc(&c.low_dev, sizeof(LowModes)), "hipMalloc low modes") ||
      check_hip(c, hipMalloc(&c.low_rows, (size_t)c.N[1] * c.N[2] * 16 * c.esz), "hipMalloc low rows"))
    return CUP_E_NOMEM;
  hipMemcpy(c.low_dev, &c.low, sizeof(LowModes), hipMemcpyHostToDevice);
  hipMemsetAsync(c.low_counter, 0, sizeof(unsigned), c.stream);
  hipMemsetAsync(c.low_coef, 0, (size_t)c.nlow * 8 * sizeof(double), c.stream);
  return CUP_OK;
}

// Fused forcing, stage 2 (plane DFTs in y, forcing_planes_wg) as a kernel of its own:
// grid (Nz, parts).
template <typename T>
__global__ __launch_bounds__(NT) void k_forcing_planes(const double* __restrict__ rowbuf, double* __restrict__ plane,
                                                       int Ny, const C2<T>* __restrict__ twy)
{
  __shared__ double acc[FORCING_PLANES_LDS / 8];
  forcing_planes_wg<T>(rowbuf, plane, Ny, twy, blockIdx.x, blockIdx.y, gridDim.y, acc);
}

// Fused forcing, stage 3 (the z sums): one workgroup per column owner, (mode m, component q)
// for o < 3n and the sum |u*|^2 for o = 3n. For mode m = (kx, ky, kz) and component q, the
// plane value is U(z) = Pr + i sg Pi, with Pr, Pi the plane DFTs (at ky, conjugated for
// ky < 0) of the slots holding re and im of the x-DFT at |kx| (sg = -1 for kx < 0, the
// conjugate; Pi = 0 at kx = 0); the owner sums U(z) e^{-i kz z} over z, threads over planes
// and a fixed shuffle tree (deterministic), into red[6m + 2q + {re, im}] (red[W-1] for
// |u*|^2). The last workgroup to finish computes, on the device, SpectralForcing's injection
// and factor (SpectralForcing.cpp:37-77) from tke = <|u*|^2>/2 (Parseval) and tke_low = sum
// over the low modes, and the coefficient tables: the forcing's pressure-RHS modes
// fac*(h^2/dt)*i*sum_d sin(2 pi k_d/N_d) u_d(k) (added in the z pass) and the velocity modes
// fac/N^3 u(k) (added in the gradient pass).
template <typename T>
__global__ __launch_bounds__(NT) void k_forcing_reduce(const double* __restrict__ plane, int W,
                                                       double* __restrict__ red, unsigned* __restrict__ counter,
                                                       const LowModes* __restrict__ lmp, DevScalars* __restrict__ sc,
                                                       double* __restrict__ coef, double tke_target, double inj_rate,
                                                       double h, double ncell, int N0, int N1, int N2,
                                                       const C2<T>* __restrict__ twz, int parts)
{
  __shared__ double wsum[NT / 64][2];
  __shared__ double fin[6 * LOW_MAX + 1];
  __shared__ double facs[1];
  __shared__ bool last;
  const LowModes& lm = *lmp;
  const int t = threadIdx.x, o = blockIdx.x, nq = 3 * lm.n;
  double ar = 0, ai = 0;
  // plane partials zp = z*parts + part
  const int nzp = N2 * parts;
  if (o == nq) {
    for (int zp = t; zp < nzp; zp += NT) ar += plane[((size_t)zp * 16 + 15) * 5];
  } else {
    const int m = o / 3, q = o % 3;
    const int kx = lm.k[m][0], ky = lm.k[m][1], kz = lm.k[m][2];
    const int ak = kx < 0 ? -kx : kx, aky = ky < 0 ? -ky : ky;
    const int sre = ak == 0 ? 5 * q : 5 * q + 2 * ak - 1;
    const double sg = kx < 0 ? -1.0 : 1.0, cy = ky < 0 ? -1.0 : 1.0;
    for (int zp = t; zp < nzp; zp += NT) {
      const int z = zp / parts;
      const double* P = plane + (size_t)zp * 16 * 5;
      // plane DFT of slot s at ky: (P[s][0], 0) for ky = 0, else (P[s][2|ky|-1], cy P[s][2|ky|])
      const double rr = aky == 0 ? P[sre * 5] : P[sre * 5 + 2 * aky - 1];
      const double ri = aky == 0 ? 0.0 : cy * P[sre * 5 + 2 * aky];
      double ir = 0, ii = 0;
      if (ak != 0) {
        ir = aky == 0 ? P[(sre + 1) * 5] : P[(sre + 1) * 5 + 2 * aky - 1];
        ii = aky == 0 ? 0.0 : cy * P[(sre + 1) * 5 + 2 * aky];
      }
      // U = (rr + i ri) + i sg (ir + i ii)
      const double Ur = rr - sg * ii, Ui = ri + sg * ir;
      const C2<T> e = twz[((kz * z) % N2 + N2) % N2];
      const double er = (double)e.x, ei = (double)e.y;
      ar += Ur * er - Ui * ei;
      ai += Ur * ei + Ui * er;
    }
  }
  ar = wave_sum(ar);
  ai = wave_sum(ai);
  if ((t & 63) == 0) {
    wsum[t >> 6][0] = ar;
    wsum[t >> 6][1] = ai;
  }
  __syncthreads();
  if (t == 0) {
    double sr = 0, si = 0;
    for (int w = 0; w < NT / 64; ++w) {
      sr += wsum[w][0];
      si += wsum[w][1];
    }
    if (o == nq) {
      red[W - 1] = sr;
    } else {
      red[2 * o] = sr;  // column 6m + 2q = 2o
      red[2 * o + 1] = si;
    }
    __threadfence();
    last = atomicAdd(counter, 1u) == (unsigned)(gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // acquire: the other workgroups' sums are visible to plain loads
  if (t < W) fin[t] = red[t];
  __syncthreads();
  const double* r = fin;
  const double dt = sc->dt;
  const double norm = (double)N0 * N1 * N2;
  if (t < 64) {
    // sum over the low modes' 3n columns of |c|^2, lanes over columns and a fixed shuffle tree
    double e = 0;
    for (int o2 = t; o2 < nq; o2 += 64) e += r[2 * o2] * r[2 * o2] + r[2 * o2 + 1] * r[2 * o2 + 1];
    e = wave_sum(e);
    if (t == 0) {
      *counter = 0u;
      const double tke = 0.5 * r[W - 1] / ncell;
      const double tke_low = e / (norm * norm);
      double inj = 0;
      if (tke_target > 0) inj = (tke_target - tke) / dt;
      else if (inj_rate > 0) inj = inj_rate;
      const double fac = dt * inj / (2 * tke_low);
      sc->tke = tke;
      sc->tke_low = tke_low;
      sc->forcing_fac = fac;
      facs[0] = fac;
    }
  }
  __syncthreads();
  const double fac = facs[0];
  if (t >= lm.n) return;
  const int m = t;  // one thread per mode
  double* cm = coef + 8 * m;
  if (!(fac > 0)) {
    for (int q = 0; q < 8; ++q) cm[q] = 0;
    return;
  }
  const double s0 = lm.sk[m][0], s1 = lm.sk[m][1], s2 = lm.sk[m][2];
  const double dr = s0 * r[6 * m] + s1 * r[6 * m + 2] + s2 * r[6 * m + 4];
  const double di = s0 * r[6 * m + 1] + s1 * r[6 * m + 3] + s2 * r[6 * m + 5];
  const double gg = fac * h * h / dt;
  cm[0] = -gg * di;  // i * (dr + i di) = -di + i dr
  cm[1] = gg * dr;
  for (int q = 0; q < 3; ++q) {
    cm[2 + 2 * q] = fac / norm * r[6 * m + 2 * q];
    cm[3 + 2 * q] = fac / norm * r[6 * m + 2 * q + 1];
  }
}

ColExtra forcing_planes_extra(Ctx& c)
{
  ColExtra e{};
  e.rowbuf = c.low_part;
  e.plane = c.low_red;
  e.twy = c.tw[1];
  e.twz = c.tw[2];
  e.rows = c.low_rows;
  e.parts = forcing_plane_parts(c.N[1]);
  e.nz = c.N[2];
  return e;
}

int launch_forcing_reduce(Ctx& c, double tke_target, double inj_rate, bool planes)
{
  const int W = 6 * c.nlow + 1;
  const int nz = c.N[2];
  const int parts = forcing_plane_parts(c.N[1]);
  double* plane = c.low_red;  // [Nz][parts][16][5] plane DFTs, then the W column sums
  double* red = plane + (size_t)nz * FR_PARTS_MAX * 16 * 5;
  if (c.precision == 4) {
    if (planes)
      hipLaunchKernelGGL((k_forcing_planes<float>), dim3(nz, parts), dim3(NT), 0, c.stream,
                         (const double*)c.low_part, plane, c.N[1], (const C2<float>*)c.tw[1]);
    hipLaunchKernelGGL((k_forcing_reduce<float>), dim3(3 * c.nlow + 1), dim3(NT), 0, c.stream, (const double*)plane, W,
                       red, c.low_counter, c.low_dev, c.sc, c.low_coef, tke_target, inj_rate, c.h, (double)c.ncell,
                       c.N[0], c.N[1], c.N[2], (const C2<float>*)c.tw[2], parts);
  } else {
    if (planes)
      hipLaunchKernelGGL((k_forcing_planes<double>), dim3(nz, parts), dim3(NT), 0, c.stream,
                         (const double*)c.low_part, plane, c.N[1], (const C2<double>*)c.tw[1]);
    hipLaunchKernelGGL((k_forcing_reduce<double>), dim3(3 * c.nlow + 1), dim3(NT), 0, c.stream, (const double*)plane,
                       W, red, c.low_counter, c.low_dev, c.sc, c.low_coef, tke_target, inj_rate, c.h,
                       (double)c.ncell, c.N[0], c.N[1], c.N[2], (const C2<double>*)c.tw[2], parts);
  }
  return check_hip(c, hipGetLastError(), "k_forcing_reduce");
}

// Fused forcing, stage 4 (the per-row synthesis coefficients, forcing_rows_wg) as a kernel of
// its own: one workgroup per z plane.
template <typename T>
__global__ __launch_bounds__(NT) void k_forcing_rows(const LowModes* __restrict__ lmp,
                                                     const double* __restrict__ coef, const C2<T>* __restrict__ twy,
                                                     const C2<T>* __restrict__ twz, T* __restrict__ rows, int Ny,
                                                     int Nz)
{
  __shared__ double lds[(FORCING_ROWS_LDS + 7) / 8];
  forcing_rows_wg<T>(lmp, coef, twy, twz, rows, Ny, Nz, blockIdx.x, lds);
}

int launch_forcing_rows(Ctx& c)
{
  const dim3 grid(c.N[2]);
  if (c.precision == 4)
    hipLaunchKernelGGL((k_forcing_rows<float>), grid, dim3(NT), 0, c.stream, (const LowModes*)c.low_dev,
                       (const double*)c.low_coef, (const C2<float>*)c.tw[1], (const C2<float>*)c.tw[2],
                       (float*)c.low_rows, c.N[1], c.N[2]);
  else
    hipLaunchKernelGGL((k_forcing_rows<double>), grid, dim3(NT), 0, c.stream, (const LowModes*)c.low_dev,
                       (const double*)c.low_coef, (const C2<double>*)c.tw[1], (const C2<double>*)c.tw[2],
                       (double*)c.low_rows, c.N[1], c.N[2]);
  return check_hip(c, hipGetLastError(), "k_forcing_rows");
}

// ------------------------------------------------------------ HIT analysis
// SpectralManipFFTW::_compute_analysis (SpectralManipFFTW.cpp:110-207) on the packed half
// spectra of u, v, w (x halved): one thread per (ky, kz) line walks kx = 0 .. Nx/2. Slots
// kx = 1 .. Nx/2-1 stand for kx and -kx (weight 2); slot 0 packs the kx = 0 and kx = Nx/2
// planes as A + iB, unpacked with the mirrored entry Z(-ky, -kz): A = (Z + conj Zm)/2,
// B = -i (Z - conj Zm)/2 (weight 1 each). Per mode, with E = w (|U|^2+|V|^2+|W|^2)/2:
// tke += E, eps += w |Omega|^2/2 (Omega from the central-difference symbols 2 sin(2 pi k/N)),
// lint += E/|k|, and E into shell floor(sqrt(kx^2+ky^2+kz^2) (nyq-1)/nyq) when
// kx^2+ky^2+kz^2 < nyq^2 (nyq = maxN/2). Along a line the shells are non-decreasing, so
// the thread writes run sums: ana[(3 + shell)*nlines + line], and ana[q*nlines + line] for
// the three totals; the rows are then summed line by line in order (deterministic).
template <typename T>
__global__ __launch_bounds__(NT) void k_hit_analysis_lines(const C2<T>* __restrict__ su, const C2<T>* __restrict__ sv,
                                                          const C2<T>* __restrict__ sw, double* __restrict__ ana,
                                                          int N0, int N1, int N2, double wfx, double wfy,
                                                          double wfz, int nyq)
{
  const int nlines = N1 * N2;
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  const int ly = line % N1, lz = line / N1;
  const int jj = ly <= N1 / 2 ? ly : ly - N1, kk = lz <= N2 / 2 ? lz : lz - N2;
  const int ncols = N0 / 2;
  const size_t row = ((size_t)lz * N1 + ly) * ncols;
  const size_t mrow = ((size_t)((N2 - lz) % N2) * N1 + (N1 - ly) % N1) * ncols;
  const double dY = 2 * sin(2 * M_PI * jj / N1), dZ = 2 * sin(2 * M_PI * kk / N2);
  const double nyq_scale = (nyq - 1) / (double)nyq;
  const long nyq2 = (long)nyq * nyq;
  double tke = 0, eps = 0, lint = 0, run = 0;
  int cur = -1;
  // mode m = 0 .. ncols: ii = m (m = ncols is kx = Nx/2), weight 1 at m = 0 and m = ncols
  for (int m = 0; m <= ncols; ++m) {
    double U[3][2];
    const C2<T>* S[3] = {su, sv, sw};
    double wgt;
    if (m == 0 || m == ncols) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const C2<T> Z = S[q][row], Zm = S[q][mrow];
        if (m == 0) {  // A = (Z + conj Zm)/2
          U[q][0] = 0.5 * ((double)Z.x + (double)Zm.x);
          U[q][1] = 0.5 * ((double)Z.y - (double)Zm.y);
        } else {  // B = -i (Z - conj Zm)/2
          U[q][0] = 0.5 * ((double)Z.y + (double)Zm.y);
          U[q][1] = -0.5 * ((double)Z.x - (double)Zm.x);
        }
      }
      wgt = 1;
    } else {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const C2<T> Z = S[q][row + m];
        U[q][0] = Z.x;
        U[q][1] = Z.y;
      }
      wgt = 2;
    }
    const int ii = m;
    const double kx = ii * wfx, ky = jj * wfy, kz = kk * wfz;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double dX = 2 * sin(2 * M_PI * ii / N0);
    // Omega = i (dY W - dZ V, dZ U - dX W, dX V - dY U); |i z|^2 = |z|^2
    const double ox0 = dY * U[2][0] - dZ * U[1][0], ox1 = dY * U[2][1] - dZ * U[1][1];
    const double oy0 = dZ * U[0][0] - dX * U[2][0], oy1 = dZ * U[0][1] - dX * U[2][1];
    const double oz0 = dX * U[1][0] - dY * U[0][0], oz1 = dX * U[1][1] - dY * U[0][1];
    const double e = wgt * 0.5 *
                     (U[0][0] * U[0][0] + U[0][1] * U[0][1] + U[1][0] * U[1][0] + U[1][1] * U[1][1] +
                      U[2][0] * U[2][0] + U[2][1] * U[2][1]);
    tke += e;
    eps += wgt * 0.5 * (ox0 * ox0 + ox1 * ox1 + oy0 * oy0 + oy1 * oy1 + oz0 * oz0 + oz1 * oz1);
    lint += k2 > 0 ? e / sqrt(k2) : 0.0;
    const long kind = (long)ii * ii + (long)jj * jj + (long)kk * kk;
    if (kind < nyq2) {
      const int b = (int)floor(sqrt((double)kind) * nyq_scale);
      if (b != cur) {
        if (cur >= 0) ana[(size_t)(3 + cur) * nlines + line] = run;
        cur = b;
        run = 0;
      }
      run += e;
    }
  }
  if (cur >= 0) ana[(size_t)(3 + cur) * nlines + line] = run;
  ana[line] = tke;
  ana[(size_t)nlines + line] = eps;
  ana[2 * (size_t)nlines + line] = lint;
}

// Row r of a [nrows][len] array summed in a fixed order (thread strides, shuffle tree,
// waves in order) into out[r].
__global__ __launch_bounds__(NT) void k_rows_sum(const double* __restrict__ a, int len, double* __restrict__ out)
{
  __shared__ double ws[NT / 64];
  const int r = blockIdx.x, t = threadIdx.x;
  double acc = 0;
  for (int i = t; i < len; i += NT) acc += a[(size_t)r * len + i];
  acc = wave_sum(acc);
  if ((t & 63) == 0) ws[t >> 6] = acc;
  __syncthreads();
  if (t == 0) {
    double v = 0;
    for (int w = 0; w < NT / 64; ++w) v += ws[w];
    out[r] = v;
  }
}

template <typename T>
static int hit_analysis_t(Ctx& c, double* sums, double* E, int nbin)
{
  if (int rc = alloc_extra_spectra(c)) return rc;
  const int nlines = c.N[1] * c.N[2];
  const size_t nrows = 3 + (size_t)nbin;
  if (!c.ana && check_hip(c, hipMalloc(&c.ana, (nrows * nlines + nrows) * sizeof(double)), "hipMalloc analysis"))
    return CUP_E_NOMEM;
  if (int rc = launch_fft_forward_field(c, c.f[CUP_U], c.spec)) return rc;
  if (int rc = launch_fft_forward_field(c, c.f[CUP_V], c.spec2)) return rc;
  if (int rc = launch_fft_forward_field(c, c.f[CUP_W], c.spec3)) return rc;
  if (int rc = check_hip(c, hipMemsetAsync(c.ana, 0, nrows * nlines * sizeof(double), c.stream), "memset analysis"))
    return rc;
  int maxN = std::max(c.N[0], std::max(c.N[1], c.N[2]));
  hipLaunchKernelGGL((k_hit_analysis_lines<T>), dim3((nlines + NT - 1) / NT), dim3(NT), 0, c.stream,
                     (const C2<T>*)c.spec, (const C2<T>*)c.spec2, (const C2<T>*)c.spec3, c.ana, c.N[0], c.N[1],
                     c.N[2], 2 * M_PI / c.extent[0], 2 * M_PI / c.extent[1], 2 * M_PI / c.extent[2], maxN / 2);
  double* tot = c.ana + nrows * nlines;
  hipLaunchKernelGGL(k_rows_sum, dim3((unsigned)nrows), dim3(NT), 0, c.stream, (const double*)c.ana, nlines, tot);
  if (int rc = check_hip(c, hipGetLastError(), "k_hit_analysis")) return rc;
  std::vector<double> h(nrows);
  if (int rc = check_hip(c, hipMemcpyAsync(h.data(), tot, nrows * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                         "hipMemcpy analysis"))
    return rc;
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "synchronize")) return rc;
  for (int q = 0; q < 3; ++q) sums[q] = h[q];
  for (int b = 0; b < nbin; ++b) E[b] = h[3 + b];
  return CUP_OK;
}

int launch_hit_analysis(Ctx& c, double* tke_eps_lint, double* E, int nbin, double grad[2])
{
  int rc = c.precision == 4 ? hit_analysis_t<float>(c, tke_eps_lint, E, nbin)
                            : hit_analysis_t<double>(c, tke_eps_lint, E, nbin);
  if (rc) return rc;
  return launch_grad_stats(c, grad);
}

}  // namespace cup
