// cup_fft.hip -- hand-written 3-D real-to-complex FFT for the periodic Poisson solve.
//
// Spectral layout ("packed half spectrum"): complex S[z][y][kx], kx in [0, Nx/2). The
// x transform is real-to-complex; its kx = 0 and kx = Nx/2 outputs are both real and are
// packed into slot 0 as (X[0], X[Nx/2]). The y and z transforms then act on Nx/2 complex
// columns. Slot 0 carries A + iB with A, B the kx = 0 and kx = Nx/2 planes; because the
// grid-consistent Poisson symbol is real, even, and equal at kx = 0 and kx = Nx/2
// (cos(4 pi k/N) = 1 at both), dividing the packed column by the kx = 0 symbol solves both
// planes at once.
//
// Passes (each one HBM read + write of S; x FFTs run in LDS, radix 8/4/2 Stockham; y and z
// FFTs of 64 to 256 points in registers with one LDS transpose, others as the x FFTs):
//   x-forward : 8 rows (one block row of a z plane) per workgroup, two real rows packed
//               into one complex FFT; the source is either a real block-layout field or,
//               fused, the central divergence of the advected velocity (PressureRHS).
//   y         : TK consecutive kx columns x all y of one z plane per workgroup.
//   z         : TK columns x all z of one y row; fused: forward, spectral division, inverse.
//   x-inverse : 8 rows per workgroup, complex-to-real, written into a block-layout field.
// Unnormalized in both directions like FFTW's r2c/c2r; the 1/N^3 and h factors live in the
// division (PoissonSolverPeriodic.cpp:23-35).
#include "cup_internal.h"
#include "cup_forcing_dev.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace cup {

template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
  return C2<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return C2<T>{a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return C2<T>{a.x - b.x, a.y - b.y}; }
// a * e^{DIR i pi/2}: DIR=-1 -> -i a, DIR=+1 -> +i a
template <int DIR, typename T>
__device__ __forceinline__ C2<T> rot90(C2<T> a)
{
  return DIR < 0 ? C2<T>{a.y, -a.x} : C2<T>{-a.y, a.x};
}

template <int DIR, typename T>
__device__ __forceinline__ void dft2(C2<T>* v)
{
  const C2<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(C2<T>& v0, C2<T>& v1, C2<T>& v2, C2<T>& v3)
{
  const C2<T> t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = rot90<DIR>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft8(C2<T>* v)
{
  C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  C2<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  const T c = (T)0.70710678118654752440;
  // W^k, W = e^{DIR 2 pi i / 8}
  const C2<T> w1o = cmul(o1, C2<T>{c, (T)DIR * c});
  const C2<T> w2o = rot90<DIR>(o2);
  const C2<T> w3o = cmul(o3, C2<T>{-c, (T)DIR * c});
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1o);
  v[5] = csub(e1, w1o);
  v[2] = cadd(e2, w2o);
  v[6] = csub(e2, w2o);
  v[3] = cadd(e3, w3o);
  v[7] = csub(e3, w3o);
}

// 16-point DFT in registers, natural order in and out: radix 4 x 4 with n = q + 4p,
// k = r + 4s, X[r + 4s] = sum_q W16^{qr} W4^{qs} sum_p v[q + 4p] W4^{pr}, W16 = e^{DIR 2 pi i/16}.
template <int DIR, typename T>
__device__ __forceinline__ void dft16(C2<T>* v)
{
  C2<T> b[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    b[q][0] = v[q];
    b[q][1] = v[q + 4];
    b[q][2] = v[q + 8];
    b[q][3] = v[q + 12];
    dft4<DIR>(b[q][0], b[q][1], b[q][2], b[q][3]);
  }
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
  const T d = (T)DIR;
  // W16^m for m = q r: 1, 2, 3, 4 (a quarter turn), 6, 9
  b[1][1] = cmul(b[1][1], C2<T>{c1, d * s1});
  b[1][2] = cmul(b[1][2], C2<T>{h, d * h});
  b[1][3] = cmul(b[1][3], C2<T>{s1, d * c1});
  b[2][1] = cmul(b[2][1], C2<T>{h, d * h});
  b[2][2] = rot90<DIR>(b[2][2]);
  b[2][3] = cmul(b[2][3], C2<T>{-h, d * h});
  b[3][1] = cmul(b[3][1], C2<T>{s1, d * c1});
  b[3][2] = cmul(b[3][2], C2<T>{-h, d * h});
  b[3][3] = cmul(b[3][3], C2<T>{-c1, -d * s1});
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    dft4<DIR>(b[0][r], b[1][r], b[2][r], b[3][r]);
    v[r] = b[0][r];
    v[r + 4] = b[1][r];
    v[r + 8] = b[2][r];
    v[r + 12] = b[3][r];
  }
}

// 3- and 5-point DFTs, W = e^{DIR 2 pi i / R}: X_q = sum_r v_r W^{q r}. The conjugate pairs
// W^r + W^{-r} and W^r - W^{-r} are formed once (for R = 5 with c_m = cos(2 pi m/5),
// s_m = DIR sin(2 pi m/5)).
template <int DIR, typename T>
__device__ __forceinline__ void dft3(C2<T>* v)
{
  const T s = (T)DIR * (T)0.86602540378443864676;  // sin(2 pi/3)
  const C2<T> a = v[0], t1 = cadd(v[1], v[2]), t3 = csub(v[1], v[2]);
  const C2<T> m = C2<T>{a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y};
  const C2<T> n = C2<T>{-s * t3.y, s * t3.x};  // i s (v1 - v2)
  v[0] = cadd(a, t1);
  v[1] = cadd(m, n);
  v[2] = csub(m, n);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft5(C2<T>* v)
{
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)DIR * (T)0.95105651629515357212, s2 = (T)DIR * (T)0.58778525229247312917;
  const C2<T> a = v[0];
  const C2<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const C2<T> m1 = C2<T>{a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
  const C2<T> m2 = C2<T>{a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
  // i (s1 t3 + s2 t4) and i (s2 t3 - s1 t4)
  const C2<T> n1 = C2<T>{-(s1 * t3.y + s2 * t4.y), s1 * t3.x + s2 * t4.x};
  const C2<T> n2 = C2<T>{-(s2 * t3.y - s1 * t4.y), s2 * t3.x - s1 * t4.x};
  v[0] = cadd(a, cadd(t1, t2));
  v[1] = cadd(m1, n1);
  v[4] = csub(m1, n1);
  v[2] = cadd(m2, n2);
  v[3] = csub(m2, n2);
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dftR(C2<T>* v)
{
  if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5) dft5<DIR>(v);
  else if constexpr (R == 3) dft3<DIR>(v);
  else dft2<DIR>(v);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// One Stockham stage (radix R, Ns = product of the earlier radices) over NFFT transforms
// of length N held in LDS at buf[f + n*NFFT] (columns interleaved: the thread map
// f-fastest is conflict free). tw[t] = e^{-2 pi i t/N}.
template <int N, int NFFT, int R, int Ns, int DIR, typename T>
__device__ __forceinline__ void fft_stage(C2<T>* buf, const C2<T>* __restrict__ tw)
{
  constexpr int NB = NFFT * (N / R);
  constexpr int IT = (NB + NT - 1) / NT;
  C2<T> v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int t = threadIdx.x + it * NT;
    if (NB % NT == 0 || t < NB) {
      const int f = t % NFFT, j = t / NFFT, k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = buf[f + (j + r * (N / R)) * NFFT];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C2<T> w = tw[r * k * (N / (Ns * R))];
          if (DIR > 0) w.y = -w.y;
          v[it][r] = cmul(v[it][r], w);
        }
      }
      dftR<R, DIR>(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int t = threadIdx.x + it * NT;
    if (NB % NT == 0 || t < NB) {
      const int f = t % NFFT, j = t / NFFT, k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[f + (base + r * Ns) * NFFT] = v[it][r];
    }
  }
  __syncthreads();
}

template <int N, int NFFT, int Ns, int DIR, typename T>
__device__ __forceinline__ void fft_stages(C2<T>* buf, const C2<T>* __restrict__ tw)
{
  if constexpr (Ns < N) {
    constexpr int rem = N / Ns;
    constexpr int R = rem >= 8 ? 8 : rem;
    fft_stage<N, NFFT, R, Ns, DIR>(buf, tw);
    fft_stages<N, NFFT, Ns * R, DIR>(buf, tw);
  }
}

// NFFT complex transforms of length N in LDS, in place; caller synchronizes before.
template <int N, int NFFT, int DIR, typename T>
__device__ __forceinline__ void lds_fft(C2<T>* buf, const C2<T>* __restrict__ tw)
{
  fft_stages<N, NFFT, 1, DIR>(buf, tw);
}

// The same Stockham stages for a length known at run time (the sizes that are not powers of
// two: N = 8 x blocks with factors 2, 3 and 5, up to FFT_MAX_N), radices from the plan.
// A stage's butterflies are held in registers across the barrier, as in fft_stage: at most
// ITM = ceil(NFFT FFT_MAX_N / (R NT)) per thread, predicated on the actual count.
template <int R, int NFFT, int DIR, typename T>
__device__ __forceinline__ void fft_stage_rt(C2<T>* buf, const C2<T>* __restrict__ tw, int n, int Ns)
{
  constexpr int ITM = (NFFT * FFT_MAX_N / R + NT - 1) / NT;
  const int NB = NFFT * (n / R), m = n / R, st = n / (Ns * R);
  C2<T> v[ITM][R];
#pragma unroll
  for (int it = 0; it < ITM; ++it) {
    const int t = threadIdx.x + it * NT;
    if (t < NB) {
      const int f = t % NFFT, j = t / NFFT, k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = buf[f + (j + r * m) * NFFT];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C2<T> w = tw[r * k * st];
        if (DIR > 0) w.y = -w.y;
        v[it][r] = cmul(v[it][r], w);
      }
      dftR<R, DIR>(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITM; ++it) {
    const int t = threadIdx.x + it * NT;
    if (t < NB) {
      const int f = t % NFFT, j = t / NFFT, k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[f + (base + r * Ns) * NFFT] = v[it][r];
    }
  }
  __syncthreads();
}

template <int NFFT, int DIR, typename T>
__device__ __forceinline__ void lds_fft_rt(C2<T>* buf, const C2<T>* __restrict__ tw, const FftPlan& p)
{
  int Ns = 1;
  for (int s = 0; s < p.ns; ++s) {
    const int R = p.r[s];
    if (R == 8) fft_stage_rt<8, NFFT, DIR>(buf, tw, p.n, Ns);
    else if (R == 4) fft_stage_rt<4, NFFT, DIR>(buf, tw, p.n, Ns);
    else if (R == 2) fft_stage_rt<2, NFFT, DIR>(buf, tw, p.n, Ns);
    else if (R == 3) fft_stage_rt<3, NFFT, DIR>(buf, tw, p.n, Ns);
    else fft_stage_rt<5, NFFT, DIR>(buf, tw, p.n, Ns);
    Ns *= R;
  }
}

// The LDS tile of a pass: a static array of COUNT elements, or for COUNT == 0 (a length known
// at run time) the dynamic shared memory the launch sizes.
template <typename T, int COUNT>
__device__ __forceinline__ C2<T>* lds_tile()
{
  if constexpr (COUNT > 0) {
    __shared__ __align__(32) C2<T> tile[COUNT > 0 ? COUNT : 1];
    return tile;
  } else {
    extern __shared__ __align__(32) unsigned char dyn_lds[];
    return reinterpret_cast<C2<T>*>(dyn_lds);
  }
}

// --------------------------------------------------------------------- x passes

constexpr int XROWS = 8;         // rows per workgroup = one block row of a z plane
constexpr int XFFT = XROWS / 2;  // complex FFTs per workgroup (two real rows each)

enum XSrc { XSRC_FIELD = 0, XSRC_DIV = 1 };

// Real rows -> packed half spectrum. grid = (Ny/8, Nz).
// FORCE (fused linear forcing): the same pass also forms, per grid row, the x-DFT at
// kx = 0, 1, 2 of each advected velocity component and sum |u*|^2: 16 values per row
// (component q's kx = 0, then re/im of kx = 1, 2 at 5q..5q+4; |u*|^2 at 15), written to
// rowbuf[(z*Ny + y)*16 + slot]. The divergence already loads v on row y+1 and w on plane
// z+1, so the workgroup of row (y, z) takes u of (y, z), v of (y+1, z) and w of (y, z+1);
// every row of every component is summed by exactly one workgroup.
// Each thread's cells lie in one row (t -> bx = t/64, iy, ix), so row sums are lane
// shuffles over ix plus an LDS sum over the 4 waves.
template <typename T, int N, int SRC, bool FORCE>
__global__ __launch_bounds__(NT) void k_fft_x_fwd(const T* __restrict__ src0, const T* __restrict__ src1,
                                                  const T* __restrict__ src2, C2<T>* __restrict__ S,
                                                  const C2<T>* __restrict__ tw, GridArgs g,
                                                  const DevScalars* __restrict__ sc, const int* __restrict__ ob_slot,
                                                  const T* __restrict__ ob_rhs, double* __restrict__ rowbuf,
                                                  FftPlan pl)
{
  const int n = N > 0 ? N : pl.n;
  C2<T>* buf = lds_tile<T, N * XFFT>();
  constexpr int NA = 1 + 2 * LOW_KX;  // row DFT values per component: kx=0 (real), kx=1..: re, im
  static_assert(LOW_KX == 2, "the row DFT below evaluates kx = 1, 2");
  static_assert(3 * NA + 1 == 16, "16 row values");
  __shared__ double rowacc[FORCE ? NT / 64 : 1][FORCE ? XROWS : 1][FORCE ? 3 * NA + 1 : 1];
  const int by = blockIdx.x, z = blockIdx.y;
  const int bz = z >> 3, iz = z & 7;
  T fac = 0;
  if (SRC == XSRC_DIV) fac = (T)(0.5 * g.h * g.h / sc->dt);
  T* rb = (T*)buf;
  T acc[FORCE ? 3 * NA + 1 : 1];
  if (FORCE)
#pragma unroll
    for (int q = 0; q < 3 * NA + 1; ++q) acc[q] = 0;
  // Each thread takes quads of 4 x-adjacent cells (half a block row): q -> bx = q/16,
  // iy = (q/2)%8, x0 = 8 bx + 4 (q%2), with 16-byte loads of u, of v at y +- 1, of w at
  // z +- 1, and single loads of u across the quad's two x ends.
  using V4 = typename Vec4<T>::type;
  for (int q4 = threadIdx.x; q4 < n * XROWS / 4; q4 += NT) {
    const int bx = q4 >> 4, iy = (q4 >> 1) & 7, half = q4 & 1, ix0 = 4 * half;
    const int x0 = bx * BS + ix0;
    const int b = bx + g.nbx * (by + g.nby * bz);
    const size_t o = (size_t)b * BS3 + (iz * BS + iy) * BS + ix0;
    T val[4];
    if (SRC == XSRC_FIELD) {
      const V4 a = *reinterpret_cast<const V4*>(src0 + o);
      val[0] = a.x; val[1] = a.y; val[2] = a.z; val[3] = a.w;
    } else {
      const int slot = ob_slot ? ob_slot[b] : -1;
      // neighbour offsets inside the block, or into the periodic neighbour block
      const long sx = BS3, sy = (long)g.nbx * BS3, sz = (long)g.nbx * g.nby * BS3;
      const long on = iy < 7 ? BS : (by + 1 < g.nby ? sy : sy - (long)g.nby * sy) - 7 * BS;
      const long os = iy > 0 ? -BS : (by > 0 ? -sy : (long)(g.nby - 1) * sy) + 7 * BS;
      const long ob = iz < 7 ? BS * BS : (bz + 1 < g.nbz ? sz : sz - (long)g.nbz * sz) - 7 * BS * BS;
      const long of = iz > 0 ? -BS * BS : (bz > 0 ? -sz : (long)(g.nbz - 1) * sz) + 7 * BS * BS;
      // u west of x0 and east of x0 + 3
      const long ow = half ? -1 : (bx > 0 ? -sx : (long)(g.nbx - 1) * sx) + 7;
      const long oe = half ? 3 + (bx + 1 < g.nbx ? sx : sx - (long)g.nbx * sx) - 7 : 4;
      const V4 uc = *reinterpret_cast<const V4*>(src0 + o);
      const T uu[6] = {src0[o + ow], uc.x, uc.y, uc.z, uc.w, src0[o + oe]};
      const V4 vn = *reinterpret_cast<const V4*>(src1 + o + on), wb = *reinterpret_cast<const V4*>(src2 + o + ob);
      const T vN[4] = {vn.x, vn.y, vn.z, vn.w}, wB[4] = {wb.x, wb.y, wb.z, wb.w};
      if (slot >= 0) {
        const V4 r = *reinterpret_cast<const V4*>(ob_rhs + (size_t)slot * BS3 + (iz * BS + iy) * BS + ix0);
        val[0] = r.x; val[1] = r.y; val[2] = r.z; val[3] = r.w;
      } else {
        const V4 vs = *reinterpret_cast<const V4*>(src1 + o + os), wf = *reinterpret_cast<const V4*>(src2 + o + of);
        const T vS[4] = {vs.x, vs.y, vs.z, vs.w}, wF[4] = {wf.x, wf.y, wf.z, wf.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) val[k] = fac * (uu[k + 2] - uu[k] + vN[k] - vS[k] + wB[k] - wF[k]);
      }
      if (FORCE) {
        // u of this row, v of row y+1, w of plane z+1 (see above)
        const T* cv = vN;
        const T* cw = wB;
        // e^{-2 pi i x/Nx} at x0, rotated by the table's first entries for x0 + 1..3
        const C2<T> e0 = tw[x0];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          C2<T> e1 = e0;
          if (k > 0) {
            const C2<T> r = tw[k];
            e1 = C2<T>{e0.x * r.x - e0.y * r.y, e0.x * r.y + e0.y * r.x};
          }
          const C2<T> e2 = {e1.x * e1.x - e1.y * e1.y, 2 * e1.x * e1.y};  // e^{-4 pi i x/Nx}
          const T cu[3] = {uu[k + 1], cv[k], cw[k]};
          acc[3 * NA] += cu[0] * cu[0] + cu[1] * cu[1] + cu[2] * cu[2];
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            acc[q * NA] += cu[q];
            acc[q * NA + 1] += cu[q] * e1.x;
            acc[q * NA + 2] += cu[q] * e1.y;
            acc[q * NA + 3] += cu[q] * e2.x;
            acc[q * NA + 4] += cu[q] * e2.y;
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) rb[2 * ((x0 + k) * XFFT + (iy >> 1)) + (iy & 1)] = val[k];
  }
  if (FORCE) {
    // rows: lanes with the same iy = (lane >> 1) & 7 differ in bits 0 (half) and 4-5 (bx)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, iy = (lane >> 1) & 7;
#pragma unroll
    for (int q = 0; q < 3 * NA + 1; ++q) {
      T v = acc[q];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if ((lane & 0x31) == 0) rowacc[wv][iy][q] = (double)v;
    }
  }
  __syncthreads();
  if (FORCE && threadIdx.x < XROWS * (3 * NA + 1)) {
    // the rows' sums over the 4 waves, in wave order, to the row each component belongs to
    const int r = threadIdx.x >> 4, q = threadIdx.x & 15;
    const double v = rowacc[0][r][q] + rowacc[1][r][q] + rowacc[2][r][q] + rowacc[3][r][q];
    const int y = by * BS + r;
    const int yy = (q >= NA && q < 2 * NA) ? (y + 1 < g.Ny ? y + 1 : 0) : y;
    const int zz = (q >= 2 * NA && q < 3 * NA) ? (z + 1 < g.Nz ? z + 1 : 0) : z;
    rowbuf[((size_t)zz * g.Ny + yy) * 16 + q] = v;
  }
  if constexpr (N > 0) lds_fft<N, XFFT, -1>(buf, tw);
  else lds_fft_rt<XFFT, -1>(buf, tw, pl);
  // Separate the two real transforms of each complex FFT and pack kx = N/2 into slot 0.
  for (int t = threadIdx.x; t < XROWS * (n / 2); t += NT) {
    const int row = t / (n / 2), k = t % (n / 2);
    const int q = row >> 1, odd = row & 1;
    const C2<T> Z = buf[k * XFFT + q];
    C2<T> out;
    if (k == 0) {
      const C2<T> Zn = buf[(n / 2) * XFFT + q];
      out = odd ? C2<T>{Z.y, Zn.y} : C2<T>{Z.x, Zn.x};
    } else {
      const C2<T> Zm = buf[(n - k) * XFFT + q];
      // Xa = (Z + conj Zm)/2 ; Xb = -i (Z - conj Zm)/2
      out = odd ? C2<T>{(T)0.5 * (Z.y + Zm.y), (T)0.5 * (Zm.x - Z.x)}
                : C2<T>{(T)0.5 * (Z.x + Zm.x), (T)0.5 * (Z.y - Zm.y)};
    }
    const int y = by * BS + row;
    S[((size_t)z * g.Ny + y) * (n / 2) + k] = out;
  }
}

// Packed half spectrum -> real rows of a block-layout field. grid = (Ny/8, Nz).
// XF: also write each block's x faces (ix = 0 and 7) to xf[(b*2 + side)*64 + iz*8 + iy], so that
// the gradient reads them as 64 contiguous values instead of one value per 32-byte sector.
template <typename T, int N, bool XF = false>
__global__ __launch_bounds__(NT) void k_fft_x_inv(const C2<T>* __restrict__ S, T* __restrict__ dst,
                                                  const C2<T>* __restrict__ tw, GridArgs g, FftPlan pl,
                                                  T* __restrict__ xf = nullptr)
{
  const int n = N > 0 ? N : pl.n;
  C2<T>* buf = lds_tile<T, N * XFFT>();
  const int by = blockIdx.x, z = blockIdx.y;
  const int bz = z >> 3, iz = z & 7;
  for (int t = threadIdx.x; t < XFFT * (n / 2); t += NT) {
    const int q = t / (n / 2), k = t % (n / 2);
    const int ya = by * BS + 2 * q;
    const C2<T> A = S[((size_t)z * g.Ny + ya) * (n / 2) + k];
    const C2<T> B = S[((size_t)z * g.Ny + ya + 1) * (n / 2) + k];
    if (k == 0) {
      buf[0 * XFFT + q] = C2<T>{A.x, B.x};
      buf[(n / 2) * XFFT + q] = C2<T>{A.y, B.y};
    } else {
      buf[k * XFFT + q] = C2<T>{A.x - B.y, A.y + B.x};
      buf[(n - k) * XFFT + q] = C2<T>{A.x + B.y, B.x - A.y};
    }
  }
  __syncthreads();
  if constexpr (N > 0) lds_fft<N, XFFT, +1>(buf, tw);
  else lds_fft_rt<XFFT, +1>(buf, tw, pl);
  const T* rb = (const T*)buf;
  for (int t = threadIdx.x; t < n * XROWS; t += NT) {
    const int bx = t >> 6, iy = (t >> 3) & 7, ix = t & 7;
    const int x = bx * BS + ix;
    const int b = bx + g.nbx * (by + g.nby * bz);
    // streamed past the caches: the gradient pass reads p with its own halo order
    const T v = rb[2 * (x * XFFT + (iy >> 1)) + (iy & 1)];
    __builtin_nontemporal_store(v, &dst[(size_t)b * BS3 + (iz * BS + iy) * BS + ix]);
    if (XF && (ix == 0 || ix == BS - 1)) xf[((size_t)b * 2 + (ix != 0)) * 64 + iz * BS + iy] = v;
  }
}

// ------------------------------------------------------------------ y / z passes

template <typename T, int N>
struct ColTile {
  // columns per workgroup: LDS tile at most 64 KiB, at most 16 columns; a length known at run
  // time (N = 0, at most FFT_MAX_N) takes 4 (FP32) or 2 (FP64) columns, 32 KiB at most
  static constexpr int raw = N > 0 ? 65536 / (N * (int)sizeof(C2<T>)) : (sizeof(T) == 4 ? 4 : 2);
  static constexpr int TK = raw >= 16 ? 16 : (raw < 1 ? 1 : raw);
};

enum ZMode { ZM_FWD = 0, ZM_INV = 1, ZM_POISSON = 2 };

// Column FFT along y (AXIS=1) or z (AXIS=2). grid = (ceil(ncols/TK), other axis).
// MODE: forward, inverse, or (z only) forward + grid-consistent division + inverse.
template <typename T, int N, int AXIS, int MODE, bool FORCE = false>
__global__ __launch_bounds__(NT) void k_fft_col(C2<T>* __restrict__ S, const C2<T>* __restrict__ tw, int ncols,
                                                int Ny, int Nz, double norm, const double* __restrict__ cx,
                                                const double* __restrict__ cy, const double* __restrict__ cz,
                                                const LowModes* __restrict__ lmp, const double* __restrict__ coef,
                                                FftPlan pl)
{
  constexpr int TK = ColTile<T, N>::TK;
  static_assert(TK % 2 == 0, "columns move in pairs");
  const int n = N > 0 ? N : pl.n;
  C2<T>* buf = lds_tile<T, N * TK>();
  const int kx0 = blockIdx.x * TK, other = blockIdx.y;
  const int nc = min(TK, ncols - kx0);
  // element (f, n): AXIS=1: S[(other*Ny + n)*ncols + kx0 + f]; AXIS=2: S[(n*Ny + other)*ncols + kx0 + f].
  // Columns move in pairs (ncols, TK and so nc are even): one 2-element vector access each.
  using V4 = typename Vec4<T>::type;
  for (int t = threadIdx.x; t < n * TK / 2; t += NT) {
    const int f = 2 * (t % (TK / 2)), e = t / (TK / 2);
    V4 a = {0, 0, 0, 0};
    if (f < nc) {
      const size_t gi = AXIS == 1 ? ((size_t)other * Ny + e) * ncols + kx0 + f : ((size_t)e * Ny + other) * ncols + kx0 + f;
      a = *reinterpret_cast<const V4*>(S + gi);
    }
    *reinterpret_cast<V4*>(buf + f + e * TK) = a;
  }
  __syncthreads();
  if (MODE == ZM_INV) {
    if constexpr (N > 0) lds_fft<N, TK, +1>(buf, tw);
    else lds_fft_rt<TK, +1>(buf, tw, pl);
  } else {
    if constexpr (N > 0) lds_fft<N, TK, -1>(buf, tw);
    else lds_fft_rt<TK, -1>(buf, tw, pl);
  }
  if (FORCE) {
    // fused forcing: add the forcing's RHS modes (coefficients from k_forcing_reduce)
    const LowModes& lm = *lmp;
    for (int e = threadIdx.x; e < lm.npos; e += NT) {
      const int slot = lm.pos[e][0], ky = lm.pos[e][1], kz = lm.pos[e][2], mc = lm.pos[e][3];
      if (ky != other || slot < kx0 || slot >= kx0 + nc) continue;
      const double* cm = coef + 8 * (mc >> 1);
      const double ci = (mc & 1) ? -cm[1] : cm[1];
      C2<T>& v = buf[(slot - kx0) + kz * TK];
      v.x = (T)((double)v.x + cm[0]);
      v.y = (T)((double)v.y + ci);
    }
    __syncthreads();
  }
  if (MODE == ZM_POISSON) {
    // PoissonSolverPeriodic::_solve grid-consistent branch (PoissonSolverPeriodic.cpp:23-61):
    // multiply by norm/(2D-6), D = cos(4 pi kx/Nx) + cos(4 pi ky/Ny) + cos(4 pi kz/Nz),
    // in FP64; zero the modes with every index in {0, N/2} (kx slot 0 carries kx=0 and Nx/2).
    const int ky = other;
    for (int t = threadIdx.x; t < n * TK; t += NT) {
      const int f = t % TK, kz = t / TK;
      const int kx = kx0 + f;
      if (f < nc) {
        const double D = cx[kx] + cy[ky] + cz[kz];
        double fac = norm / (2 * D - 6);
        const bool zero = kx == 0 && (ky == 0 || 2 * ky == Ny) && (kz == 0 || 2 * kz == n);
        if (zero) fac = 0;
        C2<T> v = buf[f + kz * TK];
        v.x = (T)((double)v.x * fac);
        v.y = (T)((double)v.y * fac);
        if (zero) v = C2<T>{0, 0};
        buf[f + kz * TK] = v;
      }
    }
    __syncthreads();
    if constexpr (N > 0) lds_fft<N, TK, +1>(buf, tw);
    else lds_fft_rt<TK, +1>(buf, tw, pl);
  }
  for (int t = threadIdx.x; t < n * TK / 2; t += NT) {
    const int f = 2 * (t % (TK / 2)), n = t / (TK / 2);
    if (f < nc) {
      const size_t gi = AXIS == 1 ? ((size_t)other * Ny + n) * ncols + kx0 + f : ((size_t)n * Ny + other) * ncols + kx0 + f;
      *reinterpret_cast<V4*>(S + gi) = *reinterpret_cast<const V4*>(buf + f + n * TK);
    }
  }
}

// ------------------------------------------- y / z passes, register four-step form
// N = 16 NTH (NTH = 4, 8, 16 threads per column), TK = NT/NTH columns per workgroup.
// Layout A: thread (f, j) holds v[m] = x[j + NTH m], m < 16. Layout B: thread (f, j) holds
// v[i NTH + k2] = X[k1 + 16 k2] for its RPT = 16/NTH rows k1 = j RPT + i. With
// W_N^{(j + NTH m)(k1 + 16 k2)} = W_N^{j k1} W_NTH^{j k2} W_16^{m k1}, a transform is a register
// DFT16 over m, the twiddle W_N^{j k1}, one LDS transpose, and register DFTs of length NTH
// over j (A -> B); B -> A runs the same steps in reverse order. Column f is the workgroup's
// f-th kx column; lanes run f fastest, so each global access covers 16 columns of a row.
template <typename T, int N>
struct ColReg {
  static constexpr bool ok = N >= 64 && N <= 256;
  static constexpr int NTH = ok ? N / 16 : 1, RPT = 16 / NTH, TK = NT / NTH;
  // LDS stride of one k1 row. Unpadded, the tile is 32 KiB (FP32), so five workgroups share
  // a CU; the transposes' writes are conflict free and their reads cost less than a lost slot.
  static constexpr int ROWP = NTH * TK;
};

template <int L, int DIR, typename T>
__device__ __forceinline__ void dftL(C2<T>* v)
{
  if constexpr (L == 16) dft16<DIR>(v);
  else if constexpr (L == 8) dft8<DIR>(v);
  else dft4<DIR>(v[0], v[1], v[2], v[3]);
}

// A -> B. The caller synchronizes before when buf was read since its last barrier.
template <typename T, int N, int DIR>
__device__ __forceinline__ void fft_a_to_b(C2<T> (&v)[16], C2<T>* buf, const C2<T>* __restrict__ tw, int f, int j)
{
  using CR = ColReg<T, N>;
  dft16<DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) {
    C2<T> w = tw[(j * k1) & (N - 1)];
    if (DIR > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) buf[k1 * CR::ROWP + j * CR::TK + f] = v[k1];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < CR::RPT; ++i)
#pragma unroll
    for (int jp = 0; jp < CR::NTH; ++jp) v[i * CR::NTH + jp] = buf[(j * CR::RPT + i) * CR::ROWP + jp * CR::TK + f];
#pragma unroll
  for (int i = 0; i < CR::RPT; ++i) dftL<CR::NTH, DIR>(v + i * CR::NTH);
}

// B -> A. The caller synchronizes before when buf was read since its last barrier.
template <typename T, int N, int DIR>
__device__ __forceinline__ void fft_b_to_a(C2<T> (&v)[16], C2<T>* buf, const C2<T>* __restrict__ tw, int f, int j)
{
  using CR = ColReg<T, N>;
#pragma unroll
  for (int i = 0; i < CR::RPT; ++i) {
    dftL<CR::NTH, DIR>(v + i * CR::NTH);
    const int k1 = j * CR::RPT + i;
#pragma unroll
    for (int jn = 1; jn < CR::NTH; ++jn) {
      C2<T> w = tw[(jn * k1) & (N - 1)];
      if (DIR > 0) w.y = -w.y;
      v[i * CR::NTH + jn] = cmul(v[i * CR::NTH + jn], w);
    }
#pragma unroll
    for (int jn = 0; jn < CR::NTH; ++jn) buf[k1 * CR::ROWP + jn * CR::TK + f] = v[i * CR::NTH + jn];
  }
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) v[k1] = buf[k1 * CR::ROWP + j * CR::TK + f];
  dft16<DIR>(v);
}

// The column pass of k_fft_col in the register form (same arguments and modes). EXTRA: the
// grid's first ex.ey rows are workgroups of the forcing's plane stage (1, beside the y forward
// pass) or row stage (2, beside the z pass; coef is the forcing's coefficient table), in the
// column tile's LDS.
template <typename T, int N, int AXIS, int MODE, bool FORCE = false, int EXTRA = 0>
__global__ __launch_bounds__(NT) void k_fft_colr(C2<T>* __restrict__ S, const C2<T>* __restrict__ tw, int ncols,
                                                 int Ny, int Nz, double norm, const double* __restrict__ cx,
                                                 const double* __restrict__ cy, const double* __restrict__ cz,
                                                 const LowModes* __restrict__ lmp, const double* __restrict__ coef,
                                                 ColExtra ex)
{
  using CR = ColReg<T, N>;
  constexpr int TK = CR::TK, NTH = CR::NTH, RPT = CR::RPT;
  __shared__ __align__(16) C2<T> buf[16 * CR::ROWP];
  static_assert(EXTRA == 0 || (sizeof(buf) >= FORCING_PLANES_LDS && sizeof(buf) >= FORCING_ROWS_LDS),
                "the forcing stages fit the column tile's LDS");
  if constexpr (EXTRA != 0) {
    if ((int)blockIdx.y < ex.ey) {
      const int w = blockIdx.y * gridDim.x + blockIdx.x;  // uniform over the workgroup
      if (EXTRA == 1 && w < ex.nz * ex.parts)
        forcing_planes_wg<T>(ex.rowbuf, ex.plane, Ny, (const C2<T>*)ex.twy, w / ex.parts, w % ex.parts, ex.parts, buf);
      else if (EXTRA == 2 && w < ex.nz)
        forcing_rows_wg<T>(lmp, coef, (const C2<T>*)ex.twy, (const C2<T>*)ex.twz, (T*)ex.rows, Ny, Nz, w, buf);
      return;
    }
  }
  const int kx0 = blockIdx.x * TK, other = blockIdx.y - (EXTRA != 0 ? ex.ey : 0);
  const int f = threadIdx.x % TK, j = threadIdx.x / TK;
  const bool live = kx0 + f < ncols;
  // element n of column kx0 + f
  auto gi = [&](int n) -> size_t {
    return AXIS == 1 ? ((size_t)other * Ny + n) * ncols + kx0 + f : ((size_t)n * Ny + other) * ncols + kx0 + f;
  };
  C2<T> v[16];
  if (MODE == ZM_INV) {
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
      for (int k2 = 0; k2 < NTH; ++k2) v[i * NTH + k2] = live ? S[gi(j * RPT + i + 16 * k2)] : C2<T>{0, 0};
    fft_b_to_a<T, N, +1>(v, buf, tw, f, j);
    if (live)
#pragma unroll
      for (int m = 0; m < 16; ++m) S[gi(j + NTH * m)] = v[m];
    return;
  }
#pragma unroll
  for (int m = 0; m < 16; ++m) v[m] = live ? S[gi(j + NTH * m)] : C2<T>{0, 0};
  fft_a_to_b<T, N, -1>(v, buf, tw, f, j);
  if (FORCE && kx0 == 0) {
    // fused forcing: add the forcing's RHS modes (coefficients from k_forcing_reduce) to the
    // entries this thread holds (column f, k = k1 + 16 k2 with k1 / RPT = j)
    const LowModes& lm = *lmp;
    for (int e = 0; e < lm.npos; ++e) {
      const int slot = lm.pos[e][0], ky = lm.pos[e][1], kz = lm.pos[e][2], mc = lm.pos[e][3];
      if (ky != other || slot != f || (kz & 15) / RPT != j) continue;
      const int r = ((kz & 15) % RPT) * NTH + (kz >> 4);
      const double* cm = coef + 8 * (mc >> 1);
      const double ci = (mc & 1) ? -cm[1] : cm[1];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q == r) {
          v[q].x = (T)((double)v[q].x + cm[0]);
          v[q].y = (T)((double)v[q].y + ci);
        }
    }
  }
  if (MODE == ZM_FWD) {
    if (live)
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int k2 = 0; k2 < NTH; ++k2) S[gi(j * RPT + i + 16 * k2)] = v[i * NTH + k2];
    return;
  }
  // ZM_POISSON (z): PoissonSolverPeriodic::_solve's grid-consistent branch
  // (PoissonSolverPeriodic.cpp:23-61), multiply by norm/(2D-6) in FP64 (also for FP32 fields:
  // an FP32 factor, even formed without cancellation as -norm/(4 sum sin^2), adds rounding to
  // the lowest modes that long FP32 trajectories and their adjoints amplify), zero the modes
  // with every index in {0, N/2}; then the inverse transform.
  {
    const int kx = kx0 + f, ky = other;
    const double dxy = live ? cx[kx] + cy[ky] : 0.0;
    const bool zxy = kx == 0 && (ky == 0 || 2 * ky == Ny);
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
      for (int k2 = 0; k2 < NTH; ++k2) {
        const int kz = j * RPT + i + 16 * k2;
        const double D = dxy + cz[kz];
        double fac = norm / (2 * D - 6);
        if (zxy && (kz == 0 || 2 * kz == N)) fac = 0;
        C2<T>& q = v[i * NTH + k2];
        q.x = (T)((double)q.x * fac);
        q.y = (T)((double)q.y * fac);
      }
  }
  __syncthreads();  // the forward transpose's reads of buf are done
  fft_b_to_a<T, N, +1>(v, buf, tw, f, j);
  if (live)
#pragma unroll
    for (int m = 0; m < 16; ++m) S[gi(j + NTH * m)] = v[m];
}

// Separate spectral division (the PLAIN decomposition's "FFTW solve" pass).
template <typename T>
__global__ void k_divide(C2<T>* __restrict__ S, int ncols, int Ny, int Nz, double norm, const double* __restrict__ cx,
                         const double* __restrict__ cy, const double* __restrict__ cz)
{
  const size_t n = (size_t)ncols * Ny * Nz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(i % ncols);
    const int ky = (int)((i / ncols) % Ny);
    const int kz = (int)(i / ((size_t)ncols * Ny));
    const double D = cx[kx] + cy[ky] + cz[kz];
    const double fac = norm / (2 * D - 6);
    const bool zero = kx == 0 && (ky == 0 || 2 * ky == Ny) && (kz == 0 || 2 * kz == Nz);
    C2<T> v = S[i];
    v.x = (T)((double)v.x * fac);
    v.y = (T)((double)v.y * fac);
    if (zero) v = C2<T>{0, 0};
    S[i] = v;
  }
}

// ------------------------------------------------------------------ dispatch

template <typename T, int N>
struct XPass {
  // dynamic LDS of the passes whose length is known at run time (N = 0)
  static size_t lds(const Ctx& c) { return N > 0 ? 0 : (size_t)c.N[0] * XFFT * sizeof(C2<T>); }
  static void fwd(Ctx& c, const T* s0, const T* s1, const T* s2, int src, bool force)
  {
    dim3 grid(c.N[1] / BS, c.N[2]);
    const int* slot = c.body.present && !c.plain_rhs ? c.ob_slot_all : nullptr;
    if (src == XSRC_FIELD)
      hipLaunchKernelGGL((k_fft_x_fwd<T, N, XSRC_FIELD, false>), grid, dim3(NT), lds(c), c.stream, s0, s1, s2,
                         (C2<T>*)c.spec, (const C2<T>*)c.tw[0], c.grid(), c.sc, (const int*)nullptr,
                         (const T*)nullptr, (double*)nullptr, c.plan[0]);
    else if (force)
      hipLaunchKernelGGL((k_fft_x_fwd<T, N, XSRC_DIV, true>), grid, dim3(NT), lds(c), c.stream, s0, s1, s2, (C2<T>*)c.spec,
                         (const C2<T>*)c.tw[0], c.grid(), c.sc, slot, (const T*)c.ob_rhs_all, c.low_part, c.plan[0]);
    else
      hipLaunchKernelGGL((k_fft_x_fwd<T, N, XSRC_DIV, false>), grid, dim3(NT), lds(c), c.stream, s0, s1, s2,
                         (C2<T>*)c.spec, (const C2<T>*)c.tw[0], c.grid(), c.sc, slot, (const T*)c.ob_rhs_all,
                         (double*)nullptr, c.plan[0]);
  }
  static void fwd_to(Ctx& c, const T* s0, C2<T>* spec)
  {
    dim3 grid(c.N[1] / BS, c.N[2]);
    hipLaunchKernelGGL((k_fft_x_fwd<T, N, XSRC_FIELD, false>), grid, dim3(NT), lds(c), c.stream, s0, s0, s0, spec,
                       (const C2<T>*)c.tw[0], c.grid(), c.sc, (const int*)nullptr, (const T*)nullptr,
                       (double*)nullptr, c.plan[0]);
  }
  static void inv(Ctx& c, const C2<T>* spec, T* dst, T* xf = nullptr)
  {
    dim3 grid(c.N[1] / BS, c.N[2]);
    if (xf)
      hipLaunchKernelGGL((k_fft_x_inv<T, N, true>), grid, dim3(NT), lds(c), c.stream, spec, dst, (const C2<T>*)c.tw[0],
                         c.grid(), c.plan[0], xf);
    else
      hipLaunchKernelGGL((k_fft_x_inv<T, N>), grid, dim3(NT), lds(c), c.stream, spec, dst, (const C2<T>*)c.tw[0],
                         c.grid(), c.plan[0], (T*)nullptr);
  }
};

template <typename T, int N, int AXIS>
struct ColPass {
  // extra: 1 (y forward) or 2 (z pass with the forcing) runs the forcing's plane or row stage
  // beside the transforms; only the register form takes it (col_extra_ok)
  static void run(Ctx& c, C2<T>* spec, int mode, double norm, bool force = false, int extra = 0)
  {
    if constexpr (ColReg<T, N>::ok) {
      constexpr int TK = ColReg<T, N>::TK;
      const int ncols = c.N[0] / 2;
      const int other = AXIS == 1 ? c.N[2] : c.N[1];
      const int gx = (ncols + TK - 1) / TK;
      ColExtra ex{};
      if (extra) {
        ex = forcing_planes_extra(c);
        const int work = extra == 1 ? ex.nz * ex.parts : ex.nz;
        ex.ey = (work + gx - 1) / gx;
      }
      dim3 grid(gx, other + ex.ey);
      const C2<T>* tw = (const C2<T>*)c.tw[AXIS];
      if (mode == ZM_FWD && extra == 1)
        hipLaunchKernelGGL((k_fft_colr<T, N, AXIS, ZM_FWD, false, 1>), grid, dim3(NT), 0, c.stream, spec, tw, ncols,
                           c.N[1], c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)nullptr,
                           ex);
      else if (mode == ZM_FWD)
        hipLaunchKernelGGL((k_fft_colr<T, N, AXIS, ZM_FWD>), grid, dim3(NT), 0, c.stream, spec, tw, ncols, c.N[1],
                           c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)nullptr, ex);
      else if (mode == ZM_INV)
        hipLaunchKernelGGL((k_fft_colr<T, N, AXIS, ZM_INV>), grid, dim3(NT), 0, c.stream, spec, tw, ncols, c.N[1],
                           c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)nullptr, ex);
      else if (AXIS == 2 && force && extra == 2)
        hipLaunchKernelGGL((k_fft_colr<T, N, 2, ZM_POISSON, true, 2>), grid, dim3(NT), 0, c.stream, spec, tw, ncols,
                           c.N[1], c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev,
                           (const double*)c.low_coef, ex);
      else if (AXIS == 2 && force)
        hipLaunchKernelGGL((k_fft_colr<T, N, 2, ZM_POISSON, true>), grid, dim3(NT), 0, c.stream, spec, tw, ncols,
                           c.N[1], c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev,
                           (const double*)c.low_coef, ex);
      else if (AXIS == 2)
        hipLaunchKernelGGL((k_fft_colr<T, N, 2, ZM_POISSON, false>), grid, dim3(NT), 0, c.stream, spec, tw, ncols,
                           c.N[1], c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)nullptr,
                           ex);
      return;
    }
    constexpr int TK = ColTile<T, N>::TK;
    const int ncols = c.N[0] / 2;
    const int other = AXIS == 1 ? c.N[2] : c.N[1];
    dim3 grid((ncols + TK - 1) / TK, other);
    const C2<T>* tw = (const C2<T>*)c.tw[AXIS];
    const size_t lds = N > 0 ? 0 : (size_t)c.N[AXIS] * TK * sizeof(C2<T>);
    if (mode == ZM_FWD)
      hipLaunchKernelGGL((k_fft_col<T, N, AXIS, ZM_FWD>), grid, dim3(NT), lds, c.stream, spec, tw, ncols, c.N[1], c.N[2],
                         norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)nullptr, c.plan[AXIS]);
    else if (mode == ZM_INV)
      hipLaunchKernelGGL((k_fft_col<T, N, AXIS, ZM_INV>), grid, dim3(NT), lds, c.stream, spec, tw, ncols, c.N[1], c.N[2],
                         norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)nullptr, c.plan[AXIS]);
    else if (AXIS == 2 && force)
      hipLaunchKernelGGL((k_fft_col<T, N, 2, ZM_POISSON, true>), grid, dim3(NT), lds, c.stream, spec, tw, ncols, c.N[1],
                         c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)c.low_coef, c.plan[AXIS]);
    else if (AXIS == 2)
      hipLaunchKernelGGL((k_fft_col<T, N, 2, ZM_POISSON, false>), grid, dim3(NT), lds, c.stream, spec, tw, ncols, c.N[1],
                         c.N[2], norm, c.cos4[0], c.cos4[1], c.cos4[2], c.low_dev, (const double*)nullptr, c.plan[AXIS]);
  }
};

// Runtime N -> template N.
template <template <typename, int> class F, typename T, typename... A>
static bool dispatch_n(int n, A&&... a)
{
  switch (n) {
    case 8: F<T, 8>::go(a...); return true;
    case 16: F<T, 16>::go(a...); return true;
    case 32: F<T, 32>::go(a...); return true;
    case 64: F<T, 64>::go(a...); return true;
    case 128: F<T, 128>::go(a...); return true;
    case 256: F<T, 256>::go(a...); return true;
    case 512: F<T, 512>::go(a...); return true;
    case 1024: F<T, 1024>::go(a...); return true;
    default: F<T, 0>::go(a...); return true;  // the plan's radices, fft_setup checked the length
  }
}

template <typename T, int N>
struct GoXFwd {
  static void go(Ctx& c, const T* s0, const T* s1, const T* s2, int src, bool force)
  {
    XPass<T, N>::fwd(c, s0, s1, s2, src, force);
  }
};
template <typename T, int N>
struct GoXFwdTo {
  static void go(Ctx& c, const T* s0, C2<T>* spec) { XPass<T, N>::fwd_to(c, s0, spec); }
};
template <typename T, int N>
struct GoXInv {
  static void go(Ctx& c, const C2<T>* spec, T* dst, T* xf = nullptr) { XPass<T, N>::inv(c, spec, dst, xf); }
};
template <typename T, int N>
struct GoY {
  static void go(Ctx& c, C2<T>* spec, int mode, double norm) { ColPass<T, N, 1>::run(c, spec, mode, norm); }
};
template <typename T, int N>
struct GoZ {
  static void go(Ctx& c, C2<T>* spec, int mode, double norm) { ColPass<T, N, 2>::run(c, spec, mode, norm); }
};
template <typename T, int N>
struct GoZF {
  static void go(Ctx& c, C2<T>* spec, int mode, double norm) { ColPass<T, N, 2>::run(c, spec, mode, norm, true); }
};
// the same passes with the forcing's plane stage (y forward) or row stage (z) beside them
template <typename T, int N>
struct GoYPlanes {
  static void go(Ctx& c, C2<T>* spec, int mode, double norm) { ColPass<T, N, 1>::run(c, spec, mode, norm, false, 1); }
};
template <typename T, int N>
struct GoZFRows {
  static void go(Ctx& c, C2<T>* spec, int mode, double norm) { ColPass<T, N, 2>::run(c, spec, mode, norm, true, 2); }
};

// The column lengths whose passes take the register form (ColReg::ok over dispatch_n's cases).
static bool col_extra_ok(int n) { return n == 64 || n == 128 || n == 256; }

static double poisson_norm(const Ctx& c)
{
  return 4.0 / ((double)c.N[0] * c.h * (double)c.N[1] * (double)c.N[2]);
}

template <typename T>
static int poisson_fused_t(Ctx& c, bool force, double tke_target, double inj_rate, bool xface)
{
  C2<T>* S = (C2<T>*)c.spec;
  const double norm = poisson_norm(c);
  stage_mark(c, ST_XFWD, false);
  dispatch_n<GoXFwd, T>(c.N[0], c, (const T*)c.f[c.ustar], (const T*)c.f[c.ustar + 1],
                        (const T*)c.f[c.ustar + 2], (int)XSRC_DIV, force);
  stage_mark(c, ST_XFWD, true);
  // With the register form of the column passes, the forcing's plane stage runs as extra
  // workgroups of the y forward pass and its row stage beside the z pass, leaving the mode sums
  // (k_forcing_reduce) the only kernel of the forcing between them. Otherwise the three stages
  // are small kernels in order (a second stream's fork and join cost more here).
  const bool extra = force && col_extra_ok(c.N[1]) && col_extra_ok(c.N[2]);
  if (force && !extra) {
    stage_mark(c, ST_FREDUCE, false);
    if (int rc = launch_forcing_reduce(c, tke_target, inj_rate)) return rc;
    if (int rc = launch_forcing_rows(c)) return rc;
    stage_mark(c, ST_FREDUCE, true);
  }
  stage_mark(c, ST_YFWD, false);
  if (extra) dispatch_n<GoYPlanes, T>(c.N[1], c, S, (int)ZM_FWD, norm);
  else dispatch_n<GoY, T>(c.N[1], c, S, (int)ZM_FWD, norm);
  stage_mark(c, ST_YFWD, true);
  if (extra) {
    stage_mark(c, ST_FREDUCE, false);
    if (int rc = launch_forcing_reduce(c, tke_target, inj_rate, false)) return rc;
    stage_mark(c, ST_FREDUCE, true);
  }
  stage_mark(c, ST_ZPASS, false);
  if (extra) dispatch_n<GoZFRows, T>(c.N[2], c, S, (int)ZM_POISSON, norm);
  else if (force) dispatch_n<GoZF, T>(c.N[2], c, S, (int)ZM_POISSON, norm);
  else dispatch_n<GoZ, T>(c.N[2], c, S, (int)ZM_POISSON, norm);
  stage_mark(c, ST_ZPASS, true);
  stage_mark(c, ST_YINV, false);
  dispatch_n<GoY, T>(c.N[1], c, S, (int)ZM_INV, norm);
  stage_mark(c, ST_YINV, true);
  stage_mark(c, ST_XINV, false);
  dispatch_n<GoXInv, T>(c.N[0], c, (const C2<T>*)S, (T*)c.f[CUP_P], xface ? (T*)c.xface : (T*)nullptr);
  stage_mark(c, ST_XINV, true);
  c.xface_ok = xface;
  return check_hip(c, hipGetLastError(), "poisson_fused");
}

template <typename T>
static int poisson_plain_t(Ctx& c)
{
  C2<T>* S = (C2<T>*)c.spec;
  const double norm = poisson_norm(c);
  dispatch_n<GoXFwdTo, T>(c.N[0], c, (const T*)c.rbuf, S);
  dispatch_n<GoY, T>(c.N[1], c, S, (int)ZM_FWD, norm);
  dispatch_n<GoZ, T>(c.N[2], c, S, (int)ZM_FWD, norm);
  const size_t n = (size_t)(c.N[0] / 2) * c.N[1] * c.N[2];
  size_t gsz = (n + NT - 1) / NT;
  if (gsz > 65536) gsz = 65536;
  hipLaunchKernelGGL((k_divide<T>), dim3((unsigned)gsz), dim3(NT), 0, c.stream, S, c.N[0] / 2, c.N[1], c.N[2], norm,
                     (const double*)c.cos4[0], (const double*)c.cos4[1], (const double*)c.cos4[2]);
  dispatch_n<GoZ, T>(c.N[2], c, S, (int)ZM_INV, norm);
  dispatch_n<GoY, T>(c.N[1], c, S, (int)ZM_INV, norm);
  dispatch_n<GoXInv, T>(c.N[0], c, (const C2<T>*)S, (T*)c.rbuf);
  return check_hip(c, hipGetLastError(), "poisson_plain");
}

int launch_poisson_fused(Ctx& c, bool force, double tke_target, double inj_rate, bool xface)
{
  c.xface_ok = false;
  if (force && !c.low_fused_ok) return set_error(c, CUP_E_INVALID, "fused forcing: low-mode set not supported");
  return c.precision == 4 ? poisson_fused_t<float>(c, force, tke_target, inj_rate, xface)
                          : poisson_fused_t<double>(c, force, tke_target, inj_rate, xface);
}

int launch_poisson_plain(Ctx& c)
{
  return c.precision == 4 ? poisson_plain_t<float>(c) : poisson_plain_t<double>(c);
}

template <typename T>
static int fwd_field_t(Ctx& c, const void* real_blocks, void* spec)
{
  C2<T>* S = (C2<T>*)spec;
  dispatch_n<GoXFwdTo, T>(c.N[0], c, (const T*)real_blocks, S);
  dispatch_n<GoY, T>(c.N[1], c, S, (int)ZM_FWD, 0.0);
  dispatch_n<GoZ, T>(c.N[2], c, S, (int)ZM_FWD, 0.0);
  return check_hip(c, hipGetLastError(), "fft_forward_field");
}

template <typename T>
static int inv_field_t(Ctx& c, void* spec, void* real_blocks)
{
  C2<T>* S = (C2<T>*)spec;
  dispatch_n<GoZ, T>(c.N[2], c, S, (int)ZM_INV, 0.0);
  dispatch_n<GoY, T>(c.N[1], c, S, (int)ZM_INV, 0.0);
  dispatch_n<GoXInv, T>(c.N[0], c, (const C2<T>*)S, (T*)real_blocks);
  return check_hip(c, hipGetLastError(), "fft_inverse_field");
}

int launch_fft_forward_field(Ctx& c, const void* real_blocks, void* spec)
{
  return c.precision == 4 ? fwd_field_t<float>(c, real_blocks, spec) : fwd_field_t<double>(c, real_blocks, spec);
}

int launch_fft_inverse_field(Ctx& c, void* spec, void* real_blocks)
{
  return c.precision == 4 ? inv_field_t<float>(c, spec, real_blocks) : inv_field_t<double>(c, spec, real_blocks);
}

// ------------------------------------------------------- PoissonSolverMixed
// PoissonSolverMixed::solve (PoissonSolverMixed.cpp:66-103, _solve PoissonSolverMixed.h:27-80):
// FFTW r2r transforms per axis, R2HC / HC2R on periodic axes and REDFT10 / REDFT01 (DCT-II /
// DCT-III) on the others (PoissonSolverMixed.cpp:37-42), here as dense N x N matrices applied
// to each line in LDS (N^2 work per line: the path for non-periodic boxes, not a benchmark
// path); the coefficient (i, j, k) times norm / (denX_i + denY_j + denZ_k) in FP64, the first
// one zeroed. Works on the RHS buffer in block layout.
template <typename T>
__global__ __launch_bounds__(NT) void k_r2r_axis(T* __restrict__ a, const T* __restrict__ mt, GridArgs g, int axis)
{
  extern __shared__ __align__(16) unsigned char r2r_lds[];
  T* line = reinterpret_cast<T*>(r2r_lds);
  const int n = axis == 0 ? g.Nx : (axis == 1 ? g.Ny : g.Nz);
  const int n1 = axis == 0 ? g.Ny : g.Nx;  // the first of the other two axes
  const int p = blockIdx.x, o1 = p % n1, o2 = p / n1;
  auto at = [&](int k) -> size_t {
    return axis == 0 ? cell_index(g, k, o1, o2) : (axis == 1 ? cell_index(g, o1, k, o2) : cell_index(g, o1, o2, k));
  };
  for (int j = threadIdx.x; j < n; j += NT) line[j] = a[at(j)];
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += NT) {
    T acc = 0;
    for (int j = 0; j < n; ++j) acc += mt[(size_t)j * n + k] * line[j];
    a[at(k)] = acc;
  }
}

template <typename T>
__global__ void k_mixed_divide(T* __restrict__ a, GridArgs g, double norm, const double* __restrict__ dx,
                               const double* __restrict__ dy, const double* __restrict__ dz)
{
  const size_t n = (size_t)g.nb * BS3;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(i / BS3), c = (int)(i % BS3);
    const int x = (b % g.nbx) * BS + (c & 7), y = ((b / g.nbx) % g.nby) * BS + ((c >> 3) & 7),
              z = (b / (g.nbx * g.nby)) * BS + (c >> 6);
    if (x == 0 && y == 0 && z == 0) a[i] = 0;
    else a[i] = (T)((double)a[i] * (norm / (dx[x] + dy[y] + dz[z])));
  }
}

template <typename T>
static int mixed_setup(Ctx& c)
{
  const double tol = 0.01;  // PoissonSolverMixed.h:50
  for (int d = 0; d < 3; ++d) {
    if (c.mden[d]) continue;
    const int n = c.N[d];
    const bool per = c.bc[d] == CUP_BC_PERIODIC;
    std::vector<T> fw((size_t)n * n), bw((size_t)n * n);  // transposed: m[j][k] = M[k][j]
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) {
        double f, b;
        if (per) {
          // R2HC (halfcomplex out: r_0..r_{n/2}, then i_{n/2-1}..i_1) and HC2R
          f = k <= n / 2 ? cos(2 * M_PI * (double)j * k / n) : -sin(2 * M_PI * (double)j * (n - k) / n);
          if (j == 0) b = 1;
          else if (2 * j == n) b = cos(M_PI * (double)k);
          else if (j < n / 2) b = 2 * cos(2 * M_PI * (double)j * k / n);
          else b = -2 * sin(2 * M_PI * (double)(n - j) * k / n);
        } else {
          f = 2 * cos(M_PI * (j + 0.5) * k / n);                      // REDFT10
          b = j == 0 ? 1 : 2 * cos(M_PI * (double)j * (k + 0.5) / n);  // REDFT01
        }
        fw[(size_t)j * n + k] = (T)f;
        bw[(size_t)j * n + k] = (T)b;
      }
    std::vector<double> den((size_t)n);
    const double wf = (per ? 2 : 1) * M_PI / n;
    for (int i = 0; i < n; ++i) {
      const int kk = per ? (i <= n / 2 ? i : n - i) : i;
      const double rk = (kk + (per ? 0.0 : 0.5)) * wf;
      den[i] = (1 - tol) * (cos(2 * wf * i) - 1) / 2 - tol * rk * rk;
    }
    for (int q = 0; q < 2; ++q) {
      if (check_hip(c, hipMalloc(&c.r2r[d][q], fw.size() * sizeof(T)), "hipMalloc r2r")) return CUP_E_NOMEM;
      hipMemcpy(c.r2r[d][q], q ? bw.data() : fw.data(), fw.size() * sizeof(T), hipMemcpyHostToDevice);
    }
    if (check_hip(c, hipMalloc(&c.mden[d], den.size() * sizeof(double)), "hipMalloc mixed symbol")) return CUP_E_NOMEM;
    hipMemcpy(c.mden[d], den.data(), den.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  return CUP_OK;
}

template <typename T>
static int poisson_mixed_t(Ctx& c)
{
  if (int rc = mixed_setup<T>(c)) return rc;
  T* a = (T*)c.rbuf;
  const GridArgs g = c.grid();
  auto axis_pass = [&](int d, int q) {
    const int n = c.N[d], lines = (int)(c.ncell / (size_t)n);
    hipLaunchKernelGGL((k_r2r_axis<T>), dim3(lines), dim3(NT), (size_t)n * sizeof(T), c.stream, a,
                       (const T*)c.r2r[d][q], g, d);
  };
  for (int d = 0; d < 3; ++d) axis_pass(d, 0);
  double norm = 1 / c.h;
  for (int d = 0; d < 3; ++d) norm *= (c.bc[d] == CUP_BC_PERIODIC ? 1.0 : 0.5) / c.N[d];
  size_t gsz = (c.ncell + NT - 1) / NT;
  if (gsz > 65536) gsz = 65536;
  hipLaunchKernelGGL((k_mixed_divide<T>), dim3((unsigned)gsz), dim3(NT), 0, c.stream, a, g, norm,
                     (const double*)c.mden[0], (const double*)c.mden[1], (const double*)c.mden[2]);
  for (int d = 0; d < 3; ++d) axis_pass(d, 1);
  return check_hip(c, hipGetLastError(), "poisson_mixed");
}

int launch_poisson_mixed(Ctx& c)
{
  return c.precision == 4 ? poisson_mixed_t<float>(c) : poisson_mixed_t<double>(c);
}

// Twiddle tables e^{-2 pi i t/N} (t < N) per axis and the Poisson cosine tables, built in
// FP64 on the host.
// ------------------------------------------------------ PoissonSolverUnbounded
// PoissonSolverUnbounded::solve (PoissonSolverUnbounded.cpp:126-157): Eastwood and
// Brownrigg's cyclic convolution of the zero-padded RHS with the free-space Green's function
// of the uniform mesh, -h^2/(4 pi r) (r in cells; r = 0: -0.1924173658 h^2, :216-218), times
// 1/h^3 (:50). The reference pads to 2N-1 points per axis; here the period is 2N (the FFT
// passes' lengths), with the Green's function mirrored about N: on the physical points the
// cyclic convolution of either period is the linear one (distances stay below N, and the
// entry at distance N never meets a non-zero RHS value), so the solution is the same. The
// padded grid is a periodic context of its own (2 bpd blocks per axis) that shares the
// parent's stream; its Green's function spectrum (real: the function is even) is computed
// once.

// RHS blocks into the padded grid (the padding blocks stay zero), or the solution back.
template <typename T>
__global__ __launch_bounds__(NT) void k_pad_blocks(const T* __restrict__ src, T* __restrict__ dst, GridArgs g, int out)
{
  const int b = blockIdx.x;
  const int bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
  const size_t pb = (size_t)bx + 2 * g.nbx * (by + 2 * (size_t)g.nby * bz);
  using V4 = typename Vec4<T>::type;
  if (threadIdx.x >= BS3 / 4) return;
  const size_t o = (size_t)b * BS3 + 4 * threadIdx.x, po = pb * BS3 + 4 * threadIdx.x;
  if (out) *reinterpret_cast<V4*>(dst + o) = *reinterpret_cast<const V4*>(src + po);
  else *reinterpret_cast<V4*>(dst + po) = *reinterpret_cast<const V4*>(src + o);
}

// The free-space Green's function on the padded grid (block layout), distances mirrored
// about N per axis, times 1/h^3: -1/(4 pi h r), r = 0: -0.1924173658/h.
template <typename T>
__global__ __launch_bounds__(NT) void k_green(T* __restrict__ G, GridArgs pg, double h)
{
  const int b = blockIdx.x;
  const int bx = b % pg.nbx, by = (b / pg.nbx) % pg.nby, bz = b / (pg.nbx * pg.nby);
  for (int c = threadIdx.x; c < BS3; c += NT) {
    const int X = bx * BS + (c & 7), Y = by * BS + ((c >> 3) & 7), Z = bz * BS + (c >> 6);
    const double xi = X <= pg.Nx / 2 ? X : pg.Nx - X, yi = Y <= pg.Ny / 2 ? Y : pg.Ny - Y,
                 zi = Z <= pg.Nz / 2 ? Z : pg.Nz - Z;
    const double r = sqrt(xi * xi + yi * yi + zi * zi);
    G[(size_t)b * BS3 + c] = (T)(r > 0 ? -1.0 / (4 * M_PI * h * r) : -0.1924173658 / h);
  }
}

// The spectrum times the Green's function's (FP64, normalization included). Slots kx >= 1:
// S *= G. Slot 0 packs the kx = 0 and kx = Nx/2 planes as A + iB; with G0 = Re Gs, GN = Im Gs
// there (both real and even), the product is a S(k) + b conj S(-k), a = (G0 + GN)/2,
// b = (G0 - GN)/2, so each thread takes a (k, -k) pair of slot 0.
template <typename T>
__global__ __launch_bounds__(NT) void k_green_mul(C2<T>* __restrict__ S, const C2<double>* __restrict__ Gs,
                                                  int ncols, int Ny, int Nz)
{
  const size_t n = (size_t)ncols * Ny * Nz;
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
    const int kx = (int)(i % ncols);
    if (kx) {
      const double gg = Gs[i].x;
      C2<T> v = S[i];
      S[i] = C2<T>{(T)(v.x * gg), (T)(v.y * gg)};
      continue;
    }
    const int ky = (int)((i / ncols) % Ny), kz = (int)(i / ((size_t)ncols * Ny));
    const int my = (Ny - ky) % Ny, mz = (Nz - kz) % Nz;
    const size_t m = ((size_t)mz * Ny + my) * ncols;
    if (m < i) continue;  // the pair's lower index does both
    const double a = 0.5 * (Gs[i].x + Gs[i].y), bq = 0.5 * (Gs[i].x - Gs[i].y);
    const C2<T> z = S[i], zm = S[m];
    // a z + b conj(zm), and at the mirror a zm + b conj(z) (G is even: the same a, b)
    const C2<T> o1{(T)(a * z.x + bq * zm.x), (T)(a * z.y - bq * zm.y)};
    const C2<T> o2{(T)(a * zm.x + bq * z.x), (T)(a * zm.y - bq * z.y)};
    S[i] = o1;
    if (m != i) S[m] = o2;
  }
}

void unbounded_free(Ctx& c)
{
  if (!c.unb) return;
  Ctx& p = *c.unb;
  void* bufs[] = {p.spec, p.rbuf, p.f[CUP_P], p.ghat, p.tw[0], p.tw[1], p.tw[2]};
  for (void* b : bufs)
    if (b) hipFree(b);
  for (auto* q : p.cos4)
    if (q) hipFree(q);
  delete c.unb;
  c.unb = nullptr;
}

template <typename T>
static int unbounded_setup_t(Ctx& c)
{
  Ctx* p = new (std::nothrow) Ctx();
  if (!p) return set_error(c, CUP_E_NOMEM, "poisson_unbounded: out of host memory");
  c.unb = p;
  p->dev = c.dev;
  p->stream = c.stream;
  p->precision = c.precision;
  p->esz = c.esz;
  p->decomposition = CUP_PLAIN;
  for (int a = 0; a < 3; ++a) {
    p->bpd[a] = 2 * c.bpd[a];
    p->N[a] = 2 * c.N[a];
    p->extent[a] = 2 * c.extent[a];
  }
  p->nb = 8 * c.nb;
  p->ncell = 8 * c.ncell;
  p->h = c.h;
  p->maxextent = 2 * c.maxextent;
  p->sc = c.sc;
  if (int rc = fft_setup(*p)) return set_error(c, rc, "poisson_unbounded: " + p->err);
  const size_t nspec = (size_t)(p->N[0] / 2) * p->N[1] * p->N[2];
  if (check_hip(c, hipMalloc(&p->spec, nspec * sizeof(C2<T>)), "hipMalloc unbounded spectrum") ||
      check_hip(c, hipMalloc(&p->rbuf, p->ncell * sizeof(T)), "hipMalloc unbounded rhs") ||
      check_hip(c, hipMalloc(&p->f[CUP_P], p->ncell * sizeof(T)), "hipMalloc unbounded solution") ||
      check_hip(c, hipMalloc(&p->ghat, nspec * sizeof(C2<double>)), "hipMalloc unbounded kernel"))
    return CUP_E_NOMEM;
  // the Green's function's spectrum, normalized by the padded size, kept in FP64
  hipLaunchKernelGGL((k_green<T>), dim3(p->nb), dim3(NT), 0, c.stream, (T*)p->rbuf, p->grid(), c.h);
  if (int rc = launch_fft_forward_field(*p, p->rbuf, p->spec)) return set_error(c, rc, "poisson_unbounded: " + p->err);
  const double norm = 1.0 / ((double)p->N[0] * p->N[1] * p->N[2]);
  std::vector<C2<T>> hs(nspec);
  std::vector<C2<double>> gd(nspec);
  if (check_hip(c, hipMemcpyAsync(hs.data(), p->spec, nspec * sizeof(C2<T>), hipMemcpyDeviceToHost, c.stream),
                "unbounded kernel") ||
      check_hip(c, hipStreamSynchronize(c.stream), "synchronize"))
    return CUP_E_DEVICE;
  for (size_t i = 0; i < nspec; ++i) gd[i] = C2<double>{(double)hs[i].x * norm, (double)hs[i].y * norm};
  if (check_hip(c, hipMemcpy(p->ghat, gd.data(), nspec * sizeof(C2<double>), hipMemcpyHostToDevice), "unbounded kernel") ||
      check_hip(c, hipMemsetAsync(p->rbuf, 0, p->ncell * sizeof(T), c.stream), "unbounded zero"))
    return CUP_E_DEVICE;
  return CUP_OK;
}

template <typename T>
static int poisson_unbounded_t(Ctx& c)
{
  if (!c.unb)
    if (int rc = unbounded_setup_t<T>(c)) {
      unbounded_free(c);
      return rc;
    }
  Ctx& p = *c.unb;
  p.stream = c.stream;
  hipLaunchKernelGGL((k_pad_blocks<T>), dim3(c.nb), dim3(NT), 0, c.stream, (const T*)c.rbuf, (T*)p.rbuf, c.grid(), 0);
  if (int rc = launch_fft_forward_field(p, p.rbuf, p.spec)) return set_error(c, rc, "poisson_unbounded: " + p.err);
  const size_t n = (size_t)(p.N[0] / 2) * p.N[1] * p.N[2];
  const unsigned gsz = (unsigned)std::min<size_t>(65536, (n + NT - 1) / NT);
  hipLaunchKernelGGL((k_green_mul<T>), dim3(gsz), dim3(NT), 0, c.stream, (C2<T>*)p.spec, (const C2<double>*)p.ghat,
                     p.N[0] / 2, p.N[1], p.N[2]);
  if (int rc = launch_fft_inverse_field(p, p.spec, p.f[CUP_P])) return set_error(c, rc, "poisson_unbounded: " + p.err);
  hipLaunchKernelGGL((k_pad_blocks<T>), dim3(c.nb), dim3(NT), 0, c.stream, (const T*)p.f[CUP_P], (T*)c.f[CUP_P],
                     c.grid(), 1);
  return check_hip(c, hipGetLastError(), "poisson_unbounded");
}

int launch_poisson_unbounded(Ctx& c)
{
  if (2 * std::max(c.N[0], std::max(c.N[1], c.N[2])) > FFT_MAX_N)
    return set_error(c, CUP_E_INVALID, "poisson_unbounded: the padded grid needs at most 512 cells per axis");
  return c.precision == 4 ? poisson_unbounded_t<float>(c) : poisson_unbounded_t<double>(c);
}

int fft_setup(Ctx& c)
{
  for (int a = 0; a < 3; ++a) {
    const int n = c.N[a];
    // powers of two take the kernels specialised for their length; other lengths the
    // run-time radix plan (8, 4, 2, then 3 and 5), as FFTW plans any size
    // (PoissonSolverPeriodic.cpp:108-121)
    FftPlan& pl = c.plan[a];
    pl = FftPlan{};
    pl.n = n;
    int rem = n;
    for (int r : {8, 4, 2, 3, 5})
      while (rem % r == 0 && pl.ns < 16) {
        pl.r[pl.ns++] = r;
        rem /= r;
      }
    if (n < 8 || n > FFT_MAX_N || rem != 1)
      return set_error(c, CUP_E_INVALID,
                       "FFT sizes must be between 8 and 1024 cells per axis with no prime factor above 5");
    std::vector<double> tw(2 * (size_t)n), cs((size_t)n);
    for (int t = 0; t < n; ++t) {
      const double ang = 2.0 * M_PI * (double)t / (double)n;
      tw[2 * t] = cos(ang);
      tw[2 * t + 1] = -sin(ang);
      cs[t] = cos(4.0 * M_PI * (double)t / (double)n);
    }
    if (check_hip(c, hipMalloc(&c.tw[a], (size_t)n * 2 * c.esz), "hipMalloc twiddles")) return CUP_E_DEVICE;
    if (c.precision == 4) {
      std::vector<float> twf(2 * (size_t)n);
      for (size_t i = 0; i < twf.size(); ++i) twf[i] = (float)tw[i];
      hipMemcpy(c.tw[a], twf.data(), twf.size() * sizeof(float), hipMemcpyHostToDevice);
    } else {
      hipMemcpy(c.tw[a], tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    if (check_hip(c, hipMalloc(&c.cos4[a], (size_t)n * sizeof(double)), "hipMalloc cos4")) return CUP_E_DEVICE;
    hipMemcpy(c.cos4[a], cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  return CUP_OK;
}

}  // namespace cup
