// cup_forcing_dev.h -- the fused forcing's plane and row stages as workgroup bodies. They run as
// kernels of their own (k_forcing_planes, k_forcing_rows) or as extra workgroups of the y and z
// column passes, which do not read what they write: the y pass runs the plane stage beside its
// transforms and the z pass the row stage, so neither stage is a launch on the step's critical
// path. No reference counterpart (SpectralForcing.cpp:37-77 forms the same quantities with six
// full FFTs).
#ifndef CUP_FORCING_DEV_H
#define CUP_FORCING_DEV_H

#include "cup_internal.h"

namespace cup {

// The extra workgroups of a column pass: the first ey rows of its grid (every x index) are
// workgroups of the plane stage (kind 1, nz * parts of them, the rest idle) or of the row stage
// (kind 2, nz of them); the column transforms take the rows after them.
struct ColExtra {
  const double* rowbuf;  // plane stage: the x pass's row values
  double* plane;         // plane stage: output
  const void* twy;       // twiddles e^{-2 pi i t/Ny} (context precision)
  const void* twz;       // twiddles e^{-2 pi i t/Nz} (row stage)
  void* rows;            // row stage: output (context precision)
  int parts, nz, ey;
};

// LDS bytes the two stages need (carved from the column pass's tile).
constexpr int FORCING_PLANES_LDS = 16 * 16 * 5 * 8;
constexpr int FORCING_ROWS_LDS = LOW_MAX * 3 * 2 * 8 + 5 * 3 * 3 * 2 * 8 + LOW_MAX * 3 * 4;

// Fused forcing, stage 2 (plane DFTs in y): workgroup (z, part) over the x pass's row values
// R(y, z, s) (16 real values per row, rowbuf[(z*Ny + y)*16 + s]) of rows y in part's share of
// the parts. For each slot s it forms P(z, part, s) = sum_y R e^{-i ky y} at ky = 0, 1, 2 (5 real
// values: the ky = 0 sum, then re/im at ky = 1 and 2; ky < 0 are their conjugates, the slots
// being real) into plane[((z*parts + part)*16 + s)*5 + k]. Thread t takes slot t % 16 and every
// 16th row from t/16; the 16 row groups are added in a fixed order. lds: FORCING_PLANES_LDS bytes.
template <typename T>
__device__ __forceinline__ void forcing_planes_wg(const double* __restrict__ rowbuf, double* __restrict__ plane,
                                                  int Ny, const C2<T>* __restrict__ twy, int z, int part, int parts,
                                                  void* lds)
{
  static_assert(NT == 256, "16 slots x 16 row groups");
  double(*acc)[16][5] = reinterpret_cast<double(*)[16][5]>(lds);
  const int t = threadIdx.x, sl = t & 15, grp = t >> 4;
  const int ny = Ny / parts, y0 = part * ny;
  double a[5] = {0, 0, 0, 0, 0};
#pragma unroll 4
  for (int y = y0 + grp; y < y0 + ny; y += 16) {
    const double R = rowbuf[((size_t)z * Ny + y) * 16 + sl];
    const C2<T> e1 = twy[y], e2 = twy[2 * y < Ny ? 2 * y : 2 * y - Ny];  // e^{-i a y}, e^{-2 i a y}
    a[0] += R;
    a[1] += R * (double)e1.x;
    a[2] += R * (double)e1.y;
    a[3] += R * (double)e2.x;
    a[4] += R * (double)e2.y;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) acc[grp][sl][k] = a[k];
  __syncthreads();
  if (t < 16 * 5) {
    const int s2 = t / 5, k = t % 5;
    double v = 0;
    for (int g2 = 0; g2 < 16; ++g2) v += acc[g2][s2][k];
    plane[(((size_t)z * parts + part) * 16 + s2) * 5 + k] = v;
  }
}

// Fused forcing, stage 4: for every grid row (y, z) the coefficients of the velocity modes'
// synthesis along x, fac*u_low(x) = R0 + 2 Re(R1 e^{2 pi i x/Nx}) + 2 Re(R2 e^{4 pi i x/Nx})
// per component (5 values x 3 = 15, padded to 16), so the gradient pass evaluates the forcing
// with a handful of FMAs per cell. Mode m (velocity coefficient c, component by component)
// contributes c e^{2 pi i (ky y/Ny + kz z/Nz)} to the entries of |kx|: 2 Re at kx = 0, Re and
// (sign kx) Im at |kx| = 1, 2. Taking a mode with kx < 0 as its conjugate (conj c, -ky, -kz; the
// real parts are unchanged), the row value is a sum over ky of A_ky(z) e^{2 pi i ky y/Ny} with
// A_ky(z) = sum over the modes of that ky of c e^{2 pi i kz z/Nz}: the workgroup of plane z forms
// B_m = c_m e^{...z} per mode, then the 5 x 3 x 3 sums A (in mode order), then each thread the 16
// entries of one row from two y twiddles. All in FP64; the phases are conjugates of the twiddles
// e^{-2 pi i t/N}. lds: FORCING_ROWS_LDS bytes.
template <typename T>
__device__ __forceinline__ void forcing_rows_wg(const LowModes* __restrict__ lmp, const double* __restrict__ coef,
                                                const C2<T>* __restrict__ twy, const C2<T>* __restrict__ twz,
                                                T* __restrict__ rows, int Ny, int Nz, int z, void* lds)
{
  double(*B)[3][2] = reinterpret_cast<double(*)[3][2]>(lds);                      // [LOW_MAX][3][2]
  double(*A)[3][3][2] = reinterpret_cast<double(*)[3][3][2]>(B + LOW_MAX);        // [5][3][3][2]
  int(*mk)[3] = reinterpret_cast<int(*)[3]>(reinterpret_cast<double*>(A + 5));   // [LOW_MAX][3]
  const LowModes& lm = *lmp;
  const int n = lm.n, t = threadIdx.x;
  if (t < n) {
    const int kx = lm.k[t][0], sg = kx < 0 ? -1 : 1;
    const int ky = sg * lm.k[t][1], kz = sg * lm.k[t][2];
    mk[t][0] = sg * kx;
    mk[t][1] = ky;
    mk[t][2] = kz;
    const C2<T> e = twz[((kz * z) % Nz + Nz) % Nz];
    const double ec = (double)e.x, es = -(double)e.y;  // e^{+2 pi i kz z/Nz}
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double cr = coef[8 * t + 2 + 2 * q], ci = sg * coef[8 * t + 3 + 2 * q];
      B[t][q][0] = cr * ec - ci * es;
      B[t][q][1] = cr * es + ci * ec;
    }
  }
  __syncthreads();
  if (t < 5 * 3 * 3) {
    const int ky = t / 9 - 2, q = (t / 3) % 3, ak = t % 3;
    double ar = 0, ai = 0;
    for (int m = 0; m < n; ++m)
      if (mk[m][1] == ky && mk[m][0] == ak) {
        ar += B[m][q][0];
        ai += B[m][q][1];
      }
    A[ky + 2][q][ak][0] = ar;
    A[ky + 2][q][ak][1] = ai;
  }
  __syncthreads();
  // one (component, |kx|) pair at a time, stored as it is formed: the stage keeps few registers
  // when it runs beside the z pass's transforms, whose occupancy is set by the larger count
  for (int y = t; y < Ny; y += NT) {
    const C2<T> t1 = twy[y], t2 = twy[(2 * y) % Ny];
    // e^{+2 pi i ky y/Ny} for ky = -2 .. 2
    const double er[5] = {(double)t2.x, (double)t1.x, 1.0, (double)t1.x, (double)t2.x};
    const double ei[5] = {(double)t2.y, (double)t1.y, 0.0, -(double)t1.y, -(double)t2.y};
    T* o = rows + ((size_t)z * Ny + y) * 16;
#pragma unroll 1
    for (int qa = 0; qa < 9; ++qa) {
      const int q = qa / 3, ak = qa % 3;
      double xr = 0, xi = 0;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        xr += A[k][q][ak][0] * er[k] - A[k][q][ak][1] * ei[k];
        xi += A[k][q][ak][0] * ei[k] + A[k][q][ak][1] * er[k];
      }
      if (ak == 0) {
        o[5 * q] = (T)(2 * xr);
      } else {
        o[5 * q + 2 * ak - 1] = (T)xr;
        o[5 * q + 2 * ak] = (T)xi;
      }
    }
    o[15] = (T)0;
  }
}

}  // namespace cup

#endif
