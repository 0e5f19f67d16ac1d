// Blocked Gauss-Jordan inverse without pivoting, bitwise equal to MATINV
// (PREP/template/GLOBAL/src/prep_GLOBAL.f90:516-568) for matrices too large
// for LDS (order n up to a few thousand), batched over `count` matrices.
//
// MATINV gives every element one fixed sequence of operations.  Forward
// step i (rows below i) and backward step i (rows above i) each replace
// element (r, j) by  B(r,j) - b(r) * B(i,j)  with the multiplier b(r) the
// value element (r, i) held just before the step (and element (r, i) itself
// first set to zero).  The pivot row i is final when the step reads it.  So
// the steps can be grouped in panels of KB consecutive pivots, provided each
// element still receives its updates one at a time in pivot order:
//
//   forward, panel P = rows/pivots [p0, p1):
//     diag_kernel      rows of P, columns of P: steps of P in order (scale,
//                      eliminate), recording each step's reciprocal pivot
//                      and multipliers
//     cols_kernel      rows of P, every other column: the same steps, one
//                      column per thread (a column's updates need only its
//                      own values and the recorded pivots and multipliers)
//     (copy the final rows of P to Rq)
//     tri_kernel       each later row: the KB steps on its P columns, which
//                      yields its KB multipliers (the running values)
//     trail_kernel     each later row, every column outside P: KB ordered
//                      updates  acc = acc - m(r,k) * Rq(k,j)
//   backward, panel Q (increasing), rows above the panel first:
//     (copy rows of Q to Rq: these are the forward results the steps read)
//     tri_kernel / trail_kernel on rows [0, q0)
//     diag_kernel / cols_kernel  rows of Q: steps of Q above their own index
//
// No operation is reassociated or fused (the file is built with
// -ffp-contract=off), so the result is bitwise MATINV's.
#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

constexpr int MB_KB = 32;

#define BM(i, j) B[(long)(i) + (long)(j) * n]  // 0-based, column-major

// Steps of one panel on its own columns [p0, p1), one workgroup per matrix
// with the kb x kb block in LDS.  Forward: for each pivot i, scale row i
// (recording 1/B(i,i)), then rows below i in the panel take
// B(r,c) <- (c == i ? 0 : B(r,c)) - m(r,i) * B(i,c) with m(r,i) = B(r,i) just
// before the step (recorded).  Backward: for each i, rows above i in the panel
// take B(r,c) <- (c == i ? 0 : B(r,c)) - m(r,i) * Rq(i,c), Rq the panel's
// forward rows.  pm per matrix: [0, MB_KB) reciprocal pivots, then
// m(r, i) at MB_KB + r * MB_KB + i.
constexpr int MB_PM = MB_KB + MB_KB * MB_KB;

__global__ __launch_bounds__(256) void mb_diag_kernel(double *mats, long ms, int n, int p0, int p1,
                                                      int backward, const double *rqs, int *zp,
                                                      double *pms, long pm_stride) {
  __shared__ double Dg[MB_KB][MB_KB + 1];
  __shared__ double Rd[MB_KB][MB_KB + 1];
  __shared__ double b1[MB_KB];
  __shared__ double piv;
  double *B = mats + (long)blockIdx.x * ms;
  const double *Rq = rqs + (long)blockIdx.x * MB_KB * n;
  double *P = pms + (long)blockIdx.x * pm_stride;
  const int t = threadIdx.x, kb = p1 - p0;
  for (int idx = t; idx < kb * kb; idx += 256) {
    const int r = idx % kb, c = idx / kb;
    Dg[r][c] = BM(p0 + r, p0 + c);
    if (backward) Rd[r][c] = Rq[(long)r * n + p0 + c];
  }
  __syncthreads();
  if (!backward) {
    for (int i = 0; i < kb; ++i) {
      if (t == 0) {
        if (Dg[i][i] == 0.0 && zp) atomicAdd(zp, 1);
        piv = 1.0 / Dg[i][i];
        Dg[i][i] = 1.0;
        P[i] = piv;
      }
      __syncthreads();
      const double bp = piv;
      if (t < kb) Dg[i][t] = Dg[i][t] * bp;
      if (t > i && t < kb) {
        b1[t] = Dg[t][i];
        P[MB_KB + t * MB_KB + i] = Dg[t][i];
      }
      __syncthreads();
      const int rows = kb - i - 1;
      for (int idx = t; idx < rows * kb; idx += 256) {
        const int r = i + 1 + idx % rows, c = idx / rows;
        const double cur = (c == i) ? 0.0 : Dg[r][c];
        Dg[r][c] = cur - b1[r] * Dg[i][c];
      }
      __syncthreads();
    }
  } else {
    for (int i = 1; i < kb; ++i) {
      if (t < i) {
        b1[t] = Dg[t][i];
        P[MB_KB + t * MB_KB + i] = Dg[t][i];
      }
      __syncthreads();
      for (int idx = t; idx < i * kb; idx += 256) {
        const int r = idx % i, c = idx / i;
        const double cur = (c == i) ? 0.0 : Dg[r][c];
        Dg[r][c] = cur - b1[r] * Rd[i][c];
      }
      __syncthreads();
    }
  }
  for (int idx = t; idx < kb * kb; idx += 256) {
    const int r = idx % kb, c = idx / kb;
    BM(p0 + r, p0 + c) = Dg[r][c];
  }
}

// The same steps on every column outside [p0, p1): one thread per column,
// the kb panel rows in registers, updates in step order per element.
__global__ __launch_bounds__(256) void mb_cols_kernel(double *mats, long ms, int n, int p0, int p1,
                                                      int backward, const double *rqs,
                                                      const double *pms, long pm_stride) {
  __shared__ double Pm[MB_PM];
  double *B = mats + (long)blockIdx.y * ms;
  const double *Rq = rqs + (long)blockIdx.y * MB_KB * n;
  const double *P = pms + (long)blockIdx.y * pm_stride;
  const int kb = p1 - p0;
  for (int idx = threadIdx.x; idx < MB_PM; idx += 256) Pm[idx] = P[idx];
  __syncthreads();
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n || (j >= p0 && j < p1)) return;
  double x[MB_KB];
#pragma unroll
  for (int k = 0; k < MB_KB; ++k) x[k] = (k < kb) ? BM(p0 + k, j) : 0.0;
  if (!backward) {
#pragma unroll
    for (int i = 0; i < MB_KB; ++i) {
      if (i < kb) {
        x[i] = x[i] * Pm[i];
#pragma unroll
        for (int r = i + 1; r < MB_KB; ++r)
          if (r < kb) x[r] = x[r] - Pm[MB_KB + r * MB_KB + i] * x[i];
      }
    }
  } else {
#pragma unroll
    for (int i = 1; i < MB_KB; ++i) {
      if (i < kb) {
        const double rv = Rq[(long)i * n + j];
#pragma unroll
        for (int r = 0; r < i; ++r) x[r] = x[r] - Pm[MB_KB + r * MB_KB + i] * rv;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < MB_KB; ++k)
    if (k < kb) BM(p0 + k, j) = x[k];
}

// Rq(k, j) = B(p0 + k, j), k < kb (row-major, ld n)
__global__ void mb_copy_rows_kernel(const double *mats, long ms, int n, int p0, int kb,
                                    double *rqs) {
  const double *B = mats + (long)blockIdx.y * ms;
  double *Rq = rqs + (long)blockIdx.y * MB_KB * n;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)kb * n) return;
  const int k = (int)(t / n), j = (int)(t % n);
  Rq[(long)k * n + j] = BM(p0 + k, j);
}

// Rows [r_lo, r_hi): apply the kb steps p0.. to the panel columns, recording
// the multipliers m(r, k) at M[k * n + r] (consecutive rows adjacent, so the
// stores of a wavefront are contiguous).
__global__ __launch_bounds__(256) void mb_tri_kernel(double *mats, long ms, int n, int p0, int kb,
                                                     int r_lo, int r_hi, const double *rqs,
                                                     double *mult) {
  __shared__ double D[MB_KB][MB_KB + 1];  // D[k][j] = Rq(k, p0 + j)
  double *B = mats + (long)blockIdx.y * ms;
  const double *Rq = rqs + (long)blockIdx.y * MB_KB * n;
  double *M = mult + (long)blockIdx.y * (long)n * MB_KB;
  for (int idx = threadIdx.x; idx < kb * kb; idx += blockDim.x) {
    const int k = idx / kb, j = idx % kb;
    D[k][j] = Rq[(long)k * n + p0 + j];
  }
  __syncthreads();
  const int r = r_lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= r_hi) return;
  double cur[MB_KB];
#pragma unroll
  for (int j = 0; j < MB_KB; ++j) cur[j] = (j < kb) ? BM(r, p0 + j) : 0.0;
  // guarded, not broken out of, so that the loops unroll and cur stays in registers
#pragma unroll
  for (int k = 0; k < MB_KB; ++k) {
    if (k < kb) {
      const double b = cur[k];
      M[(long)k * n + r] = b;
#pragma unroll
      for (int j = 0; j < MB_KB; ++j) {
        if (j < kb) {
          const double c = (j == k) ? 0.0 : cur[j];
          cur[j] = c - b * D[k][j];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < MB_KB; ++j)
    if (j < kb) BM(r, p0 + j) = cur[j];
}

// Rows [r_lo, r_hi), columns outside [p0, p0+kb): B(r,j) <- kb ordered updates.
constexpr int MT_R = 64, MT_C = 64;
__global__ __launch_bounds__(256) void mb_trail_kernel(double *mats, long ms, int n, int p0,
                                                       int kb, int r_lo, int r_hi,
                                                       const double *rqs, const double *mult) {
  __shared__ double Ms[MT_R][MB_KB + 1];
  __shared__ double Rs[MB_KB][MT_C + 1];
  double *B = mats + (long)blockIdx.z * ms;
  const double *Rq = rqs + (long)blockIdx.z * MB_KB * n;
  const double *M = mult + (long)blockIdx.z * (long)n * MB_KB;
  const int r0 = r_lo + blockIdx.x * MT_R;
  const int c0 = blockIdx.y * MT_C;
  if (r0 >= r_hi) return;
  const int t = threadIdx.x;
  for (int idx = t; idx < MT_R * MB_KB; idx += 256) {
    const int rr = idx % MT_R, k = idx / MT_R;
    const int r = r0 + rr;
    Ms[rr][k] = (r < r_hi && k < kb) ? M[(long)k * n + r] : 0.0;
  }
  for (int idx = t; idx < MB_KB * MT_C; idx += 256) {
    const int cc = idx % MT_C, k = idx / MT_C;
    const int c = c0 + cc;
    Rs[k][cc] = (c < n && k < kb) ? Rq[(long)k * n + c] : 0.0;
  }
  __syncthreads();
  const int tr = t % 16, tc = t / 16;  // rows tr + 16 i, columns 4 tc + jj
  double acc[4][4];
  bool okr[4], okc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + tr + 16 * i;
    okr[i] = r < r_hi;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int c = c0 + 4 * tc + jj;
      okc[jj] = c < n && (c < p0 || c >= p0 + kb);
      acc[i][jj] = (okr[i] && okc[jj]) ? BM(r, c) : 0.0;
    }
  }
  for (int k = 0; k < kb; ++k) {
    double mv[4], rv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) mv[i] = Ms[tr + 16 * i][k];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) rv[jj] = Rs[k][4 * tc + jj];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) acc[i][jj] = acc[i][jj] - mv[i] * rv[jj];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
      if (okr[i] && okc[jj]) BM(r0 + tr + 16 * i, c0 + 4 * tc + jj) = acc[i][jj];
}
#undef BM

size_t matinv_blocked_scratch(int n, int count) {
  // Rq (count * KB * n) + multipliers (count * n * KB), doubles
  return (size_t)2 * count * MB_KB * n;
}

hipError_t launch_matinv_blocked(double *mats, long ms, int n, int count, int *zp,
                                 double *scratch, hipStream_t st) {
  if (n < MB_PM / MB_KB + 1) return hipErrorInvalidValue;  // pm must fit a matrix's multiplier rows
  double *rq = scratch;
  double *mult = scratch + (size_t)count * MB_KB * n;
  auto copy_rows = [&](int p0, int kb) {
    const long tot = (long)kb * n;
    hipLaunchKernelGGL(mb_copy_rows_kernel, dim3((unsigned)((tot + 255) / 256), count), dim3(256),
                       0, st, mats, ms, n, p0, kb, rq);
  };
  auto tri = [&](int p0, int kb, int r_lo, int r_hi) {
    if (r_hi <= r_lo) return;
    hipLaunchKernelGGL(mb_tri_kernel, dim3((r_hi - r_lo + 255) / 256, count), dim3(256), 0, st,
                       mats, ms, n, p0, kb, r_lo, r_hi, rq, mult);
  };
  auto trail = [&](int p0, int kb, int r_lo, int r_hi) {
    if (r_hi <= r_lo) return;
    dim3 grid((r_hi - r_lo + MT_R - 1) / MT_R, (n + MT_C - 1) / MT_C, count);
    hipLaunchKernelGGL(mb_trail_kernel, grid, dim3(256), 0, st, mats, ms, n, p0, kb, r_lo, r_hi,
                       rq, mult);
  };
  // a panel's pivots and multipliers live in the multiplier scratch, whose
  // tri / trail use of the same panel is over before they are written
  const long pm_stride = (long)n * MB_KB;
  auto panel = [&](int p0, int p1, int backward) {
    hipLaunchKernelGGL(mb_diag_kernel, dim3(count), dim3(256), 0, st, mats, ms, n, p0, p1, backward,
                       rq, backward ? nullptr : zp, mult, pm_stride);
    hipLaunchKernelGGL(mb_cols_kernel, dim3((n + 255) / 256, count), dim3(256), 0, st, mats, ms, n,
                       p0, p1, backward, rq, mult, pm_stride);
  };
  for (int p0 = 0; p0 < n; p0 += MB_KB) {
    const int p1 = min(n, p0 + MB_KB), kb = p1 - p0;
    panel(p0, p1, 0);
    if (p1 < n) {
      copy_rows(p0, kb);
      tri(p0, kb, p1, n);
      trail(p0, kb, p1, n);
    }
  }
  for (int q0 = 0; q0 < n; q0 += MB_KB) {
    const int q1 = min(n, q0 + MB_KB), kb = q1 - q0;
    copy_rows(q0, kb);
    tri(q0, kb, 0, q0);
    trail(q0, kb, 0, q0);
    panel(q0, q1, 1);
  }
  return hipGetLastError();
}

}  // namespace sevp
