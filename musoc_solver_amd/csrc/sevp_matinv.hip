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
//     panel_kernel     rows of P: steps of P in order (scale, eliminate)
//     (copy the final rows of P to Rq)
//     tri_kernel       each later row: the KB steps on its P columns, which
//                      yields its KB multipliers (the running values)
//     trail_kernel     each later row, every column outside P: KB ordered
//                      updates  acc = acc - m(r,k) * Rq(k,j)
//   backward, panel Q (increasing), rows above the panel first:
//     (copy rows of Q to Rq: these are the forward results the steps read)
//     tri_kernel / trail_kernel on rows [0, q0)
//     panel_kernel     rows of Q: steps of Q above their own index
//
// No operation is reassociated or fused (the file is built with
// -ffp-contract=off), so the result is bitwise MATINV's.
#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

constexpr int MB_KB = 32;

#define BM(i, j) B[(long)(i) + (long)(j) * n]  // 0-based, column-major

// Forward: rows [p0,p1) all columns; scale each pivot row then eliminate the
// panel rows below it.  Backward: rows [q0,q1), steps i in (row, q1) using the
// forward rows kept in Rq.
__global__ __launch_bounds__(256) void mb_panel_kernel(double *mats, long ms, int n, int p0,
                                                       int p1, int backward, const double *rqs,
                                                       int *zp) {
  __shared__ double b1[MB_KB];
  __shared__ double piv;
  double *B = mats + (long)blockIdx.x * ms;
  const double *Rq = rqs + (long)blockIdx.x * MB_KB * n;
  const int t = threadIdx.x, nt = blockDim.x;
  const int kb = p1 - p0;
  if (!backward) {
    for (int i = p0; i < p1; ++i) {
      if (t == 0) {
        if (BM(i, i) == 0.0 && zp) atomicAdd(zp, 1);
        piv = 1.0 / BM(i, i);
        BM(i, i) = 1.0;
      }
      __syncthreads();
      const double bp = piv;
      for (int j = t; j < n; j += nt) BM(i, j) = BM(i, j) * bp;
      for (int r = i + 1 + t; r < p1; r += nt) b1[r - p0] = BM(r, i);
      __syncthreads();
      const int rows = p1 - i - 1;
      for (long idx = t; idx < (long)rows * n; idx += nt) {
        const int r = i + 1 + (int)(idx % rows), j = (int)(idx / rows);
        const double cur = (j == i) ? 0.0 : BM(r, j);
        BM(r, j) = cur - b1[r - p0] * BM(i, j);
      }
      __syncthreads();
    }
  } else {
    for (int i = p0 + 1; i < p1; ++i) {
      for (int r = p0 + t; r < i; r += nt) b1[r - p0] = BM(r, i);
      __syncthreads();
      const int rows = i - p0;
      const double *Ri = Rq + (long)(i - p0) * n;
      for (long idx = t; idx < (long)rows * n; idx += nt) {
        const int r = p0 + (int)(idx % rows), j = (int)(idx / rows);
        const double cur = (j == i) ? 0.0 : BM(r, j);
        BM(r, j) = cur - b1[r - p0] * Ri[j];
      }
      __syncthreads();
    }
  }
  (void)kb;
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
// the multipliers m(r, k) (row-major, ld MB_KB).
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
#pragma unroll
  for (int k = 0; k < MB_KB; ++k) {
    if (k >= kb) break;
    const double b = cur[k];
    M[(long)r * MB_KB + k] = b;
#pragma unroll
    for (int j = 0; j < MB_KB; ++j) {
      if (j >= kb) break;
      const double c = (j == k) ? 0.0 : cur[j];
      cur[j] = c - b * D[k][j];
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
    const int k = idx % MB_KB, rr = idx / MB_KB;
    const int r = r0 + rr;
    Ms[rr][k] = (r < r_hi && k < kb) ? M[(long)r * MB_KB + k] : 0.0;
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
  for (int p0 = 0; p0 < n; p0 += MB_KB) {
    const int p1 = min(n, p0 + MB_KB), kb = p1 - p0;
    hipLaunchKernelGGL(mb_panel_kernel, dim3(count), dim3(256), 0, st, mats, ms, n, p0, p1, 0,
                       rq, zp);
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
    hipLaunchKernelGGL(mb_panel_kernel, dim3(count), dim3(256), 0, st, mats, ms, n, q0, q1, 1,
                       rq, zp);
  }
  return hipGetLastError();
}

}  // namespace sevp
