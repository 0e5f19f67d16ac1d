// Mixed-precision influence-matrix corrections on the matrix cores.
//
// D(m, n) = sum_k A(m, k) B(k, n) with A the FP64 error vectors (one row per
// right-hand side) rounded to FP32 as they are loaded, B an FP32 copy of
// RINV / RINV1 (column-major), and FP32 accumulation in
// v_mfma_f32_16x16x4f32.  A wavefront owns 16 output columns and every
// right-hand side (up to 64 rows = four 16x16 accumulator blocks), so each
// matrix element is read once per product.  K is split over `splits`
// workgroups; the last workgroup to finish a column tile (an atomic counter
// per tile) adds the partial sums in split order and applies the same fused
// epilogue as the FP64 product, so the result does not depend on timing.
//
// Inside one wavefront the k order is permuted: lane group q = lane/16 loads
// k = k0 + 4q .. k0 + 4q + 3 as one run for both operands and feeds them to
// four successive MFMA steps.  A and B use the same permutation, so the
// product is unchanged.
#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MX_MAXM = 64;  // right-hand sides per product
constexpr int MX_WAVES = 4;  // column tiles per workgroup

__global__ __launch_bounds__(256) void gemm_f32_mfma_kernel(GemmF32Args p, const double *a64,
                                                            int splits, int kchunk, float *partial,
                                                            int *counters) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ntile = blockIdx.x * MX_WAVES + wave;
  const int s = blockIdx.y;
  const int n0 = ntile * 16;
  if (n0 >= p.N) return;
  const int K = p.K, M = p.M, N = p.N;
  const int col = lane & 15, q = lane >> 4;
  const int n = n0 + col;
  const int k_lo = s * kchunk, k_hi = min(K, k_lo + kchunk);
  f32x4 acc[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) acc[mb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const float *bcol = p.b + (long)min(n, N - 1) * K;
  for (int k0 = k_lo; k0 < k_hi; k0 += 16) {
    const int kb = k0 + 4 * q;
    float bv[4], av[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = kb + t;
      bv[t] = (k < k_hi && n < N) ? bcol[k] : 0.f;
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const int m = mb * 16 + col;
      const double *arow = a64 + (long)m * K;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = kb + t;
        av[mb][t] = (m < M && k < k_hi) ? (float)arow[k] : 0.f;
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mb][t], bv[t], acc[mb], 0, 0, 0);
    }
  }
  // D block (mb): lane holds rows mb*16 + 4q + i, column col
  float *part = partial + (long)s * M * N;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = mb * 16 + 4 * q + i;
      if (m < M && n < N) part[(long)m * N + n] = acc[mb][i];
    }
  if (splits > 1) {
    __threadfence();
    int prev = 0;
    if (lane == 0) prev = atomicAdd(&counters[ntile], 1);
    prev = __shfl(prev, 0);
    if (prev != splits - 1) return;
    __threadfence();
    if (lane == 0) counters[ntile] = 0;
  }
  const int i0 = p.i2 + 2;
  for (int idx = lane; idx < 16 * M; idx += 64) {
    const int m = idx / 16, nn = n0 + idx % 16;
    if (nn >= N) continue;
    float sum = 0.f;
    for (int ss = 0; ss < splits; ++ss) sum += partial[(long)ss * M * N + (long)m * N + nn];
    const double d = (double)sum;
    switch (p.ep) {
      case EP_FWD: {
        double *x = p.x + (long)m * p.x_stride;
        const long xi = f2(i0, nn + 2, p.xrow);
        const double xv = x[xi];
        p.dum0[(long)m * p.dum0_stride + nn] = xv;
        x[xi] = xv - d;
        break;
      }
      case EP_BWD: {
        double *x = p.x + (long)m * p.x_stride;
        const long xi = f2(i0, nn + 2, p.xrow);
        x[xi] = x[xi] + d;
        p.out[(long)m * p.out_ld + nn] = d;
        if (p.g) p.g[(long)m * p.out_ld + nn] = d * (double)p.cy[(long)nn + (long)(p.cyrow - 1) * p.i2];
        break;
      }
      default:
        p.out[(long)m * p.out_ld + nn] = d;
        break;
    }
  }
}

__global__ void cvt_f32_kernel(const double *src, float *dst, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (float)src[t];
}

hipError_t launch_cvt_f32(const double *src, float *dst, long n, hipStream_t st) {
  hipLaunchKernelGGL(cvt_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

// partial: splits * M * N floats; counters: ceil(N/16) ints, zero on entry
// and left zero on exit.
hipError_t launch_gemm_mixed(const GemmF32Args &a, const double *a64, float *partial,
                             int *counters, int splits, hipStream_t st) {
  if (a.M > MX_MAXM || a.M < 1) return hipErrorInvalidValue;
  int kchunk = (a.K + splits - 1) / splits;
  kchunk = (kchunk + 15) / 16 * 16;
  const int ntiles = (a.N + 15) / 16;
  dim3 grid((ntiles + MX_WAVES - 1) / MX_WAVES, splits);
  hipLaunchKernelGGL(gemm_f32_mfma_kernel, grid, dim3(64 * MX_WAVES), 0, st, a, a64, splits,
                     kchunk, partial, counters);
  return hipGetLastError();
}

}  // namespace sevp
