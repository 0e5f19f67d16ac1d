// Mixed-precision influence-matrix corrections on the matrix cores.
//
// D(m, n) = sum_k A(m, k) B(k, n) with A the error vectors (one row per
// right-hand side) rounded to FP32 by the kernels that produce them, B an
// FP32 copy of RINV / RINV1 (column-major), and FP32 accumulation in
// v_mfma_f32_16x16x4f32.  Operands are read as float4 runs of k and the next
// step's operands are loaded while the current step's MFMAs run.  A wavefront owns 16 output columns and every
// right-hand side (up to 64 rows = four 16x16 accumulator blocks), so each
// matrix element is read once per product.  K is split over the wavefronts of
// the tile's workgroup, whose partial tiles are added through LDS in a fixed
// order before the same fused epilogue as the FP64 product, so the result
// does not depend on timing.
//
// Inside one wavefront the k order is permuted: lane group q = lane/16 loads
// k = k0 + 4q .. k0 + 4q + 3 as one run for both operands and feeds them to
// four successive MFMA steps.  A and B use the same permutation, so the
// product is unchanged.
#include <algorithm>

#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MX_MAXM = 64;  // right-hand sides per product

// Four consecutive k of one row (or column) starting at k, zero past k_hi.
__device__ __forceinline__ f32x4 load4(const float *base, int k, int k_hi, bool vec) {
  if (vec && k + 3 < k_hi) return *reinterpret_cast<const f32x4 *>(base + k);
  f32x4 v;
  v.x = (k < k_hi) ? base[k] : 0.f;
  v.y = (k + 1 < k_hi) ? base[k + 1] : 0.f;
  v.z = (k + 2 < k_hi) ? base[k + 2] : 0.f;
  v.w = (k + 3 < k_hi) ? base[k + 3] : 0.f;
  return v;
}

// One workgroup per (16-column tile, MBW x 16 rows); its NW wavefronts take
// consecutive K slices (16-wide MFMA steps), then sum their partial tiles
// through LDS in wavefront order and apply the epilogue.  The workgroups that
// share a column tile (and so the same RINV columns) get block indices equal
// modulo 8, which places them on the same XCD and its L2.
template <int NW, int MBW, int STEPS>
__global__ __launch_bounds__(64 * NW) void gemm_f32_mfma_kernel(GemmF32Args p, int kq, int mgroups) {
  __shared__ float red[NW][16 * MBW][17];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wg = blockIdx.x;
  const int ntile = (wg / (8 * mgroups)) * 8 + (wg % 8);
  const int mgroup = (wg / 8) % mgroups;
  const int n0 = ntile * 16;
  const int mbase = mgroup * 16 * MBW;
  const int K = p.K, M = p.M, N = p.N;
  if (n0 >= N || mbase >= M) return;
  const int col = lane & 15, q = lane >> 4;
  const int n = n0 + col;
  const int k_lo = wave * kq, k_hi = min(K, k_lo + kq);
  const bool vec = (K % 4) == 0;
  f32x4 acc[MBW];
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb) acc[mb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nc = min(n, N - 1);
  const float *bcol = (p.b2 && nc >= p.nsplit) ? p.b2 + (long)(nc - p.nsplit) * K : p.b + (long)nc * K;
  const float *arow[MBW];
  bool mok[MBW];
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb) {
    const int m = mbase + mb * 16 + col;
    mok[mb] = m < M;
    arow[mb] = p.a + (long)min(m, M - 1) * K;
  }
  auto fetch = [&](int k0, f32x4 &b_, f32x4 (&a_)[MBW]) {
    const int kb = k0 + 4 * q;
    b_ = (n < N) ? load4(bcol, kb, k_hi, vec) : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb)
      a_[mb] = mok[mb] ? load4(arow[mb], kb, k_hi, vec) : f32x4{0.f, 0.f, 0.f, 0.f};
  };
  if (STEPS > 0) {
    // the whole K slice (STEPS x 16) is loaded before the first MFMA, so the
    // wavefront waits for memory once
    f32x4 bq[STEPS > 0 ? STEPS : 1], aq[STEPS > 0 ? STEPS : 1][MBW];
#pragma unroll
    for (int st = 0; st < STEPS; ++st) fetch(k_lo + 16 * st, bq[st], aq[st]);
#pragma unroll
    for (int st = 0; st < STEPS; ++st)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mb = 0; mb < MBW; ++mb)
          acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aq[st][mb][t], bq[st][t], acc[mb], 0, 0, 0);
  } else {
    f32x4 bq, aq[MBW];
    if (k_lo < k_hi) fetch(k_lo, bq, aq);
    for (int k0 = k_lo; k0 < k_hi; k0 += 16) {
      f32x4 bn = bq, an[MBW];
#pragma unroll
      for (int mb = 0; mb < MBW; ++mb) an[mb] = aq[mb];
      if (k0 + 16 < k_hi) fetch(k0 + 16, bn, an);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mb = 0; mb < MBW; ++mb)
          acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aq[mb][t], bq[t], acc[mb], 0, 0, 0);
      bq = bn;
#pragma unroll
      for (int mb = 0; mb < MBW; ++mb) aq[mb] = an[mb];
    }
  }
  // D block mb: lane holds rows mb*16 + 4q + i, column col
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][mb * 16 + 4 * q + i][col] = acc[mb][i];
  __syncthreads();
  const int i0 = p.i2 + 2;
  for (int idx = threadIdx.x; idx < 16 * 16 * MBW; idx += 64 * NW) {
    const int ml = idx / 16, c = idx % 16, nn = n0 + c, m = mbase + ml;
    if (nn >= N || m >= M) continue;
    float sum = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) sum += red[w][ml][c];
    const double d = (double)sum;
    if (p.b2 && nn >= p.nsplit) {
      p.out2[(long)m * p.nsplit + (nn - p.nsplit)] = d;
      continue;
    }
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
        if (p.g) {
          const double gv = d * (double)p.cy[(long)nn + (long)(p.cyrow - 1) * p.i2];
          p.g[(long)m * p.out_ld + nn] = gv;
          p.g32[(long)m * p.out_ld + nn] = (float)gv;
        }
        break;
      }
      default:
        p.out[(long)m * p.out_ld + nn] = d;
        break;
    }
  }
}

// ------------------------------------------------ split-K partial products
//
// Workgroup (g, s): K slice s = [s*KS, min(K, (s+1)*KS)) and four
// 16-column tiles, one per wavefront, each over every right-hand side (up to
// four 16-row accumulator blocks).  The A slice is staged once in LDS with a
// row pitch of KS+4 floats (an odd number of 16-byte units, so the 16
// rows one ds_read_b128 group touches fall in distinct banks) and shared by
// the four wavefronts; B columns come from HBM as float4 runs of k, the whole
// slice held in registers.  The FP32 slice sums go to P[s][m][n] and the
// consumers add them in slice order in FP64.
constexpr int GP_WAVES = 4;

// Four consecutive k at k (a multiple of 4) of a run of klen (a multiple of
// 4) floats, zero at or past klen: one unconditional float4 load (from k = 0
// when out of range) and a select, so the load never branches.
__device__ __forceinline__ f32x4 load4m(const float *base, int k, int klen) {
  const bool ok = k < klen;
  const f32x4 v = *reinterpret_cast<const f32x4 *>(base + (ok ? k : 0));
  return ok ? v : f32x4{0.f, 0.f, 0.f, 0.f};
}

// KS (the slice length, a multiple of 16) is a compile-time constant, so the
// loads, the LDS staging and the 4 x KS/4 MFMA steps are straight-line code.
// Every global load of the workgroup is issued before the first wait: the B
// columns of the slice into registers, then the A slice, which goes to LDS.
// MB (1, 2 or 4) 16-row accumulator blocks cover the right-hand sides; LDS
// rows past M are zero.  Requires K % 4 == 0 (float4 runs of k are aligned).
template <int KS, int MB>
__global__ __launch_bounds__(64 * GP_WAVES) void gemm_part_kernel(GemmPartArgs p) {
  constexpr int LDA = KS + 4;
  constexpr int KQ = KS / 4;
  constexpr int NCH = KS / 16;
  constexpr int ROWS = 16 * MB;
  constexpr int ALD = (ROWS * KQ + 64 * GP_WAVES - 1) / (64 * GP_WAVES);
  __shared__ float sa[ROWS * LDA];
  const int s = blockIdx.y;
  const int k_lo = s * KS;
  const int klen = min(p.K - k_lo, KS);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n0 = (blockIdx.x * GP_WAVES + wave) * 16;
  const int col = lane & 15, q = lane >> 4;
  const int n = n0 + col;
  const bool nok = n < p.N;
  const int nc = nok ? n : p.N - 1;
  const float *bcol =
      ((p.b2 && nc >= p.nsplit) ? p.b2 + (long)(nc - p.nsplit) * p.K : p.b + (long)nc * p.K) + k_lo;
  const f32x4 zero = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 bq[NCH];
#pragma unroll
  for (int i = 0; i < NCH; ++i) bq[i] = load4m(bcol, 16 * i + 4 * q, klen);
  f32x4 at[ALD];
#pragma unroll
  for (int j = 0; j < ALD; ++j) {
    const int idx = threadIdx.x + j * 64 * GP_WAVES;
    const int m = idx / KQ, k = (idx % KQ) * 4;
    const bool ok = idx < ROWS * KQ && m < p.M;
    at[j] = load4m(p.a + (long)(ok ? m : 0) * p.K + k_lo, k, ok ? klen : 0);
  }
#pragma unroll
  for (int j = 0; j < ALD; ++j) {
    const int idx = threadIdx.x + j * 64 * GP_WAVES;
    if (idx < ROWS * KQ) *reinterpret_cast<f32x4 *>(&sa[(idx / KQ) * LDA + (idx % KQ) * 4]) = at[j];
  }
  __syncthreads();
  if (n0 >= p.N) return;
  f32x4 acc[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) acc[mb] = zero;
  const float *arow = sa + col * LDA + 4 * q;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    f32x4 av[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) av[mb] = *reinterpret_cast<const f32x4 *>(arow + mb * 16 * LDA + 16 * ch);
    const f32x4 bv = nok ? bq[ch] : zero;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mb][t], bv[t], acc[mb], 0, 0, 0);
  }
  if (!nok) return;
  float *out = p.part + (long)s * p.M * p.N + n;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = mb * 16 + 4 * q + i;
      if (m < p.M) out[(long)m * p.N] = acc[mb][i];
    }
}

// Slice lengths with a compiled kernel, and the one used for a product: the
// shortest (most workgroups) that leaves at most 16 partial products per
// output, since every consumer reads all of them.  For 64 x 1440 x N this is
// 96 (15 slices; 345 workgroups at N = 1440, 675 at N = 2880).
static const int kGpSlices[] = {32, 48, 64, 96, 128};
constexpr int GP_MAX_SLICES = 16;

int gemm_part_slices(int N, int K, int *kslice) {
  (void)N;
  int ks = kGpSlices[4];
  for (int c : kGpSlices)
    if ((K + c - 1) / c <= GP_MAX_SLICES) {
      ks = c;
      break;
    }
  if (kslice) *kslice = ks;
  return (K + ks - 1) / ks;
}

hipError_t launch_gemm_part(const GemmPartArgs &a, hipStream_t st) {
  if (a.M > MX_MAXM || a.M < 1 || a.N < 1 || a.K < 1 || a.K % 4 != 0) return hipErrorInvalidValue;
  int ks = 0;
  const int ns = gemm_part_slices(a.N, a.K, &ks);
  const int groups = (a.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES);
  dim3 grid(groups, ns), block(64 * GP_WAVES);
  const int mb = (a.M + 15) / 16;
#define GP_LAUNCH(KS_)                                                                             \
  do {                                                                                             \
    if (mb == 1) hipLaunchKernelGGL((gemm_part_kernel<KS_, 1>), grid, block, 0, st, a);             \
    else if (mb == 2) hipLaunchKernelGGL((gemm_part_kernel<KS_, 2>), grid, block, 0, st, a);        \
    else hipLaunchKernelGGL((gemm_part_kernel<KS_, 4>), grid, block, 0, st, a);                     \
  } while (0)
  switch (ks) {
    case 32: GP_LAUNCH(32); break;
    case 48: GP_LAUNCH(48); break;
    case 64: GP_LAUNCH(64); break;
    case 96: GP_LAUNCH(96); break;
    default: GP_LAUNCH(128); break;
  }
#undef GP_LAUNCH
  return hipGetLastError();
}

__global__ void cvt_f32_kernel(const double *src, float *dst, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (float)src[t];
}

hipError_t launch_cvt_f32(const double *src, float *dst, long n, hipStream_t st) {
  hipLaunchKernelGGL(cvt_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

// K is split over the wavefronts of one workgroup; the right-hand sides are
// split over workgroups in 32-row groups so that a 64 x 1440 product spreads
// over ~180 workgroups.
hipError_t launch_gemm_mixed(const GemmF32Args &a, hipStream_t st) {
  if (a.M > MX_MAXM || a.M < 1) return hipErrorInvalidValue;
  constexpr int MBW = 2;
  const int mgroups = (a.M + 16 * MBW - 1) / (16 * MBW);
  const int ntiles = (a.N + 15) / 16;
  const int ntiles8 = (ntiles + 7) / 8 * 8;  // whole groups of 8 for the XCD mapping
  dim3 grid(ntiles8 * mgroups);
  if (a.K >= 1024) {
    const int kq = ((a.K + 15) / 16 + 15) / 16 * 16;
    const int steps = kq / 16;
    if (steps <= 4)
      hipLaunchKernelGGL((gemm_f32_mfma_kernel<16, MBW, 4>), grid, dim3(64 * 16), 0, st, a, kq, mgroups);
    else if (steps <= 6)
      hipLaunchKernelGGL((gemm_f32_mfma_kernel<16, MBW, 6>), grid, dim3(64 * 16), 0, st, a, kq, mgroups);
    else if (steps <= 8)
      hipLaunchKernelGGL((gemm_f32_mfma_kernel<16, MBW, 8>), grid, dim3(64 * 16), 0, st, a, kq, mgroups);
    else
      hipLaunchKernelGGL((gemm_f32_mfma_kernel<16, MBW, 0>), grid, dim3(64 * 16), 0, st, a, kq, mgroups);
  } else {
    const int kq = ((a.K + 3) / 4 + 15) / 16 * 16;
    hipLaunchKernelGGL((gemm_f32_mfma_kernel<4, MBW, 0>), grid, dim3(64 * 4), 0, st, a, kq, mgroups);
  }
  return hipGetLastError();
}

}  // namespace sevp
