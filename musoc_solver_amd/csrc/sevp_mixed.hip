// Mixed-precision influence-matrix corrections on the matrix cores.
//
// D(m, n) = sum_k A(m, k) B(k, n) with A the error rows (one per right-hand
// side) rounded to FP32 by the kernels that produce them, B an FP32 copy of
// RINV / RINV1 / W (column-major), and FP32 accumulation in
// v_mfma_f32_16x16x4f32.  K is split into slices over workgroups; each
// workgroup writes its FP32 slice sums and the kernels that consume the
// correction add the slices in slice order in FP64, so the result does not
// depend on timing and no cross-workgroup reduction is needed here.
//
// Inside one wavefront the k order is permuted: lane group q = lane/16 loads
// k = k0 + 4q .. k0 + 4q + 3 as one run for both operands and feeds them to
// four successive MFMA steps.  A and B use the same permutation, so the
// product is unchanged.
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MX_MAXM = 64;  // right-hand sides per product

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
// Every global load of the workgroup is issued before the first wait: the A
// slice (to LDS), then the B chunks of the slice into registers.
// MB (1 or 2) 16-row accumulator blocks cover right-hand sides m0 ..
// m0 + 16 MB - 1; LDS rows past M are zero.  Requires K % 4 == 0 (float4
// runs of k are aligned).
template <int KS, int MB>
__device__ __forceinline__ void gemm_part_body(const GemmPartArgs &p, int bx, int s, int m0) {
  constexpr int LDA = KS + 4;
  constexpr int KQ = KS / 4;
  constexpr int NCH = KS / 16;
  constexpr int ROWS = 16 * MB;
  constexpr int ALD = (ROWS * KQ + 64 * GP_WAVES - 1) / (64 * GP_WAVES);
  __shared__ float sa[ROWS * LDA];
  span_begin(p.tspan);
  const int k_lo = s * KS;
  const int klen = min(p.K - k_lo, KS);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n0 = (bx * GP_WAVES + wave) * 16;
  const int col = lane & 15, q = lane >> 4;
  const int n = n0 + col;
  const bool nok = n < p.N;
  const int nc = nok ? n0 : ((p.N - 1) / 16) * 16;
  // A slice first (the LDS staging waits only for these), then the B
  // chunks, which the MFMA loop consumes in order as they arrive
  f32x4 at[ALD];
#pragma unroll
  for (int j = 0; j < ALD; ++j) {
    const int idx = threadIdx.x + j * 64 * GP_WAVES;
    const int m = m0 + idx / KQ, k = (idx % KQ) * 4;
    const bool ok = idx < ROWS * KQ && m < p.M;
    at[j] = load4m(p.a + (long)(ok ? m : 0) * p.K + k_lo, k, ok ? klen : 0);
  }
  // B in lane tiles: this wavefront's 16 columns are tile nt of b (or of b2
  // past nsplit); chunk k_lo/16 + i is one 1 KiB block, this lane's float4 at
  // lane position `lane`
  const int nchk = (p.K + 15) / 16;
  const int nt = nc / 16;
  const bool second = p.b2 && nt >= p.nsplit / 16;
  const f32x4 *btile = reinterpret_cast<const f32x4 *>(second ? p.b2 : p.b) +
                       (long)(second ? nt - p.nsplit / 16 : nt) * nchk * 64;
  const int ch0 = k_lo / 16;
  const f32x4 zero = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 bq[NCH];
#pragma unroll
  for (int i = 0; i < NCH; ++i) {
    const bool ok = ch0 + i < nchk;
    const f32x4 v = btile[(long)(ok ? ch0 + i : 0) * 64 + lane];
    bq[i] = ok ? v : zero;
  }
#pragma unroll
  for (int j = 0; j < ALD; ++j) {
    const int idx = threadIdx.x + j * 64 * GP_WAVES;
    if (idx < ROWS * KQ) *reinterpret_cast<f32x4 *>(&sa[(idx / KQ) * LDA + (idx % KQ) * 4]) = at[j];
  }
  __syncthreads();
  if (n0 < p.N) {
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
  float *out = p.part + (long)s * p.M * p.N + n;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + mb * 16 + 4 * q + i;
      if (nok && m < p.M) out[(long)m * p.N] = acc[mb][i];
    }
  }
  span_end(p.tspan);
}

// XCD-aware placement of a product's (column group, slice, row group)
// tiles over a 1-D grid: the dispatcher deals workgroups to the 8 XCDs in
// turn, and the row groups of one (column group, slice) get ids that land on
// the same XCD one after another, so its matrix tile comes from HBM once and
// from that XCD's L2 for the other row groups.  Ids past the last tile only
// record their span.
__device__ __forceinline__ bool gp_tile_of(int wg, int ncs, int rg, int &cs, int &z) {
  const int xcd = wg & 7, slot = wg >> 3;
  cs = (slot / rg) * 8 + xcd;
  z = slot % rg;
  return cs < ncs;
}

template <int KS, int MB>
__global__ __launch_bounds__(64 * GP_WAVES) void gemm_part_kernel(GemmPartArgs p, int groups, int ncs,
                                                                int rg) {
  int cs, z;
  if (gp_tile_of(blockIdx.x, ncs, rg, cs, z)) {
    gemm_part_body<KS, MB>(p, cs % groups, cs / groups, z * 16 * MB);
  } else {
    span_begin(p.tspan);
    span_end(p.tspan);
  }
}

// Slice lengths with a compiled kernel, and the one used for a product: the
// shortest (most workgroups) of at least 96 that leaves at most 16 partial
// products per output, since every consumer reads all of them.  For
// 64 x 1440 x N this is 96 (15 slices; 345 workgroups at N = 1440, 675 at
// N = 2880).  Shorter slices give a narrow product more workgroups but
// little work each: grouped launches over many basins (I2 <= 512) ran 17%
// faster per pass with 96 than with 32, and a single product of that size
// is latency-bound either way.
static const int kGpSlices[] = {32, 48, 64, 96, 128};
constexpr int GP_MAX_SLICES = 16;
constexpr int GP_MIN_SLICE = 96;

int gemm_part_slices(int N, int K, int *kslice) {
  (void)N;
  int ks = kGpSlices[4];
  for (int c : kGpSlices)
    if (c >= GP_MIN_SLICE && (K + c - 1) / c <= GP_MAX_SLICES) {
      ks = c;
      break;
    }
  if (kslice) *kslice = ks;
  return (K + ks - 1) / ks;
}

// Workgroups of the 1-D placement: 8 XCDs x row groups x ceil(tiles / 8).
static long gp_placed(long ncs, int rg) { return 8L * rg * ((ncs + 7) / 8); }

template <int KS, int MB>
static void gp_launch(const GemmPartArgs &a, dim3 grid, hipStream_t st, hipEvent_t ev0,
                      hipEvent_t ev1) {
  const dim3 block(64 * GP_WAVES);
  const int groups = (int)grid.x, ncs = (int)(grid.x * grid.y), rg = (int)grid.z;
  const dim3 g1((unsigned)gp_placed(ncs, rg));
  if (ev0) hipExtLaunchKernelGGL(gemm_part_kernel<KS, MB>, g1, block, 0, st, ev0, ev1, 0, a, groups, ncs, rg);
  else hipLaunchKernelGGL((gemm_part_kernel<KS, MB>), g1, block, 0, st, a, groups, ncs, rg);
}

// Right-hand sides per workgroup: up to 32 (two 16-row MFMA blocks); more
// rows go to further workgroups (grid z), which read the same B tiles while
// they are in the cache and leave each wavefront's MFMA chain as short as
// at 32.  Each output row's MFMA sequence does not depend on the split.
static int gp_row_blocks(int M) { return M <= 16 ? 1 : 2; }
static int gp_row_groups(int M) { return (M + 31) / 32; }

template <int KS>
static void gp_launch_mb(const GemmPartArgs &a, dim3 grid, hipStream_t st, hipEvent_t ev0,
                         hipEvent_t ev1) {
  if (gp_row_blocks(a.M) == 1) gp_launch<KS, 1>(a, grid, st, ev0, ev1);
  else gp_launch<KS, 2>(a, grid, st, ev0, ev1);
}

// Workgroups of one entry of a grouped launch (batch_gemm_part_kernel's own
// dense numbering, no placement).
long batch_gemm_part_workgroups(const GemmPartArgs &a) {
  return (long)((a.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES)) * gemm_part_slices(a.N, a.K, nullptr) *
         gp_row_groups(a.M);
}
long batch_gemm_part64_workgroups(const GemmPart64Args &a) {
  return (long)((a.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES)) * gemm_part_slices(a.N, a.K, nullptr) *
         gp_row_groups(a.M);
}

long gemm_part_workgroups(const GemmPartArgs &a) {
  return gp_placed((long)((a.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES)) * gemm_part_slices(a.N, a.K, nullptr),
                   gp_row_groups(a.M));
}

hipError_t launch_gemm_part(const GemmPartArgs &a, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  if (a.M > MX_MAXM || a.M < 1 || a.N < 1 || a.K < 1 || a.K % 4 != 0) return hipErrorInvalidValue;
  int ks = 0;
  const int ns = gemm_part_slices(a.N, a.K, &ks);
  const int groups = (a.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES);
  const dim3 grid(groups, ns, gp_row_groups(a.M));
  switch (ks) {
    case 32: gp_launch_mb<32>(a, grid, st, ev0, ev1); break;
    case 48: gp_launch_mb<48>(a, grid, st, ev0, ev1); break;
    case 64: gp_launch_mb<64>(a, grid, st, ev0, ev1); break;
    case 96: gp_launch_mb<96>(a, grid, st, ev0, ev1); break;
    default: gp_launch_mb<128>(a, grid, st, ev0, ev1); break;
  }
  return hipGetLastError();
}

// ------------------------------------------- FP64 split-K partial products
//
// The same decomposition on the FP64 matrix cores (v_mfma_f64_16x16x4f64,
// whose A and B lane layout is the FP32 instruction's; D differs, see the
// store below): A (the FP64 error
// rows) staged in LDS with a row pitch of KS+2 doubles (an odd number of
// 16-byte units), B from FP64 lane tiles (a 2 KiB block per chunk of 16 k,
// 32 bytes per lane), FP64 slice sums.  The products are exact FP64
// arithmetic in a blocked order: FP64 accuracy, not the reference's order.
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 load4d(const double *base, int k, int klen) {
  const bool ok = k < klen;
  const f64x4 v = *reinterpret_cast<const f64x4 *>(base + (ok ? k : 0));
  return ok ? v : f64x4{0.0, 0.0, 0.0, 0.0};
}

template <int KS, int MB>
__device__ __forceinline__ void gemm_part64_body(const GemmPart64Args &p, int bx, int s, int m0) {
  constexpr int LDA = KS + 2;
  constexpr int KQ = KS / 4;
  constexpr int NCH = KS / 16;
  constexpr int ROWS = 16 * MB;
  constexpr int ALD = (ROWS * KQ + 64 * GP_WAVES - 1) / (64 * GP_WAVES);
  __shared__ double sa[ROWS * LDA];
  span_begin(p.tspan);
  const int k_lo = s * KS;
  const int klen = min(p.K - k_lo, KS);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n0 = (bx * GP_WAVES + wave) * 16;
  const int col = lane & 15, q = lane >> 4;
  const int n = n0 + col;
  const bool nok = n < p.N;
  const int nc = nok ? n0 : ((p.N - 1) / 16) * 16;
  f64x4 at[ALD];
#pragma unroll
  for (int j = 0; j < ALD; ++j) {
    const int idx = threadIdx.x + j * 64 * GP_WAVES;
    const int m = m0 + idx / KQ, k = (idx % KQ) * 4;
    const bool ok = idx < ROWS * KQ && m < p.M;
    at[j] = load4d(p.a + (long)(ok ? m : 0) * p.K + k_lo, k, ok ? klen : 0);
  }
  const int nchk = (p.K + 15) / 16;
  const int nt = nc / 16;
  const bool second = p.b2 && nt >= p.nsplit / 16;
  const f64x4 *btile = reinterpret_cast<const f64x4 *>(second ? p.b2 : p.b) +
                       (long)(second ? nt - p.nsplit / 16 : nt) * nchk * 64;
  const int ch0 = k_lo / 16;
  const f64x4 zero = f64x4{0.0, 0.0, 0.0, 0.0};
  f64x4 bq[NCH];
#pragma unroll
  for (int i = 0; i < NCH; ++i) {
    const bool ok = ch0 + i < nchk;
    const f64x4 v = btile[(long)(ok ? ch0 + i : 0) * 64 + lane];
    bq[i] = ok ? v : zero;
  }
#pragma unroll
  for (int j = 0; j < ALD; ++j) {
    const int idx = threadIdx.x + j * 64 * GP_WAVES;
    if (idx < ROWS * KQ) *reinterpret_cast<f64x4 *>(&sa[(idx / KQ) * LDA + (idx % KQ) * 4]) = at[j];
  }
  __syncthreads();
  if (n0 < p.N) {
    f64x4 acc[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) acc[mb] = zero;
    const double *arow = sa + col * LDA + 4 * q;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      f64x4 av[MB];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) av[mb] = *reinterpret_cast<const f64x4 *>(arow + mb * 16 * LDA + 16 * ch);
      const f64x4 bv = nok ? bq[ch] : zero;
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          acc[mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mb][t], bv[t], acc[mb], 0, 0, 0);
    }
    // the FP64 instruction's D layout: lane group q holds rows q, q+4, q+8,
    // q+12 of its column (the FP32 one holds 4q .. 4q+3)
    double *out = p.part + (long)s * p.M * p.N + n;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = m0 + mb * 16 + q + 4 * i;
        if (nok && m < p.M) out[(long)m * p.N] = acc[mb][i];
      }
  }
  span_end(p.tspan);
}

template <int KS, int MB>
__global__ __launch_bounds__(64 * GP_WAVES) void gemm_part64_kernel(GemmPart64Args p, int groups,
                                                                  int ncs, int rg) {
  int cs, z;
  if (gp_tile_of(blockIdx.x, ncs, rg, cs, z)) {
    gemm_part64_body<KS, MB>(p, cs % groups, cs / groups, z * 16 * MB);
  } else {
    span_begin(p.tspan);
    span_end(p.tspan);
  }
}

template <int KS, int MB>
static void gp64_launch(const GemmPart64Args &a, dim3 grid, hipStream_t st, hipEvent_t ev0,
                        hipEvent_t ev1) {
  const dim3 block(64 * GP_WAVES);
  const int groups = (int)grid.x, ncs = (int)(grid.x * grid.y), rg = (int)grid.z;
  const dim3 g1((unsigned)gp_placed(ncs, rg));
  if (ev0) hipExtLaunchKernelGGL(gemm_part64_kernel<KS, MB>, g1, block, 0, st, ev0, ev1, 0, a, groups, ncs, rg);
  else hipLaunchKernelGGL((gemm_part64_kernel<KS, MB>), g1, block, 0, st, a, groups, ncs, rg);
}

template <int KS>
static void gp64_launch_mb(const GemmPart64Args &a, dim3 grid, hipStream_t st, hipEvent_t ev0,
                           hipEvent_t ev1) {
  if (gp_row_blocks(a.M) == 1) gp64_launch<KS, 1>(a, grid, st, ev0, ev1);
  else gp64_launch<KS, 2>(a, grid, st, ev0, ev1);
}

long gemm_part64_workgroups(const GemmPart64Args &a) {
  return gp_placed((long)((a.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES)) * gemm_part_slices(a.N, a.K, nullptr),
                   gp_row_groups(a.M));
}

hipError_t launch_gemm_part64(const GemmPart64Args &a, hipStream_t st, hipEvent_t ev0,
                              hipEvent_t ev1) {
  if (a.M > MX_MAXM || a.M < 1 || a.N < 1 || a.K < 1 || a.K % 4 != 0) return hipErrorInvalidValue;
  int ks = 0;
  const int ns = gemm_part_slices(a.N, a.K, &ks);
  const int groups = (a.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES);
  const dim3 grid(groups, ns, gp_row_groups(a.M));
  switch (ks) {
    case 32: gp64_launch_mb<32>(a, grid, st, ev0, ev1); break;
    case 48: gp64_launch_mb<48>(a, grid, st, ev0, ev1); break;
    case 64: gp64_launch_mb<64>(a, grid, st, ev0, ev1); break;
    case 96: gp64_launch_mb<96>(a, grid, st, ev0, ev1); break;
    default: gp64_launch_mb<128>(a, grid, st, ev0, ev1); break;
  }
  return hipGetLastError();
}

// FP32 operand layout of the products ("lane tiles"): an (K x N) column-major
// matrix becomes tiles of 16 columns; tile t holds ceil(K/16) chunks of 16 k,
// each a 1 KiB block in the order the 64 lanes of a wavefront load it, lane
// (q, c) = (l / 16, l % 16) holding k = 16 ch + 4q .. 4q + 3 of column
// 16 t + c as one float4.  One chunk is one coalesced 1 KiB load, and the
// chunks of a tile are consecutive; entries past K or N are zero.
template <class T>
__global__ void cvt_lane_tiles_kernel(const double *src, T *dst, int K, int N, long nmat,
                                      long src_stride, long dst_stride) {
  const int nchk = (K + 15) / 16, ntile = (N + 15) / 16;
  const long per = (long)ntile * nchk * 256;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * nmat) return;
  const long mat = i / per, r = i % per;
  const int t = (int)(r & 3), l = (int)((r >> 2) & 63);
  const long blk = r >> 8;
  const int ch = (int)(blk % nchk), tile = (int)(blk / nchk);
  const int k = 16 * ch + 4 * (l >> 4) + t, n = 16 * tile + (l & 15);
  dst[mat * dst_stride + r] =
      (k < K && n < N) ? (T)src[mat * src_stride + (long)k + (long)n * K] : (T)0;
}

long lane_tiles_floats(int K, int N) { return (long)((N + 15) / 16) * ((K + 15) / 16) * 256; }

hipError_t launch_cvt_lane_tiles(const double *src, float *dst, int K, int N, long nmat,
                                 hipStream_t st) {
  const long per = lane_tiles_floats(K, N), n = per * nmat;
  hipLaunchKernelGGL(cvt_lane_tiles_kernel<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     st, src, dst, K, N, nmat, (long)K * N, per);
  return hipGetLastError();
}

hipError_t launch_cvt_lane_tiles64(const double *src, double *dst, int K, int N, long nmat,
                                   hipStream_t st) {
  const long per = lane_tiles_floats(K, N), n = per * nmat;
  hipLaunchKernelGGL(cvt_lane_tiles_kernel<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     st, src, dst, K, N, nmat, (long)K * N, per);
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

// Grouped products (one launch for the same block step of many problems):
// entry e is one problem's product with gemm_part_workgroups(e) workgroups,
// n groups fastest, then slices, then 32-row groups: exactly the workgroups
// of its own launch.  Every entry of a launch shares the slice length KS (a
// problem's own choice, so each output is summed as in its own launch); MB
// is 2 when any entry has more than 16 rows (rows past an entry's M are zero
// and never stored, and each row's MFMA sequence does not depend on MB).
template <int KS, int MB>
__global__ __launch_bounds__(64 * GP_WAVES) void batch_gemm_part_kernel(const GemmPartArgs *tab,
                                                                      const int *start, int n) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const GemmPartArgs &p = tab[e];
  const int groups = (p.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES);
  const int gs = groups * ((p.K + KS - 1) / KS);
  const int l = b - start[e];
  gemm_part_body<KS, MB>(p, l % groups, (l % gs) / groups, (l / gs) * 32);
}

template <int KS, int MB>
__global__ __launch_bounds__(64 * GP_WAVES) void batch_gemm_part64_kernel(const GemmPart64Args *tab,
                                                                        const int *start, int n) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const GemmPart64Args &p = tab[e];
  const int groups = (p.N + 16 * GP_WAVES - 1) / (16 * GP_WAVES);
  const int gs = groups * ((p.K + KS - 1) / KS);
  const int l = b - start[e];
  gemm_part64_body<KS, MB>(p, l % groups, (l % gs) / groups, (l / gs) * 32);
}

template <class T, int KS, int MB>
static void bgp_launch(const T *tab, const int *start, int n, long nwg, hipStream_t st) {
  const dim3 grid((unsigned)nwg), block(64 * GP_WAVES);
  if constexpr (std::is_same<T, GemmPartArgs>::value)
    hipLaunchKernelGGL((batch_gemm_part_kernel<KS, MB>), grid, block, 0, st, tab, start, n);
  else
    hipLaunchKernelGGL((batch_gemm_part64_kernel<KS, MB>), grid, block, 0, st, tab, start, n);
}

template <class T, int KS>
static void bgp_launch_mb(int mb, const T *tab, const int *start, int n, long nwg, hipStream_t st) {
  if (mb == 1) bgp_launch<T, KS, 1>(tab, start, n, nwg, st);
  else bgp_launch<T, KS, 2>(tab, start, n, nwg, st);
}

template <class T>
static hipError_t bgp_dispatch(int ks, int mb, const T *tab, const int *start, int n, long nwg,
                               hipStream_t st) {
  if (n < 1 || nwg < 1 || nwg > 0x7fffffffL || (mb != 1 && mb != 2))
    return hipErrorInvalidValue;
  switch (ks) {
    case 32: bgp_launch_mb<T, 32>(mb, tab, start, n, nwg, st); break;
    case 48: bgp_launch_mb<T, 48>(mb, tab, start, n, nwg, st); break;
    case 64: bgp_launch_mb<T, 64>(mb, tab, start, n, nwg, st); break;
    case 96: bgp_launch_mb<T, 96>(mb, tab, start, n, nwg, st); break;
    case 128: bgp_launch_mb<T, 128>(mb, tab, start, n, nwg, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_gemm_part(int ks, int mb, const GemmPartArgs *tab, const int *start, int n,
                                  long nwg, hipStream_t st) {
  return bgp_dispatch(ks, mb, tab, start, n, nwg, st);
}

hipError_t launch_batch_gemm_part64(int ks, int mb, const GemmPart64Args *tab, const int *start,
                                    int n, long nwg, hipStream_t st) {
  return bgp_dispatch(ks, mb, tab, start, n, nwg, st);
}

}  // namespace sevp
