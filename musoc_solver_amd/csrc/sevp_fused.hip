// One launch per block step: the correction product of a full-width SEVP
// pass fused with what depends on it, for the MFMA corrections (FP32 or BF16
// with FP64 refinement, or FP64).
//
// The split-K pass (sevp_mixed.hip) takes four launches per block and writes
// every product as slice partials that the next launch adds up.  Here a
// workgroup computes whole dot products: its W wavefronts split K, each keeps
// one or two 16 x 16 accumulators of the MFMA instruction (two 16-row tiles
// share the matrix operand), and the W partial tiles are added in FP64 in
// wavefront order through LDS, so every output is one fixed sum.  The
// epilogue then does the work that depends only on that output's column:
//
//   forward, block b:  d = e_b RINV1_b.  Block b+1's first march row uses the
//     corrected south row X(IE(b)) - d (never written back: the backward pass
//     restores the uncorrected value anyway, REP1's DUM0), its closure
//     residual e_(b+1) needs the marched rows at the neighbouring columns, so
//     a tile computes 16 columns of d and marches them with DPP row shifts
//     (16 lanes = 16 columns of one right-hand side).  The first row
//     operation takes the free row's neighbours from memory and is exact in
//     every lane; each later one (march rows, then the closure) loses a lane
//     per side, so a tile writes the 16 - 2*steps columns that stay exact
//     (the trapezoid of the march kernels; recomputed columns repeat the
//     same operations on the same operands, so they agree bitwise with the
//     tile that owns them).
//   backward, block b: [c | cpl] = e_b [RINV_b | W_b].  X(JS+1) += c and the
//     closure residual of block b-1, which reads that row only at its own
//     column, are done at once; c and cpl are stored for the homogeneous
//     march of block b, which needs c at the neighbouring columns and runs in
//     the next launch (its workgroups come first), when every column is
//     complete.
//
// A pass is 2*NB0 + 1 launches instead of 4*NB0 + 2, each influence matrix is
// read from HBM once for all right-hand sides, and no partial products go
// through memory.  Both operands are in lane tiles, so a wavefront's operand
// for 16 rows (or columns) and 16 k is one contiguous 1 KiB (FP32) run; every
// load address is clamped in bounds instead of masked, so a wavefront issues
// all its loads of a K range before its first wait.  The workgroups of one
// column tile are placed on the same XCD so the tile's matrix columns are
// fetched from HBM once.
#include <hip/hip_ext.h>

#include <type_traits>

#include "sevp_dev.h"
#include "sevp_internal.h"
#include "sevp_march.h"

namespace sevp {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

// Lane i of each 16-lane row receives lane i-1's (row_shr:1) or lane i+1's
// (row_shl:1) value; the row's edge lane keeps its own.  The edge lanes are
// trapezoid halo columns whose results are never stored.
__device__ __forceinline__ double row_shr1(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x111, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x111, 0xf, 0xf, false));
}
__device__ __forceinline__ double row_shl1(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x101, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x101, 0xf, 0xf, false));
}

// BF16 bits of a float, rounded to nearest even (NaN stays a quiet NaN).
__device__ __forceinline__ short bf16_rne(float f) {
  const unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40);
  return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// The three product arithmetics.  Per chunk of 16 k a lane holds k = 4q ..
// 4q+3 (q = lane / 16) of its A row and B column: FP32 and FP64 take four
// 16x16x4 MFMA steps (one k each; A and B use the same permutation of k, so
// the product is unchanged); BF16 rounds the FP32 error row to BF16 in
// registers and takes one 16x16x16 step with FP32 accumulation.
struct KF32 {
  typedef float A;
  typedef float B;
  typedef float Acc;
  typedef f32x4 AV;
  typedef f32x4 BV;
  typedef f32x4 CV;
  __device__ static int row(int q, int i) { return 4 * q + i; }
  __device__ static CV mma(AV a, BV b, CV c) {
#pragma unroll
    for (int t = 0; t < 4; ++t) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], b[t], c, 0, 0, 0);
    return c;
  }
};
struct KF64 {
  typedef double A;
  typedef double B;
  typedef double Acc;
  typedef f64x4 AV;
  typedef f64x4 BV;
  typedef f64x4 CV;
  __device__ static int row(int q, int i) { return q + 4 * i; }  // the FP64 instruction's D layout
  __device__ static CV mma(AV a, BV b, CV c) {
#pragma unroll
    for (int t = 0; t < 4; ++t) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[t], c, 0, 0, 0);
    return c;
  }
};
struct KB16 {
  typedef float A;            // the error rows stay FP32 in memory
  typedef unsigned short B;   // BF16 bits
  typedef float Acc;
  typedef f32x4 AV;
  typedef s16x4 BV;
  typedef f32x4 CV;
  __device__ static int row(int q, int i) { return 4 * q + i; }
  __device__ static CV mma(AV a, BV b, CV c) {
    const s16x4 ab = {bf16_rne(a[0]), bf16_rne(a[1]), bf16_rne(a[2]), bf16_rne(a[3])};
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ab, b, c, 0, 0, 0);
  }
};

// Two workgroups per CU for the FP32 product with 16 wavefronts and one row
// tile (the mixed precision default): at most 64 VGPRs (8 waves per SIMD),
// which the kernels reach without spills when their load groups are halved
// (3 chunks instead of 6: the same loads in flight per CU).  A backward
// launch of two 32-RHS row tiles has about 400 workgroups, more than one per
// CU; with two per CU it runs in one round.  The sums are unchanged.
template <class KT, int W, int MBW>
constexpr bool two_per_cu() {
  return std::is_same<KT, KF32>::value && W == 16 && MBW == 1;
}
template <class KT, int W, int MBW>
constexpr int min_waves_per_simd() {
  return two_per_cu<KT, W, MBW>() ? 8 : 1;
}

// Partial tile of one wavefront over the chunks [c_lo, c_hi) of 16 k: MBW
// 16-row tiles of A (row tiles mb0 .. mb0+MBW-1, clamped to the last one,
// whose rows are then never stored) times the 16 columns of B around this
// lane's column n.  Chunks are loaded G at a time; a chunk past c_hi loads
// the last one again and is skipped by the MFMA steps (the range is
// wave-uniform), so no load is masked and none waits on another.
template <class KT, int G, int MBW>
__device__ __forceinline__ void part_tile(const typename KT::A *alt, int mb0, int mtiles,
                                          const typename KT::B *blt, int n, int nchk, int c_lo,
                                          int c_hi, typename KT::CV (&acc)[MBW]) {
  typedef typename KT::AV AV;
  typedef typename KT::BV BV;
  const int lane = threadIdx.x & 63, q = lane >> 4;
  const typename KT::B *bb = blt + ((long)(n >> 4) * nchk * 64 + (n & 15) + 16 * q) * 4;
  const typename KT::A *ab[MBW];
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb) ab[mb] = alt + ((long)min(mb0 + mb, mtiles - 1) * nchk * 64 + lane) * 4;
  const int clast = max(c_hi - 1, 0);
  auto group = [&](int c0) {
    AV av[G][MBW];
    BV bv[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const long ch = min(c0 + g, clast);
      bv[g] = *reinterpret_cast<const BV *>(bb + ch * 256);
#pragma unroll
      for (int mb = 0; mb < MBW; ++mb) av[g][mb] = *reinterpret_cast<const AV *>(ab[mb] + ch * 256);
    }
    // every load of the group is issued before the first MFMA step, which
    // then waits only for its own chunk (the FP64 operands' group takes too
    // many registers for that: the compiler interleaves them)
    if constexpr (!std::is_same<KT, KF64>::value) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (c0 + g < c_hi)
#pragma unroll
        for (int mb = 0; mb < MBW; ++mb) acc[mb] = KT::mma(av[g][mb], bv[g], acc[mb]);
  };
  // the first group outside the loop: straight-line code, so its loads
  // follow the caller's earlier loads without waiting for them (the FP64
  // operands keep the plain loop, which needs fewer registers)
  int c0 = c_lo;
  if constexpr (!std::is_same<KT, KF64>::value) {
    group(c0);
    c0 += G;
  }
  for (; c0 < c_hi; c0 += G) group(c0);
}

// The product tile of a workgroup: W wavefronts, each its share of the K
// chunks (96 / W chunks per load group: one group at K <= 1536; half that
// for FP64 with two row tiles and for the two-per-CU FP32 kernels); their
// partial tiles go to LDS and thread t < 256 MBW receives the FP64 sum over
// the wavefronts, in wavefront order, of row t / 16 (of the MBW row tiles)
// and column t % 16 of the tile.  Every thread of the workgroup must call it.
template <class KT, int W, int MBW>
__device__ __forceinline__ double product_tile(const typename KT::A *alt, int mb0, int mtiles,
                                               const typename KT::B *blt, int n, int K,
                                               typename KT::Acc *red) {
  const int lane = threadIdx.x & 63, q = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nchk = (K + 15) >> 4;
  const int c_lo = nchk * wave / W, c_hi = nchk * (wave + 1) / W;
  typename KT::CV acc[MBW];
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb) acc[mb] = typename KT::CV{0, 0, 0, 0};
  constexpr int G =
      ((std::is_same<KT, KF64>::value && MBW == 2) || two_per_cu<KT, W, MBW>()) ? 48 / W : 96 / W;
  part_tile<KT, G, MBW>(alt, mb0, mtiles, blt, n, nchk, c_lo, c_hi, acc);
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[(wave * MBW + mb) * 256 + KT::row(q, i) * 16 + (lane & 15)] = acc[mb][i];
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x < 256 * MBW)
#pragma unroll
    for (int w = 0; w < W; ++w) s += (double)red[w * 256 * MBW + threadIdx.x];
  return s;
}

// XCD-aware placement: workgroup id -> (column tile, 16-row block).  The
// dispatcher deals workgroups to the 8 XCDs in turn; the row blocks of one
// column tile get ids that land on the same XCD one after another.
__device__ __forceinline__ bool tile_of(int wg, int ntiles, int nmb, int &tile, int &mb) {
  const int xcd = wg & 7, slot = wg >> 3;
  tile = (slot / nmb) * 8 + xcd;
  mb = slot % nmb;
  return tile < ntiles;
}

// The right-hand sides of a launch: FP32 (a solve's s) or FP64 (a
// refinement residual), a template parameter so that their loads need no
// branch.
template <class FT>
__device__ __forceinline__ const FT *rhs_ptr(const float *f, const double *f64) {
  if constexpr (std::is_same<FT, double>::value) return f64;
  else return f;
}

// ------------------------------------------------------------------ forward
// The epilogue operands that do not depend on the product are loaded before
// it, from addresses valid for every thread (threads without an output read
// right-hand side 0), so no load is predicated and the product's loads are
// issued without waiting for them: the X rows always, and, with 8 wavefronts
// of FP32 or BF16 operands (16 wavefronts leave no registers for them, and
// the FP64 operands take them), the right-hand side and coefficients of the
// first march row and of the closure row.
template <class KT, int W, int MBW, class FT>
__global__ __launch_bounds__(64 * W, (min_waves_per_simd<KT, W, MBW>())) void fused_fwd_kernel(FusedFwdArgs p) {
  __shared__ typename KT::Acc red[W * 256 * MBW];
  constexpr bool PF = !std::is_same<KT, KF64>::value && W == 8;
  span_begin(p.tspan);
  int tile, mbg;
  if (tile_of(blockIdx.x, p.ntiles, p.nmb, tile, mbg)) {
    const int t = threadIdx.x;
    const int col = t & 15;
    const int m = 16 * MBW * mbg + (t >> 4);
    const int steps = p.steps;
    const int i2 = p.i2, i0 = i2 + 2, nchk = (i2 + 15) >> 4;
    // interior column (1-based, before wrapping) of this lane's tile column
    const int c = tile * (16 - 2 * steps) - steps + 1 + col;
    int pc;
    bool halo = false;
    if (p.periodic) {
      pc = ((c - 1) % i2 + i2) % i2 + 1;
    } else {
      halo = c < 1 || c > i2;  // Dirichlet columns: X holds the boundary values
      pc = min(max(c, 1), i2);
    }
    const int xc = halo ? (c < 1 ? 1 : i0) : pc + 1;  // X column
    // X columns of the west and east neighbours (wrapped, or the Dirichlet
    // column at the edge)
    int xcw, xce;
    if (p.periodic) {
      xcw = (pc == 1 ? i2 : pc - 1) + 1;
      xce = (pc == i2 ? 1 : pc + 1) + 1;
    } else {
      xcw = max(min(c - 1, i2 + 1), 0) + 1;
      xce = max(min(c + 1, i2 + 1), 0) + 1;
    }
    const bool out = col >= steps && col < 16 - steps && c >= 1 && c <= i2;
    const bool act = t < 256 * MBW && m < p.M;
    const int mm = act ? m : 0;
    double *x = p.x + (long)mm * p.x_stride;
    const FT *f = rhs_ptr<FT>(p.f, p.f64) + (long)mm * p.f_stride;
    const int js = p.js, jf = p.jf, ieb = p.ieb;
    const int jr = ieb - 1;                  // closure row
    const int j1 = steps > 0 ? js : jr;     // first march row (a valid row either way)
    const long co = pc - 1;
    auto cf = [&](const float *arr, int j) -> double { return (double)arr[co + (long)(j - 1) * i2]; };
    auto fv = [&](int j) -> double { return (double)f[co + (long)(j - 1) * i2]; };
    // the free row's neighbours come from memory (that row is not written
    // in the forward pass), so the first row operation is exact in every
    // lane and each later one loses one lane per side
    double vs = x[f2(i0, xc, js)];
    double vc = x[f2(i0, xc, js + 1)];
    const double fw = x[f2(i0, xcw, js + 1)];
    const double fe = x[f2(i0, xce, js + 1)];
    const double xnorth = x[f2(i0, xc, ieb + 1)];
    FT f1 = 0, fr = 0;
    float a1 = 0.f, b1 = 0.f, c1 = 0.f, d1 = 0.f, e1 = 0.f;
    float ar = 0.f, br = 0.f, cr = 0.f, dr = 0.f, er = 0.f;
    if constexpr (PF) {  // kept in their own precision until used
      const long o1 = co + (long)(j1 - 1) * i2, orr = co + (long)(jr - 1) * i2;
      f1 = f[o1];
      a1 = p.ax[o1]; b1 = p.ay[o1]; c1 = p.bb[o1]; d1 = p.cx[o1]; e1 = p.cy[o1];
      fr = f[orr];
      ar = p.ax[orr]; br = p.ay[orr]; cr = p.bb[orr]; dr = p.cx[orr]; er = p.cy[orr];
    }
    const double d = product_tile<KT, W, MBW>((const typename KT::A *)p.a, MBW * mbg, (p.M + 15) >> 4,
                                              (const typename KT::B *)p.b, pc - 1, i2, red);
    if (act) {
      if (!halo) vs = vs - d;  // the south row as block b's correction leaves it
      if (p.d_out && out) p.d_out[(long)m * i2 + pc - 1] = d;
      for (int j = js; j <= jf; ++j) {
        const double sw = row_shr1(vc), se = row_shl1(vc);
        const double xw = j == js ? fw : sw, xe = j == js ? fe : se;
        double r;
        if (PF && j == js)
          r = ((double)f1 - (double)a1 * xw - (double)b1 * vs - (double)c1 * vc - (double)d1 * xe) /
              (double)e1;
        else
          r = (fv(j) - cf(p.ax, j) * xw - cf(p.ay, j) * vs - cf(p.bb, j) * vc - cf(p.cx, j) * xe) /
              cf(p.cy, j);
        if (halo) r = x[f2(i0, xc, j + 2)];
        if (out) x[f2(i0, xc, j + 2)] = r;
        vs = vc;
        vc = r;
      }
      const double sw = row_shr1(vc), se = row_shl1(vc);
      const double xw = steps == 0 ? fw : sw, xe = steps == 0 ? fe : se;
      double e;
      if constexpr (PF)
        e = (double)fr - (double)ar * xw - (double)br * vs - (double)cr * vc - (double)dr * xe -
            (double)er * xnorth;
      else
        e = fv(jr) - cf(p.ax, jr) * xw - cf(p.ay, jr) * vs - cf(p.bb, jr) * vc - cf(p.cx, jr) * xe -
            cf(p.cy, jr) * xnorth;
      if (out) {
        const long eo = lt_index(m, (int)co, nchk);
        if (p.e_out) p.e_out[eo] = e;
        if (p.e32_out) p.e32_out[eo] = (float)e;
      }
    }
  }
  span_end(p.tspan);
}

// ----------------------------------------------------------------- backward
// Workgroups [0, nhom_wg): the homogeneous march of block b+1, one wavefront
// per (tile, RHS), first so that the launch's longest dependent chain starts
// at once; then the product tiles of RINV_b (with the epilogue), then, when
// W_b is given, those of W_b (the coupling row, only stored).  The
// epilogue's X operands (and, where the forward kernel loads them early, its
// right-hand side and coefficients) are loaded before the product from
// clamped addresses (column I2 for lanes past it, row 1 for the closure row
// of the first block, which has none).
template <class KT, int W, int MBW, class FT>
__global__ __launch_bounds__(64 * W, (min_waves_per_simd<KT, W, MBW>())) void fused_bwd_kernel(FusedBwdArgs p) {
  __shared__ typename KT::Acc red[W * 256 * MBW];
  constexpr bool PF = !std::is_same<KT, KF64>::value && W == 8;
  span_begin(p.tspan);
  const int wg = blockIdx.x;
  int tile, mbg;
  if (wg < p.nhom_wg) {
    const long wave = (long)wg * W + (threadIdx.x >> 6);
    if (wave < (long)p.hom_tiles * p.hom.nrhs)
      march_body<1, MODE_HOM>(p.hom, (int)(wave % p.hom_tiles), (int)(wave / p.hom_tiles), 0, 1,
                              threadIdx.x & 63);
  } else {
    const int wp = wg - p.nhom_wg;
    const bool cpl = wp >= p.nprod_half;
    if (tile_of(cpl ? wp - p.nprod_half : wp, p.ntiles, p.nmb, tile, mbg)) {
      const int t = threadIdx.x;
      const int m = 16 * MBW * mbg + (t >> 4);
      const int i2 = p.i2, i0 = i2 + 2, nchk = (i2 + 15) >> 4;
      const int c = 16 * tile + 1 + (t & 15);
      const bool ok = c <= i2;
      const bool act = t < 256 * MBW && m < p.M && ok;
      const int mm = act ? m : 0, cc = min(c, i2);
      const int js = p.js, jrow = max(js - 1, 1);
      double *x = p.x + (long)mm * p.x_stride;
      const bool resid = !cpl && p.do_resid && act;
      const long co = (long)(cc - 1) + (long)(jrow - 1) * i2;  // coefficient row js-1
      const int xcw = p.periodic ? (cc == 1 ? i2 : cc - 1) + 1 : cc;
      const int xce = p.periodic ? (cc == i2 ? 1 : cc + 1) + 1 : cc + 2;
      const double *dsp = p.dsouth ? p.dsouth : p.hfree_out;  // any valid (M, i2) array
      const FT *fp = rhs_ptr<FT>(p.f, p.f64) + (long)mm * p.f_stride + co;
      const double *dp = dsp + (long)mm * i2 + cc - 1;
      // closure row js-1 of block b-1; row js = IE(b-1)
      const double xfree = x[f2(i0, cc + 1, js + 1)];
      double xsouth = x[f2(i0, cc + 1, jrow)];
      const double xcen = x[f2(i0, cc + 1, js)];
      const double xw = x[f2(i0, xcw, js)];
      const double xe = x[f2(i0, xce, js)];
      FT fr = 0;
      float ar = 0.f, br = 0.f, cr = 0.f, dr = 0.f, er = 0.f;
      double ds = 0.0;
      if constexpr (PF) {
        fr = *fp;
        ar = p.ax[co]; br = p.ay[co]; cr = p.bb[co]; dr = p.cx[co]; er = p.cy[co];
        ds = *dp;
      }
      // lanes past I2 (last tile) read the zero-padded matrix columns
      const double v = product_tile<KT, W, MBW>((const typename KT::A *)p.a, MBW * mbg, (p.M + 15) >> 4,
                                                (const typename KT::B *)(cpl ? p.bw : p.b), c - 1, i2, red);
      if (act) {
        const long ro = (long)m * i2 + c - 1;
        if (cpl) {
          p.hcpl_out[ro] = v;
        } else {
          p.hfree_out[ro] = v;
          const double north = xfree + v;
          x[f2(i0, c + 1, js + 1)] = north;
          if (resid) {
            if constexpr (!PF) {
              fr = *fp;
              ar = p.ax[co]; br = p.ay[co]; cr = p.bb[co]; dr = p.cx[co]; er = p.cy[co];
              ds = *dp;
            }
            if (p.dsouth) xsouth = xsouth - ds;
            const double e = (double)fr - (double)ar * xw - (double)br * xsouth - (double)cr * xcen -
                             (double)dr * xe - (double)er * north;
            const long eo = lt_index(m, c - 1, nchk);
            if (p.e_out) p.e_out[eo] = e;
            if (p.e32_out) p.e32_out[eo] = (float)e;
          }
        }
      }
    }
  }
  span_end(p.tspan);
}

// ---------------------------------------------------------------- launchers
int fused_tiles_fwd(int i2, int steps) {
  const int tout = 16 - 2 * steps;
  return tout > 0 ? (i2 + tout - 1) / tout : 0;
}

static long placed_workgroups(int ntiles, int nmb) { return 8L * nmb * ((ntiles + 7) / 8); }

// Row tiles per workgroup: the requested 1 or 2 (0: two for FP32 and BF16,
// which share each matrix load, one for FP64), one when M <= 16.  Two row
// tiles halve the matrix reads and the workgroups; FP64 then loads its
// chunks in groups of half the size, within the registers of 16 wavefronts.
static int rows_per_group(int arith, int M, int rows) {
  if (M <= 16) return 1;
  if (rows == 1 || rows == 2) return rows;
  return arith != FUSED_F64 ? 2 : 1;
}

long fused_fwd_workgroups(const FusedFwdArgs &a) { return placed_workgroups(a.ntiles, a.nmb); }

long fused_bwd_workgroups(const FusedBwdArgs &a) {
  return a.nhom_wg + (a.bw ? 2L : 1L) * a.nprod_half;
}

static int waves_of(int w) { return w == 0 ? FUSED_WAVES : w; }

// Fill the launcher-set fields (tiles, row-tile groups) of the arguments.
int fused_fwd_prepare(FusedFwdArgs &a, int arith) {
  if (a.steps < 0 || 16 - 2 * a.steps < 2 || a.M < 1 || a.M > 64 || a.i2 % 4 != 0) return -1;
  const int w = waves_of(a.waves);
  if (w != 8 && w != 16) return -1;
  if (a.rows < 0 || a.rows > 2) return -1;
  a.ntiles = fused_tiles_fwd(a.i2, a.steps);
  a.mbw = rows_per_group(arith, a.M, a.rows);
  a.nmb = ((a.M + 15) / 16 + a.mbw - 1) / a.mbw;
  return 0;
}

int fused_bwd_prepare(FusedBwdArgs &a, int arith) {
  if (a.M < 1 || a.M > 64 || a.i2 % 4 != 0) return -1;
  const int w = waves_of(a.waves);
  if (w != 8 && w != 16) return -1;
  if (a.rows < 0 || a.rows > 2) return -1;
  a.ntiles = (a.i2 + 15) / 16;
  a.mbw = rows_per_group(arith, a.M, a.rows);
  a.nmb = ((a.M + 15) / 16 + a.mbw - 1) / a.mbw;
  a.nprod_half = (int)placed_workgroups(a.ntiles, a.nmb);
  a.hom_tiles = 0;
  a.nhom_wg = 0;
  if (a.has_hom && a.hom.jf >= a.hom.js) {  // a one-row block has nothing to march
    // one lane per column (any lane width whose tile is positive gives the
    // same values; one keeps the march's registers within the product's)
    a.hom_tiles = march_tiles(a.hom, 1);
    if (a.hom_tiles <= 0) return -1;
    a.nhom_wg = (int)(((long)a.hom_tiles * a.hom.nrhs + w - 1) / w);
  }
  return 0;
}

template <class KT, int W, int MBW, class FT>
static void fwd_go(const FusedFwdArgs &a, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  const dim3 grid((unsigned)fused_fwd_workgroups(a)), block(64 * W);
  if (ev0) hipExtLaunchKernelGGL((fused_fwd_kernel<KT, W, MBW, FT>), grid, block, 0, st, ev0, ev1, 0, a);
  else hipLaunchKernelGGL((fused_fwd_kernel<KT, W, MBW, FT>), grid, block, 0, st, a);
}

template <class KT, int W, int MBW, class FT>
static void bwd_go(const FusedBwdArgs &a, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  const dim3 grid((unsigned)fused_bwd_workgroups(a)), block(64 * W);
  if (ev0) hipExtLaunchKernelGGL((fused_bwd_kernel<KT, W, MBW, FT>), grid, block, 0, st, ev0, ev1, 0, a);
  else hipLaunchKernelGGL((fused_bwd_kernel<KT, W, MBW, FT>), grid, block, 0, st, a);
}

// the right-hand sides' precision -> the kernel that loads them
template <class KT, int W, int MBW>
static void go(const FusedFwdArgs &a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  if (a.f64) fwd_go<KT, W, MBW, double>(a, st, e0, e1);
  else fwd_go<KT, W, MBW, float>(a, st, e0, e1);
}
template <class KT, int W, int MBW>
static void go(const FusedBwdArgs &a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  if (a.f64) bwd_go<KT, W, MBW, double>(a, st, e0, e1);
  else bwd_go<KT, W, MBW, float>(a, st, e0, e1);
}

// (arithmetic, waves, row tiles per workgroup) -> the compiled kernel
template <class Args>
static void dispatch(const Args &a, int arith, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  const bool w16 = waves_of(a.waves) == 16, two = a.mbw == 2;
  if (arith == FUSED_F64) {
    if (w16) two ? go<KF64, 16, 2>(a, st, e0, e1) : go<KF64, 16, 1>(a, st, e0, e1);
    else two ? go<KF64, 8, 2>(a, st, e0, e1) : go<KF64, 8, 1>(a, st, e0, e1);
  } else if (arith == FUSED_B16) {
    if (w16) two ? go<KB16, 16, 2>(a, st, e0, e1) : go<KB16, 16, 1>(a, st, e0, e1);
    else two ? go<KB16, 8, 2>(a, st, e0, e1) : go<KB16, 8, 1>(a, st, e0, e1);
  } else {
    if (w16) two ? go<KF32, 16, 2>(a, st, e0, e1) : go<KF32, 16, 1>(a, st, e0, e1);
    else two ? go<KF32, 8, 2>(a, st, e0, e1) : go<KF32, 8, 1>(a, st, e0, e1);
  }
}

hipError_t launch_fused_fwd(const FusedFwdArgs &a0, int arith, hipStream_t st, hipEvent_t ev0,
                            hipEvent_t ev1) {
  FusedFwdArgs a = a0;
  if (fused_fwd_prepare(a, arith)) return hipErrorInvalidValue;
  dispatch(a, arith, st, ev0, ev1);
  return hipGetLastError();
}

hipError_t launch_fused_bwd(const FusedBwdArgs &a0, int arith, hipStream_t st, hipEvent_t ev0,
                            hipEvent_t ev1) {
  FusedBwdArgs a = a0;
  if (fused_bwd_prepare(a, arith)) return hipErrorInvalidValue;
  dispatch(a, arith, st, ev0, ev1);
  return hipGetLastError();
}

// FP64 column-major (K, N) matrices to BF16 lane tiles (round to nearest
// even via FP32; zero past K or N).
__global__ void cvt_lane_tiles_bf16_kernel(const double *src, unsigned short *dst, int K, int N,
                                           long nmat) {
  const int nchk = (K + 15) / 16;
  const long per = lane_tile_elems(N, K);
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * nmat) return;
  const long mat = i / per, r = i % per;
  const int t = (int)(r & 3), l = (int)((r >> 2) & 63);
  const long blk = r >> 8;
  const int ch = (int)(blk % nchk), tile = (int)(blk / nchk);
  const int k = 16 * ch + 4 * (l >> 4) + t, n = 16 * tile + (l & 15);
  dst[i] = (k < K && n < N) ? (unsigned short)bf16_rne((float)src[mat * K * (long)N + k + (long)n * K]) : 0;
}

hipError_t launch_cvt_lane_tiles_bf16(const double *src, unsigned short *dst, int K, int N, long nmat,
                                     hipStream_t st) {
  const long n = lane_tile_elems(N, K) * nmat;
  hipLaunchKernelGGL(cvt_lane_tiles_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     src, dst, K, N, nmat);
  return hipGetLastError();
}

}  // namespace sevp
