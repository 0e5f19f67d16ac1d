// Full-width SEVP building blocks (REP1, REP, PRE1, PRE and the strip PRE),
// batched over independent right-hand sides / unit vectors.
//
// March kernels use trapezoid tiles: a wavefront marches 64*K consecutive
// columns (K per lane, east/west neighbours through DPP wave shifts), and
// after s rows only the columns at least s away from the tile edge are still
// exact, so each tile carries `halo` extra columns on each side and writes
// only its central `tile_out` columns.  Recomputing the halo columns repeats
// exactly the same operations on exactly the same operands, so the written
// values are bitwise those of a march over the whole width.  Periodic
// problems index columns modulo the width; this reproduces the reference's
// halo copies X(1)=X(I0-1), X(I0)=X(2), which only ever duplicate interior
// values, and a final pass (halo_fix_kernel) leaves the halo columns as the
// reference leaves them.
//
// The correction products use a tiled FP64 GEMM whose every output is one
// thread's sequential sum over k from zero with separately rounded multiply
// and add: the reference's DUM2(N)=DUM2(N)+DUM1(M)*RINV(M,N) loop order.
//
// MATINV is the reference's Gauss-Jordan without pivoting
// (PREP/template/GLOBAL/src/prep_GLOBAL.f90:516-568) with the matrix held in
// LDS when it fits (larger orders use the blocked form in sevp_matinv.hip);
// each element receives the same sequence of operations as in the reference,
// only the independent element updates of one step run in parallel.
#include <hip/hip_ext.h>

#include "sevp_dev.h"
#include "sevp_internal.h"
#include "sevp_march.h"

namespace sevp {

template <int K, int MODE>
__global__ __launch_bounds__(64) void march_kernel(MarchArgs p) {
  march_body<K, MODE>(p, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z, threadIdx.x);
}

// Choose lanes-per-column K and the tile size for a march of `steps` rows.
static void march_shape(int w, int steps, int &K, int &tile, int &halo, int &ntiles) {
  halo = steps + 1;
  int best = -1;
  long bestcost = 0;
  for (int k : {1, 2, 4}) {
    const int t = 64 * k - 2 * halo;
    if (t <= 0) continue;
    const long n = (w + t - 1) / t;
    const long cost = n * k;
    if (best < 0 || cost < bestcost) {
      best = k;
      bestcost = cost;
      tile = t;
      ntiles = (int)n;
    }
  }
  K = best;
}

template <int MODE>
static hipError_t launch_march_mode(const MarchArgs &a0, int nbatch, int nstrips, hipStream_t st) {
  MarchArgs a = a0;
  int K, tile, halo, ntiles;
  march_shape(a.w, a.jf - a.js + 1, K, tile, halo, ntiles);
  if (K < 0) return hipErrorInvalidValue;
  a.tile_out = tile;
  a.halo = halo;
  dim3 grid(ntiles, nbatch, nstrips);
  if (K == 1) hipLaunchKernelGGL((march_kernel<1, MODE>), grid, dim3(64), 0, st, a);
  else if (K == 2) hipLaunchKernelGGL((march_kernel<2, MODE>), grid, dim3(64), 0, st, a);
  else hipLaunchKernelGGL((march_kernel<4, MODE>), grid, dim3(64), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_march(int mode, const MarchArgs &a, int nbatch, hipStream_t st) {
  const int nstrips = (mode == MODE_PRE) ? a.nrhs : 1;  // PRE: nrhs carries the strip count
  if (mode == MODE_PART) return launch_march_mode<MODE_PART>(a, nbatch, 1, st);
  if (mode == MODE_HOM) return launch_march_mode<MODE_HOM>(a, nbatch, 1, st);
  return launch_march_mode<MODE_PRE>(a, nbatch, nstrips, st);
}

// Closure residual of equation row ieb-1 at column c (1-based) for RHS r,
// given row ieb's west/centre/east values and row ieb+1's value at c.
__device__ __forceinline__ double closure_resid(const ResidArgs &p, int r, int c, const double *x,
                                                double west, double center, double east,
                                                double north) {
  const int i0 = p.i2 + 2, jr = p.ieb - 1;
  const long co = (long)(c - 1) + (long)(jr - 1) * p.i2;
  const double fv = p.f64 ? p.f64[(long)r * p.f_stride + co] : (double)p.f[(long)r * p.f_stride + co];
  return fv - (double)p.ax[co] * west - (double)p.ay[co] * x[f2(i0, c + 1, jr)] -
         (double)p.bb[co] * center - (double)p.cx[co] * east - (double)p.cy[co] * north;
}

// West and east neighbours in row ieb of column c, reading row values
// through rowv (which applies the DUM0 restore where there is one).
template <class RowV>
__device__ __forceinline__ void row_nbrs(const ResidArgs &p, const double *x, int c, RowV rowv,
                                         double &west, double &east) {
  const int i2 = p.i2, i0 = i2 + 2, ieb = p.ieb;
  if (p.periodic) {
    west = rowv(c == 1 ? i2 : c - 1);
    east = rowv(c == i2 ? 1 : c + 1);
  } else {
    west = (c == 1) ? x[f2(i0, 1, ieb)] : rowv(c - 1);
    east = (c == i2) ? x[f2(i0, i0, ieb)] : rowv(c + 1);
  }
}

// Backward restore of X(:, ieb) from DUM0 and closure residual at row ieb-1.
__device__ void resid_body(const ResidArgs &p, long t) {
  if (t >= (long)p.nrhs * p.i2) return;
  const int r = (int)(t / p.i2);
  const int c = (int)(t % p.i2) + 1;
  const int i2 = p.i2, i0 = i2 + 2, ieb = p.ieb;
  double *x = p.x + (long)r * p.x_stride;
  const double *d0 = p.dum0 ? p.dum0 + (long)r * p.dum0_stride + (long)(p.nb - 1) * i2 : nullptr;
  auto rowv = [&](int cc) -> double {  // value of row ieb at column cc (1..i2) after the restore
    return d0 ? d0[cc - 1] : x[f2(i0, cc + 1, ieb)];
  };
  const double center = rowv(c);
  double west, east;
  row_nbrs(p, x, c, rowv, west, east);
  const double e = closure_resid(p, r, c, x, west, center, east, x[f2(i0, c + 1, ieb + 1)]);
  if (p.e) p.e[(long)r * i2 + c - 1] = e;
  if (p.e32) p.e32[(long)r * i2 + c - 1] = (float)e;
  if (d0) x[f2(i0, c + 1, ieb)] = center;
}

// Mixed-precision backward, per (rhs, column): X(apply_row) += the free-row
// correction of the block above (summed from its partial products), then,
// when do_resid, the closure residual of the block below (row ieb-1 =
// apply_row-2), whose only read of apply_row is this column's.  X(ieb) was
// never overwritten by the forward pass here, so no restore is needed.
__device__ void apply_resid_body(const ResidArgs &p, long t) {
  if (t >= (long)p.nrhs * p.i2) return;
  const int r = (int)(t / p.i2);
  const int c = (int)(t % p.i2) + 1;
  const int i2 = p.i2, i0 = i2 + 2, ieb = p.ieb;
  double *x = p.x + (long)r * p.x_stride;
  const long xa = f2(i0, c + 1, p.apply_row);
  const double north = x[xa] + p.corr.sum(r, c - 1);
  x[xa] = north;
  if (!p.do_resid) return;
  auto rowv = [&](int cc) -> double { return x[f2(i0, cc + 1, ieb)]; };
  double west, east;
  row_nbrs(p, x, c, rowv, west, east);
  const double e = closure_resid(p, r, c, x, west, rowv(c), east, north);
  if (p.e) p.e[(long)r * i2 + c - 1] = e;
  if (p.e32) p.e32[(long)r * i2 + c - 1] = (float)e;
}

__global__ void resid_kernel(ResidArgs p) {
  resid_body(p, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// The backward H march of block b and the restore + closure residual of
// block b-1 touch disjoint rows and read nothing the other writes: one
// launch runs both (march tiles first, then 64-column residual groups).
template <int K>
__device__ __forceinline__ void hom_resid_body(const MarchArgs &m, const ResidArgs &r, int ntiles,
                                               int has_resid, long b) {
  const long nm = (long)ntiles * m.nrhs;
  if (b < nm) {
    march_body<K, MODE_HOM>(m, (int)(b % ntiles), (int)(b / ntiles), 0, 1, threadIdx.x);
    return;
  }
  if (!has_resid) return;
  if (r.corr.on()) apply_resid_body(r, (b - nm) * 64 + threadIdx.x);
  else resid_body(r, (b - nm) * 64 + threadIdx.x);
}

template <int K>
__global__ __launch_bounds__(64) void hom_resid_kernel(MarchArgs m, ResidArgs r, int ntiles,
                                                       int has_resid) {
  hom_resid_body<K>(m, r, ntiles, has_resid, blockIdx.x);
}

int march_tiles(MarchArgs &a, int K) {
  const int steps = a.jf - a.js + 1;
  int k, tile, halo, ntiles;
  march_shape(a.w, steps, k, tile, halo, ntiles);
  if (k < 0) return 0;
  if (K > 0) {  // a given lane width (any width whose tile is positive gives the same values)
    tile = 64 * K - 2 * halo;
    if (tile <= 0) return 0;
    ntiles = (a.w + tile - 1) / tile;
  }
  a.tile_out = tile;
  a.halo = halo;
  return ntiles;
}

int march_lanes(int w, int steps) {
  int k, tile, halo, ntiles;
  march_shape(w, steps, k, tile, halo, ntiles);
  return k;
}

long hom_resid_prepare(const MarchArgs &a0, const ResidArgs *r, int K, HomResidEnt &ent) {
  ent.m = a0;
  const int ntiles = march_tiles(ent.m, K);
  if (ntiles <= 0) return -1;
  ent.ntiles = ntiles;
  const long nm = (long)ntiles * a0.nrhs;
  const long nr = r ? ((long)r->nrhs * r->i2 + 63) / 64 : 0;
  const bool march = a0.jf >= a0.js;
  ent.r = r ? *r : ResidArgs{};
  ent.has_resid = r != nullptr;
  if (!march) ent.m.nrhs = 0;  // no march blocks: every block is a residual group
  return (march ? nm : 0) + nr;
}

hipError_t launch_hom_resid(const MarchArgs &a0, const ResidArgs *r, hipStream_t st) {
  const int K = march_lanes(a0.w, a0.jf - a0.js + 1);
  if (K < 0) return hipErrorInvalidValue;
  HomResidEnt e;
  const long nwg = hom_resid_prepare(a0, r, K, e);
  if (nwg < 0) return hipErrorInvalidValue;
  if (nwg == 0) return hipSuccess;
  const dim3 grid((unsigned)nwg);
  if (K == 1) hipLaunchKernelGGL((hom_resid_kernel<1>), grid, dim3(64), 0, st, e.m, e.r, e.ntiles, e.has_resid);
  else if (K == 2) hipLaunchKernelGGL((hom_resid_kernel<2>), grid, dim3(64), 0, st, e.m, e.r, e.ntiles, e.has_resid);
  else hipLaunchKernelGGL((hom_resid_kernel<4>), grid, dim3(64), 0, st, e.m, e.r, e.ntiles, e.has_resid);
  return hipGetLastError();
}

hipError_t launch_resid(const ResidArgs &a, hipStream_t st) {
  const long n = (long)a.nrhs * a.i2;
  hipLaunchKernelGGL(resid_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sequential-order FP64 GEMM with fused epilogues.
//   D(m, n) = sum_{k=0..K-1} A(m,k) * B(k,n)   (EP_NEGSUB: 0 - sum ... as
//   successive subtractions)
// One output per thread, so that the independent sequential chains fill the
// GPU (a 64 x 1440 product has 92160 of them).  Tile 16 (m) x 16 (n), K in
// chunks of 64 staged through LDS from a ring of register slots filled three
// chunks ahead.
constexpr int GT_M = 16, GT_N = 16, GT_K = 64;

__global__ __launch_bounds__(256) void gemm_seq_kernel(GemmArgs p) {
  span_begin(p.tspan);
  // both tiles k-contiguous (rows of GT_K + 2 doubles, 16-byte aligned) so a
  // thread reads two consecutive k of each operand with one 128-bit load
  __shared__ __align__(16) double As[GT_M][GT_K + 2];
  __shared__ __align__(16) double Bs[GT_N][GT_K + 2];
  const int t = threadIdx.x;
  // XCD-aware placement: the dispatcher deals workgroups to the 8 XCDs in
  // turn, so the row tiles of one column tile get ids on the same XCD, one
  // after another, and read that tile of B from HBM once (the others from
  // the XCD's L2); ids past the last column tile have no tile
  const int ntn = (p.N + GT_N - 1) / GT_N, nmb = (p.M + GT_M - 1) / GT_M;
  const int slot = blockIdx.x >> 3, tcol = (slot / nmb) * 8 + (blockIdx.x & 7);
  if (tcol >= ntn) {
    span_end(p.tspan);
    return;
  }
  const int n0 = tcol * GT_N, m0 = (slot % nmb) * GT_M;
  const int tn = t % GT_N, tm = t / GT_N;
  // staging: each thread moves 4 elements of A and 4 of B per chunk
  double ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = t + 256 * q;
      {  // A tile: kk fastest (contiguous when a_sk == 1)
        const int kk = (p.a_sk == 1) ? idx % GT_K : idx / GT_M;
        const int mm = (p.a_sk == 1) ? idx / GT_K : idx % GT_M;
        const int m = m0 + mm, k = k0 + kk;
        ra[q] = (m < p.M && k < p.K) ? p.a[(long)m * p.a_sm + (long)k * p.a_sk] : 0.0;
      }
      {  // B tile: kk fastest when b_sk == 1
        const int kk = (p.b_sk == 1) ? idx % GT_K : idx / GT_N;
        const int nn = (p.b_sk == 1) ? idx / GT_K : idx % GT_N;
        const int n = n0 + nn, k = k0 + kk;
        rb[q] = (n < p.N && k < p.K) ? p.b[(long)k * p.b_sk + (long)n * p.b_sn] : 0.0;
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = t + 256 * q;
      const int kka = (p.a_sk == 1) ? idx % GT_K : idx / GT_M;
      const int mma = (p.a_sk == 1) ? idx / GT_K : idx % GT_M;
      As[mma][kka] = ra[q];
      const int kkb = (p.b_sk == 1) ? idx % GT_K : idx / GT_N;
      const int nnb = (p.b_sk == 1) ? idx / GT_K : idx % GT_N;
      Bs[nnb][kkb] = rb[q];
    }
  };
  const bool neg = (p.ep == EP_NEGSUB);
  double acc = 0.0;
  // chunks are staged in a ring of PF register slots, fetched PF chunks
  // ahead of their use, so the global latency overlaps PF chunks of compute
  constexpr int PF = 3;
  double ra_r[PF][4], rb_r[PF][4];
  const int nchunk = (p.K + GT_K - 1) / GT_K;
  auto fetch_to = [&](int c, double (&ra_)[4], double (&rb_)[4]) {
    fetch(c * GT_K);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra_[q] = ra[q];
      rb_[q] = rb[q];
    }
  };
#pragma unroll
  for (int d = 0; d < PF; ++d)
    if (d < nchunk) fetch_to(d, ra_r[d], rb_r[d]);
  for (int c0 = 0; c0 < nchunk; c0 += PF) {
#pragma unroll
    for (int d = 0; d < PF; ++d) {
      const int c = c0 + d;
      if (c >= nchunk) break;
      const int k0 = c * GT_K;
      const int kc = min(GT_K, p.K - k0);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ra[q] = ra_r[d][q];
        rb[q] = rb_r[d][q];
      }
      stash();
      __syncthreads();
      if (c + PF < nchunk) fetch_to(c + PF, ra_r[d], rb_r[d]);
      const double2 *ar = reinterpret_cast<const double2 *>(&As[tm][0]);
      const double2 *br = reinterpret_cast<const double2 *>(&Bs[tn][0]);
      const int kc2 = kc / 2;
      if (kc == GT_K) {
        // a whole chunk in groups of 16 k: the group's LDS reads are issued
        // one group ahead and its 16 products (independent of the sum) are
        // formed before the 16 dependent additions, which then follow one
        // another at the add latency instead of each waiting for its
        // multiply.  Same products, same order of additions.
        constexpr int GP = 8;  // k pairs per group
        double2 a_c[GP], b_c[GP], a_n[GP], b_n[GP];
#pragma unroll
        for (int j = 0; j < GP; ++j) {
          a_c[j] = ar[j];
          b_c[j] = br[j];
        }
#pragma unroll
        for (int g = 0; g < GT_K / 2 / GP; ++g) {
          if (g + 1 < GT_K / 2 / GP) {
#pragma unroll
            for (int j = 0; j < GP; ++j) {
              a_n[j] = ar[(g + 1) * GP + j];
              b_n[j] = br[(g + 1) * GP + j];
            }
          }
          double pr[2 * GP];
#pragma unroll
          for (int j = 0; j < GP; ++j) {
            pr[2 * j] = a_c[j].x * b_c[j].x;
            pr[2 * j + 1] = a_c[j].y * b_c[j].y;
          }
          if (neg) {
#pragma unroll
            for (int j = 0; j < 2 * GP; ++j) acc = acc - pr[j];
          } else {
#pragma unroll
            for (int j = 0; j < 2 * GP; ++j) acc = acc + pr[j];
          }
#pragma unroll
          for (int j = 0; j < GP; ++j) {
            a_c[j] = a_n[j];
            b_c[j] = b_n[j];
          }
        }
      } else if (neg) {
        for (int k2 = 0; k2 < kc2; ++k2) {
          const double2 av = ar[k2], bv = br[k2];
          acc = acc - av.x * bv.x;
          acc = acc - av.y * bv.y;
        }
        if (kc & 1) acc = acc - As[tm][kc - 1] * Bs[tn][kc - 1];
      } else {
        for (int k2 = 0; k2 < kc2; ++k2) {
          const double2 av = ar[k2], bv = br[k2];
          acc = acc + av.x * bv.x;
          acc = acc + av.y * bv.y;
        }
        if (kc & 1) acc = acc + As[tm][kc - 1] * Bs[tn][kc - 1];
      }
    }
  }
  const int n = n0 + tn, m = m0 + tm;
  const int i0 = p.i2 + 2;
  const double d = acc;
  if (n < p.N && m < p.M) switch (p.ep) {
    case EP_FWD: {
      double *x = p.x + (long)m * p.x_stride;
      const long xi = f2(i0, n + 2, p.xrow);
      const double xv = x[xi];
      p.dum0[(long)m * p.dum0_stride + n] = xv;
      x[xi] = xv - d;
      break;
    }
    case EP_BWD: {
      double *x = p.x + (long)m * p.x_stride;
      const long xi = f2(i0, n + 2, p.xrow);
      x[xi] = x[xi] + d;
      p.out[(long)m * p.out_ld + n] = d;
      if (p.g) p.g[(long)m * p.out_ld + n] = d * (double)p.cy[(long)n + (long)(p.cyrow - 1) * p.i2];
      break;
    }
    case EP_CPL:
      p.out[(long)m * p.out_ld + n] = d;
      break;
    default:  // EP_NEGSUB, EP_COLMAJ: column-major output (m fastest)
      p.out[(long)m + (long)n * p.out_ld] = d;
      break;
  }
  span_end(p.tspan);
}

// T(i, k) = R(i, k) * CY(k, cyrow): columns of an n x n column-major matrix
// scaled by one row of the north coefficient.
__global__ void scale_cols_kernel(const double *r, const float *cy, int cyrow, int i2, double *t) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)i2 * i2) return;
  const int k = (int)(q / i2);
  t[q] = r[q] * (double)cy[(long)k + (long)(cyrow - 1) * i2];
}

hipError_t launch_scale_cols(const double *r, const float *cy, int cyrow, int i2, double *t,
                             hipStream_t st) {
  const long n = (long)i2 * i2;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, r, cy,
                     cyrow, i2, t);
  return hipGetLastError();
}

long gemm_workgroups(const GemmArgs &a) {
  return 8L * ((a.M + GT_M - 1) / GT_M) * (((a.N + GT_N - 1) / GT_N + 7) / 8);
}

hipError_t launch_gemm(const GemmArgs &a, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  const dim3 grid((unsigned)gemm_workgroups(a));
  if (ev0) hipExtLaunchKernelGGL(gemm_seq_kernel, grid, dim3(256), 0, st, ev0, ev1, 0, a);
  else hipLaunchKernelGGL(gemm_seq_kernel, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// MATINV in LDS for n <= MATINV_LDS_MAX, one workgroup per matrix.
constexpr int MATINV_LDS_MAX = 128;

__global__ __launch_bounds__(256) void matinv_lds_kernel(double *mats, long stride, int n, int *zp) {
  extern __shared__ double sm[];
  double *B = sm;                 // n*n, column-major
  double *b1 = sm + (long)n * n;  // n
  __shared__ double piv;
  __shared__ int zcount;
  double *g = mats + (long)blockIdx.x * stride;
  const int t = threadIdx.x, nt = blockDim.x;
  const int nn = n * n;
  for (int i = t; i < nn; i += nt) B[i] = g[i];
  if (t == 0) zcount = 0;
  __syncthreads();
#define BM(i, j) B[((i) - 1) + ((j) - 1) * n]
  for (int i = 1; i <= n - 1; ++i) {
    if (t == 0) {
      if (BM(i, i) == 0.0) ++zcount;
      piv = 1.0 / BM(i, i);
      BM(i, i) = 1.0;
    }
    __syncthreads();
    const double bp = piv;
    for (int j = 1 + t; j <= n; j += nt) BM(i, j) = BM(i, j) * bp;
    for (int ia = i + 1 + t; ia <= n; ia += nt) b1[ia - 1] = BM(ia, i);
    __syncthreads();
    // rows ia > i: column i is zeroed then updated like every other column
    const int rows = n - i;
    for (int idx = t; idx < rows * n; idx += nt) {
      const int ia = i + 1 + idx % rows, j = 1 + idx / rows;
      const double cur = (j == i) ? 0.0 : BM(ia, j);
      BM(ia, j) = cur - b1[ia - 1] * BM(i, j);
    }
    __syncthreads();
  }
  if (t == 0) {
    if (BM(n, n) == 0.0) ++zcount;
    piv = 1.0 / BM(n, n);
    BM(n, n) = 1.0;
  }
  __syncthreads();
  {
    const double bp = piv;
    for (int j = 1 + t; j <= n; j += nt) BM(n, j) = BM(n, j) * bp;
  }
  __syncthreads();
  for (int i = 2; i <= n; ++i) {
    for (int ib = 1 + t; ib <= i - 1; ib += nt) b1[ib - 1] = BM(ib, i);
    __syncthreads();
    const int rows = i - 1;
    for (int idx = t; idx < rows * n; idx += nt) {
      const int ib = 1 + idx % rows, j = 1 + idx / rows;
      const double cur = (j == i) ? 0.0 : BM(ib, j);
      BM(ib, j) = cur - b1[ib - 1] * BM(i, j);
    }
    __syncthreads();
  }
#undef BM
  for (int i = t; i < nn; i += nt) g[i] = B[i];
  if (t == 0 && zp && zcount) atomicAdd(zp, zcount);
}

bool matinv_uses_lds(int n) {
  return n <= MATINV_LDS_MAX && sizeof(double) * ((size_t)n * n + n) <= 64 * 1024;
}

// LDS-resident elimination when the matrix fits, else the blocked form
// (sevp_matinv.hip) with `scratch` of matinv_blocked_scratch(n, count) doubles.
hipError_t launch_matinv(double *mats, long stride, int n, int count, int *zp, double *scratch,
                         hipStream_t st) {
  if (matinv_uses_lds(n)) {
    const size_t lds = sizeof(double) * ((size_t)n * n + n);
    hipLaunchKernelGGL(matinv_lds_kernel, dim3(count), dim3(256), lds, st, mats, stride, n, zp);
    return hipGetLastError();
  }
  if (!scratch) return hipErrorInvalidValue;
  return launch_matinv_blocked(mats, stride, n, count, zp, scratch, st);
}

// Periodic halo columns at the end of REP1 (MAIN/template/GLOBAL/src/sevp.f90:103-106).
__global__ void halo_fix_kernel(double *x, long xs, int i2, int j2, int nrhs) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nrhs * j2) return;
  const int r = (int)(t / j2), j = (int)(t % j2) + 2;
  const int i0 = i2 + 2;
  double *xr = x + (long)r * xs;
  xr[f2(i0, 1, j)] = xr[f2(i0, i0 - 1, j)];
  xr[f2(i0, i0, j)] = xr[f2(i0, 2, j)];
}

hipError_t launch_halo_fix(double *x, long xs, int i2, int j2, int nrhs, hipStream_t st) {
  const long n = (long)nrhs * j2;
  hipLaunchKernelGGL(halo_fix_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, xs, i2, j2, nrhs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grouped launches of the mixed pass: one launch runs the same step of many
// independent problems (group_entry finds a workgroup's problem), each entry
// with exactly the workgroups and arguments of its own launch, so every value
// written is the one the problem's own launch writes.
template <int K>
__global__ __launch_bounds__(64) void batch_march_kernel(const MarchEnt *__restrict__ tab,
                                                         const int *__restrict__ start, int n) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const MarchEnt &t = tab[e];
  const int l = b - start[e];
  march_body<K, MODE_PART>(t.m, l % t.ntiles, l / t.ntiles, 0, 1, threadIdx.x);
}

template <int K>
__global__ __launch_bounds__(64) void batch_hom_resid_kernel(const HomResidEnt *__restrict__ tab,
                                                             const int *__restrict__ start, int n) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const HomResidEnt &t = tab[e];
  hom_resid_body<K>(t.m, t.r, t.ntiles, t.has_resid, b - start[e]);
}

__global__ __launch_bounds__(256) void batch_resid_kernel(const ResidArgs *__restrict__ tab,
                                                          const int *__restrict__ start, int n) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  resid_body(tab[e], (long)(b - start[e]) * 256 + threadIdx.x);
}

// Zero fill of n 32-bit words per entry (the correction arrays of a pass).
__global__ __launch_bounds__(256) void batch_fill_kernel(const FillEnt *__restrict__ tab,
                                                         const int *__restrict__ start, int n) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const long t = (long)(b - start[e]) * 256 + threadIdx.x;
  if (t < tab[e].n) tab[e].p[t] = 0u;
}

// X += D over the interior of entries whose problem is still refining
// (act[problem] != 0, or every entry when act is null), then, in a second
// launch, the periodic halo columns.
__global__ __launch_bounds__(256) void batch_axpy_kernel(const AxpyEnt *__restrict__ tab,
                                                         const int *__restrict__ start, int n,
                                                         const int *__restrict__ act) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const AxpyEnt &a = tab[e];
  if (act && !act[a.problem]) return;
  const long t = (long)(b - start[e]) * 256 + threadIdx.x;
  const long np = (long)a.i2 * a.j2;
  if (t >= np * a.nrhs) return;
  const int r = (int)(t / np);
  const long q = t % np;
  const int i = (int)(q % a.i2) + 2, j = (int)(q / a.i2) + 2;
  const long o = f2(a.i2 + 2, i, j);
  a.x[(long)r * a.x_stride + o] += a.d[(long)r * a.d_stride + o];
}

__global__ __launch_bounds__(256) void batch_halo_kernel(const AxpyEnt *__restrict__ tab,
                                                         const int *__restrict__ start, int n,
                                                         const int *__restrict__ act) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const AxpyEnt &a = tab[e];
  if (act && !act[a.problem]) return;
  const long t = (long)(b - start[e]) * 256 + threadIdx.x;
  if (t >= (long)a.nrhs * a.j2) return;
  const int r = (int)(t / a.j2), j = (int)(t % a.j2) + 2;
  const int i0 = a.i2 + 2;
  double *xr = a.x + (long)r * a.x_stride;
  xr[f2(i0, 1, j)] = xr[f2(i0, i0 - 1, j)];
  xr[f2(i0, i0, j)] = xr[f2(i0, 2, j)];
}

static bool batch_grid_ok(int n, long nwg) { return n >= 1 && nwg >= 1 && nwg <= 0x7fffffffL; }

hipError_t launch_batch_march(int K, const MarchEnt *tab, const int *start, int n, long nwg,
                              hipStream_t st) {
  if (!batch_grid_ok(n, nwg)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nwg);
  if (K == 1) hipLaunchKernelGGL((batch_march_kernel<1>), grid, dim3(64), 0, st, tab, start, n);
  else if (K == 2) hipLaunchKernelGGL((batch_march_kernel<2>), grid, dim3(64), 0, st, tab, start, n);
  else if (K == 4) hipLaunchKernelGGL((batch_march_kernel<4>), grid, dim3(64), 0, st, tab, start, n);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_batch_hom_resid(int K, const HomResidEnt *tab, const int *start, int n, long nwg,
                                  hipStream_t st) {
  if (!batch_grid_ok(n, nwg)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nwg);
  if (K == 1) hipLaunchKernelGGL((batch_hom_resid_kernel<1>), grid, dim3(64), 0, st, tab, start, n);
  else if (K == 2) hipLaunchKernelGGL((batch_hom_resid_kernel<2>), grid, dim3(64), 0, st, tab, start, n);
  else if (K == 4) hipLaunchKernelGGL((batch_hom_resid_kernel<4>), grid, dim3(64), 0, st, tab, start, n);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_batch_resid(const ResidArgs *tab, const int *start, int n, long nwg, hipStream_t st) {
  if (!batch_grid_ok(n, nwg)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_resid_kernel, dim3((unsigned)nwg), dim3(256), 0, st, tab, start, n);
  return hipGetLastError();
}

hipError_t launch_batch_fill(const FillEnt *tab, const int *start, int n, long nwg, hipStream_t st) {
  if (!batch_grid_ok(n, nwg)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_fill_kernel, dim3((unsigned)nwg), dim3(256), 0, st, tab, start, n);
  return hipGetLastError();
}

hipError_t launch_batch_axpy(const AxpyEnt *tab, const int *start, int n, long nwg, const int *act,
                             bool halo, hipStream_t st) {
  if (!batch_grid_ok(n, nwg)) return hipErrorInvalidValue;
  if (halo) hipLaunchKernelGGL(batch_halo_kernel, dim3((unsigned)nwg), dim3(256), 0, st, tab, start, n, act);
  else hipLaunchKernelGGL(batch_axpy_kernel, dim3((unsigned)nwg), dim3(256), 0, st, tab, start, n, act);
  return hipGetLastError();
}

}  // namespace sevp
