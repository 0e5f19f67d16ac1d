// Host/device argument blocks and launchers shared by the SEVP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "sevp_records.h"  // the strip kernels' record order (bir_record_table, BirKind)

namespace sevp {

// Record msg as the calling thread's sevp_last_error text; returns code.
int set_error(int code, const std::string &msg);

// Largest block count of the strip kernels (REPBIR, BIR sweeps): their block
// rows IE are held in LDS for the whole launch.
constexpr int BIR_MAX_BLOCKS = 4096;

// BIR multiplicative sweeps, one workgroup per right-hand side.
struct BirArgs {
  const float *al, *ab, *ac, *ar, *at;  // (i2, j2)
  const float *s;                        // (i2, j2) per RHS
  const double *rinv, *rinv1;            // (ibir^2, nb0, nbir), (ibir^2, nb0-1, nbir)
  double *x;                             // (i2+2, j2+2) per RHS
  double *xx;                            // (ibir+2, j2+2) per RHS, persistent seam scratch
  double *dum0;                          // (ibir, nb0-1) per RHS
  const int *ie;                         // (nb0)
  const int *recs;                       // bir_record_table, fwd_records forward entries first
  int fwd_records;
  int i2, j2, nbir, ibir, nb0, nsweep, periodic, seam_fix;
  long s_stride, x_stride, xx_stride, dum0_stride;
};

// One REPBIR call with the reference's argument conventions.
struct RepbirArgs {
  const float *ax, *ay, *bb, *cx, *cy;  // ld i2, already offset to the strip
  const float *f;                        // ld m2
  double *x;                             // ld i0, already offset to the strip
  const double *rinv, *rinv1;
  double *dum0;
  const int *ie;
  const int *recs;  // bir_record_table
  int fwd_records;
  int i0, i2, m2, nblk;
  long f_stride, x_stride, dum0_stride;
};

// Column selection of a march over a strip of the coefficient arrays.
// Physical coefficient column (0-based) of strip column c (1-based, 1..w):
//   seam == 0 : il - 1 + (c - 1)
//   seam == 1 : c <= inc ? il - 1 + (c - 1) : c - inc - 1
struct ColMap {
  int il, inc, seam;
  __host__ __device__ int phys(int c) const {
    return (seam && c > inc) ? (c - inc - 1) : (il - 2 + c);
  }
};

// A correction held as split-K partial products of the mixed-precision
// GEMM: value(m, n) = sum over slices s of P[s][m][n], added in FP64 in slice
// order (so every reader of one element computes the same double).
struct PartRef {
  const float *p;   // FP32 slice sums; null with pd null: no correction
  const double *pd; // FP64 slice sums (the FP64 MFMA products), else null
  long sstride;     // elements per slice (M * ld)
  int ns;           // number of slices
  int ld;           // N of the product (elements per right-hand side)
  // eight loads in flight per group; the additions stay in slice order
  __device__ double sum(int m, int n) const {
    double acc = 0.0;
    if (pd) {
      const double *q = pd + (long)m * ld + n;
      for (int s0 = 0; s0 < ns; s0 += 8) {
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = q[(long)min(s0 + i, ns - 1) * sstride];
#pragma unroll
        for (int i = 0; i < 8; ++i)
          if (s0 + i < ns) acc += v[i];
      }
      return acc;
    }
    const float *q = p + (long)m * ld + n;
    for (int s0 = 0; s0 < ns; s0 += 8) {
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = q[(long)min(s0 + i, ns - 1) * sstride];  // no branch
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (s0 + i < ns) acc += (double)v[i];
    }
    return acc;
  }
  __device__ bool on() const { return p || pd; }
};

// Full-width (or strip) march over one block, batched over independent
// right-hand sides.  Mode selects the recurrence:
//   MODE_PART : X(j+2) = (F - AX Xw - AY Xs - BB Xc - CX Xe) / CY        (REP1/REP forward)
//   MODE_HOM  : H(j+2) = (-AX Hw - AY Hs - BB Hc - CX He) / CY, X += H    (REP1/REP backward)
//   MODE_PRE  : H(j+2) = -(AX Hw + AY Hs + BB Hc + CX He) / CY            (PRE/PRE1)
enum { MODE_PART = 0, MODE_HOM = 1, MODE_PRE = 2 };

struct MarchArgs {
  const float *ax, *ay, *bb, *cx, *cy;  // (cld, j2)
  int cld;
  const float *f;   // MODE_PART: (i2, j2) per RHS, ld i2
  const double *f64;  // MODE_PART: FP64 right-hand side used instead of f when non-null
  long f_stride;
  double *x;        // (i2+2, j2+2) per RHS (MODE_PART / MODE_HOM)
  long x_stride;
  // MODE_HOM: starting rows H(js) (hcpl, may be null => 0) and H(js+1) (hfree): (nrhs, i2)
  const double *hcpl, *hfree;
  // MODE_PRE: coupling rows RINV1(g, :, b-1) (w x w col-major) scaled by CY(g, js-1)
  const double *rinv1_prev;  // null for the first block
  // MODE_PRE outputs: RINV(g, i, b) and T(g, n) (both w x w col-major)
  double *rinv_b, *tmat;
  long rinv_sstride, rinv1_sstride;  // MODE_PRE: per-strip strides of RINV and RINV1
  // residual out (MODE_PART, when resid != 0): (nrhs, i2); e32 (may be null)
  // receives the same values rounded to FP32 for the mixed-precision product
  double *e;
  float *e32;
  int e_lt;         // nonzero: e / e32 are in lane tiles (lt_index, I2 columns), else (nrhs, i2)
  int i2, j2, periodic, js, jf, ieb, resid;
  // mixed precision, split-K corrections read directly from the partials:
  //   MODE_PART: subtracted from row js (the row the previous block corrected;
  //              the X array keeps the uncorrected value)
  //   MODE_HOM : H(js+1) at column n, and, when cpl_off >= 0, H(js) at
  //              column cpl_off + n (used instead of hfree / hcpl)
  PartRef corr;
  int cpl_off;
  int w;            // number of columns marched (i2 for full width)
  int tile_out;     // output columns per wave
  int halo;         // extra columns each side
  ColMap map;       // coefficient column map (identity for full width)
  int nrhs;
};

struct ResidArgs {  // backward restore + closure residual, one thread per (rhs, column)
  const float *ax, *ay, *bb, *cx, *cy;
  const float *f;
  const double *f64;   // used instead of f when non-null
  long f_stride;
  double *x;
  long x_stride;
  const double *dum0;  // (i2, nb0-1) per RHS, null for the top block
  long dum0_stride;
  double *e;           // (nrhs, i2), may be null
  float *e32;          // FP32 copy of e (may be null)
  int i2, j2, periodic, nb, ieb, nrhs;
  // mixed precision backward: X(apply_row) += corr (the free-row correction
  // of the block above) before the residual, which reads that row as its
  // north neighbour; do_resid == 0 applies the correction only
  PartRef corr;
  int apply_row, do_resid;
};

enum { EP_FWD = 0, EP_BWD = 1, EP_CPL = 2, EP_NEGSUB = 3, EP_COLMAJ = 4 };

struct GemmArgs {  // D(m, n) = sum_k A(m, k) B(k, n), sequential k
  const double *a;
  long a_sm, a_sk;
  const double *b;
  long b_sk, b_sn;
  int M, N, K;
  int ep;
  // EP_FWD / EP_BWD : X row update, per RHS m
  double *x;
  long x_stride;
  int xrow;           // 1-based X row updated (ieb for FWD, js+1 for BWD)
  double *dum0;       // EP_FWD: DUM0(:, nb) per RHS
  long dum0_stride;
  double *out;        // EP_BWD: hfree (M, N); EP_CPL: hcpl (M, N); EP_NEGSUB: (ld out_ld)
  long out_ld;
  double *g;          // EP_BWD: coupling input g = c * CY(n, cyrow) (may be null)
  const float *cy;
  int cyrow;          // 1-based coefficient row of CY for the coupling
  int i2;
  unsigned long long *tspan;  // profiling: kernel span slot (null: off)
};

// Split-K partial products of the mixed-precision corrections:
// P[s][m][n] = sum over k in slice s of A(m, k) B(k, n) in FP32 (MFMA).
struct GemmPartArgs {
  const float *a;   // (M, K) row-major, ld K
  const float *b;   // (K, nsplit) in lane tiles (launch_cvt_lane_tiles)
  const float *b2;  // columns n >= nsplit: lane tiles, column n - nsplit (may be null)
  int nsplit;       // a multiple of 16 when b2 is set
  int M, N, K;
  float *part;      // (S, M, N)
  unsigned long long *tspan;  // profiling: kernel span slot (null: off)
};
// The FP64 counterpart: A the FP64 error rows, B FP64 lane tiles
// (launch_cvt_lane_tiles64), FP64 slice sums.
struct GemmPart64Args {
  const double *a;
  const double *b;
  const double *b2;
  int nsplit;
  int M, N, K;
  double *part;  // (S, M, N)
  unsigned long long *tspan;
};
hipError_t launch_gemm_part64(const GemmPart64Args &a, hipStream_t st, hipEvent_t ev0 = nullptr,
                              hipEvent_t ev1 = nullptr);
long gemm_part64_workgroups(const GemmPart64Args &a);
hipError_t launch_cvt_lane_tiles64(const double *src, double *dst, int K, int N, long nmat,
                                   hipStream_t st);
// Slice length and number of slices used for an (N, K) product.
int gemm_part_slices(int N, int K, int *kslice);
// ev0 / ev1 (optional): events recorded with the kernel's own start and end
// timestamps (hipExtLaunchKernelGGL), for profiling outside graph capture.
hipError_t launch_gemm_part(const GemmPartArgs &a, hipStream_t st, hipEvent_t ev0 = nullptr,
                            hipEvent_t ev1 = nullptr);
long gemm_part_workgroups(const GemmPartArgs &a);  // grid size of launch_gemm_part
// workgroups of one entry of a grouped product launch
long batch_gemm_part_workgroups(const GemmPartArgs &a);
long batch_gemm_part64_workgroups(const GemmPart64Args &a);

// One launch per block step (sevp_fused.hip): a product whose workgroups
// sum whole dot products (W wavefronts split K, FP64 sums of their partial
// tiles in wavefront order) with the work that depends on it in the
// epilogue.  Both operands are in lane tiles (lt_index): A the error rows of
// one RHS group (FP32, or FP64 for the FP64 MFMA; rows past M and columns
// past I2 of the last tiles are finite, zero until written), B an influence
// matrix (the split-K copies: FP32, FP64 or BF16 lane tiles).
constexpr int FUSED_WAVES = 16;  // the default; 8 is the other compiled choice

// Forward step: d = e_b RINV1_b, then block b+1's march from the corrected
// south row X(js) - d and its closure residual e_(b+1).
struct FusedFwdArgs {
  const void *a;      // e_b, lane tiles of (M, i2)
  const void *b;      // RINV1_b, lane tiles of (i2, i2)
  int M, i2, periodic;
  // block b+1: march rows js .. jf (steps = jf - js + 1), closure row ieb - 1
  int js, jf, ieb, steps;
  const float *ax, *ay, *bb, *cx, *cy;  // (i2, j2)
  const float *f;      // right-hand sides (i2, j2) per RHS, or
  const double *f64;   // FP64 right-hand sides (the refinement residual)
  long f_stride;
  double *x;           // (i2+2, j2+2) per RHS
  long x_stride;
  double *e_out;       // e_(b+1), FP64 lane tiles (may be null)
  float *e32_out;      // e_(b+1) rounded to FP32, lane tiles (may be null)
  double *d_out;       // d itself (M, i2) row-major, kept when block b+1 is one row high (may be null)
  int waves;           // wavefronts per workgroup (8 or 16; 0: FUSED_WAVES)
  int rows;            // 16-row tiles per workgroup (1 or 2; 0: by arithmetic)
  int ntiles, nmb, mbw;  // set by the launcher: column tiles, row-tile groups, row tiles per group
  unsigned long long *tspan;  // profiling: kernel span slot (null: off)
};

// Backward step for block b: [c | cpl] = e_b [RINV_b | W_b]; X(js+1) += c;
// closure residual of block b-1 (do_resid); c and cpl stored for the
// homogeneous march of block b, which the next backward launch runs (hom).
struct FusedBwdArgs {
  const void *a;       // e_b, lane tiles
  const void *b;       // RINV_b, lane tiles
  const void *bw;      // W_b, lane tiles (null for the first block)
  int M, i2, periodic;
  int js;              // JS of block b (IE(b-1), or 1)
  int do_resid;        // b > 1: closure residual of block b-1 at row js-1
  const float *ax, *ay, *bb, *cx, *cy;
  const float *f;
  const double *f64;
  long f_stride;
  double *x;
  long x_stride;
  double *e_out;       // e_(b-1), FP64 lane tiles (may be null)
  float *e32_out;      // FP32 lane tiles (may be null)
  // block b-1 one row high: its closure row's south neighbour is row
  // IE(b-2), which REP1 still holds corrected by d_(b-2) (restored from DUM0
  // only at block b-2's step); d_(b-2) (M, i2), else null
  const double *dsouth;
  double *hfree_out, *hcpl_out;  // c and cpl of block b (M, i2) row-major
  MarchArgs hom;       // MODE_HOM march of block b+1 (hfree / hcpl of the previous launch)
  int has_hom;
  int waves;           // wavefronts per workgroup (8 or 16; 0: FUSED_WAVES)
  int rows;            // 16-row tiles per workgroup (1 or 2; 0: by arithmetic)
  int ntiles, nmb, mbw, nprod_half, hom_tiles, nhom_wg;  // set by the launcher
  unsigned long long *tspan;
};
// Elements of the lane tiles of an (m, k) operand: 16-row by 16-column
// blocks of 256 elements.
__host__ __device__ inline long lane_tile_elems(int m, int k) {
  return (long)((m + 15) / 16) * ((k + 15) / 16) * 256;
}
int fused_tiles_fwd(int i2, int steps);
// Set the launcher's fields (tiles, row blocks; K: the march lane width of
// the backward launch's homogeneous march); nonzero if the shape is refused.
int fused_fwd_prepare(FusedFwdArgs &a, int arith);
int fused_bwd_prepare(FusedBwdArgs &a, int arith);
long fused_fwd_workgroups(const FusedFwdArgs &a);  // after the prepare call
long fused_bwd_workgroups(const FusedBwdArgs &a);
// Product arithmetic of the fused kernels: FP32 operands (FP32 MFMA), FP64
// (FP64 MFMA), or FP32 error rows against BF16 matrices, the rows rounded to
// BF16 in registers (BF16 MFMA, FP32 accumulation).
enum { FUSED_F32 = 0, FUSED_F64 = 1, FUSED_B16 = 2 };
// ev0 / ev1 (optional): events with the kernel's own start and end times.
hipError_t launch_fused_fwd(const FusedFwdArgs &a, int arith, hipStream_t st,
                            hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_fused_bwd(const FusedBwdArgs &a, int arith, hipStream_t st,
                            hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// FP64 column-major (K, N) matrices to BF16 lane tiles (round to nearest even via FP32).
hipError_t launch_cvt_lane_tiles_bf16(const double *src, unsigned short *dst, int K, int N, long nmat,
                                     hipStream_t st);

// Step 14 of FS_GLOBAL (the barotropic correction around the solve); arrays
// have the reference shapes, Fortran order (see include/sevp.h).
struct Step14Args {
  int i2, j2, k1;
  int mask, lopenw, harness, free_surface, periodic_we;
  float odt;
  const short *kb, *iu0, *iv0, *iu, *iv;
  const float *odx, *ody, *ocs, *csv, *odyv, *odz;
  float *u, *v, *w, *p0, *scr, *s;
  const double *x;
  float *tmp_part;  // step14_partials(i2, j2) per-workgroup maxima
};
long step14_partials(int i2, int j2);
// masking (if mask), harness override (if harness), W integration and S
hipError_t launch_step14_prepare(const Step14Args &p, hipStream_t st);
// P0, U, V, SCR corrections; the largest top-layer land speed to *tmp_dev
hipError_t launch_step14_correct(const Step14Args &p, float *tmp_dev, hipStream_t st);
hipError_t launch_step14_zero_scr(float *scr, long n, hipStream_t st);

hipError_t launch_bir_sweeps(const BirArgs &a, int nrhs, hipStream_t st);
hipError_t launch_repbir(const RepbirArgs &a, int nrhs, hipStream_t st);
hipError_t launch_march(int mode, const MarchArgs &a, int nbatch, hipStream_t st);
hipError_t launch_resid(const ResidArgs &a, hipStream_t st);
// backward H march of one block fused with the restore + residual of the
// block below (r may be null for the bottom block)
hipError_t launch_hom_resid(const MarchArgs &m, const ResidArgs *r, hipStream_t st);
hipError_t launch_gemm(const GemmArgs &a, hipStream_t st, hipEvent_t ev0 = nullptr,
                       hipEvent_t ev1 = nullptr);
long gemm_workgroups(const GemmArgs &a);  // grid size of launch_gemm
hipError_t launch_cvt_f32(const double *src, float *dst, long n, hipStream_t st);
// FP32 "lane tiles" of nmat consecutive (K x N) column-major FP64 matrices
// (the B layout of launch_gemm_part); lane_tiles_floats is one matrix's size.
long lane_tiles_floats(int K, int N);
hipError_t launch_cvt_lane_tiles(const double *src, float *dst, int K, int N, long nmat,
                                 hipStream_t st);
hipError_t launch_matinv(double *mats, long mat_stride, int n, int count, int *zero_pivots,
                         double *scratch, hipStream_t st);
bool matinv_uses_lds(int n);
size_t matinv_blocked_scratch(int n, int count);
hipError_t launch_matinv_blocked(double *mats, long ms, int n, int count, int *zp,
                                 double *scratch, hipStream_t st);
hipError_t launch_scale_cols(const double *r, const float *cy, int cyrow, int i2, double *t,
                             hipStream_t st);
hipError_t launch_halo_fix(double *x, long x_stride, int i2, int j2, int nrhs, hipStream_t st);

// ------------------------------- grouped launches (many problems per launch)
// A grouped launch takes a table of n entries and the first workgroup of
// each (start[0] = 0, increasing, every entry at least one workgroup);
// nwg is the total.  Each entry runs with the workgroups and arguments its
// own launch would have.
struct MarchEnt {     // forward march (MODE_PART): ntiles * m.nrhs workgroups of 64
  MarchArgs m;
  int ntiles;
};
struct HomResidEnt {  // backward march + residual: hom_resid_prepare's count, 64 each
  MarchArgs m;
  ResidArgs r;
  int ntiles, has_resid;
};
struct FillEnt {      // zero n words: ceil(n / 256) workgroups of 256
  unsigned int *p;
  long n;
};
struct AxpyEnt {      // x += d over the interior (or the periodic halo fix), problem `problem`
  double *x;
  long x_stride;
  const double *d;
  long d_stride;
  int i2, j2, nrhs, problem;
};
// Tiles of a march with K columns per lane (K <= 0: the launch's own choice);
// sets tile_out and halo, 0 when K is too narrow for the block.
int march_tiles(MarchArgs &a, int K);
int march_lanes(int w, int steps);  // the single launch's K (-1: none fits)
// Fills ent for the backward step (launch_hom_resid's arguments at lane
// width K); returns the workgroup count (-1: K too narrow).
long hom_resid_prepare(const MarchArgs &a, const ResidArgs *r, int K, HomResidEnt &ent);
hipError_t launch_batch_march(int K, const MarchEnt *tab, const int *start, int n, long nwg,
                              hipStream_t st);
hipError_t launch_batch_hom_resid(int K, const HomResidEnt *tab, const int *start, int n, long nwg,
                                  hipStream_t st);
hipError_t launch_batch_resid(const ResidArgs *tab, const int *start, int n, long nwg, hipStream_t st);
hipError_t launch_batch_fill(const FillEnt *tab, const int *start, int n, long nwg, hipStream_t st);
// act[problem] != 0 selects the entries applied (act null: all); halo: the
// periodic halo fix instead
hipError_t launch_batch_axpy(const AxpyEnt *tab, const int *start, int n, long nwg, const int *act,
                             bool halo, hipStream_t st);
// ks: the shared slice length; mb: 1 when every M <= 16, else 2 (32-row groups)
hipError_t launch_batch_gemm_part(int ks, int mb, const GemmPartArgs *tab, const int *start, int n,
                                  long nwg, hipStream_t st);
hipError_t launch_batch_gemm_part64(int ks, int mb, const GemmPart64Args *tab, const int *start,
                                    int n, long nwg, hipStream_t st);

// ------------------------------------------------ Krylov solvers (sevp_krylov.hip)
constexpr int KRY_MAXB = 8192;  // most workgroups of a Krylov kernel (partial sums per dot)

struct KryOp {
  const float *al, *ab, *ac, *ar, *at;  // (i2, j2)
  int i2, j2, periodic;
};

// The vectors of one solve, each (i2+2, j2+2) FP64; p and v alternate.
struct KryVecs {
  const float *b;  // (i2, j2) right-hand side
  double *x, *r, *rh, *p[2], *v[2], *s, *t;
};

// Scalars of a BiCGSTAB or P-CSI solve, resident on the device.
struct KryState {
  double rho, alpha, w, beta, rhn, tp, res, tol, rtol;
  int iter, stop, restart, max_iter;  // stop: 1 inner loop ended, 2 converged at a restart
};

int kry_workgroups(int i2, int j2);
hipError_t kry_bicg_restart(const KryOp &op, const KryVecs &q, KryState *st, double *part, int fast,
                            hipStream_t s);
hipError_t kry_bicg_iter(const KryOp &op, const KryVecs &q, KryState *st, double *part, int fast,
                         hipStream_t s);
hipError_t kry_bnd(double *x, int i2, int j2, hipStream_t s);
hipError_t kry_fill(double *a, long n, double v, hipStream_t s);
hipError_t kry_pcsi_qx(int i2, int j2, int periodic, const double *r, double *q, double *x,
                       const double *omega, const double *cq, double c0, int first,
                       const KryState *st, hipStream_t s);
hipError_t kry_pcsi_r(const KryOp &op, const float *b, const double *x, double *r, KryState *st,
                      double *part, int with_dot, int fast, hipStream_t s);
hipError_t kry_spmv(const KryOp &op, const double *v, double *out, int transposed, hipStream_t s);
hipError_t kry_lincomb(double *out, const double *sv, double alpha, const double *v, double beta,
                       const double *vp, long n, hipStream_t s);
hipError_t kry_scale(double *v, double sig, double *w, double bet, long n, hipStream_t s);
hipError_t kry_dot(int i2, int j2, const double *a, const double *b, double *out, double *part, int fast,
                   hipStream_t s);
// smallest eigenvalue of the symmetric tridiagonal (d, e[1..m-1]), bisection
double kry_tridiag_min_eig(const double *d, const double *e, int m);

}  // namespace sevp
