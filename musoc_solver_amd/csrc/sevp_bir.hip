// Plain FP64 strip SEVP (REPBIR) and the BIR multiplicative strip sweeps.
//
// One wavefront owns one right-hand side and runs the whole chain the
// reference runs on one CPU core: nsweep sweeps x NBIR overlapping strips x
// (forward + backward block march).  The chain is a true data dependency
// (each strip reads the halo columns and interior guess written by the one
// before it, and each block's march starts from the corrected last row of the
// block below), so the only parallelism inside one solve is the strip's
// columns: lane l marches strip column l+1.  Independent right-hand sides run
// as independent wavefronts.
//
// The march keeps the two live rows in registers and exchanges the stencil's
// east/west neighbours with DPP wave shifts.  The influence-matrix columns a
// lane needs (column n of RINV/RINV1 for lane n) are loaded into registers at
// the start of the block, so their latency hides behind the march.  The
// reference's scratch arrays H, DUM1 and DUM2 live in registers; DUM0 (the
// saved last rows) is per-RHS global scratch.
//
// Arithmetic follows MAIN/template/GLOBAL/src/sevp.f90:218-328 operation for
// operation (see sevp_dev.h), so results are bitwise equal to the reference.
#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

struct StripCtx {
  // Coefficients: value for lane's column at row j is base[col + (j-1)*ld].
  const float *ax, *ay, *bb, *cx, *cy;
  long ccol;
  int cld;
  // Right-hand side F, same convention.
  const float *f;
  long fcol;
  int fld;
  // Strip view of X: X(i,j) = xv[(i-1) + (j-1)*ldx], i = 1..w+2.
  double *xv;
  int ldx;
  // Influence matrices of this strip: RINV(w,w,nblk), RINV1(w,w,nblk-1).
  const double *rinv, *rinv1;
  double *dum0;  // (w, nblk-1)
  const int *ie;
  int nblk;
  int w;
};

template <int WMAX>
__device__ __forceinline__ void load_col(double (&r)[WMAX], const double *mat, int w, int n) {
#pragma unroll
  for (int m = 0; m < WMAX; ++m) r[m] = (m < w) ? mat[m + (long)n * w] : 0.0;
}

// sum_{m=1..w} e(m) * R(m, n), sequential from zero (lane n holds column n).
template <int WMAX>
__device__ __forceinline__ double dot_col(double e, const double (&r)[WMAX], int w) {
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < WMAX; ++m)
    if (m < w) s = s + read_lane(e, m) * r[m];
  return s;
}

#define LD_C(p, j) ((double)(p)[c.ccol + (long)((j) - 1) * c.cld])
#define LD_F(j) ((double)c.f[c.fcol + (long)((j) - 1) * c.fld])
#define XV(i, j) c.xv[(long)((i) - 1) + (long)((j) - 1) * c.ldx]
#define LDC32(p, j) ((p)[c.ccol + (long)((j) - 1) * c.cld])

// REPBIR on one strip for the calling wavefront.
//
// Each march reads its row operands from a ring of D rows loaded D rows
// ahead, so no row waits for a load issued in the row before it.  Measured
// at 720x360 with 24-row blocks, rings of 2, 4 and 8 rows all take about
// 0.35 us per row step (a one-row prefetch took 0.43 us): the row's own
// dependent chain (wave shifts, four products and differences, the
// correctly rounded FP64 division) sets the pace; with 2-row blocks the
// shallower rings are faster, and D = 4 is used throughout.  The forward pass
// consumes rows 1 .. jlast strictly in order (each a march row of its block
// or the block's closure row), so one ring runs across the block boundaries;
// the next block's RINV1 columns and first guess row are requested at each
// closure, a whole block before they are needed.  The backward H march fills
// a ring per block, after the block's restore of row IE (which its last row
// reads: same lane, so program order makes the store visible).
template <int WMAX, int D>
__device__ __forceinline__ void repbir_wave(const StripCtx &c) {
  const int lane = threadIdx.x;
  const int w = c.w;
  const bool act = lane < w;
  const int li = act ? lane + 2 : 2;  // this lane's X column (1-based, strip view)
  const int wi = w + 2;               // east halo column
  const int n = act ? lane : 0;
  const int nblk = c.nblk;
  double rc[WMAX], rc2[WMAX];

  // ---------------- forward ----------------
  const int ieb_last = c.ie[nblk - 1];
  const int js_last = nblk > 1 ? c.ie[nblk - 2] : 1;
  const int jlast = (ieb_last - 2 >= js_last) ? ieb_last - 2 : js_last - 1;
  float fa[D], fb[D], fc[D], fd[D], fe[D], ff_[D];
  double fhw[D], fhe[D];
#define FFETCH(k, jj)                                                                  \
  do {                                                                                 \
    const int j_ = (jj);                                                               \
    fa[k] = LDC32(c.ax, j_); fb[k] = LDC32(c.ay, j_); fc[k] = LDC32(c.bb, j_);         \
    fd[k] = LDC32(c.cx, j_); fe[k] = LDC32(c.cy, j_);                                  \
    ff_[k] = c.f[c.fcol + (long)(j_ - 1) * c.fld];                                     \
    fhw[k] = XV(1, j_ + 1); fhe[k] = XV(wi, j_ + 1);                                   \
  } while (0)
#pragma unroll
  for (int k = 0; k < D; ++k) FFETCH(k, min(1 + k, max(jlast, 1)));
  int nb = 1, ieb = c.ie[0];
  if (nblk > 1) load_col<WMAX>(rc, c.rinv1, w, n);
  double xs = XV(li, 1), xc = XV(li, 2);
  double xnn = (nblk > 1) ? XV(li, ieb + 1) : 0.0;
  for (int j0 = 1; j0 <= jlast; j0 += D) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int j = j0 + k;
      if (j <= jlast) {
        const double a = fa[k], b = fb[k], cc = fc[k], d = fd[k], e = fe[k], ff = ff_[k];
        const double hw = fhw[k], he = fhe[k];
        FFETCH(k, min(j + D, jlast));
        double xw = wave_shr1(xc, hw);
        double xe = wave_shl1(xc, he);
        if (lane == w - 1) xe = he;
        if (j == ieb - 1) {
          // closure residual of block nb at row ieb-1, correction of row ieb
          const double r = ff - a * xw - b * xs - cc * xc - d * xe - e * xnn;
          const double dd = dot_col<WMAX>(r, rc, w);
          if (act) c.dum0[n + (long)(nb - 1) * w] = xc;
          const double xcor = xc - dd;
          if (act) XV(li, ieb) = xcor;
          xs = xcor;
          xc = xnn;
          ++nb;
          ieb = c.ie[nb - 1];
          if (nb < nblk) {
            load_col<WMAX>(rc, c.rinv1 + (long)(nb - 1) * w * w, w, n);
            xnn = XV(li, ieb + 1);
          }
        } else {
          const double xn = (ff - a * xw - b * xs - cc * xc - d * xe) / e;
          if (act && j + 2 < ieb) XV(li, j + 2) = xn;
          xs = xc;
          xc = xn;
        }
      }
    }
  }
  if (act) XV(li, ieb) = xc;  // the last block's row IE(NBLK)
#undef FFETCH

  // ---------------- backward ----------------
  float ba[D], bb_[D], bc[D], bd[D], be[D];
  double bxo[D];
#define BFETCH(k, jj)                                                                  \
  do {                                                                                 \
    const int j_ = (jj);                                                               \
    ba[k] = LDC32(c.ax, j_); bb_[k] = LDC32(c.ay, j_); bc[k] = LDC32(c.bb, j_);        \
    bd[k] = LDC32(c.cx, j_); be[k] = LDC32(c.cy, j_); bxo[k] = XV(li, j_ + 2);         \
  } while (0)
  for (int nbs = 1; nbs <= nblk; ++nbs) {
    const int nbk = nblk - nbs + 1;
    const int iebk = c.ie[nbk - 1];
    const int js = (nbk != 1) ? c.ie[nbk - 2] : 1;
    const int jf = iebk - 2;
    load_col<WMAX>(rc, c.rinv + (long)(nbk - 1) * w * w, w, n);
    if (nbk != 1) load_col<WMAX>(rc2, c.rinv1 + (long)(nbk - 2) * w * w, w, n);
    double xrow;  // X(li, ieb) after the restore
    if (nbk != nblk) {
      xrow = c.dum0[n + (long)(nbk - 1) * w];
      if (act) XV(li, iebk) = xrow;
    } else {
      xrow = XV(li, iebk);
    }
    // the H march's operands, requested with the closure's
    if (jf >= js) {
#pragma unroll
      for (int k = 0; k < D; ++k) BFETCH(k, min(js + k, jf));
    }
    const int jr = iebk - 1;
    const double xsr = XV(li, jr), xnn2 = XV(li, iebk + 1);
    const double hw = XV(1, iebk), he = XV(wi, iebk);
    double xw = wave_shr1(xrow, hw);
    double xe = wave_shl1(xrow, he);
    if (lane == w - 1) xe = he;
    const double r = LD_F(jr) - LD_C(c.ax, jr) * xw - LD_C(c.ay, jr) * xsr -
                     LD_C(c.bb, jr) * xrow - LD_C(c.cx, jr) * xe - LD_C(c.cy, jr) * xnn2;
    const double cv = dot_col<WMAX>(r, rc, w);
    const double xf = XV(li, js + 1);
    if (act) XV(li, js + 1) = xf + cv;
    double hs = 0.0;
    if (nbk != 1) {
      const double g = cv * LD_C(c.cy, js - 1);
      hs = dot_col<WMAX>(g, rc2, w);
    }
    double hc = cv;
    for (int j0 = js; j0 <= jf; j0 += D) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const int j = j0 + k;
        if (j <= jf) {
          const double a = ba[k], b = bb_[k], cc = bc[k], d = bd[k], e = be[k], xo = bxo[k];
          BFETCH(k, min(j + D, jf));
          double hwv = wave_shr1(hc, 0.0);
          double hev = wave_shl1(hc, 0.0);
          if (lane == w - 1) hev = 0.0;
          const double hn = (-(a * hwv) - b * hs - cc * hc - d * hev) / e;
          if (act) XV(li, j + 2) = xo + hn;
          hs = hc;
          hc = hn;
        }
      }
    }
  }
#undef BFETCH
}
#undef LDC32
#undef XV

// Whole FS_SOLVER BIR loop (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1902-1966)
// for one RHS per wavefront.
template <int WMAX>
__global__ __launch_bounds__(64) void bir_sweeps_kernel(BirArgs p) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int w = p.ibir, inc = p.ibir / 2;
  const int i0 = p.i2 + 2, m0 = w + 2;
  const int j1 = p.j2 + 1;
  double *x = p.x + (long)r * p.x_stride;
  double *xx = p.xx + (long)r * p.xx_stride;
  const float *s = p.s + (long)r * p.s_stride;
  StripCtx c;
  c.ax = p.al; c.ay = p.ab; c.bb = p.ac; c.cx = p.ar; c.cy = p.at;
  c.cld = p.i2;
  c.f = s;
  c.fld = p.i2;
  c.ie = p.ie;
  c.nblk = p.nb0;
  c.w = w;
  c.dum0 = p.dum0 + (long)r * p.dum0_stride;
  const long ibr2 = (long)w * w;
  for (int it = 0; it < p.nsweep; ++it) {
    for (int n = 1; n <= p.nbir; ++n) {
      const int il = (n - 1) * inc + 1;
      c.rinv = p.rinv + (long)(n - 1) * ibr2 * p.nb0;
      c.rinv1 = p.rinv1 + (long)(n - 1) * ibr2 * (p.nb0 - 1);
      const int lc = lane < w ? lane : 0;
      if (n != p.nbir) {
        c.ccol = il - 1 + lc;
        c.fcol = c.ccol;
        c.xv = x + f2(i0, il, 1);
        c.ldx = i0;
        repbir_wave<WMAX, 4>(c);
        __threadfence_block();
      } else {
        // seam strip: halo columns of XX for rows 1..J2 as the reference does
        // (row J1 keeps its old value), or 1..J1 with seam_fix
        const int jg = p.seam_fix ? j1 : p.j2;
        for (int j = 1 + lane; j <= jg; j += SEVP_WAVE) {
          xx[f2(m0, 1, j)] = x[f2(i0, il, j)];
          xx[f2(m0, w + 2, j)] = x[f2(i0, inc + 2, j)];
        }
        __threadfence_block();
        c.ccol = (lc < inc) ? (il - 1 + lc) : (lc - inc);
        c.fcol = c.ccol;
        c.xv = xx;
        c.ldx = m0;
        repbir_wave<WMAX, 4>(c);
        __threadfence_block();
        for (long t = lane; t < (long)w * (j1 - 1); t += SEVP_WAVE) {
          const int i = (int)(t % w) + 1, j = (int)(t / w) + 2;
          const double v = xx[f2(m0, i + 1, j)];
          if (i <= inc) x[f2(i0, il + i, j)] = v;
          else x[f2(i0, i - inc + 1, j)] = v;
        }
        __threadfence_block();
        if (p.periodic) {
          for (int j = 2 + lane; j <= j1; j += SEVP_WAVE) {
            x[f2(i0, 1, j)] = x[f2(i0, i0 - 1, j)];
            x[f2(i0, i0, j)] = x[f2(i0, 2, j)];
          }
          __threadfence_block();
        }
      }
    }
  }
}

// One REPBIR call with the reference's argument conventions, batched over
// independent right-hand sides (F and X strided by f_stride / x_stride).
template <int WMAX>
__global__ __launch_bounds__(64) void repbir_kernel(RepbirArgs p) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  StripCtx c;
  c.ax = p.ax; c.ay = p.ay; c.bb = p.bb; c.cx = p.cx; c.cy = p.cy;
  c.ccol = lane < p.m2 ? lane : 0;
  c.cld = p.i2;
  c.f = p.f + (long)r * p.f_stride;
  c.fcol = c.ccol;
  c.fld = p.m2;
  c.xv = p.x + (long)r * p.x_stride;
  c.ldx = p.i0;
  c.rinv = p.rinv;
  c.rinv1 = p.rinv1;
  c.dum0 = p.dum0 + (long)r * p.dum0_stride;
  c.ie = p.ie;
  c.nblk = p.nblk;
  c.w = p.m2;
  repbir_wave<WMAX, 4>(c);
}

hipError_t launch_bir_sweeps(const BirArgs &a, int nrhs, hipStream_t st) {
  if (a.ibir <= 32)
    hipLaunchKernelGGL(bir_sweeps_kernel<32>, dim3(nrhs), dim3(64), 0, st, a);
  else
    hipLaunchKernelGGL(bir_sweeps_kernel<64>, dim3(nrhs), dim3(64), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_repbir(const RepbirArgs &a, int nrhs, hipStream_t st) {
  if (a.m2 <= 32)
    hipLaunchKernelGGL(repbir_kernel<32>, dim3(nrhs), dim3(64), 0, st, a);
  else
    hipLaunchKernelGGL(repbir_kernel<64>, dim3(nrhs), dim3(64), 0, st, a);
  return hipGetLastError();
}

}  // namespace sevp
