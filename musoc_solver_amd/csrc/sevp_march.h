// The row march of one block over one tile of columns and one right-hand
// side (or unit vector), on one wavefront: the body of the march kernels in
// sevp_full.hip and of the homogeneous march inside the fused backward
// launch (sevp_fused.hip).  Device code is not linked across translation
// units, so both include it.
#pragma once
#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

template <int K>
__device__ __forceinline__ void nbr(const double (&v)[K], int k, double &w_, double &e_) {
  if (k > 0) w_ = v[k - 1];
  else w_ = wave_shr1(v[K - 1], 0.0);
  if (k < K - 1) e_ = v[k + 1];
  else e_ = wave_shl1(v[0], 0.0);
}

template <int K, int MODE>
__device__ void march_body(const MarchArgs &p, int tile, int bidx, int strip, int nstrips,
                           int lane) {
  // bidx: RHS (PART/HOM) or unit-vector column g-1 (PRE)
  const int w = p.w;
  const int T = p.tile_out;
  const int out_lo = 1 + tile * T;
  const int out_hi = min(out_lo + T, w + 1);  // exclusive
  ColMap map = p.map;
  if (MODE == MODE_PRE) {
    map.il = strip * map.inc + 1;
    map.seam = (map.seam && strip == nstrips - 1) ? 1 : 0;
  }
  const int js = p.js, jf = p.jf, ieb = p.ieb;

  int c[K], pc[K], cofs[K];
  bool out[K], halo_w[K], halo_e[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    c[k] = out_lo - p.halo + lane * K + k;
    if (p.periodic) {
      int m = (c[k] - 1) % w;
      if (m < 0) m += w;
      pc[k] = m + 1;
      halo_w[k] = halo_e[k] = false;
    } else {
      pc[k] = min(max(c[k], 1), w);
      halo_w[k] = c[k] <= 0;
      halo_e[k] = c[k] >= w + 1;
    }
    out[k] = c[k] >= out_lo && c[k] < out_hi;
    cofs[k] = map.phys(pc[k]);
  }
#define COEF(arr, k, j) ((double)(arr)[(long)cofs[k] + (long)((j) - 1) * p.cld])

  double *x = nullptr;
  const float *f = nullptr;
  const double *f64 = nullptr;
  const int i0 = p.i2 + 2;
  if (MODE != MODE_PRE) {
    x = p.x + (long)bidx * p.x_stride;
    if (MODE == MODE_PART) {
      if (p.f64) f64 = p.f64 + (long)bidx * p.f_stride;
      else f = p.f + (long)bidx * p.f_stride;
    }
  }
  // X index (1-based) of a lane's column: halo lanes read the Dirichlet column.
  auto xcol = [&](int k) -> int { return halo_w[k] ? 1 : (halo_e[k] ? i0 : pc[k] + 1); };

  double vs[K], vc[K];
  const int g = bidx + 1;  // PRE: unit column
  double *rinv_b = nullptr, *tmat = nullptr;
  const double *r1p = nullptr;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool h = halo_w[k] || halo_e[k];
    if (MODE == MODE_PART) {
      vs[k] = x[f2(i0, xcol(k), js)];
      vc[k] = x[f2(i0, xcol(k), js + 1)];
    } else if (MODE == MODE_HOM && !p.corr.on()) {
      vs[k] = (h || !p.hcpl) ? 0.0 : p.hcpl[(long)bidx * p.i2 + pc[k] - 1];
      vc[k] = h ? 0.0 : p.hfree[(long)bidx * p.i2 + pc[k] - 1];
    }
  }
  if (MODE == MODE_PRE) {
    rinv_b = p.rinv_b + (long)strip * p.rinv_sstride;
    tmat = p.tmat ? p.tmat + (long)strip * p.rinv_sstride : nullptr;
    r1p = p.rinv1_prev ? p.rinv1_prev + (long)strip * p.rinv1_sstride : nullptr;
    double cyg = 0.0;
    if (r1p) cyg = (double)p.cy[(long)map.phys(g) + (long)(js - 2) * p.cld];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool h = halo_w[k] || halo_e[k];
      vs[k] = (h || !r1p) ? 0.0 : r1p[(g - 1) + (long)(pc[k] - 1) * w] * cyg;
      vc[k] = (!h && pc[k] == g) ? 1.0 : 0.0;
    }
  }

  double ca[K], cb[K], cc[K], cd[K], ce[K], cf[K], cx[K];
  auto load_row = [&](int j, double (&a_)[K], double (&b_)[K], double (&c_)[K],
                      double (&d_)[K], double (&e_)[K], double (&f_)[K], double (&x_)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      a_[k] = COEF(p.ax, k, j);
      b_[k] = COEF(p.ay, k, j);
      c_[k] = COEF(p.bb, k, j);
      d_[k] = COEF(p.cx, k, j);
      e_[k] = COEF(p.cy, k, j);
      if (MODE == MODE_PART) {
        const long fo = (long)(pc[k] - 1) + (long)(j - 1) * p.i2;
        f_[k] = f64 ? f64[fo] : (double)f[fo];
      } else {
        f_[k] = 0.0;
      }
      // next-row X operand: PART halo lanes need the Dirichlet value of row j+2,
      // HOM lanes need X(j+2) to accumulate into.
      if (MODE == MODE_PART) x_[k] = (halo_w[k] || halo_e[k]) ? x[f2(i0, xcol(k), j + 2)] : 0.0;
      else if (MODE == MODE_HOM) x_[k] = x[f2(i0, pc[k] + 1, j + 2)];
      else x_[k] = 0.0;
    }
  };

  const int jlast = jf + 1;  // closure row, operands needed after the march
  load_row(js, ca, cb, cc, cd, ce, cf, cx);
  // the closure residual's north row (the next block's free row) does not
  // depend on the march: load it with the first rows
  double xnorth[K];
  if (MODE == MODE_PART && p.resid) {
#pragma unroll
    for (int k = 0; k < K; ++k) xnorth[k] = x[f2(i0, pc[k] + 1, ieb + 1)];
  }
  // split-K corrections, read after the row loads above are in flight
  if (MODE != MODE_PRE && p.corr.on()) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool h = halo_w[k] || halo_e[k];
      if (MODE == MODE_PART) {
        if (!h) vs[k] = vs[k] - p.corr.sum(bidx, pc[k] - 1);
      } else {
        vs[k] = (h || p.cpl_off < 0) ? 0.0 : p.corr.sum(bidx, p.cpl_off + pc[k] - 1);
        vc[k] = h ? 0.0 : p.corr.sum(bidx, pc[k] - 1);
      }
    }
  }
  for (int j = js; j <= jf; ++j) {
    double na[K], nb_[K], nc[K], nd[K], ne[K], nf[K], nx[K];
    if (j + 1 <= jlast && (MODE != MODE_HOM || j + 1 <= jf)) load_row(j + 1, na, nb_, nc, nd, ne, nf, nx);
    double vn[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double xw, xe;
      nbr<K>(vc, k, xw, xe);
      double r;
      if (MODE == MODE_PART) {
        r = (cf[k] - ca[k] * xw - cb[k] * vs[k] - cc[k] * vc[k] - cd[k] * xe) / ce[k];
        if (halo_w[k] || halo_e[k]) r = cx[k];
      } else if (MODE == MODE_HOM) {
        r = (-(ca[k] * xw) - cb[k] * vs[k] - cc[k] * vc[k] - cd[k] * xe) / ce[k];
        if (halo_w[k] || halo_e[k]) r = 0.0;
      } else {
        r = -((ca[k] * xw + cb[k] * vs[k] + cc[k] * vc[k] + cd[k] * xe) / ce[k]);
        if (halo_w[k] || halo_e[k]) r = 0.0;
      }
      vn[k] = r;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (out[k]) {
        if (MODE == MODE_PART) x[f2(i0, pc[k] + 1, j + 2)] = vn[k];
        else if (MODE == MODE_HOM) x[f2(i0, pc[k] + 1, j + 2)] = cx[k] + vn[k];
      }
      vs[k] = vc[k];
      vc[k] = vn[k];
    }
    if (j + 1 <= jlast && (MODE != MODE_HOM || j + 1 <= jf)) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        ca[k] = na[k]; cb[k] = nb_[k]; cc[k] = nc[k]; cd[k] = nd[k]; ce[k] = ne[k];
        cf[k] = nf[k]; cx[k] = nx[k];
      }
    }
  }
  // row operands now hold the closure row jlast = ieb - 1
  if (MODE == MODE_PART && p.resid) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double xw, xe;
      nbr<K>(vc, k, xw, xe);
      const double r = cf[k] - ca[k] * xw - cb[k] * vs[k] - cc[k] * vc[k] - cd[k] * xe - ce[k] * xnorth[k];
      if (out[k]) {
        const long eo = p.e_lt ? lt_index(bidx, pc[k] - 1, p.e_lt) : (long)bidx * p.i2 + pc[k] - 1;
        if (p.e) p.e[eo] = r;
        if (p.e32) p.e32[eo] = (float)r;
      }
    }
  }
  if (MODE == MODE_PRE) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double xw, xe;
      nbr<K>(vc, k, xw, xe);
      const double r = ca[k] * xw + cb[k] * vs[k] + cc[k] * vc[k] + cd[k] * xe;
      if (out[k]) {
        rinv_b[(g - 1) + (long)(pc[k] - 1) * w] = r;
        if (tmat) tmat[(g - 1) + (long)(pc[k] - 1) * w] = vc[k];
      }
    }
  }
#undef COEF
}

}  // namespace sevp
