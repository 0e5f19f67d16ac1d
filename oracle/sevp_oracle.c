/*
 * sevp_oracle.c -- CPU restatement of the TIMCOM SEVP barotropic solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; nothing in the product path links or calls it.
 *
 * Every routine follows one reference routine statement by statement, with
 * the reference's arithmetic: REAL(4) coefficients and right-hand sides are
 * promoted exactly to double, every operation is a separately rounded IEEE
 * double operation evaluated left to right (build with -ffp-contract=off so no
 * multiply-add is fused), and every dot product is a sequential sum that
 * starts from zero.
 *
 * Array arguments follow the Fortran conventions of the reference: column
 * major, 1-based in the index macros below, and the caller passes pointers
 * already offset to the first element of the strip (as the Fortran caller
 * passes AL(IL,1) or X(IL,1)).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   orc_repbir, orc_rep1, orc_rep, orc_repbir_period: bitwise against the
 *     reference routines compiled from MAIN/template/GLOBAL/src/sevp.f90 into
 *     oracle/_ref and against the golden vectors in tests/golden/.
 *   orc_bir_solve: bitwise against a driver that calls the compiled reference
 *     REPBIR in the reference strip order (tests/golden).
 *   orc_matinv, orc_pre, orc_pre1, orc_bir_prep: the reference PREP program is
 *     not buildable here (its host scope needs a header the reference does not
 *     ship), so these are pinned by mathematics: exact rational Gauss-Jordan on
 *     small matrices, B*inv(B) = I, brute-force influence matrices, and
 *     end-to-end solves that must reproduce scipy's sparse direct solution.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX2(ld, i, j) (((long)(i) - 1) + ((long)(j) - 1) * (long)(ld))

/* ------------------------------------------------------------------------ */
/* REPBIR: strip SEVP solve.  Follows MAIN/template/GLOBAL/src/sevp.f90:218-328.
 * AX..CY have leading dimension i2, F has leading dimension m2, X leading
 * dimension i0, H leading dimension m0.  RINV(m2,m2,nblk), RINV1(m2,m2,nblk-1),
 * DUM0(m2,nblk-1), DUM1(m2), DUM2(m2) are caller-owned. */
void orc_repbir(const float *ax, const float *ay, const float *bb,
                const float *cx, const float *cy, const double *rinv,
                const double *rinv1, double *dum0, double *dum1, double *dum2,
                const float *f, double *h, double *x, const int *ie, int i0,
                int i2, int m0, int m2, int nblk)
{
#define AX(i, j) ((double)ax[IDX2(i2, i, j)])
#define AY(i, j) ((double)ay[IDX2(i2, i, j)])
#define BB(i, j) ((double)bb[IDX2(i2, i, j)])
#define CX(i, j) ((double)cx[IDX2(i2, i, j)])
#define CY(i, j) ((double)cy[IDX2(i2, i, j)])
#define F(i, j) ((double)f[IDX2(m2, i, j)])
#define X(i, j) x[IDX2(i0, i, j)]
#define H(i, j) h[IDX2(m0, i, j)]
#ifndef ORC_PLANT_TRANSPOSED_RINV
#define RINV(m, n, b) rinv[((long)(m) - 1) + ((long)(n) - 1) * m2 + ((long)(b) - 1) * (long)m2 * m2]
#else /* deliberate mistake, used only by the test that checks the pins catch it */
#define RINV(m, n, b) rinv[((long)(n) - 1) + ((long)(m) - 1) * m2 + ((long)(b) - 1) * (long)m2 * m2]
#endif
#define RINV1(m, n, b) rinv1[((long)(m) - 1) + ((long)(n) - 1) * m2 + ((long)(b) - 1) * (long)m2 * m2]
#define DUM0(n, b) dum0[((long)(n) - 1) + ((long)(b) - 1) * m2]
  int js = 1;
  for (int nb = 1; nb <= nblk; ++nb) {
    int jf = ie[nb - 1] - 2;
    for (int j = js; j <= jf; ++j)
      for (int i = 1; i <= m2; ++i)
        X(i + 1, j + 2) = (F(i, j) - AX(i, j) * X(i, j + 1) - AY(i, j) * X(i + 1, j) -
                           BB(i, j) * X(i + 1, j + 1) - CX(i, j) * X(i + 2, j + 1)) / CY(i, j);
    if (nb != nblk) {
      int j = ie[nb - 1] - 1;
      for (int i = 1; i <= m2; ++i)
        dum1[i - 1] = F(i, j) - AX(i, j) * X(i, j + 1) - AY(i, j) * X(i + 1, j) -
                      BB(i, j) * X(i + 1, j + 1) - CX(i, j) * X(i + 2, j + 1) -
                      CY(i, j) * X(i + 1, j + 2);
      j = ie[nb - 1];
      for (int n = 1; n <= m2; ++n) {
        dum2[n - 1] = 0.0;
        for (int m = 1; m <= m2; ++m)
          dum2[n - 1] = dum2[n - 1] + dum1[m - 1] * RINV1(m, n, nb);
        DUM0(n, nb) = X(n + 1, j);
        X(n + 1, j) = X(n + 1, j) - dum2[n - 1];
      }
    }
    js = ie[nb - 1];
  }
  for (int nbs = 1; nbs <= nblk; ++nbs) {
    int nb = nblk - nbs + 1;
    js = (nb != 1) ? ie[nb - 2] : 1;
    int jf = ie[nb - 1] - 2;
    if (nb != nblk) {
      int j = ie[nb - 1];
      for (int n = 1; n <= m2; ++n) X(n + 1, j) = DUM0(n, nb);
    }
    for (int j = js; j <= ie[nb - 1]; ++j)
      for (int i = 1; i <= m0; ++i) H(i, j) = 0.0;
    int j = ie[nb - 1] - 1;
    for (int i = 1; i <= m2; ++i)
      dum1[i - 1] = F(i, j) - AX(i, j) * X(i, j + 1) - AY(i, j) * X(i + 1, j) -
                    BB(i, j) * X(i + 1, j + 1) - CX(i, j) * X(i + 2, j + 1) -
                    CY(i, j) * X(i + 1, j + 2);
    for (int n = 1; n <= m2; ++n) {
      dum2[n - 1] = 0.0;
      for (int m = 1; m <= m2; ++m)
        dum2[n - 1] = dum2[n - 1] + dum1[m - 1] * RINV(m, n, nb);
      H(n + 1, js + 1) = dum2[n - 1];
      X(n + 1, js + 1) = X(n + 1, js + 1) + dum2[n - 1];
    }
    if (nb != 1) {
      for (int m = 1; m <= m2; ++m) dum1[m - 1] = H(m + 1, js + 1) * CY(m, js - 1);
      j = ie[nb - 2];
      for (int n = 1; n <= m2; ++n) {
        dum2[n - 1] = 0.0;
        for (int m = 1; m <= m2; ++m)
          dum2[n - 1] = dum2[n - 1] + dum1[m - 1] * RINV1(m, n, nb - 1);
        H(n + 1, j) = dum2[n - 1];
      }
    }
    for (j = js; j <= jf; ++j)
      for (int i = 1; i <= m2; ++i) {
        H(i + 1, j + 2) = (-AX(i, j) * H(i, j + 1) - AY(i, j) * H(i + 1, j) -
                           BB(i, j) * H(i + 1, j + 1) - CX(i, j) * H(i + 2, j + 1)) / CY(i, j);
        X(i + 1, j + 2) = X(i + 1, j + 2) + H(i + 1, j + 2);
      }
  }
}

/* REP: full-width non-periodic SEVP, MAIN/template/GLOBAL/src/sevp.f90:335-438.
 * Statement for statement it is REPBIR with M0=I0 and M2=I2 (the H zeroing
 * runs over I=1..I0, F has leading dimension I2); the tests pin this identity
 * against both compiled reference routines. */
void orc_rep(const float *ax, const float *ay, const float *bb, const float *cx,
             const float *cy, const double *rinv, const double *rinv1,
             double *dum0, double *dum1, double *dum2, const float *f,
             double *h, double *x, const int *ie, int i0, int i2, int nblk)
{
  orc_repbir(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie,
             i0, i2, i0, i2, nblk);
}

/* Periodic full-width SEVP.  REP1 (MAIN/template/GLOBAL/src/sevp.f90:11-107)
 * with m0 = i0 and m2 = i2; REPBIR_PERIOD (sevp.f90:109-217) is the same
 * statement sequence with the march over M2 columns, F of leading dimension
 * M2 and H of leading dimension M0, while the periodic copies index column I0
 * and I0-1 of both X and H.  Both are expressed by this one routine. */
static void rep_periodic(const float *ax, const float *ay, const float *bb,
                         const float *cx, const float *cy, const double *rinv,
                         const double *rinv1, double *dum0, double *dum1,
                         double *dum2, const float *f, double *h, double *x,
                         const int *ie, int i0, int i2, int m0, int m2, int nblk)
{
  const int i1 = i0 - 1;
  int js = 1;
  for (int nb = 1; nb <= nblk; ++nb) {
    int jf = ie[nb - 1] - 2;
    for (int j = js; j <= jf; ++j) {
      X(1, j + 1) = X(i1, j + 1);
      X(i0, j + 1) = X(2, j + 1);
      for (int i = 1; i <= m2; ++i)
        X(i + 1, j + 2) = (F(i, j) - AX(i, j) * X(i, j + 1) - AY(i, j) * X(i + 1, j) -
                           BB(i, j) * X(i + 1, j + 1) - CX(i, j) * X(i + 2, j + 1)) / CY(i, j);
    }
    X(1, jf + 2) = X(i1, jf + 2);
    X(i0, jf + 2) = X(2, jf + 2);
    if (nb != nblk) {
      int j = ie[nb - 1] - 1;
      for (int i = 1; i <= m2; ++i)
        dum1[i - 1] = F(i, j) - AX(i, j) * X(i, j + 1) - AY(i, j) * X(i + 1, j) -
                      BB(i, j) * X(i + 1, j + 1) - CX(i, j) * X(i + 2, j + 1) -
                      CY(i, j) * X(i + 1, j + 2);
      j = ie[nb - 1];
      for (int n = 1; n <= m2; ++n) {
        dum2[n - 1] = 0.0;
        for (int m = 1; m <= m2; ++m)
          dum2[n - 1] = dum2[n - 1] + dum1[m - 1] * RINV1(m, n, nb);
        DUM0(n, nb) = X(n + 1, j);
        X(n + 1, j) = X(n + 1, j) - dum2[n - 1];
      }
      X(1, j) = X(i1, j);
      X(i0, j) = X(2, j);
    }
    js = ie[nb - 1];
  }
  for (int nbs = 1; nbs <= nblk; ++nbs) {
    int nb = nblk - nbs + 1;
    js = (nb != 1) ? ie[nb - 2] : 1;
    int jf = ie[nb - 1] - 2;
    if (nb != nblk) {
      int j = ie[nb - 1];
      for (int n = 1; n <= m2; ++n) X(n + 1, j) = DUM0(n, nb);
    }
    for (int j = js; j <= ie[nb - 1]; ++j)
      for (int i = 1; i <= m0; ++i) H(i, j) = 0.0;
    int j = ie[nb - 1] - 1;
    X(1, j + 1) = X(i1, j + 1);
    X(i0, j + 1) = X(2, j + 1);
    for (int i = 1; i <= m2; ++i)
      dum1[i - 1] = F(i, j) - AX(i, j) * X(i, j + 1) - AY(i, j) * X(i + 1, j) -
                    BB(i, j) * X(i + 1, j + 1) - CX(i, j) * X(i + 2, j + 1) -
                    CY(i, j) * X(i + 1, j + 2);
    for (int n = 1; n <= m2; ++n) {
      dum2[n - 1] = 0.0;
      for (int m = 1; m <= m2; ++m)
        dum2[n - 1] = dum2[n - 1] + dum1[m - 1] * RINV(m, n, nb);
      H(n + 1, js + 1) = dum2[n - 1];
      X(n + 1, js + 1) = X(n + 1, js + 1) + dum2[n - 1];
    }
    H(1, js + 1) = H(i1, js + 1);
    H(i0, js + 1) = H(2, js + 1);
    X(1, js + 1) = X(i1, js + 1);
    X(i0, js + 1) = X(2, js + 1);
    if (nb != 1) {
      for (int m = 1; m <= m2; ++m) dum1[m - 1] = H(m + 1, js + 1) * CY(m, js - 1);
      j = ie[nb - 2];
      for (int n = 1; n <= m2; ++n) {
        dum2[n - 1] = 0.0;
        for (int m = 1; m <= m2; ++m)
          dum2[n - 1] = dum2[n - 1] + dum1[m - 1] * RINV1(m, n, nb - 1);
        H(n + 1, j) = dum2[n - 1];
      }
      H(1, j) = H(i1, j);
      H(i0, j) = H(2, j);
    }
    for (j = js; j <= jf; ++j) {
      H(1, j + 1) = H(i1, j + 1);
      H(i0, j + 1) = H(2, j + 1);
      for (int i = 1; i <= m2; ++i) {
        H(i + 1, j + 2) = (-AX(i, j) * H(i, j + 1) - AY(i, j) * H(i + 1, j) -
                           BB(i, j) * H(i + 1, j + 1) - CX(i, j) * H(i + 2, j + 1)) / CY(i, j);
        X(i + 1, j + 2) = X(i + 1, j + 2) + H(i + 1, j + 2);
      }
      X(1, j + 2) = X(i1, j + 2);
      X(i0, j + 2) = X(2, j + 2);
    }
  }
  int j1 = ie[nblk - 1];
  for (int j = 2; j <= j1; ++j) {
    X(1, j) = X(i1, j);
    X(i0, j) = X(2, j);
  }
}

void orc_rep1(const float *ax, const float *ay, const float *bb,
              const float *cx, const float *cy, const double *rinv,
              const double *rinv1, double *dum0, double *dum1, double *dum2,
              const float *f, double *h, double *x, const int *ie, int i0,
              int i2, int nblk)
{
  rep_periodic(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie,
               i0, i2, i0, i2, nblk);
}

void orc_repbir_period(const float *ax, const float *ay, const float *bb,
                       const float *cx, const float *cy, const double *rinv,
                       const double *rinv1, double *dum0, double *dum1,
                       double *dum2, const float *f, double *h, double *x,
                       const int *ie, int i0, int i2, int m0, int m2, int nblk)
{
  rep_periodic(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie,
               i0, i2, m0, m2, nblk);
}
#undef F
#undef X

/* ------------------------------------------------------------------------ */
/* MATINV: in-place Gauss-Jordan inverse without pivoting.
 * Follows PREP/template/GLOBAL/src/prep_GLOBAL.f90:516-568: the pivot row is
 * scaled by the reciprocal of its pivot, the pivot column below is eliminated
 * with separately rounded multiply and subtract, then the same is done above
 * the diagonal in a second sweep.  Returns the number of zero pivots met (the
 * reference divides by them silently). */
int orc_matinv(double *b, int n)
{
#define B(i, j) b[IDX2(n, i, j)]
  int zero_pivots = 0;
  double *b1 = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *b2 = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  for (int i = 1; i <= n - 1; ++i) {
    if (B(i, i) == 0.0) ++zero_pivots;
    b1[0] = 1.0 / B(i, i);
    B(i, i) = 1.0;
    for (int j = 1; j <= n; ++j) B(i, j) = B(i, j) * b1[0];
    for (int ia = i + 1; ia <= n; ++ia) b1[ia - 1] = B(ia, i);
    for (int ia = i + 1; ia <= n; ++ia) B(ia, i) = 0.0;
    for (int j = 1; j <= n; ++j) b2[j - 1] = B(i, j);
    for (int ia = i + 1; ia <= n; ++ia)
      for (int j = 1; j <= n; ++j) B(ia, j) = B(ia, j) - b1[ia - 1] * b2[j - 1];
  }
  if (B(n, n) == 0.0) ++zero_pivots;
  b1[0] = 1.0 / B(n, n);
  B(n, n) = 1.0;
  for (int j = 1; j <= n; ++j) B(n, j) = B(n, j) * b1[0];
  for (int i = 2; i <= n; ++i) {
    for (int ib = 1; ib <= i; ++ib) b1[ib - 1] = B(ib, i);
    for (int ib = 1; ib <= i - 1; ++ib) B(ib, i) = 0.0;
    for (int j = 1; j <= n; ++j) b2[j - 1] = B(i, j);
    for (int ib = 1; ib <= i - 1; ++ib)
      for (int j = 1; j <= n; ++j) B(ib, j) = B(ib, j) - b1[ib - 1] * b2[j - 1];
  }
  free(b1);
  free(b2);
  return zero_pivots;
#undef B
}

/* PRE: influence matrices of one strip.
 * Follows PREP/template/GLOBAL/src/prep_GLOBAL.f90:457-513.  AX..CY have
 * leading dimension i2, H leading dimension m0; RINV(m2,m2,nblk) and
 * RINV1(m2,m2,nblk-1) are written.  The last-row responses T of block nb are
 * kept in the RINV(:,:,nblk) slot until that block is processed, as the
 * reference does.  Returns the number of zero pivots met in MATINV. */
int orc_pre(const float *ax, const float *ay, const float *bb, const float *cx,
            const float *cy, double *rinv, double *rinv1, double *h,
            const int *ie, int i2, int m0, int m2, int nblk)
{
  int zp = 0;
  int jl = 1;
  for (int nb = 1; nb <= nblk; ++nb) {
    int jh = ie[nb - 1];
    int jhp = jh + 1, jhm = jh - 2, jg = jl + 1;
    for (int ng = 1; ng <= m2; ++ng) {
      int ig = ng + 1;
      for (int j = jl; j <= jhp; ++j)
        for (int i = 1; i <= m0; ++i) H(i, j) = 0.0;
      H(ig, jg) = 1.0;
      if (nb != 1)
        for (int n = 1; n <= m2; ++n) H(n + 1, jl) = RINV1(ng, n, nb - 1) * CY(ig - 1, jg - 2);
      for (int j = jl; j <= jhm; ++j)
        for (int i = 1; i <= m2; ++i)
          H(i + 1, j + 2) = -(AX(i, j) * H(i, j + 1) + AY(i, j) * H(i + 1, j) +
                              BB(i, j) * H(i + 1, j + 1) + CX(i, j) * H(i + 2, j + 1)) / CY(i, j);
      int j = jh - 1;
      for (int i = 1; i <= m2; ++i)
        RINV(ng, i, nb) = AX(i, j) * H(i, j + 1) + AY(i, j) * H(i + 1, j) +
                          BB(i, j) * H(i + 1, j + 1) + CX(i, j) * H(i + 2, j + 1);
      if (nb != nblk) {
        j = ie[nb - 1];
        for (int n = 1; n <= m2; ++n) RINV(ng, n, nblk) = H(n + 1, j);
      }
      H(ig, jg) = 0.0;
    }
    zp += orc_matinv(&RINV(1, 1, nb), m2);
    if (nb == nblk) return zp;
    for (int i = 1; i <= m2; ++i)
      for (int j = 1; j <= m2; ++j) {
        RINV1(i, j, nb) = 0.0;
        for (int k = 1; k <= m2; ++k)
          RINV1(i, j, nb) = RINV1(i, j, nb) - RINV(i, k, nb) * RINV(k, j, nblk);
      }
    jl = jh;
  }
  return zp;
}

/* PRE1: periodic full-width influence matrices.
 * Follows PREP/template/GLOBAL/src/prep_GLOBAL.f90:327-386: as PRE with the
 * periodic copies H(1)=H(I0-1), H(I0)=H(2) of each marched row. */
int orc_pre1(const float *ax, const float *ay, const float *bb, const float *cx,
             const float *cy, double *rinv, double *rinv1, double *h,
             const int *ie, int i0, int i2, int nblk)
{
  const int m0 = i0, m2 = i2, i1 = i0 - 1;
  int zp = 0;
  int jl = 1;
  for (int nb = 1; nb <= nblk; ++nb) {
    int jh = ie[nb - 1];
    int jhp = jh + 1, jhm = jh - 2, jg = jl + 1;
    for (int ng = 1; ng <= i2; ++ng) {
      int ig = ng + 1;
      for (int j = jl; j <= jhp; ++j)
        for (int i = 1; i <= i0; ++i) H(i, j) = 0.0;
      H(ig, jg) = 1.0;
      H(1, jg) = H(i1, jg);
      H(i0, jg) = H(2, jg);
      if (nb != 1)
        for (int n = 1; n <= i2; ++n) H(n + 1, jl) = RINV1(ng, n, nb - 1) * CY(ig - 1, jg - 2);
      for (int j = jl; j <= jhm; ++j) {
        H(1, j + 1) = H(i1, j + 1);
        H(i0, j + 1) = H(2, j + 1);
        for (int i = 1; i <= i2; ++i)
          H(i + 1, j + 2) = -(AX(i, j) * H(i, j + 1) + AY(i, j) * H(i + 1, j) +
                              BB(i, j) * H(i + 1, j + 1) + CX(i, j) * H(i + 2, j + 1)) / CY(i, j);
      }
      int j = jh - 1;
      H(1, j + 1) = H(i1, j + 1);
      H(i0, j + 1) = H(2, j + 1);
      for (int i = 1; i <= i2; ++i)
        RINV(ng, i, nb) = AX(i, j) * H(i, j + 1) + AY(i, j) * H(i + 1, j) +
                          BB(i, j) * H(i + 1, j + 1) + CX(i, j) * H(i + 2, j + 1);
      if (nb != nblk) {
        j = ie[nb - 1];
        for (int n = 1; n <= i2; ++n) RINV(ng, n, nblk) = H(n + 1, j);
        H(1, jg) = 0.0;
        H(i0, jg) = 0.0;
      }
      H(ig, jg) = 0.0;
    }
    zp += orc_matinv(&RINV(1, 1, nb), i2);
    if (nb == nblk) return zp;
    for (int i = 1; i <= i2; ++i)
      for (int j = 1; j <= i2; ++j) {
        RINV1(i, j, nb) = 0.0;
        for (int k = 1; k <= i2; ++k)
          RINV1(i, j, nb) = RINV1(i, j, nb) - RINV(i, k, nb) * RINV(k, j, nblk);
      }
    jl = jh;
  }
  return zp;
}
#undef AX
#undef AY
#undef BB
#undef CX
#undef CY
#undef H
#undef RINV
#undef RINV1
#undef DUM0

/* ------------------------------------------------------------------------ */
/* Seam strip gather: the last BIR strip starts at column il and wraps the
 * periodic seam, taking inc columns il..il+inc-1 and then columns 1..inc.
 * Follows PREP/template/GLOBAL/src/prep_GLOBAL.f90:293-313 and
 * MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1925-1942 (both select the same
 * columns).  src has leading dimension i2, dst leading dimension ibir. */
static void seam_gather_f(const float *src, float *dst, int i2, int j2, int il,
                          int inc, int ibir)
{
  for (int j = 1; j <= j2; ++j) {
    for (int i = 1; i <= inc; ++i) dst[IDX2(ibir, i, j)] = src[IDX2(i2, il + i - 1, j)];
    for (int i = inc + 1; i <= ibir; ++i) dst[IDX2(ibir, i, j)] = src[IDX2(i2, i - inc, j)];
  }
}

/* INITFS strip loop of PREP: influence matrices of every BIR strip.
 * Follows PREP/template/GLOBAL/src/prep_GLOBAL.f90:286-317.  RINV has the
 * reference shape (ibir*ibir, nb0, nbir), RINV1 (ibir*ibir, nb0-1, nbir).
 * h is scratch of (ibir+2) x (j2+2).  Returns the number of zero pivots. */
int orc_bir_prep_strips(const float *al, const float *ab, const float *ac,
                        const float *ar, const float *at, double *rinv, double *rinv1,
                        double *h, const int *ie, int i2, int j2, int nbir, int nb0,
                        int n_lo, int n_hi);

int orc_bir_prep(const float *al, const float *ab, const float *ac,
                 const float *ar, const float *at, double *rinv, double *rinv1,
                 double *h, const int *ie, int i2, int j2, int nbir, int nb0)
{
  return orc_bir_prep_strips(al, ab, ac, ar, at, rinv, rinv1, h, ie, i2, j2, nbir, nb0, 1, nbir);
}

/* The same loop over strips n_lo..n_hi only (1-based, inclusive); the other
 * strips' slices of RINV / RINV1 are left as they are. */
int orc_bir_prep_strips(const float *al, const float *ab, const float *ac,
                        const float *ar, const float *at, double *rinv, double *rinv1,
                        double *h, const int *ie, int i2, int j2, int nbir, int nb0,
                        int n_lo, int n_hi)
{
  const int ibir = 2 * i2 / nbir, inc = ibir / 2;
  const long ibr2 = (long)ibir * ibir;
  int zp = 0;
  float *c[5] = {0};
  for (int n = n_lo; n <= n_hi; ++n) {
    int il = (n - 1) * inc + 1;
    double *rv = rinv + (long)(n - 1) * ibr2 * nb0;
    double *rv1 = rinv1 + (long)(n - 1) * ibr2 * (nb0 - 1);
    if (nbir == 1 || n != nbir) {
      long o = IDX2(i2, il, 1);
      zp += orc_pre(al + o, ab + o, ac + o, ar + o, at + o, rv, rv1, h, ie, i2,
                    ibir + 2, ibir, nb0);
    } else {
      const float *srcs[5] = {al, ab, ac, ar, at};
      for (int k = 0; k < 5; ++k) {
        c[k] = (float *)malloc(sizeof(float) * (size_t)ibir * j2);
        seam_gather_f(srcs[k], c[k], i2, j2, il, inc, ibir);
      }
      zp += orc_pre(c[0], c[1], c[2], c[3], c[4], rv, rv1, h, ie, ibir,
                    ibir + 2, ibir, nb0);
      for (int k = 0; k < 5; ++k) free(c[k]);
    }
  }
  return zp;
}

/* FS_SOLVER, FL_EVP_STP=1 branch: multiplicative sweeps over overlapping BIR
 * strips.  Follows MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1902-1966.
 *   x   : (i2+2) x (j2+2) solution with halos, in/out (initial guess in).
 *   xx  : (ibir+2) x (j2+2) seam scratch, in/out.  The reference keeps it as a
 *         module array that starts at zero and is never cleared between
 *         calls, and its halo entries in row j2+1 and its halo rows are never
 *         refreshed by the gather; pass the same buffer to successive calls.
 *   periodic: apply the FLAG_PERIODIC_WE halo refresh after the seam strip.
 *   seam_fix: 0 reproduces the reference, whose halo gather for the seam
 *         strip covers rows 1..J2 although REPBIR reads the halo columns on
 *         rows up to J2+1, so the seam strip sees zero there and the sweeps
 *         stall short of the solution; 1 also gathers row J2+1, after which
 *         the sweeps converge to the global periodic solution.
 * The call order of strips and sweeps is a data dependency and is kept. */
void orc_bir_solve(const float *al, const float *ab, const float *ac,
                   const float *ar, const float *at, const float *s,
                   const double *rinv, const double *rinv1, double *x,
                   double *xx, const int *ie, int i2, int j2, int nbir,
                   int nb0, int nsweep, int periodic, int seam_fix)
{
  const int ibir = 2 * i2 / nbir, inc = ibir / 2;
  const int i0 = i2 + 2, i1 = i2 + 1, j1 = j2 + 1, j0 = j2 + 2;
  const long ibr2 = (long)ibir * ibir;
  const int m0 = ibir + 2;
  float *src = (float *)malloc(sizeof(float) * (size_t)ibir * j2);
  float *c[5];
  for (int k = 0; k < 5; ++k) c[k] = (float *)malloc(sizeof(float) * (size_t)ibir * j2);
  double *h = (double *)calloc((size_t)m0 * j0, sizeof(double));
  double *dum0 = (double *)calloc((size_t)ibir * (nb0 > 1 ? nb0 - 1 : 1), sizeof(double));
  double *dum1 = (double *)calloc((size_t)ibir, sizeof(double));
  double *dum2 = (double *)calloc((size_t)ibir, sizeof(double));
  for (int it = 1; it <= nsweep; ++it) {
    for (int n = 1; n <= nbir; ++n) {
      int il = (n - 1) * inc + 1;
      const double *rv = rinv + (long)(n - 1) * ibr2 * nb0;
      const double *rv1 = rinv1 + (long)(n - 1) * ibr2 * (nb0 - 1);
      if (n != nbir) {
        for (int j = 1; j <= j2; ++j)
          for (int i = 1; i <= ibir; ++i) src[IDX2(ibir, i, j)] = s[IDX2(i2, il + i - 1, j)];
        long o = IDX2(i2, il, 1);
        orc_repbir(al + o, ab + o, ac + o, ar + o, at + o, rv, rv1, dum0, dum1,
                   dum2, src, h, x + IDX2(i0, il, 1), ie, i0, i2, m0, ibir, nb0);
      } else {
        for (int j = 1; j <= (seam_fix ? j1 : j2); ++j) {
          xx[IDX2(m0, 1, j)] = x[IDX2(i0, il, j)];
          xx[IDX2(m0, ibir + 2, j)] = x[IDX2(i0, inc + 2, j)];
        }
        seam_gather_f(s, src, i2, j2, il, inc, ibir);
        seam_gather_f(al, c[0], i2, j2, il, inc, ibir);
        seam_gather_f(ab, c[1], i2, j2, il, inc, ibir);
        seam_gather_f(ac, c[2], i2, j2, il, inc, ibir);
        seam_gather_f(ar, c[3], i2, j2, il, inc, ibir);
        seam_gather_f(at, c[4], i2, j2, il, inc, ibir);
        orc_repbir(c[0], c[1], c[2], c[3], c[4], rv, rv1, dum0, dum1, dum2, src,
                   h, xx, ie, m0, ibir, m0, ibir, nb0);
        for (int j = 2; j <= j1; ++j) {
          for (int i = 1; i <= inc; ++i) x[IDX2(i0, il + i, j)] = xx[IDX2(m0, i + 1, j)];
          for (int i = inc + 1; i <= ibir; ++i) x[IDX2(i0, i - inc + 1, j)] = xx[IDX2(m0, i + 1, j)];
        }
        if (periodic)
          for (int j = 2; j <= j1; ++j) {
            x[IDX2(i0, 1, j)] = x[IDX2(i0, i1, j)];
            x[IDX2(i0, i0, j)] = x[IDX2(i0, 2, j)];
          }
      }
    }
  }
  free(src);
  for (int k = 0; k < 5; ++k) free(c[k]);
  free(h);
  free(dum0);
  free(dum1);
  free(dum2);
}

/* 5-point product r = A x on the interior, following p_sqprod
 * (MAIN/template/GLOBAL/src/sevp.f90:696-713): the terms are summed in the
 * order AB, AL, AC, AR, AT.  x and r are (i2+2) x (j2+2); the halo of r is not
 * touched. */
void orc_sqprod(const float *al, const float *ab, const float *ac,
                const float *ar, const float *at, const double *x, double *r,
                int i2, int j2)
{
  const int i0 = i2 + 2;
  for (int j = 1; j <= j2; ++j)
    for (int i = 1; i <= i2; ++i) {
      long c = IDX2(i2, i, j);
      r[IDX2(i0, i + 1, j + 1)] =
          (double)ab[c] * x[IDX2(i0, i + 1, j)] + (double)al[c] * x[IDX2(i0, i, j + 1)] +
          (double)ac[c] * x[IDX2(i0, i + 1, j + 1)] + (double)ar[c] * x[IDX2(i0, i + 2, j + 1)] +
          (double)at[c] * x[IDX2(i0, i + 1, j + 2)];
    }
}

/* ---------------------------------------------------------------- step 14
 * One ITMASK iteration of FS_GLOBAL step 14
 * (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1154-1283), in two halves around
 * the FS_SOLVER call.  Arrays have the reference shapes (Fortran order):
 *   U (I0,J0,K1), V (I0,J1,K1), W (I0,J0,K0), P0 (I0,J0), SCR (I0,J0,>=4)
 *   REAL(4); X (I0,J0) REAL(8); KB, IU0, IV0 (I0,J0), IU (I0,J0,K1),
 *   IV (I0,J1,K1) INTEGER(2); ODX, ODY, OCS (J0), CSV, ODYV (J1), ODZ (K1)
 *   REAL(4); S (I2,J2) REAL(4).
 * Mixed-kind expressions follow Fortran promotion: an INTEGER(2) operand
 * of a REAL(4) product is converted to REAL(4); REAL(4)*REAL(8) is
 * evaluated in REAL(8) and rounded when stored into a REAL(4) array.
 * Options (the reference selects them with cpp flags and namelist values):
 *   mask          apply the land masking of U, V (:1158-1171) first
 *   lopenw        LOPENW == 1: also mask the periodic faces I = 1, I1
 *   harness       the shipped source's override U = 1, V = 0, W(:,:,K0) = 0
 *                 (:1173-1176)
 *   free_surface  FLAG_FS: integrate W upward, S = W(:,:,1) (:1178-1192);
 *                 else downward, S = -W(:,:,KB+1) (:1193-1209)
 *   periodic_we   FLAG_PERIODIC_WE: seam-face correction (:1230-1233)
 */
#define U3(a, i, j, k) (a)[((long)(i) - 1) + ((long)(j) - 1) * I0 + ((long)(k) - 1) * I0 * J0]
#define V3(a, i, j, k) (a)[((long)(i) - 1) + ((long)(j) - 1) * I0 + ((long)(k) - 1) * I0 * J1]
#define A2(a, i, j) (a)[((long)(i) - 1) + ((long)(j) - 1) * I0]

void orc_step14_prepare(int i2, int j2, int k1, int mask, int lopenw, int harness,
                        int free_surface, const short *kb, const short *iu0,
                        const short *iv0, const float *odx, const float *ody,
                        const float *ocs, const float *csv, const float *odz, float *u,
                        float *v, float *w, float *s)
{
  const int I0 = i2 + 2, J0 = j2 + 2, I1 = I0 - 1, J1 = J0 - 1, I2 = i2, J2 = j2;
  const int K1 = k1, K0 = k1 + 1;
  if (mask) {
    for (int J = 2; J <= J1; ++J) {
      for (int I = 2; I <= I2; ++I) U3(u, I, J, 1) = (float)A2(iu0, I, J) * U3(u, I, J, 1);
      if (lopenw == 1) {
        U3(u, 1, J, 1) = (float)A2(iu0, 1, J) * U3(u, 1, J, 1);
        U3(u, I1, J, 1) = (float)A2(iu0, I1, J) * U3(u, I1, J, 1);
      }
    }
    for (int J = 2; J <= J2; ++J)
      for (int I = 2; I <= I1; ++I) V3(v, I, J, 1) = (float)A2(iv0, I, J) * V3(v, I, J, 1);
  }
  if (harness) {
    for (long n = 0; n < (long)I0 * J0 * K1; ++n) u[n] = 1.0f;
    for (long n = 0; n < (long)I0 * J1 * K1; ++n) v[n] = 0.0f;
    for (int J = 1; J <= J0; ++J)
      for (int I = 1; I <= I0; ++I) U3(w, I, J, K0) = 0.0f;
  }
  if (free_surface) {
    for (int K = K1; K >= 1; --K) {
      const float temp1 = 1.0f / odz[K - 1];
      for (int J = 2; J <= J1; ++J) {
        const float temp = ocs[J - 1] * ody[J - 1];
        for (int I = 2; I <= I1; ++I) {
          const float a = (U3(u, I, J, K) - U3(u, I - 1, J, K)) * odx[J - 1];
          const float b = (csv[J - 1] * V3(v, I, J, K) - csv[J - 2] * V3(v, I, J - 1, K)) * temp;
          U3(w, I, J, K) = U3(w, I, J, K + 1) + (a + b) * temp1;
        }
      }
    }
    for (int J = 2; J <= J1; ++J)
      for (int I = 2; I <= I1; ++I) s[(I - 2) + (long)(J - 2) * I2] = U3(w, I, J, 1);
  } else {
    for (int K = 1; K <= K1; ++K) {
      const float temp1 = 1.0f / odz[K - 1];
      for (int J = 2; J <= J1; ++J) {
        const float temp = ocs[J - 1] * ody[J - 1];
        for (int I = 2; I <= I1; ++I) {
          const float a = (U3(u, I, J, K) - U3(u, I - 1, J, K)) * odx[J - 1];
          const float b = (csv[J - 1] * V3(v, I, J, K) - csv[J - 2] * V3(v, I, J - 1, K)) * temp;
          U3(w, I, J, K + 1) = U3(w, I, J, K) - (a + b) * temp1;
        }
      }
    }
    for (int J = 2; J <= J1; ++J)
      for (int I = 2; I <= I1; ++I)
        s[(I - 2) + (long)(J - 2) * I2] = -U3(w, I, J, A2(kb, I, J) + 1);
  }
}

/* The half after the solve (:1214-1283): P0 += ODT*X, the U and V
 * corrections and their running sums SCR(:,:,3:4), and the largest
 * top-layer speed on land faces (the value compared with TOLERANCE). */
float orc_step14_correct(int i2, int j2, int k1, int periodic_we, float odt,
                         const short *iu, const short *iv, const short *iu0,
                         const short *iv0, const float *odx, const float *odyv,
                         const double *x, float *p0, float *scr, float *u, float *v)
{
  const int I0 = i2 + 2, J0 = j2 + 2, I1 = I0 - 1, J1 = J0 - 1, I2 = i2, J2 = j2;
  const int K1 = k1;
  for (int J = 2; J <= J1; ++J)
    for (int I = 2; I <= I1; ++I)
      A2(p0, I, J) = (float)((double)A2(p0, I, J) + (double)odt * A2(x, I, J));
  for (int J = 2; J <= J1; ++J) {
    for (int I = 2; I <= I2; ++I)
#ifndef ORC_PLANT_STEP14_ROW
      U3(scr, I, J, 1) = (float)(-((A2(x, I + 1, J) - A2(x, I, J)) * (double)odx[J - 1]));
#else
      /* planted slip for the tests: the metric of the next row */
      U3(scr, I, J, 1) = (float)(-((A2(x, I + 1, J) - A2(x, I, J)) * (double)odx[J]));
#endif
    U3(scr, 1, J, 1) = 0.0f;
    U3(scr, I1, J, 1) = 0.0f;
    if (periodic_we) {
      U3(scr, 1, J, 1) = (float)(-((A2(x, 2, J) - A2(x, I1, J)) * (double)odx[J - 1]));
      U3(scr, I1, J, 1) = U3(scr, 1, J, 1);
    }
    for (int I = 1; I <= I1; ++I) {
      U3(scr, I, J, 3) = U3(scr, I, J, 3) + U3(scr, I, J, 1);
      U3(u, I, J, 1) = U3(u, I, J, 1) + U3(scr, I, J, 1);
      for (int K = 2; K <= K1; ++K)
        U3(u, I, J, K) = U3(u, I, J, K) + U3(scr, I, J, 1) * (float)U3(iu, I, J, K);
    }
  }
  for (int J = 2; J <= J2; ++J)
    for (int I = 2; I <= I1; ++I)
      U3(scr, I, J, 2) = (float)(-((A2(x, I, J + 1) - A2(x, I, J)) * (double)odyv[J - 1]));
  for (int I = 2; I <= I1; ++I) {
    U3(scr, I, 1, 2) = 0.0f;
    U3(scr, I, J1, 2) = 0.0f;
  }
  for (int J = 1; J <= J1; ++J)
    for (int I = 2; I <= I1; ++I) {
      U3(scr, I, J, 4) = U3(scr, I, J, 4) + U3(scr, I, J, 2);
      V3(v, I, J, 1) = V3(v, I, J, 1) + U3(scr, I, J, 2);
    }
  for (int K = 2; K <= K1; ++K)
    for (int J = 1; J <= J1; ++J)
      for (int I = 2; I <= I1; ++I)
        V3(v, I, J, K) = V3(v, I, J, K) + U3(scr, I, J, 2) * (float)V3(iv, I, J, K);
  float tmp = 0.0f;
  for (int J = 2; J <= J1; ++J)
    for (int I = 2; I <= I2; ++I) {
      const float t = (float)(1 - A2(iu0, I, J)) * fabsf(U3(u, I, J, 1));
      tmp = tmp > t ? tmp : t;
    }
  for (int J = 2; J <= J2; ++J)
    for (int I = 2; I <= I1; ++I) {
      const float t = (float)(1 - A2(iv0, I, J)) * fabsf(V3(v, I, J, 1));
      tmp = tmp > t ? tmp : t;
    }
  return tmp;
}
#undef U3
#undef V3
#undef A2

/* ---------------------------------------------------------------- Krylov
 * The cross-check solvers of MAIN/template/GLOBAL/src/sevp.f90:443-795:
 * p_bicgstab_le, p_csi and lanczos with their helpers p_sqprod_T, p_inprod
 * and solver_boundary_update.  `preconditioning` (:757-767) is a copy
 * (PH = P), so PH and SH are the arrays they copy after the periodic halo
 * update; every work array is (i2+2) x (j2+2), zero at entry, and the
 * whole-array statements update every element, halos included.  The stop
 * thresholds are the reference's REAL(4) literals promoted to double
 * (callers pass (double)1e-8f etc.).
 *
 * Where the reference is undefined the restatement picks the reading that
 * makes it defined, and says so:
 *   - lanczos writes SIGMA(M+1) and BETA(M+1) past their DIMENSION(M)
 *     arrays on its last step; here they have M+1 entries.
 *   - lanczos's local arrays S, Q and WORK are never initialised, and only
 *     their interiors are written; their halos are taken as zero.
 *   - FIGI leaves E(1) unset and lanczos reads it in its Gershgorin bound;
 *     it is taken as zero.
 *   - FIGI and RATQR are EISPACK routines the reference does not ship (its
 *     calls match neither EISPACK's nor J. Burkardt's eispack.f90 argument
 *     lists).  FIGI's symmetrisation is restated (E(I) = SQRT(T(I,1)*
 *     T(I-1,3)), D = T(:,2)); for RATQR, which returns the smallest
 *     eigenvalue of the symmetric tridiagonal matrix to a tolerance, the
 *     eigenvalue itself is computed by bisection on Sturm counts. */

static void krylov_bnd(double *x, int i2, int j2, int periodic)
{
  const int i0 = i2 + 2;
  if (!periodic) return;
  for (int j = 2; j <= j2 + 1; ++j) {
    x[IDX2(i0, 1, j)] = x[IDX2(i0, i2 + 1, j)];
    x[IDX2(i0, i2 + 2, j)] = x[IDX2(i0, 2, j)];
  }
}

/* p_inprod (:741-755): sequential, j outer, i inner, over the interior. */
double orc_inprod(const double *r, const double *p, int i2, int j2)
{
  const int i0 = i2 + 2;
  double tp = 0.0;
  for (int j = 2; j <= j2 + 1; ++j)
    for (int i = 2; i <= i2 + 1; ++i) tp = tp + r[IDX2(i0, i, j)] * p[IDX2(i0, i, j)];
  return tp;
}

/* p_sqprod_T (:714-740): the transposed 5-point product, with the
 * coefficients read from zero-padded copies TAB, TAL, TAR, TAT. */
void orc_sqprod_t(const float *al, const float *ab, const float *ac,
                  const float *ar, const float *at, const double *x, double *r,
                  int i2, int j2)
{
  const int i0 = i2 + 2;
  for (int j = 1; j <= j2; ++j)
    for (int i = 1; i <= i2; ++i) {
      const double tat = j >= 2 ? (double)at[IDX2(i2, i, j - 1)] : 0.0;      /* TAT(I+1,J) */
      const double tar = i >= 2 ? (double)ar[IDX2(i2, i - 1, j)] : 0.0;      /* TAR(I,J+1) */
      const double tal = i <= i2 - 1 ? (double)al[IDX2(i2, i + 1, j)] : 0.0; /* TAL(I+2,J+1) */
      const double tab = j <= j2 - 1 ? (double)ab[IDX2(i2, i, j + 1)] : 0.0; /* TAB(I+1,J+2) */
      r[IDX2(i0, i + 1, j + 1)] =
          tat * x[IDX2(i0, i + 1, j)] + tar * x[IDX2(i0, i, j + 1)] +
          (double)ac[IDX2(i2, i, j)] * x[IDX2(i0, i + 1, j + 1)] + tal * x[IDX2(i0, i + 2, j + 1)] +
          tab * x[IDX2(i0, i + 1, j + 2)];
    }
}

static void b_minus(const float *b, double *r, int i2, int j2)
{
  const int i0 = i2 + 2;
  for (int j = 1; j <= j2; ++j)
    for (int i = 1; i <= i2; ++i)
      r[IDX2(i0, i + 1, j + 1)] = (double)b[IDX2(i2, i, j)] - r[IDX2(i0, i + 1, j + 1)];
}

/* p_bicgstab_le (:443-552).  Restarts every `restart` iterations from the
 * true residual with RHO = ALPHA = W = 1 but P and V kept.  Returns the
 * iterations run; *res_out = SQRT(TP) at exit. */
int orc_bicgstab(const float *al, const float *ab, const float *ac, const float *ar,
                 const float *at, const float *b, double *x, int i2, int j2, int periodic,
                 int max_iter, int restart, double tol, double rtol, double *res_out)
{
  const long n = (long)(i2 + 2) * (j2 + 2);
  double *r = calloc(n, 8), *rh = calloc(n, 8), *p = calloc(n, 8), *v = calloc(n, 8);
  double *s = calloc(n, 8), *t = calloc(n, 8), *ph = calloc(n, 8), *sh = calloc(n, 8);
  int iter = 0;
  double tp = 1.0, rho, rhn, alpha, beta, w, res, tmp;
  int early = 0;
  krylov_bnd(x, i2, j2, periodic);
  while (sqrt(tp) > tol && iter < max_iter) {
    rho = 1.0;
    alpha = 1.0;
    w = 1.0;
    orc_sqprod(al, ab, ac, ar, at, x, r, i2, j2);
    b_minus(b, r, i2, j2);
    krylov_bnd(r, i2, j2, periodic);
    memcpy(rh, r, n * 8);
    tp = orc_inprod(r, r, i2, j2);
    res = sqrt(tp);
    if (res < tol) {
      early = 1;
      break;
    }
    while (sqrt(tp) / res > rtol) {
      iter = iter + 1;
      rhn = orc_inprod(r, rh, i2, j2);
      beta = (rhn / rho) * (alpha / w);
      for (long k = 0; k < n; ++k) p[k] = r[k] + beta * (p[k] - w * v[k]);
      memcpy(ph, p, n * 8);
      krylov_bnd(ph, i2, j2, periodic);
      orc_sqprod(al, ab, ac, ar, at, ph, v, i2, j2);
      krylov_bnd(v, i2, j2, periodic);
      tmp = orc_inprod(rh, v, i2, j2);
      alpha = rhn / tmp;
      for (long k = 0; k < n; ++k) s[k] = r[k] - alpha * v[k];
      memcpy(sh, s, n * 8);
      krylov_bnd(sh, i2, j2, periodic);
      orc_sqprod(al, ab, ac, ar, at, sh, t, i2, j2);
      krylov_bnd(t, i2, j2, periodic);
      tmp = orc_inprod(t, t, i2, j2);
      w = orc_inprod(t, s, i2, j2);
      w = w / tmp;
#ifdef ORC_PLANT_BICG_UPDATE /* a planted slip for the tests: SH in place of PH */
      for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * sh[k] + w * sh[k];
#else
      for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * ph[k] + w * sh[k];
#endif
      for (long k = 0; k < n; ++k) r[k] = s[k] - w * t[k];
      tp = orc_inprod(r, r, i2, j2);
      rho = rhn;
      if (iter % restart == 0) break;
    }
  }
  if (!early) krylov_bnd(x, i2, j2, periodic);
  *res_out = sqrt(tp);
  free(r); free(rh); free(p); free(v); free(s); free(t); free(ph); free(sh);
  return iter;
}

/* p_csi (:553-616): Chebyshev iteration on the interval [mineig, maxeig]. */
int orc_pcsi(const float *al, const float *ab, const float *ac, const float *ar,
             const float *at, double mineig, double maxeig, const float *b, double *x,
             int i2, int j2, int periodic, int max_iter, double tol, double *res_out)
{
  const long n = (long)(i2 + 2) * (j2 + 2);
  double *r = calloc(n, 8), *s = calloc(n, 8), *q = calloc(n, 8), *work1 = calloc(n, 8);
  int iter = 0;
  double tp = 1.0;
  const double alpha = 2.0 / (maxeig - mineig);
  const double beta = (maxeig + mineig) / (maxeig - mineig);
  const double csy = beta / alpha;
  double omega = 2.0 / csy;
  krylov_bnd(x, i2, j2, periodic);
  orc_sqprod(al, ab, ac, ar, at, x, s, i2, j2);
  krylov_bnd(s, i2, j2, periodic);
  for (int j = 1; j <= j2; ++j)
    for (int i = 1; i <= i2; ++i)
      r[IDX2(i2 + 2, i + 1, j + 1)] = (double)b[IDX2(i2, i, j)] - s[IDX2(i2 + 2, i + 1, j + 1)];
  memcpy(work1, r, n * 8);
  krylov_bnd(work1, i2, j2, periodic);
  const double c0 = 1.0 / csy;
  for (long k = 0; k < n; ++k) q[k] = c0 * work1[k];
  for (long k = 0; k < n; ++k) x[k] = x[k] + q[k];
  orc_sqprod(al, ab, ac, ar, at, x, s, i2, j2);
  krylov_bnd(s, i2, j2, periodic);
  for (int j = 1; j <= j2; ++j)
    for (int i = 1; i <= i2; ++i)
      r[IDX2(i2 + 2, i + 1, j + 1)] = (double)b[IDX2(i2, i, j)] - s[IDX2(i2 + 2, i + 1, j + 1)];
  while (sqrt(tp) > tol && iter < max_iter) {
    omega = 1.0 / (csy - omega / (4.0 * alpha * alpha));
    memcpy(work1, r, n * 8);
    krylov_bnd(work1, i2, j2, periodic);
    const double cq = csy * omega - 1.0;
    for (long k = 0; k < n; ++k) q[k] = omega * work1[k] + cq * q[k];
    for (long k = 0; k < n; ++k) x[k] = x[k] + q[k];
    orc_sqprod(al, ab, ac, ar, at, x, s, i2, j2);
    for (int j = 1; j <= j2; ++j)
      for (int i = 1; i <= i2; ++i)
        r[IDX2(i2 + 2, i + 1, j + 1)] = (double)b[IDX2(i2, i, j)] - s[IDX2(i2 + 2, i + 1, j + 1)];
    krylov_bnd(r, i2, j2, periodic);
    tp = orc_inprod(r, r, i2, j2);
    iter = iter + 1;
  }
  krylov_bnd(x, i2, j2, periodic);
  *res_out = sqrt(tp);
  free(r); free(s); free(q); free(work1);
  return iter;
}

/* Number of eigenvalues of the symmetric tridiagonal matrix (d, e) below x,
 * by the signs of the LDL^T pivots (Sturm count); e[i] couples rows i-1, i. */
static int sturm_below(const double *d, const double *e, int m, double x)
{
  int cnt = 0;
  double q = 1.0;
  for (int i = 0; i < m; ++i) {
    const double off = i > 0 ? e[i] * e[i] / q : 0.0;
    q = (d[i] - x) - off;
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

/* Smallest eigenvalue of the symmetric tridiagonal matrix by bisection
 * between its Gershgorin bounds, to the last bit. */
double orc_tridiag_min_eig(const double *d, const double *e, int m)
{
  double lo = d[0], hi = d[0];
  for (int i = 0; i < m; ++i) {
    const double rad = (i > 0 ? fabs(e[i]) : 0.0) + (i + 1 < m ? fabs(e[i + 1]) : 0.0);
    if (d[i] - rad < lo) lo = d[i] - rad;
    if (d[i] + rad > hi) hi = d[i] + rad;
  }
  for (int it = 0; it < 2200; ++it) {
    const double mid = 0.5 * lo + 0.5 * hi;
    if (mid <= lo || mid >= hi) break;
    if (sturm_below(d, e, m, mid) >= 1) hi = mid;
    else lo = mid;
  }
  return hi;
}

/* lanczos (:618-694): M steps of two-sided Lanczos from V = W = 1/||1||,
 * then the eigenvalue estimates.  alpha has m entries, beta and sigma m+1
 * (index k holds the reference's element k+1). */
void orc_lanczos(const float *al, const float *ab, const float *ac, const float *ar,
                 const float *at, int i2, int j2, int m, double *alpha, double *beta,
                 double *sigma, double *mineig, double *maxeig)
{
  const long n = (long)(i2 + 2) * (j2 + 2);
  double *v = malloc(n * 8), *w = malloc(n * 8), *vp = calloc(n, 8), *wp = calloc(n, 8);
  double *s = calloc(n, 8), *q = calloc(n, 8), *work = calloc(n, 8);
  double tp;
  for (long k = 0; k < n; ++k) v[k] = 1.0;
  for (int k = 0; k < m; ++k) alpha[k] = 0.0;
  for (int k = 0; k <= m; ++k) beta[k] = sigma[k] = 0.0;
  tp = orc_inprod(v, v, i2, j2);
  const double c = 1.0 / sqrt(tp);
  for (long k = 0; k < n; ++k) v[k] = c * v[k];
  memcpy(w, v, n * 8);
  for (int it = 1; it <= m; ++it) {
    orc_sqprod(al, ab, ac, ar, at, v, s, i2, j2);
    tp = orc_inprod(s, w, i2, j2);
    alpha[it - 1] = tp;
    for (long k = 0; k < n; ++k) work[k] = beta[it - 1] * vp[k];
    memcpy(vp, v, n * 8);
    for (long k = 0; k < n; ++k) v[k] = s[k] - alpha[it - 1] * v[k] - work[k];
    orc_sqprod_t(al, ab, ac, ar, at, w, q, i2, j2);
    for (long k = 0; k < n; ++k) work[k] = sigma[it - 1] * wp[k];
    memcpy(wp, w, n * 8);
    for (long k = 0; k < n; ++k) w[k] = q[k] - alpha[it - 1] * w[k] - work[k];
    tp = orc_inprod(v, w, i2, j2);
    sigma[it] = sqrt(fabs(tp));
    beta[it] = copysign(sigma[it], tp);
    for (long k = 0; k < n; ++k) v[k] = v[k] / sigma[it];
    for (long k = 0; k < n; ++k) w[k] = w[k] / beta[it];
  }
  orc_sqprod(al, ab, ac, ar, at, v, s, i2, j2);
  tp = orc_inprod(s, w, i2, j2);
  alpha[m - 1] = tp;
  /* FIGI: D = ALPHA, E(I) = SQRT(SIGMA(I)*SIGMA(I)) for I = 2..M, E(1) = 0 */
  double *d = malloc(m * 8), *e = malloc(m * 8);
  for (int i = 0; i < m; ++i) {
    d[i] = alpha[i];
    e[i] = i > 0 ? sqrt(sigma[i] * sigma[i]) : 0.0;
  }
  double mx = d[0] + e[0];
  for (int i = 1; i < m; ++i) {
    const double g = d[i] + e[i] + e[i - 1];
    mx = g > mx ? g : mx;
  }
  *maxeig = mx;
  *mineig = orc_tridiag_min_eig(d, e, m);
  free(d); free(e);
  free(v); free(w); free(vp); free(wp); free(s); free(q); free(work);
}
