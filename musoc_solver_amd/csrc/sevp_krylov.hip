// Krylov cross-check and fallback solvers on the device: BiCGSTAB, P-CSI and
// two-sided Lanczos of MAIN/template/GLOBAL/src/sevp.f90:443-795.
//
// Every vector is an (I2+2) x (J2+2) FP64 array, column major, halos
// included, as in the reference.  Element-wise statements are evaluated for
// every element, halos too, with the reference's operation order; 5-point
// products follow p_sqprod (AB, AL, AC, AR, AT) and p_sqprod_T.  The
// reference's periodic halo update (solver_boundary_update) is folded into
// the producers: a halo column element is computed as its interior mirror,
// which gives the same bits as computing the interior element and copying.
//
// Inner products are the only place where the order of summation is free.
// Two modes:
//   fast      each workgroup sums its contiguous chunk in a fixed order, the
//             last workgroup to finish (device-scope fence + counter) sums
//             the partials in a fixed order and advances the scalar
//             recurrences on the device; deterministic from run to run.
//   ref_order one thread sums the interior sequentially, j outer and i
//             inner, as p_inprod does: the whole iteration is then bitwise
//             the reference's (slow: for parity checks on small grids).
// BiCGSTAB's scalars (RHO, ALPHA, W, BETA) and its stop flag live in device
// memory, so iterations are enqueued without a host round trip; the host
// only checks the flag between batches of iterations.
#include <math.h>

#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

constexpr int KT = 256;  // threads per workgroup

// Workgroup b of a grid of nb (a multiple of 8) owns logical chunk
// (b % 8) * (nb / 8) + b / 8: consecutive chunks, hence neighbouring rows,
// land on the same XCD (workgroups are dealt round-robin over the 8 XCDs).
__device__ inline int kry_chunk_id() {
  const int nb = gridDim.x;
  return (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
}

// Sum over the workgroup in a fixed order; the result is valid in thread 0.
__device__ inline double kry_block_sum(double v, double *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < KT / 64; ++k) s += lds[k];
  __syncthreads();
  return s;
}

// Publishes this workgroup's partial sums; returns true in the last
// workgroup to finish, whose thread 0 then holds the totals in out[].
template <int ND>
__device__ bool kry_finish(const double (&v)[ND], double *part, unsigned *count, double (&out)[ND]) {
  __shared__ double lds[KT / 64];
  __shared__ int last;
  const int cid = kry_chunk_id();
  for (int d = 0; d < ND; ++d) {
    const double s = kry_block_sum(v[d], lds);
    if (threadIdx.x == 0) part[(long)d * KRY_MAXB + cid] = s;
  }
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(count, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  for (int d = 0; d < ND; ++d) {
    double a = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += KT) a += part[(long)d * KRY_MAXB + k];
    out[d] = kry_block_sum(a, lds);
  }
  if (threadIdx.x == 0) *count = 0u;
  return threadIdx.x == 0;
}

// ----------------------------------------------------- scalar recurrences
// The statements of p_bicgstab_le / p_csi that follow each inner product.
__device__ inline void fin_restart(KryState *st, double rr) {
  st->tp = rr;
  st->res = sqrt(rr);
  st->rhn = rr;  // RH = R, so (R, RH) is (R, R) term by term
  st->rho = 1.0;
  st->alpha = 1.0;
  st->w = 1.0;
  st->beta = (st->rhn / st->rho) * (st->alpha / st->w);
  st->stop = st->res < st->tol ? 2 : 0;
}
__device__ inline void fin_pv(KryState *st, double rv) { st->alpha = st->rhn / rv; }
__device__ inline void fin_st(KryState *st, double tt, double ts) { st->w = ts / tt; }
__device__ inline void fin_xr(KryState *st, double rr, double rrh) {
  st->tp = rr;
  st->iter += 1;
  st->rho = st->rhn;
  st->rhn = rrh;
  st->beta = (st->rhn / st->rho) * (st->alpha / st->w);
  if (!(sqrt(st->tp) / st->res > st->rtol) || st->iter % st->restart == 0) st->stop = 1;
}
__device__ inline void fin_pcsi(KryState *st, double rr) {
  st->tp = rr;
  st->iter += 1;
  if (!(sqrt(st->tp) > st->tol && st->iter < st->max_iter)) st->stop = 1;
}

// --------------------------------------------------------------- helpers
struct Cells {
  long n, chunk;
  int I0, I2, J2;
  __device__ Cells(int i2, int j2) : I0(i2 + 2), I2(i2), J2(j2) {
    n = (long)I0 * (j2 + 2);
    const long per = (n + gridDim.x - 1) / gridDim.x;
    chunk = (per + KT - 1) / KT * KT;
  }
  __device__ long begin() const { return (long)kry_chunk_id() * chunk + threadIdx.x; }
  __device__ long end() const {
    const long e = ((long)kry_chunk_id() + 1) * chunk;
    return e < n ? e : n;
  }
};

// The element whose stencil value a halo or interior element holds: the
// element itself (interior), its periodic mirror (halo column of an
// interior row, periodic), or none (-1: the value is zero).
__device__ inline long stencil_site(int i, int j, int I0, int I2, int J2, int periodic) {
  if (j < 1 || j > J2) return -1;
  if (i >= 1 && i <= I2) return i + (long)j * I0;
  if (!periodic) return -1;
  return (i == 0 ? I2 : 1) + (long)j * I0;
}

// p_sqprod at site cc (coefficients of interior element k) of the field f.
template <class F>
__device__ inline double sqprod_at(const KryOp &op, long cc, long k, int I0, F f) {
  return (double)op.ab[k] * f(cc - I0) + (double)op.al[k] * f(cc - 1) +
         (double)op.ac[k] * f(cc) + (double)op.ar[k] * f(cc + 1) + (double)op.at[k] * f(cc + I0);
}

__device__ inline long coef_index(long cc, int I0, int I2) {
  const int i = (int)(cc % I0), j = (int)(cc / I0);
  return (i - 1) + (long)(j - 1) * I2;
}

// ------------------------------------------------------------- BiCGSTAB
// R = B - A X, RH = R, (R, R) (the start of each restart cycle).
__global__ __launch_bounds__(KT) void bicg_restart_kernel(KryOp op, const float *b, const double *x,
                                                          double *r, double *rh, KryState *st,
                                                          double *part, int fast) {
  const Cells g(op.i2, op.j2);
  double acc[1] = {0.0};
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % g.I0), j = (int)(c / g.I0);
    const long cc = stencil_site(i, j, g.I0, g.I2, g.J2, op.periodic);
    double v = 0.0;
    if (cc >= 0) {
      const long k = coef_index(cc, g.I0, g.I2);
      const double ax = sqprod_at(op, cc, k, g.I0, [&](long e) { return x[e]; });
      v = (double)b[k] - ax;
    }
    r[c] = v;
    rh[c] = v;
    if (i >= 1 && i <= g.I2 && j >= 1 && j <= g.J2) acc[0] += v * v;
  }
  if (!fast) return;
  double out[1];
  if (kry_finish(acc, part, &st->count, out)) fin_restart(st, out[0]);
}

// The current P and V are p[iter & 1], v[iter & 1]: each iteration writes
// the other pair, so that no workgroup reads a neighbour another one is
// rewriting, and the parity follows the device's own iteration count (the
// host may enqueue iterations past the stop).  Every workgroup reads
// st->iter before the last one to finish advances it.

// P = R + BETA*(P - W*V); V = A P; (RH, V).
__global__ __launch_bounds__(KT) void bicg_pv_kernel(KryOp op, KryVecs q, KryState *st, double *part,
                                                     int fast) {
  if (st->stop) return;
  const double beta = st->beta, w = st->w;
  const int cur = st->iter & 1;
  const double *r = q.r, *rh = q.rh, *pin = q.p[cur], *vin = q.v[cur];
  double *pout = q.p[cur ^ 1], *vout = q.v[cur ^ 1];
  const Cells g(op.i2, op.j2);
  auto pn = [&](long e) { return r[e] + beta * (pin[e] - w * vin[e]); };
  double acc[1] = {0.0};
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % g.I0), j = (int)(c / g.I0);
    pout[c] = pn(c);
    const long cc = stencil_site(i, j, g.I0, g.I2, g.J2, op.periodic);
    double v = 0.0;
    if (cc >= 0) v = sqprod_at(op, cc, coef_index(cc, g.I0, g.I2), g.I0, pn);
    vout[c] = v;
    if (i >= 1 && i <= g.I2 && j >= 1 && j <= g.J2) acc[0] += rh[c] * v;
  }
  if (!fast) return;
  double out[1];
  if (kry_finish(acc, part, &st->count, out)) fin_pv(st, out[0]);
}

// S = R - ALPHA*V; T = A S; (T, T), (T, S).
__global__ __launch_bounds__(KT) void bicg_st_kernel(KryOp op, KryVecs q, KryState *st, double *part,
                                                     int fast) {
  if (st->stop) return;
  const double alpha = st->alpha;
  const double *r = q.r, *v = q.v[(st->iter & 1) ^ 1];
  double *s = q.s, *t = q.t;
  const Cells g(op.i2, op.j2);
  auto sn = [&](long e) { return r[e] - alpha * v[e]; };
  double acc[2] = {0.0, 0.0};
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % g.I0), j = (int)(c / g.I0);
    const double sv = sn(c);
    s[c] = sv;
    const long cc = stencil_site(i, j, g.I0, g.I2, g.J2, op.periodic);
    double tv = 0.0;
    if (cc >= 0) tv = sqprod_at(op, cc, coef_index(cc, g.I0, g.I2), g.I0, sn);
    t[c] = tv;
    if (i >= 1 && i <= g.I2 && j >= 1 && j <= g.J2) {
      acc[0] += tv * tv;
      acc[1] += tv * sv;
    }
  }
  if (!fast) return;
  double out[2];
  if (kry_finish(acc, part, &st->count, out)) fin_st(st, out[0], out[1]);
}

// X = X + ALPHA*P + W*S; R = S - W*T; (R, R), (R, RH).
__global__ __launch_bounds__(KT) void bicg_xr_kernel(KryOp op, KryVecs q, KryState *st, double *part,
                                                     int fast) {
  if (st->stop) return;
  const double alpha = st->alpha, w = st->w;
  const double *p = q.p[(st->iter & 1) ^ 1], *s = q.s, *t = q.t, *rh = q.rh;
  double *x = q.x, *r = q.r;
  const Cells g(op.i2, op.j2);
  double acc[2] = {0.0, 0.0};
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % g.I0), j = (int)(c / g.I0);
    x[c] = x[c] + alpha * p[c] + w * s[c];
    const double rv = s[c] - w * t[c];
    r[c] = rv;
    if (i >= 1 && i <= g.I2 && j >= 1 && j <= g.J2) {
      acc[0] += rv * rv;
      acc[1] += rv * rh[c];
    }
  }
  if (!fast) return;
  double out[2];
  if (kry_finish(acc, part, &st->count, out)) fin_xr(st, out[0], out[1]);
}

// p_inprod, sequentially: one thread, j outer, i inner.
__device__ inline double seq_dot(const double *a, const double *b, int i2, int j2) {
  const int I0 = i2 + 2;
  double tp = 0.0;
  for (int j = 1; j <= j2; ++j)
    for (int i = 1; i <= i2; ++i) tp = tp + a[i + (long)j * I0] * b[i + (long)j * I0];
  return tp;
}

// The reference-order counterparts of the finishing sums.  which: 0
// restart, 1 PV, 2 ST, 3 XR, 4 P-CSI.
__global__ void bicg_seq_kernel(int which, int i2, int j2, KryVecs q, KryState *st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (which == 0) {
    fin_restart(st, seq_dot(q.r, q.r, i2, j2));
    return;
  }
  if (st->stop) return;
  const int nxt = (st->iter & 1) ^ 1;
  if (which == 1) {
    fin_pv(st, seq_dot(q.rh, q.v[nxt], i2, j2));
  } else if (which == 2) {
    const double tt = seq_dot(q.t, q.t, i2, j2);
    const double ts = seq_dot(q.t, q.s, i2, j2);
    fin_st(st, tt, ts);
  } else if (which == 3) {
    const double rr = seq_dot(q.r, q.r, i2, j2);
    const double rrh = seq_dot(q.r, q.rh, i2, j2);
    fin_xr(st, rr, rrh);
  } else if (which == 4) {
    fin_pcsi(st, seq_dot(q.r, q.r, i2, j2));
  }
}

// solver_boundary_update on X (periodic only).
__global__ void kry_bnd_kernel(double *x, int i2, int j2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (j > j2) return;
  const long I0 = i2 + 2;
  x[0 + j * I0] = x[i2 + j * I0];
  x[(i2 + 1) + j * I0] = x[1 + j * I0];
}

__global__ void kry_fill_kernel(double *a, long n, double v) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}

// ----------------------------------------------------------------- P-CSI
// WORK1 = R with the periodic halo; Q = OMEGA*WORK1 + (CSY*OMEGA-1)*Q (or
// Q = (1/CSY)*WORK1 on the first step); X = X + Q.
__global__ __launch_bounds__(KT) void pcsi_qx_kernel(int i2, int j2, int periodic, const double *r,
                                                     double *q, double *x, const double *omega,
                                                     const double *cq, double c0, int first,
                                                     const KryState *st) {
  if (!first && st->stop) return;
  const Cells g(i2, j2);
  const double om = first ? 0.0 : omega[st->iter];
  const double cf = first ? 0.0 : cq[st->iter];
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % g.I0), j = (int)(c / g.I0);
    long src = c;
    if (periodic && (i == 0 || i == g.I0 - 1) && j >= 1 && j <= g.J2)
      src = (i == 0 ? g.I2 : 1) + (long)j * g.I0;
    const double w1 = r[src];
    const double qv = first ? c0 * w1 : om * w1 + cf * q[c];
    q[c] = qv;
    x[c] = x[c] + qv;
  }
}

// R = B - A X (interior, periodic halo columns), optionally (R, R).
__global__ __launch_bounds__(KT) void pcsi_r_kernel(KryOp op, const float *b, const double *x,
                                                    double *r, KryState *st, double *part,
                                                    int with_dot, int fast) {
  if (with_dot && st->stop) return;
  const Cells g(op.i2, op.j2);
  double acc[1] = {0.0};
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % g.I0), j = (int)(c / g.I0);
    const long cc = stencil_site(i, j, g.I0, g.I2, g.J2, op.periodic);
    double v = 0.0;
    if (cc >= 0) {
      const long k = coef_index(cc, g.I0, g.I2);
      v = (double)b[k] - sqprod_at(op, cc, k, g.I0, [&](long e) { return x[e]; });
    }
    r[c] = v;
    if (i >= 1 && i <= g.I2 && j >= 1 && j <= g.J2) acc[0] += v * v;
  }
  if (!with_dot || !fast) return;
  double out[1];
  if (kry_finish(acc, part, &st->count, out)) fin_pcsi(st, out[0]);
}

// --------------------------------------------------------------- Lanczos
// S = A V (transposed = 0, p_sqprod) or A^T V (p_sqprod_T) on the interior;
// the halo of S is not written.
__global__ __launch_bounds__(KT) void kry_spmv_kernel(KryOp op, const double *v, double *s,
                                                      int transposed) {
  const Cells g(op.i2, op.j2);
  const int I0 = g.I0, I2 = g.I2, J2 = g.J2;
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % I0), j = (int)(c / I0);
    if (i < 1 || i > I2 || j < 1 || j > J2) continue;
    const long k = (i - 1) + (long)(j - 1) * I2;
    if (!transposed) {
      s[c] = sqprod_at(op, c, k, I0, [&](long e) { return v[e]; });
    } else {
      // zero-padded TAT(I+1,J), TAR(I,J+1), TAL(I+2,J+1), TAB(I+1,J+2)
      const double tat = j >= 2 ? (double)op.at[k - I2] : 0.0;
      const double tar = i >= 2 ? (double)op.ar[k - 1] : 0.0;
      const double tal = i <= I2 - 1 ? (double)op.al[k + 1] : 0.0;
      const double tab = j <= J2 - 1 ? (double)op.ab[k + I2] : 0.0;
      s[c] = tat * v[c - I0] + tar * v[c - 1] + (double)op.ac[k] * v[c] + tal * v[c + 1] +
             tab * v[c + I0];
    }
  }
}

// OUT = (S - ALPHA*V) - WORK with WORK = BETA*VP; OUT may alias VP.
__global__ void kry_lincomb_kernel(double *out, const double *s, double alpha, const double *v,
                                   double beta, const double *vp, long n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double work = beta * vp[k];
  out[k] = s[k] - alpha * v[k] - work;
}

__global__ void kry_scale_kernel(double *v, double sig, double *w, double bet, long n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  v[k] = v[k] / sig;
  w[k] = w[k] / bet;
}

// One inner product over the interior into *out (device).
__global__ __launch_bounds__(KT) void kry_dot_kernel(int i2, int j2, const double *a, const double *b,
                                                     double *out, double *part, unsigned *count,
                                                     int fast) {
  if (!fast) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = seq_dot(a, b, i2, j2);
    return;
  }
  const Cells g(i2, j2);
  double acc[1] = {0.0};
  for (long c = g.begin(); c < g.end(); c += KT) {
    const int i = (int)(c % g.I0), j = (int)(c / g.I0);
    if (i >= 1 && i <= g.I2 && j >= 1 && j <= g.J2) acc[0] += a[c] * b[c];
  }
  double o[1];
  if (kry_finish(acc, part, count, o)) *out = o[0];
}

// -------------------------------------------------------------- launchers
int kry_workgroups(int i2, int j2) {
  const long n = (long)(i2 + 2) * (j2 + 2);
  long nb = (n + KT - 1) / KT;
  if (nb > KRY_MAXB) nb = KRY_MAXB;
  nb = (nb + 7) / 8 * 8;
  return (int)nb;
}

static dim3 kgrid(const KryOp &op) { return dim3((unsigned)kry_workgroups(op.i2, op.j2)); }

hipError_t kry_bicg_restart(const KryOp &op, const KryVecs &q, KryState *st, double *part, int fast,
                            hipStream_t s) {
  hipLaunchKernelGGL(bicg_restart_kernel, kgrid(op), dim3(KT), 0, s, op, q.b, q.x, q.r, q.rh, st, part,
                     fast);
  if (!fast) hipLaunchKernelGGL(bicg_seq_kernel, dim3(1), dim3(64), 0, s, 0, op.i2, op.j2, q, st);
  return hipGetLastError();
}

hipError_t kry_bicg_iter(const KryOp &op, const KryVecs &q, KryState *st, double *part, int fast,
                         hipStream_t s) {
  const dim3 g = kgrid(op), blk(KT), one(1), w64(64);
  hipLaunchKernelGGL(bicg_pv_kernel, g, blk, 0, s, op, q, st, part, fast);
  if (!fast) hipLaunchKernelGGL(bicg_seq_kernel, one, w64, 0, s, 1, op.i2, op.j2, q, st);
  hipLaunchKernelGGL(bicg_st_kernel, g, blk, 0, s, op, q, st, part, fast);
  if (!fast) hipLaunchKernelGGL(bicg_seq_kernel, one, w64, 0, s, 2, op.i2, op.j2, q, st);
  hipLaunchKernelGGL(bicg_xr_kernel, g, blk, 0, s, op, q, st, part, fast);
  if (!fast) hipLaunchKernelGGL(bicg_seq_kernel, one, w64, 0, s, 3, op.i2, op.j2, q, st);
  return hipGetLastError();
}

hipError_t kry_bnd(double *x, int i2, int j2, hipStream_t s) {
  hipLaunchKernelGGL(kry_bnd_kernel, dim3((unsigned)((j2 + 255) / 256)), dim3(256), 0, s, x, i2, j2);
  return hipGetLastError();
}

hipError_t kry_fill(double *a, long n, double v, hipStream_t s) {
  hipLaunchKernelGGL(kry_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, v);
  return hipGetLastError();
}

hipError_t kry_pcsi_qx(int i2, int j2, int periodic, const double *r, double *q, double *x,
                       const double *omega, const double *cq, double c0, int first,
                       const KryState *st, hipStream_t s) {
  hipLaunchKernelGGL(pcsi_qx_kernel, dim3((unsigned)kry_workgroups(i2, j2)), dim3(KT), 0, s, i2, j2,
                     periodic, r, q, x, omega, cq, c0, first, st);
  return hipGetLastError();
}

hipError_t kry_pcsi_r(const KryOp &op, const float *b, const double *x, double *r, KryState *st,
                      double *part, int with_dot, int fast, hipStream_t s) {
  hipLaunchKernelGGL(pcsi_r_kernel, kgrid(op), dim3(KT), 0, s, op, b, x, r, st, part, with_dot, fast);
  if (with_dot && !fast) {
    KryVecs q{};
    q.r = r;
    hipLaunchKernelGGL(bicg_seq_kernel, dim3(1), dim3(64), 0, s, 4, op.i2, op.j2, q, st);
  }
  return hipGetLastError();
}

hipError_t kry_spmv(const KryOp &op, const double *v, double *out, int transposed, hipStream_t s) {
  hipLaunchKernelGGL(kry_spmv_kernel, kgrid(op), dim3(KT), 0, s, op, v, out, transposed);
  return hipGetLastError();
}

hipError_t kry_lincomb(double *out, const double *sv, double alpha, const double *v, double beta,
                       const double *vp, long n, hipStream_t s) {
  hipLaunchKernelGGL(kry_lincomb_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, sv,
                     alpha, v, beta, vp, n);
  return hipGetLastError();
}

hipError_t kry_scale(double *v, double sig, double *w, double bet, long n, hipStream_t s) {
  hipLaunchKernelGGL(kry_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, sig, w,
                     bet, n);
  return hipGetLastError();
}

hipError_t kry_dot(int i2, int j2, const double *a, const double *b, double *out, double *part,
                   unsigned *count, int fast, hipStream_t s) {
  hipLaunchKernelGGL(kry_dot_kernel, fast ? dim3((unsigned)kry_workgroups(i2, j2)) : dim3(1), dim3(KT),
                     0, s, i2, j2, a, b, out, part, count, fast);
  return hipGetLastError();
}

}  // namespace sevp

namespace sevp {

// Number of eigenvalues of the symmetric tridiagonal matrix (d, e) below x
// (signs of the LDL^T pivots); e[i] couples rows i-1 and i, e[0] unused.
static int sturm_below(const double *d, const double *e, int m, double x) {
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

double kry_tridiag_min_eig(const double *d, const double *e, int m) {
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

}  // namespace sevp
