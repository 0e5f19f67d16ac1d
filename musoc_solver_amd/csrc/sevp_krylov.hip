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
// Each kernel that needs one stores per-workgroup partial sums (its chunk in
// a fixed order); a one-workgroup finishing kernel launched after it adds
// them in a fixed order and advances the scalar recurrences on the device,
// so results are deterministic from run to run.  (A last-workgroup-finishes
// scheme inside the kernel needs a device-scope release per workgroup, an L2
// writeback on this GPU; measured, those cost more than the extra launch.)
// With ref_order the finishing kernel instead sums the interior with one
// thread, j outer and i inner, as p_inprod does: every iterate is then
// bitwise the reference's (slow: for parity checks on small grids).
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

// Sum over a workgroup of NT threads in a fixed order; the result is valid
// in thread 0.
template <int NT = KT>
__device__ inline double kry_block_sum(double v, double *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NT / 64; ++k) s += lds[k];
  __syncthreads();
  return s;
}

// Stores this workgroup's sums (thread 0) as partials of chunk
// kry_chunk_id(); the finishing kernel that follows on the stream adds them.
template <int ND>
__device__ void kry_partials(const double (&v)[ND], double *part) {
  __shared__ double lds[KT / 64];
  const int cid = kry_chunk_id();
  for (int d = 0; d < ND; ++d) {
    const double s = kry_block_sum(v[d], lds);
    if (threadIdx.x == 0) part[(long)d * KRY_MAXB + cid] = s;
  }
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
  int n, chunk, I0, I2, J2;
  __device__ Cells(int i2, int j2) : I0(i2 + 2), I2(i2), J2(j2) {
    n = I0 * (j2 + 2);
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    chunk = (per + KT - 1) / KT * KT;
  }
  __device__ int begin() const { return kry_chunk_id() * chunk + (int)threadIdx.x; }
  __device__ int end() const {
    const int e = (kry_chunk_id() + 1) * chunk;
    return e < n ? e : n;
  }
};

// One thread's elements c = begin, begin + KT, ... of its chunk, with the
// column i and row j (0-based, halos included) stepped along without a
// division per element.
struct Walk {
  int c, e, i, j, I0;
  __device__ explicit Walk(const Cells &g) : c(g.begin()), e(g.end()), I0(g.I0) {
    i = c % I0;
    j = c / I0;
  }
  __device__ bool ok() const { return c < e; }
  __device__ void next() {
    c += KT;
    i += KT;
    while (i >= I0) {
      i -= I0;
      ++j;
    }
  }
};

// The column whose 5-point value element (i, j) holds: i itself (interior),
// its periodic mirror (halo column of an interior row, periodic), or -1 (the
// value is zero).
__device__ inline int site_col(int i, int j, int I2, int J2, int periodic) {
  if (j < 1 || j > J2) return -1;
  if (i >= 1 && i <= I2) return i;
  if (!periodic) return -1;
  return i == 0 ? I2 : 1;
}

// p_sqprod at site cc (coefficients of interior element k) of the field f.
template <class F>
__device__ inline double sqprod_at(const KryOp &op, int cc, int k, int I0, F f) {
  return (double)op.ab[k] * f(cc - I0) + (double)op.al[k] * f(cc - 1) +
         (double)op.ac[k] * f(cc) + (double)op.ar[k] * f(cc + 1) + (double)op.at[k] * f(cc + I0);
}

// ------------------------------------------------------------- BiCGSTAB
// Loads are issued unconditionally (an element without a 5-point value reads
// the valid site (1,1) and selects zero), through restrict-qualified
// pointers, so that a thread's operands are all in flight before its stores.

struct Site {
  int cc, k;  // element and coefficient index of the 5-point site
  bool on;    // whether the element holds a 5-point value
  __device__ Site(int i, int j, int I0, int I2, int J2, int periodic) {
    const int si = site_col(i, j, I2, J2, periodic);
    on = si >= 0;
    const int sc = on ? si : 1, sj = on ? j : 1;
    cc = sc + sj * I0;
    k = (sc - 1) + (sj - 1) * I2;
  }
};

__device__ inline bool interior(int i, int j, const Cells &g) {
  return i >= 1 && i <= g.I2 && j >= 1 && j <= g.J2;
}

// R = B - A X, RH = R, (R, R) (the start of each restart cycle).
__global__ __launch_bounds__(KT) void bicg_restart_kernel(KryOp op, const float *__restrict__ b,
                                                          const double *__restrict__ x,
                                                          double *__restrict__ r,
                                                          double *__restrict__ rh, KryState *st,
                                                          double *part, int fast) {
  const Cells g(op.i2, op.j2);
  double acc[1] = {0.0};
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    const Site t(i, j, g.I0, g.I2, g.J2, op.periodic);
    const double bk = (double)b[t.k];
    const double ax = sqprod_at(op, t.cc, t.k, g.I0, [&](int e) { return x[e]; });
    const double v = t.on ? bk - ax : 0.0;
    r[c] = v;
    rh[c] = v;
    if (interior(i, j, g)) acc[0] += v * v;
  }
  if (fast) kry_partials(acc, part);
}

// The current P and V are p[iter & 1], v[iter & 1]: each iteration writes
// the other pair, so that no workgroup reads a neighbour another one is
// rewriting, and the parity follows the device's own iteration count (the
// host may enqueue iterations past the stop); the XR finishing kernel
// advances it after the iteration's last kernel.

// P = R + BETA*(P - W*V); V = A P; (RH, V).
__global__ __launch_bounds__(KT) void bicg_pv_kernel(KryOp op, KryVecs q, KryState *st, double *part,
                                                     int fast) {
  if (st->stop) return;
  const double beta = st->beta, w = st->w;
  const int cur = st->iter & 1;
  const double *__restrict__ r = q.r;
  const double *__restrict__ rh = q.rh;
  const double *__restrict__ pin = q.p[cur];
  const double *__restrict__ vin = q.v[cur];
  double *__restrict__ pout = q.p[cur ^ 1];
  double *__restrict__ vout = q.v[cur ^ 1];
  const Cells g(op.i2, op.j2);
  auto pn = [&](int e) { return r[e] + beta * (pin[e] - w * vin[e]); };
  double acc[1] = {0.0};
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    const Site t(i, j, g.I0, g.I2, g.J2, op.periodic);
    const double pc = pn(c), rhc = rh[c];
    const double av = sqprod_at(op, t.cc, t.k, g.I0, pn);
    const double v = t.on ? av : 0.0;
    pout[c] = pc;
    vout[c] = v;
    if (interior(i, j, g)) acc[0] += rhc * v;
  }
  if (fast) kry_partials(acc, part);
}

// S = R - ALPHA*V; T = A S; (T, T), (T, S).
__global__ __launch_bounds__(KT) void bicg_st_kernel(KryOp op, KryVecs q, KryState *st, double *part,
                                                     int fast) {
  if (st->stop) return;
  const double alpha = st->alpha;
  const double *__restrict__ r = q.r;
  const double *__restrict__ v = q.v[(st->iter & 1) ^ 1];
  double *__restrict__ s = q.s;
  double *__restrict__ t = q.t;
  const Cells g(op.i2, op.j2);
  auto sn = [&](int e) { return r[e] - alpha * v[e]; };
  double acc[2] = {0.0, 0.0};
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    const Site z(i, j, g.I0, g.I2, g.J2, op.periodic);
    const double sv = sn(c);
    const double asv = sqprod_at(op, z.cc, z.k, g.I0, sn);
    const double tv = z.on ? asv : 0.0;
    s[c] = sv;
    t[c] = tv;
    if (interior(i, j, g)) {
      acc[0] += tv * tv;
      acc[1] += tv * sv;
    }
  }
  if (fast) kry_partials(acc, part);
}

// X = X + ALPHA*P + W*S; R = S - W*T; (R, R), (R, RH).
__global__ __launch_bounds__(KT) void bicg_xr_kernel(KryOp op, KryVecs q, KryState *st, double *part,
                                                     int fast) {
  if (st->stop) return;
  const double alpha = st->alpha, w = st->w;
  const double *__restrict__ p = q.p[(st->iter & 1) ^ 1];
  const double *__restrict__ s = q.s;
  const double *__restrict__ t = q.t;
  const double *__restrict__ rh = q.rh;
  double *__restrict__ x = q.x;
  double *__restrict__ r = q.r;
  const Cells g(op.i2, op.j2);
  double acc[2] = {0.0, 0.0};
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    const double xc = x[c], pc = p[c], sc = s[c], tc = t[c], rhc = rh[c];
    const double xn = xc + alpha * pc + w * sc;
    const double rv = sc - w * tc;
    x[c] = xn;
    r[c] = rv;
    if (interior(i, j, g)) {
      acc[0] += rv * rv;
      acc[1] += rv * rhc;
    }
  }
  if (fast) kry_partials(acc, part);
}

// p_inprod, sequentially: one thread, j outer, i inner.
__device__ inline double seq_dot(const double *a, const double *b, int i2, int j2) {
  const int I0 = i2 + 2;
  double tp = 0.0;
  for (int j = 1; j <= j2; ++j)
    for (int i = 1; i <= i2; ++i) tp = tp + a[i + (long)j * I0] * b[i + (long)j * I0];
  return tp;
}

// The sums a finishing kernel adds: the partials of the kernel before it in a
// fixed order (fast), or p_inprod's sequential sum over the vectors
// (reference order).  FT threads, each loading its (at most KRY_MAXB / FT)
// partials at once and then adding them in order.
constexpr int FT = 1024;
constexpr int FU = KRY_MAXB / FT;
__device__ inline double fin_sum(int d, int nb, const double *part, double *lds) {
  double v[FU];
#pragma unroll
  for (int u = 0; u < FU; ++u) {
    const int k = threadIdx.x + u * FT;
    v[u] = k < nb ? part[(long)d * KRY_MAXB + k] : 0.0;
  }
  double a = 0.0;
#pragma unroll
  for (int u = 0; u < FU; ++u) a += v[u];
  return kry_block_sum<FT>(a, lds);
}

// Finishes the inner products of the kernel before it and advances the
// scalar recurrences.  which: 0 restart, 1 PV, 2 ST, 3 XR, 4 P-CSI.  One
// workgroup of KT threads.
__global__ __launch_bounds__(FT) void kry_fin_kernel(int which, int i2, int j2, KryVecs q,
                                                     KryState *st, const double *part, int nb,
                                                     int fast) {
  __shared__ double lds[FT / 64];
  if (which != 0 && st->stop) return;
  double d0 = 0.0, d1 = 0.0;
  const int nd = (which == 2 || which == 3) ? 2 : 1;
  if (fast) {
    d0 = fin_sum(0, nb, part, lds);
    if (nd == 2) d1 = fin_sum(1, nb, part, lds);
  } else if (threadIdx.x == 0) {
    const int nxt = (st->iter & 1) ^ 1;
    switch (which) {
      case 0: d0 = seq_dot(q.r, q.r, i2, j2); break;
      case 1: d0 = seq_dot(q.rh, q.v[nxt], i2, j2); break;
      case 2: d0 = seq_dot(q.t, q.t, i2, j2); d1 = seq_dot(q.t, q.s, i2, j2); break;
      case 3: d0 = seq_dot(q.r, q.r, i2, j2); d1 = seq_dot(q.r, q.rh, i2, j2); break;
      default: d0 = seq_dot(q.r, q.r, i2, j2); break;
    }
  }
  if (threadIdx.x != 0) return;
  switch (which) {
    case 0: fin_restart(st, d0); break;
    case 1: fin_pv(st, d0); break;
    case 2: fin_st(st, d0, d1); break;
    case 3: fin_xr(st, d0, d1); break;
    default: fin_pcsi(st, d0); break;
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
__global__ __launch_bounds__(KT) void pcsi_qx_kernel(int i2, int j2, int periodic,
                                                     const double *__restrict__ r,
                                                     double *__restrict__ q, double *__restrict__ x,
                                                     const double *omega, const double *cq, double c0,
                                                     int first, const KryState *st) {
  if (!first && st->stop) return;
  const Cells g(i2, j2);
  const double om = first ? 0.0 : omega[st->iter];
  const double cf = first ? 0.0 : cq[st->iter];
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    int src = c;
    if (periodic && (i == 0 || i == g.I0 - 1) && j >= 1 && j <= g.J2) src = (i == 0 ? g.I2 : 1) + j * g.I0;
    const double w1 = r[src], qc = q[c], xc = x[c];
    const double qv = first ? c0 * w1 : om * w1 + cf * qc;
    q[c] = qv;
    x[c] = xc + qv;
  }
}

// R = B - A X (interior, periodic halo columns), optionally (R, R).
__global__ __launch_bounds__(KT) void pcsi_r_kernel(KryOp op, const float *__restrict__ b,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ r, KryState *st,
                                                    double *part, int with_dot, int fast) {
  if (with_dot && st->stop) return;
  const Cells g(op.i2, op.j2);
  double acc[1] = {0.0};
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    const Site t(i, j, g.I0, g.I2, g.J2, op.periodic);
    const double bk = (double)b[t.k];
    const double ax = sqprod_at(op, t.cc, t.k, g.I0, [&](int e) { return x[e]; });
    const double v = t.on ? bk - ax : 0.0;
    r[c] = v;
    if (interior(i, j, g)) acc[0] += v * v;
  }
  if (with_dot && fast) kry_partials(acc, part);
}

// --------------------------------------------------------------- Lanczos
// S = A V (transposed = 0, p_sqprod) or A^T V (p_sqprod_T) on the interior;
// the halo of S is not written.
__global__ __launch_bounds__(KT) void kry_spmv_kernel(KryOp op, const double *v, double *s,
                                                      int transposed) {
  const Cells g(op.i2, op.j2);
  const int I0 = g.I0, I2 = g.I2, J2 = g.J2;
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    if (i < 1 || i > I2 || j < 1 || j > J2) continue;
    const int k = (i - 1) + (j - 1) * I2;
    if (!transposed) {
      s[c] = sqprod_at(op, c, k, I0, [&](int e) { return v[e]; });
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

// One inner product over the interior: partials (fast), then *out from
// kry_dot_fin_kernel; in reference order the finishing kernel sums alone.
__global__ __launch_bounds__(KT) void kry_dot_kernel(int i2, int j2, const double *__restrict__ a,
                                                     const double *__restrict__ b, double *part) {
  const Cells g(i2, j2);
  double acc[1] = {0.0};
  for (Walk it(g); it.ok(); it.next()) {
    const int c = it.c, i = it.i, j = it.j;
    const double ab = a[c] * b[c];
    if (interior(i, j, g)) acc[0] += ab;
  }
  kry_partials(acc, part);
}

__global__ __launch_bounds__(FT) void kry_dot_fin_kernel(int i2, int j2, const double *a,
                                                         const double *b, const double *part, int nb,
                                                         int fast, double *out) {
  __shared__ double lds[FT / 64];
  if (fast) {
    const double d = fin_sum(0, nb, part, lds);
    if (threadIdx.x == 0) *out = d;
  } else if (threadIdx.x == 0) {
    *out = seq_dot(a, b, i2, j2);
  }
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
  const dim3 g = kgrid(op);
  hipLaunchKernelGGL(bicg_restart_kernel, g, dim3(KT), 0, s, op, q.b, q.x, q.r, q.rh, st, part, fast);
  hipLaunchKernelGGL(kry_fin_kernel, dim3(1), dim3(FT), 0, s, 0, op.i2, op.j2, q, st, part, (int)g.x, fast);
  return hipGetLastError();
}

hipError_t kry_bicg_iter(const KryOp &op, const KryVecs &q, KryState *st, double *part, int fast,
                         hipStream_t s) {
  const dim3 g = kgrid(op), blk(KT), one(1), fin(FT);
  const int nb = (int)g.x;
  hipLaunchKernelGGL(bicg_pv_kernel, g, blk, 0, s, op, q, st, part, fast);
  hipLaunchKernelGGL(kry_fin_kernel, one, fin, 0, s, 1, op.i2, op.j2, q, st, part, nb, fast);
  hipLaunchKernelGGL(bicg_st_kernel, g, blk, 0, s, op, q, st, part, fast);
  hipLaunchKernelGGL(kry_fin_kernel, one, fin, 0, s, 2, op.i2, op.j2, q, st, part, nb, fast);
  hipLaunchKernelGGL(bicg_xr_kernel, g, blk, 0, s, op, q, st, part, fast);
  hipLaunchKernelGGL(kry_fin_kernel, one, fin, 0, s, 3, op.i2, op.j2, q, st, part, nb, fast);
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
  const dim3 g = kgrid(op);
  hipLaunchKernelGGL(pcsi_r_kernel, g, dim3(KT), 0, s, op, b, x, r, st, part, with_dot, fast);
  if (with_dot) {
    KryVecs q{};
    q.r = r;
    hipLaunchKernelGGL(kry_fin_kernel, dim3(1), dim3(FT), 0, s, 4, op.i2, op.j2, q, st, part, (int)g.x,
                       fast);
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

hipError_t kry_dot(int i2, int j2, const double *a, const double *b, double *out, double *part, int fast,
                   hipStream_t s) {
  const int nb = kry_workgroups(i2, j2);
  if (fast) hipLaunchKernelGGL(kry_dot_kernel, dim3((unsigned)nb), dim3(KT), 0, s, i2, j2, a, b, part);
  hipLaunchKernelGGL(kry_dot_fin_kernel, dim3(1), dim3(FT), 0, s, i2, j2, a, b, part, nb, fast, out);
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
