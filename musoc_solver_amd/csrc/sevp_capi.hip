// C ABI of the SEVP library: plans, influence-matrix builds, solves and the
// one-to-one reference entry points.  See include/sevp.h for the contract.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <deque>
#include <vector>

#include "../../include/sevp.h"
#include "sevp_dev.h"
#include "sevp_internal.h"

using namespace sevp;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

namespace sevp {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace sevp

#define HIPC(expr)                                                                  \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      return fail(_e == hipErrorOutOfMemory ? SEVP_E_NOMEM : SEVP_E_HIP,            \
                  std::string(#expr) + ": " + hipGetErrorString(_e));               \
  } while (0)

#define CHK(expr)              \
  do {                         \
    int _r = (expr);           \
    if (_r != SEVP_OK) return _r; \
  } while (0)


// ---------------------------------------------------------------- device buffers
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    release();
    if (count == 0) return SEVP_OK;
    hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(SEVP_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    n = count;
    return SEVP_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
  ~DBuf() { release(); }
};

// Event pairs around the launches of one kernel class (measurement only).
struct KProf {
  int kclass = -1;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  size_t used = 0;
  double total_ms = 0.0;
  long launches = 0;
  hipStream_t st = nullptr;
  // while a graph is captured: the kernel nodes of the profiled class, which
  // get event-record nodes around them once the capture ends (recording an
  // event on a capturing stream only orders work, it adds no timing node)
  std::vector<hipGraphNode_t> *capture_into = nullptr;
  // events of the graph replays since the last harvest (one per RHS group)
  std::vector<const std::vector<std::pair<hipEvent_t, hipEvent_t>> *> pending;
  bool on(int k) const { return kclass == k; }
  void start(int k) {
    if (!on(k) || capture_into) return;
    if (used == pool.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
      pool.push_back({a, b});
    }
    (void)hipEventRecord(pool[used].first, st);
  }
  void stop(int k) {
    if (!on(k)) return;
    if (capture_into) {
      // the node just captured is the stream's only current dependency
      hipStreamCaptureStatus cs;
      const hipGraphNode_t *deps = nullptr;
      size_t ndeps = 0;
      if (hipStreamGetCaptureInfo_v2(st, &cs, nullptr, nullptr, &deps, &ndeps) == hipSuccess &&
          cs == hipStreamCaptureStatusActive && ndeps == 1)
        capture_into->push_back(deps[0]);
      return;
    }
    if (used >= pool.size()) return;
    (void)hipEventRecord(pool[used].second, st);
    ++used;
  }
  // Kernel spans from the kernel's own clock (self-timed launches outside
  // graph capture): a (start, end) pair of 100 MHz ticks per workgroup.
  static constexpr size_t kSpanWords = 8u << 20;
  unsigned long long *spans = nullptr;  // device, kSpanWords
  std::vector<std::pair<size_t, long>> span_runs;  // (offset, workgroups) per launch
  size_t span_next = 0;
  double span_ms = 0.0;
  long span_launches = 0;
  unsigned long long *next_span(int k, long wgs) {
    if (!on(k) || capture_into || !spans || span_next + 2 * wgs > kSpanWords) return nullptr;
    span_runs.push_back({span_next, wgs});
    unsigned long long *t = spans + span_next;
    span_next += 2 * wgs;
    return t;
  }
  // Launches that time themselves (hipExtLaunchKernelGGL): the event pair for
  // the next launch of class k outside graph capture, else nulls.  Each such
  // launch is followed by stop_ext(k).
  std::pair<hipEvent_t, hipEvent_t> start_ext(int k) {
    if (!on(k) || capture_into) return {nullptr, nullptr};
    if (used == pool.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return {nullptr, nullptr};
      pool.push_back({a, b});
    }
    return pool[used];
  }
  void stop_ext(int k, bool launched_with_events) {
    if (!on(k)) return;
    if (capture_into) {
      stop(k);  // records the captured node
      return;
    }
    if (launched_with_events) ++used;
  }
  // Accumulate the recorded pairs; the stream must be idle.
  void harvest() {
    for (size_t i = 0; i < used; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pool[i].first, pool[i].second) == hipSuccess) {
        total_ms += ms;
        ++launches;
      }
    }
    used = 0;
    for (const auto *ev : pending)
      for (auto &e : *ev) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
          total_ms += ms;
          ++launches;
        }
      }
    pending.clear();
    (void)hipGetLastError();  // a pair that never ran must not fail a later launch check
  }
  // Add up the launches timed since the last call; the stream must be idle.
  void harvest_spans() {
    if (span_runs.empty()) return;
    std::vector<unsigned long long> h(span_next);
    if (hipMemcpy(h.data(), spans, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) ==
        hipSuccess)
      for (const auto &r : span_runs) {
        unsigned long long t0 = ~0ull, t1 = 0;
        bool all = true;
        for (long w = 0; w < r.second; ++w) {
          const unsigned long long a = h[r.first + 2 * w], b = h[r.first + 2 * w + 1];
          if (!a || !b) all = false;
          t0 = std::min(t0, a);
          t1 = std::max(t1, b);
        }
        if (all && t1 > t0) {
          span_ms += 1e-5 * (double)(t1 - t0);
          ++span_launches;
        }
      }
    span_runs.clear();
    span_next = 0;
  }
  ~KProf() {
    if (spans) (void)hipFree(spans);
    for (auto &e : pool) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }
};

// One captured refinement pass (zero the correction, one full-width SEVP pass
// of the FP64 residual) for a batch size and precision.
struct PassGraph {
  int nrhs = 0, prec = 0, kind = 0, waves = 0, rows = 0;
  int r0 = 0, count = 0, ng = 0;  // the RHS group this pass covers (groups of ng)
  bool prof = false;
  hipGraphExec_t exec = nullptr;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

// The whole refinement captured as one graph: the residual of the initial
// guess, then `passes` times (the pass of every RHS group, x += dx, the
// residual, the stop test), with the stop flag on the device (the updates
// after it is set are skipped, so the result is the loop's).  The caller's s
// and x are baked in.
struct RefineGraph {
  int nrhs = 0, prec = 0, kind = 0, groups = 0, passes = 0, waves = 0, rows = 0;
  bool prof = false;
  const float *s = nullptr;
  double *x = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

// Vectors and scalars of the Krylov solvers (BiCGSTAB needs eight vectors).
struct KryWork {
  DBuf<double> vec, part, dot, omega;
  DBuf<KryState> st;
  int ensure(int i2, int j2) {
    const size_t n = (size_t)(i2 + 2) * (j2 + 2);
    if (n > ((size_t)1 << 30)) return fail(SEVP_E_ARG, "Krylov solvers take at most 2^30 grid points");
    if (vec.n < 8 * n) CHK(vec.alloc(8 * n));
    if (!part.p) CHK(part.alloc(2 * (size_t)KRY_MAXB));
    if (!dot.p) CHK(dot.alloc(1));
    if (!st.p) CHK(st.alloc(1));
    return SEVP_OK;
  }
};

struct sevp_plan_s {
  KProf prof;
  std::deque<PassGraph> graphs;  // deque: PassGraph addresses stay valid as graphs are added
  std::deque<RefineGraph> rgraphs;
  int expect_passes[4][2] = {};   // passes the last refined solve took, by precision and kernels
  DBuf<int> ctl;                  // captured refinement: passes, done, max_iter
  DBuf<double> ctl_d;             // tol, relative residuals (nrhs_max), worst
  int *ctl_h = nullptr;           // pinned: [0..2] uploaded (0, 0, max_iter), [4..5] read back,
                                  // tol as the double at [8]
  double *part_h = nullptr;
  DBuf<float> s_stage;            // the right-hand sides the captured refinement reads
  hipStream_t cap_stream = nullptr;
  bool use_graphs = true;
  sevp_desc d{};
  std::vector<int> ie;
  int ibir = 0;  // strip width (BIR) or i2 (full)
  int w = 0;     // influence-matrix order
  int nmat = 0;  // number of strips (BIR) or 1
  hipStream_t stream = nullptr;      // stream the plan's work is enqueued on
  hipStream_t own_stream = nullptr;  // the stream the plan created
  // refinement passes of independent RHS groups run concurrently: group 0 on
  // `stream`, group g > 0 on gstreams[g-1], joined through events
  int rhs_groups = 1;
  std::vector<hipStream_t> gstreams;
  std::vector<hipEvent_t> gevents;
  bool have_inv = false;
  // SEVP_STORE_FP64: RINV / RINV1 kept in FP64 (allocated on first build or
  // load); SEVP_STORE_PRODUCTS: only the FP32 product operands are kept
  int storage = SEVP_STORE_FP64;
  DBuf<float> coef;     // 5 * i2*j2 (AL, AB, AC, AR, AT)
  DBuf<int> ie_d;
  DBuf<int> brec;       // BIR: the strip record table (bir_record_table)
  int brec_fwd = 0;
  DBuf<double> rinv, rinv1;      // reference layout
  DBuf<float> rinv32, rinv1_32;  // mixed precision copies
  DBuf<float> w32;               // mixed: W_b = RINV_b diag(CY(:,IE(b-1)-1)) RINV1_(b-1), b >= 2
  DBuf<double> rinv64t, rinv1_64t, w64t;  // FP64 MFMA: the same three in FP64 lane tiles
  DBuf<unsigned short> rinv16t, rinv1_16t, w16t;  // BF16 lane tiles (fused pass only)
  // the copies above match the influence matrices; both pass kinds read them
  bool has32 = false, has64 = false, has16 = false;
  int pass_kind = SEVP_PASS_SPLITK;
  int fused_waves = FUSED_WAVES;  // wavefronts per fused workgroup (8 or 16)
  int fused_rows = 0;             // 16-row tiles per fused workgroup (1, 2; 0: by arithmetic)
  // fused pass: error rows (lane tiles, a region per RHS group, two parities;
  // finite padding, zeroed once) and c / cpl rows (nrhs, I2), two parities
  DBuf<double> fe, fh, fhc;
  DBuf<double> fd;           // fused pass: d of the blocks below one-row blocks, (NB0-1) slices
  DBuf<float> fe32;
  long fe_half = 0;          // elements per parity of fe / fe32
  // solve scratch
  DBuf<double> e, dum0, hfree, hcpl, g;  // full forms
  DBuf<float> e32;                       // mixed: FP32 copy of the error rows
  DBuf<float> kpart;                     // mixed: split-K partial products (S, nrhs, N)
  DBuf<double> kpart64;                  // FP64 MFMA: split-K partial products
  DBuf<double> xx, bdum0;                // BIR
  DBuf<double> r64, dx, part;            // refinement
  DBuf<double> b1s;                      // MATINV scratch (blocked path)
  DBuf<int> zp;
  DBuf<float> s14;                       // step 14: per-workgroup maxima + the maximum
  std::unique_ptr<struct KryWork> kry;   // Krylov solvers' vectors and scalars, on first use
  const float *al() const { return coef.p; }
  const float *ab() const { return coef.p + (size_t)d.i2 * d.j2; }
  const float *ac() const { return coef.p + 2 * (size_t)d.i2 * d.j2; }
  const float *ar() const { return coef.p + 3 * (size_t)d.i2 * d.j2; }
  const float *at() const { return coef.p + 4 * (size_t)d.i2 * d.j2; }
  size_t rinv_count() const { return (size_t)w * w * d.nb0 * nmat; }
  size_t rinv1_count() const { return (size_t)w * w * (d.nb0 > 1 ? d.nb0 - 1 : 1) * nmat; }
};

static int check_ie(const int *ie, int nb0, int j2) {
  if (!ie || nb0 < 1) return fail(SEVP_E_ARG, "IE missing or NB0 < 1");
  int prev = 1;
  for (int b = 0; b < nb0; ++b) {
    if (ie[b] <= prev) return fail(SEVP_E_ARG, "IE must be strictly increasing from 2");
    prev = ie[b];
  }
  if (ie[nb0 - 1] != j2 + 1) return fail(SEVP_E_ARG, "IE(NB0) must equal J2+1");
  return SEVP_OK;
}

// ------------------------------------------------------------------ PRE driver
// Influence matrices of `nmat` strips of width w (or one full-width periodic
// or Dirichlet problem), the block loop of PRE / PRE1.
struct PreCtx {
  const float *ax, *ay, *bb, *cx, *cy;
  int cld, i2, j2, w, nmat, periodic;
  ColMap map;
  double *rinv, *rinv1;  // reference layout, nmat strips
  const int *ie;         // host
  int nb0;
  int *zp;               // device counter
  double *b1s;           // MATINV scratch for the blocked path (null when LDS suffices)
};

static int run_pre(const PreCtx &c, hipStream_t st) {
  const long ww = (long)c.w * c.w;
  const long sstride = ww * c.nb0;
  const long s1stride = ww * (c.nb0 > 1 ? c.nb0 - 1 : 1);
  int jl = 1;
  for (int nb = 1; nb <= c.nb0; ++nb) {
    const int jh = c.ie[nb - 1];
    MarchArgs m{};
    m.ax = c.ax; m.ay = c.ay; m.bb = c.bb; m.cx = c.cx; m.cy = c.cy;
    m.cld = c.cld;
    m.i2 = c.i2; m.j2 = c.j2;
    m.periodic = c.periodic;
    m.js = jl; m.jf = jh - 2; m.ieb = jh;
    m.w = c.w;
    m.map = c.map;
    m.nrhs = c.nmat;  // strip count
    m.rinv_b = c.rinv + (long)(nb - 1) * ww;
    m.tmat = (nb != c.nb0) ? c.rinv + (long)(c.nb0 - 1) * ww : nullptr;
    m.rinv1_prev = (nb != 1) ? c.rinv1 + (long)(nb - 2) * ww : nullptr;
    m.rinv_sstride = sstride;
    m.rinv1_sstride = s1stride;
    HIPC(launch_march(MODE_PRE, m, c.w, st));
    HIPC(launch_matinv(c.rinv + (long)(nb - 1) * ww, sstride, c.w, c.nmat, c.zp, c.b1s, st));
    if (nb == c.nb0) break;
    for (int s = 0; s < c.nmat; ++s) {
      GemmArgs gm{};
      gm.a = c.rinv + s * sstride + (long)(nb - 1) * ww;
      gm.a_sm = 1; gm.a_sk = c.w;
      gm.b = c.rinv + s * sstride + (long)(c.nb0 - 1) * ww;
      gm.b_sk = 1; gm.b_sn = c.w;
      gm.M = c.w; gm.N = c.w; gm.K = c.w;
      gm.ep = EP_NEGSUB;
      gm.out = c.rinv1 + s * s1stride + (long)(nb - 1) * ww;
      gm.out_ld = c.w;
      HIPC(launch_gemm(gm, st));
    }
    jl = jh;
  }
  return SEVP_OK;
}

// ----------------------------------------------------------- full-width pass
// One REP1 (periodic) / REP pass over nrhs right-hand sides (device pointers).
struct FullCtx {
  const float *ax, *ay, *bb, *cx, *cy;  // ld i2
  int i2, j2, periodic, nb0;
  const int *ie;  // host
  const double *rinv, *rinv1;
  const float *rinv32, *rinv1_32;  // non-null selects the FP32 MFMA corrections
  const float *w32;                // mixed: coupling operators W_b (I2 x I2, b = 2..NB0)
  const double *rinv64t, *rinv1_64t, *w64t;  // FP64 MFMA: the same in FP64 lane tiles
  double *e, *dum0, *hfree, *hcpl, *g;
  float *e32;        // FP32 copy of e (mixed precision), else null
  float *kpart;      // mixed: split-K partial products
  double *kpart64;   // FP64 MFMA: split-K partial products (selects that arithmetic)
  // fused pass (one launch per block step): lane-tile operands (the FP32 or
  // FP64 ones above, or BF16) and the two-parity scratch rows: fe / fe32 in
  // lane tiles (fe_half elements per parity), the others fhalf per parity
  bool fused;
  int farith;  // FUSED_F32 / FUSED_F64 / FUSED_B16
  int fwaves;  // wavefronts per fused workgroup
  int frows;   // 16-row tiles per fused workgroup (0: by arithmetic)
  const unsigned short *rinv16t, *rinv1_16t, *w16t;
  double *fe, *fh, *fhc;
  float *fe32;
  long fhalf, fe_half;
  double *fd;  // d_b, slice b-1 of fhalf elements, for blocks b followed by a one-row block
  KProf *prof;
};

// One REP1 / REP pass with the corrections as split-K MFMA products.
// The forward march of block b reads the previous block's correction of row
// IE(b-1) straight from the partial products (the X array keeps the
// uncorrected value, which is what the backward pass restores anyway); the
// backward march of block b reads its free-row and coupling corrections the
// same way, and the launch that carries it adds the free-row correction to
// X(JS+1) and forms the closure residual of block b-1, the only reader of
// that row in the launch.
//
// mixed_pass_ops lists the launches' arguments, in order: for b = 1..NB0 the
// forward march of block b and, below the last block, its product with
// RINV1_b; the closure residual of the last block; for b = NB0..1 the
// product with [RINV_b | W_b] and the backward march of b with the residual
// of b-1; and, periodic, the halo fix.  run_mixed_pass launches them for one
// problem, the batched solve one grouped launch per step for many.
struct MixedPassOps {
  bool dbl = false;                     // FP64 MFMA products (else FP32)
  std::vector<MarchArgs> fmarch;        // [b-1], b = 1..NB0
  std::vector<GemmPartArgs> fprod;      // [b-1], b = 1..NB0-1 (FP32 products)
  std::vector<GemmPart64Args> fprod64;  // (FP64 products)
  ResidArgs mid{};
  std::vector<GemmPartArgs> bprod;      // [b-1], b = 1..NB0
  std::vector<GemmPart64Args> bprod64;
  std::vector<MarchArgs> bmarch;        // [b-1]
  std::vector<ResidArgs> bresid;        // [b-1]
};

static void mixed_pass_ops(const FullCtx &c, const float *f32, const double *f64, long f_stride,
                           double *x, long x_stride, int nrhs, MixedPassOps &o) {
  // FP64 MFMA products (kpart64) take the FP64 error rows e and FP64 lane
  // tiles; FP32 ones the FP32 rows e32 and FP32 tiles
  const bool dbl = c.kpart64 != nullptr;
  o.dbl = dbl;
  const long ww = lane_tiles_floats(c.i2, c.i2);  // one matrix in lane tiles (elements)
  const int np = (c.i2 + 15) / 16 * 16;           // W's columns start at np in [RINV | W]
  int ks = 0;
  const int ns1 = gemm_part_slices(c.i2, c.i2, &ks);
  const int ns2 = gemm_part_slices(2 * np, c.i2, &ks);
  auto part_of = [&](int n_prod) {
    PartRef r{};
    r.p = dbl ? nullptr : c.kpart;
    r.pd = dbl ? c.kpart64 : nullptr;
    r.ld = n_prod;
    r.sstride = (long)nrhs * n_prod;
    r.ns = (n_prod == 2 * np) ? ns2 : ns1;
    return r;
  };
  // block nb's product: forward (RINV1_nb, N = I2) or backward ([RINV_nb | W_nb])
  auto product = [&](bool fwd, int nb, int n_prod) {
    if (dbl) {
      GemmPart64Args g{};
      g.a = c.e;
      g.b = fwd ? c.rinv1_64t + (long)(nb - 1) * ww : c.rinv64t + (long)(nb - 1) * ww;
      g.b2 = (!fwd && nb != 1) ? c.w64t + (long)(nb - 2) * ww : nullptr;
      g.nsplit = np;
      g.M = nrhs; g.N = n_prod; g.K = c.i2;
      g.part = c.kpart64;
      (fwd ? o.fprod64 : o.bprod64).push_back(g);
    } else {
      GemmPartArgs g{};
      g.a = c.e32;
      g.b = fwd ? c.rinv1_32 + (long)(nb - 1) * ww : c.rinv32 + (long)(nb - 1) * ww;
      g.b2 = (!fwd && nb != 1) ? c.w32 + (long)(nb - 2) * ww : nullptr;
      g.nsplit = np;
      g.M = nrhs; g.N = n_prod; g.K = c.i2;
      g.part = c.kpart;
      (fwd ? o.fprod : o.bprod).push_back(g);
    }
  };
  int js = 1;
  for (int nb = 1; nb <= c.nb0; ++nb) {
    const int ieb = c.ie[nb - 1];
    MarchArgs m{};
    m.ax = c.ax; m.ay = c.ay; m.bb = c.bb; m.cx = c.cx; m.cy = c.cy;
    m.cld = c.i2;
    m.f = f32; m.f64 = f64; m.f_stride = f_stride;
    m.x = x; m.x_stride = x_stride;
    m.e = dbl ? c.e : nullptr;
    m.e32 = dbl ? nullptr : c.e32;
    m.i2 = c.i2; m.j2 = c.j2; m.periodic = c.periodic;
    m.js = js; m.jf = ieb - 2; m.ieb = ieb;
    m.resid = (nb != c.nb0);
    if (nb != 1) m.corr = part_of(c.i2);
    m.cpl_off = -1;
    m.w = c.i2;
    m.map = ColMap{1, 0, 0};
    m.nrhs = nrhs;
    o.fmarch.push_back(m);
    if (nb != c.nb0) product(true, nb, c.i2);
    js = ieb;
  }
  auto resid_args = [&](int nb) {
    ResidArgs r{};
    r.ax = c.ax; r.ay = c.ay; r.bb = c.bb; r.cx = c.cx; r.cy = c.cy;
    r.f = f32; r.f64 = f64; r.f_stride = f_stride;
    r.x = x; r.x_stride = x_stride;
    r.e = dbl ? c.e : nullptr;
    r.e32 = dbl ? nullptr : c.e32;
    r.i2 = c.i2; r.j2 = c.j2; r.periodic = c.periodic; r.nb = nb; r.ieb = c.ie[nb - 1];
    r.nrhs = nrhs;
    return r;
  };
  o.mid = resid_args(c.nb0);
  o.bmarch.resize(c.nb0);
  o.bresid.resize(c.nb0);
  if (dbl) o.bprod64.reserve(c.nb0);
  else o.bprod.reserve(c.nb0);
  // backward products are listed by block (index b-1), built from b = 1 up
  for (int nb = 1; nb <= c.nb0; ++nb) {
    const int ieb = c.ie[nb - 1];
    js = (nb != 1) ? c.ie[nb - 2] : 1;
    // [RINV_b | W_b]: free-row correction and, for b > 1, the coupling row
    const int n_prod = (nb != 1) ? 2 * np : c.i2;
    product(false, nb, n_prod);
    MarchArgs m{};
    m.ax = c.ax; m.ay = c.ay; m.bb = c.bb; m.cx = c.cx; m.cy = c.cy;
    m.cld = c.i2;
    m.x = x; m.x_stride = x_stride;
    m.corr = part_of(n_prod);
    m.cpl_off = (nb != 1) ? np : -1;
    m.i2 = c.i2; m.j2 = c.j2; m.periodic = c.periodic;
    m.js = js; m.jf = ieb - 2; m.ieb = ieb;
    m.w = c.i2;
    m.map = ColMap{1, 0, 0};
    m.nrhs = nrhs;
    ResidArgs rb = (nb != 1) ? resid_args(nb - 1) : resid_args(1);
    rb.corr = part_of(n_prod);
    rb.apply_row = js + 1;
    rb.do_resid = (nb != 1);
    o.bmarch[nb - 1] = m;
    o.bresid[nb - 1] = rb;
  }
}

static int run_mixed_pass(const FullCtx &c, const float *f32, const double *f64, long f_stride,
                          double *x, long x_stride, int nrhs, hipStream_t st) {
  MixedPassOps o;
  mixed_pass_ops(c, f32, f64, f_stride, x, x_stride, nrhs, o);
  auto product = [&](bool fwd, int nb) -> int {
    const auto ev = c.prof ? c.prof->start_ext(SEVP_KCLASS_CORR)
                           : std::pair<hipEvent_t, hipEvent_t>{nullptr, nullptr};
    if (o.dbl) {
      GemmPart64Args g = fwd ? o.fprod64[nb - 1] : o.bprod64[nb - 1];
      g.tspan = c.prof ? c.prof->next_span(SEVP_KCLASS_CORR, gemm_part64_workgroups(g)) : nullptr;
      HIPC(launch_gemm_part64(g, st, ev.first, ev.second));
    } else {
      GemmPartArgs g = fwd ? o.fprod[nb - 1] : o.bprod[nb - 1];
      g.tspan = c.prof ? c.prof->next_span(SEVP_KCLASS_CORR, gemm_part_workgroups(g)) : nullptr;
      HIPC(launch_gemm_part(g, st, ev.first, ev.second));
    }
    if (c.prof) c.prof->stop_ext(SEVP_KCLASS_CORR, ev.first != nullptr);
    return SEVP_OK;
  };
  for (int nb = 1; nb <= c.nb0; ++nb) {
    if (c.prof) c.prof->start(SEVP_KCLASS_MARCH);
    HIPC(launch_march(MODE_PART, o.fmarch[nb - 1], nrhs, st));
    if (c.prof) c.prof->stop(SEVP_KCLASS_MARCH);
    if (nb != c.nb0) CHK(product(true, nb));
  }
  if (c.prof) c.prof->start(SEVP_KCLASS_RESID);
  HIPC(launch_resid(o.mid, st));
  if (c.prof) c.prof->stop(SEVP_KCLASS_RESID);
  for (int nb = c.nb0; nb >= 1; --nb) {
    CHK(product(false, nb));
    if (c.prof) c.prof->start(SEVP_KCLASS_MARCH);
    HIPC(launch_hom_resid(o.bmarch[nb - 1], &o.bresid[nb - 1], st));
    if (c.prof) c.prof->stop(SEVP_KCLASS_MARCH);
  }
  if (c.periodic) HIPC(launch_halo_fix(x, x_stride, c.i2, c.j2, nrhs, st));
  return SEVP_OK;
}

// One REP1 / REP pass with one launch per block step (sevp_fused.hip): the
// march of block 1 with its closure residual; for b = 1..NB0-1 the product
// e_b RINV1_b with block b+1's march and closure residual; for b = NB0..1 the
// product e_b [RINV_b | W_b] with the free-row update of block b, the closure
// residual of block b-1 and the homogeneous march of block b+1; the
// homogeneous march of block 1; and, periodic, the halo fix.  The error rows
// and the c / cpl rows alternate between two buffers by block parity, so no
// launch reads what it writes.
static int run_fused_pass(const FullCtx &c, const float *f32, const double *f64, long f_stride,
                          double *x, long x_stride, int nrhs, hipStream_t st) {
  const bool dbl = c.farith == FUSED_F64;
  const long tw = lane_tile_elems(c.i2, c.i2);
  const int nchk = (c.i2 + 15) / 16;
  // lane-tile operands of block b (RINV1, RINV, W) in the product's arithmetic
  auto opnd = [&](const double *f64t, const float *f32t, const unsigned short *b16t,
                  long off) -> const void * {
    if (c.farith == FUSED_F64) return f64t + off;
    if (c.farith == FUSED_B16) return b16t + off;
    return f32t + off;
  };
  double *E[2] = {c.fe, c.fe + c.fe_half};
  float *E32[2] = {c.fe32, c.fe32 + c.fe_half};
  double *HF[2] = {c.fh, c.fh + c.fhalf}, *HC[2] = {c.fhc, c.fhc + c.fhalf};
  auto erow = [&](int b) -> const void * {
    return dbl ? (const void *)E[b & 1] : (const void *)E32[b & 1];
  };
  auto march = [&](int mode, int b) -> int {
    MarchArgs m{};
    m.ax = c.ax; m.ay = c.ay; m.bb = c.bb; m.cx = c.cx; m.cy = c.cy;
    m.cld = c.i2;
    m.i2 = c.i2; m.j2 = c.j2; m.periodic = c.periodic;
    m.x = x; m.x_stride = x_stride;
    m.js = (b != 1) ? c.ie[b - 2] : 1; m.jf = c.ie[b - 1] - 2; m.ieb = c.ie[b - 1];
    m.w = c.i2;
    m.map = ColMap{1, 0, 0};
    m.nrhs = nrhs;
    m.cpl_off = -1;
    if (mode == MODE_PART) {
      m.f = f32; m.f64 = f64; m.f_stride = f_stride;
      m.resid = 1;
      m.e = dbl ? E[b & 1] : nullptr;
      m.e32 = dbl ? nullptr : E32[b & 1];
      m.e_lt = nchk;
    } else {
      m.hfree = HF[b & 1];
      m.hcpl = (b != 1) ? HC[b & 1] : nullptr;
      if (m.jf < m.js) return SEVP_OK;  // a one-row block has no homogeneous march
    }
    if (c.prof) c.prof->start(SEVP_KCLASS_MARCH);
    HIPC(launch_march(mode, m, nrhs, st));
    if (c.prof) c.prof->stop(SEVP_KCLASS_MARCH);
    return SEVP_OK;
  };
  auto timed = [&](auto launch, long wgs, unsigned long long *&tspan) -> int {
    tspan = c.prof ? c.prof->next_span(SEVP_KCLASS_CORR, wgs) : nullptr;
    const auto ev = c.prof ? c.prof->start_ext(SEVP_KCLASS_CORR)
                           : std::pair<hipEvent_t, hipEvent_t>{nullptr, nullptr};
    HIPC(launch(ev.first, ev.second));
    if (c.prof) c.prof->stop_ext(SEVP_KCLASS_CORR, ev.first != nullptr);
    return SEVP_OK;
  };
  CHK(march(MODE_PART, 1));
  for (int b = 1; b < c.nb0; ++b) {
    FusedFwdArgs a{};
    a.a = erow(b);
    a.b = opnd(c.rinv1_64t, c.rinv1_32, c.rinv1_16t, (b - 1) * tw);
    a.M = nrhs; a.i2 = c.i2; a.periodic = c.periodic;
    a.js = c.ie[b - 1]; a.jf = c.ie[b] - 2; a.ieb = c.ie[b];
    a.steps = a.jf - a.js + 1;
    a.ax = c.ax; a.ay = c.ay; a.bb = c.bb; a.cx = c.cx; a.cy = c.cy;
    a.f = f32; a.f64 = f64; a.f_stride = f_stride;
    a.x = x; a.x_stride = x_stride;
    a.e_out = dbl ? E[(b + 1) & 1] : nullptr;
    a.e32_out = dbl ? nullptr : E32[(b + 1) & 1];
    if (a.steps == 0) a.d_out = c.fd + (long)(b - 1) * c.fhalf;  // block b+1 is one row high
    a.waves = c.fwaves;
    a.rows = c.frows;
    FusedFwdArgs sz = a;
    if (fused_fwd_prepare(sz, c.farith)) return fail(SEVP_E_ARG, "block too tall for the fused pass");
    CHK(timed([&](hipEvent_t e0, hipEvent_t e1) { return launch_fused_fwd(a, c.farith, st, e0, e1); },
              fused_fwd_workgroups(sz), a.tspan));
  }
  for (int b = c.nb0; b >= 1; --b) {
    FusedBwdArgs a{};
    a.a = erow(b);
    a.b = opnd(c.rinv64t, c.rinv32, c.rinv16t, (b - 1) * tw);
    a.bw = (b == 1) ? nullptr : opnd(c.w64t, c.w32, c.w16t, (b - 2) * tw);
    a.M = nrhs; a.i2 = c.i2; a.periodic = c.periodic;
    a.js = (b != 1) ? c.ie[b - 2] : 1;
    a.do_resid = b != 1;
    a.ax = c.ax; a.ay = c.ay; a.bb = c.bb; a.cx = c.cx; a.cy = c.cy;
    a.f = f32; a.f64 = f64; a.f_stride = f_stride;
    a.x = x; a.x_stride = x_stride;
    a.e_out = dbl ? E[(b - 1) & 1] : nullptr;
    a.e32_out = dbl ? nullptr : E32[(b - 1) & 1];
    if (b >= 3 && c.ie[b - 2] - c.ie[b - 3] == 1)  // block b-1 is one row high
      a.dsouth = c.fd + (long)(b - 3) * c.fhalf;
    a.hfree_out = HF[b & 1];
    a.hcpl_out = HC[b & 1];
    a.waves = c.fwaves;
    a.rows = c.frows;
    if (b != c.nb0) {  // the homogeneous march of block b+1, from the previous launch's rows
      MarchArgs m{};
      m.ax = c.ax; m.ay = c.ay; m.bb = c.bb; m.cx = c.cx; m.cy = c.cy;
      m.cld = c.i2;
      m.i2 = c.i2; m.j2 = c.j2; m.periodic = c.periodic;
      m.x = x; m.x_stride = x_stride;
      m.js = c.ie[b - 1]; m.jf = c.ie[b] - 2; m.ieb = c.ie[b];
      m.hfree = HF[(b + 1) & 1];
      m.hcpl = HC[(b + 1) & 1];
      m.w = c.i2;
      m.map = ColMap{1, 0, 0};
      m.nrhs = nrhs;
      m.cpl_off = -1;
      a.hom = m;
      a.has_hom = 1;
    }
    FusedBwdArgs sz = a;
    if (fused_bwd_prepare(sz, c.farith)) return fail(SEVP_E_ARG, "shape refused by the fused pass");
    CHK(timed([&](hipEvent_t e0, hipEvent_t e1) { return launch_fused_bwd(a, c.farith, st, e0, e1); },
              fused_bwd_workgroups(sz), a.tspan));
  }
  CHK(march(MODE_HOM, 1));
  if (c.periodic) HIPC(launch_halo_fix(x, x_stride, c.i2, c.j2, nrhs, st));
  return SEVP_OK;
}

static int run_full_pass(const FullCtx &c, const float *f32, const double *f64, long f_stride,
                         double *x, long x_stride, int nrhs, hipStream_t st) {
  if (c.fused) return run_fused_pass(c, f32, f64, f_stride, x, x_stride, nrhs, st);
  if (c.kpart || c.kpart64) return run_mixed_pass(c, f32, f64, f_stride, x, x_stride, nrhs, st);
  const long ww = (long)c.i2 * c.i2;
  const int i0 = c.i2 + 2, j0 = c.j2 + 2;
  (void)j0;
  const long dstride = (long)c.i2 * (c.nb0 > 1 ? c.nb0 - 1 : 1);
  // the plain FP64 pass (bitwise the reference REP1 / REP); the mixed pass is
  // run_mixed_pass
  auto corr = [&](const double *a, const double *b64, int ep, int xrow, double *dum0, double *out,
                  double *g, int cyrow) -> int {
    GemmArgs gm{};
    gm.a = a; gm.a_sm = c.i2; gm.a_sk = 1;
    gm.b = b64; gm.b_sk = 1; gm.b_sn = c.i2;
    gm.M = nrhs; gm.N = c.i2; gm.K = c.i2;
    gm.ep = ep;
    gm.x = x; gm.x_stride = x_stride; gm.xrow = xrow;
    gm.dum0 = dum0; gm.dum0_stride = dstride;
    gm.out = out; gm.out_ld = c.i2;
    gm.g = g; gm.cy = c.cy; gm.cyrow = cyrow;
    gm.i2 = c.i2;
    gm.tspan = c.prof ? c.prof->next_span(SEVP_KCLASS_CORR, gemm_workgroups(gm)) : nullptr;
    const auto ev = c.prof ? c.prof->start_ext(SEVP_KCLASS_CORR)
                           : std::pair<hipEvent_t, hipEvent_t>{nullptr, nullptr};
    HIPC(launch_gemm(gm, st, ev.first, ev.second));
    if (c.prof) c.prof->stop_ext(SEVP_KCLASS_CORR, ev.first != nullptr);
    return SEVP_OK;
  };
  (void)i0;
  int js = 1;
  for (int nb = 1; nb <= c.nb0; ++nb) {
    const int ieb = c.ie[nb - 1];
    MarchArgs m{};
    m.ax = c.ax; m.ay = c.ay; m.bb = c.bb; m.cx = c.cx; m.cy = c.cy;
    m.cld = c.i2;
    m.f = f32; m.f64 = f64; m.f_stride = f_stride;
    m.x = x; m.x_stride = x_stride;
    m.e = c.e;
    m.e32 = nullptr;
    m.i2 = c.i2; m.j2 = c.j2; m.periodic = c.periodic;
    m.js = js; m.jf = ieb - 2; m.ieb = ieb;
    m.resid = (nb != c.nb0);
    m.w = c.i2;
    m.map = ColMap{1, 0, 0};
    m.nrhs = nrhs;
    if (c.prof) c.prof->start(SEVP_KCLASS_MARCH);
    HIPC(launch_march(MODE_PART, m, nrhs, st));
    if (c.prof) c.prof->stop(SEVP_KCLASS_MARCH);
    if (nb != c.nb0)
      CHK(corr(c.e, c.rinv1 + (long)(nb - 1) * ww, EP_FWD, ieb, c.dum0 + (long)(nb - 1) * c.i2,
               nullptr, nullptr, 0));
    js = ieb;
  }
  // backward: restore + closure residual of the top block, then per block the
  // corrections and one launch for its H march fused with the next block
  // down's restore + residual
  auto resid_args = [&](int nb) {
    ResidArgs r{};
    r.ax = c.ax; r.ay = c.ay; r.bb = c.bb; r.cx = c.cx; r.cy = c.cy;
    r.f = f32; r.f64 = f64; r.f_stride = f_stride;
    r.x = x; r.x_stride = x_stride;
    r.dum0 = (nb != c.nb0) ? c.dum0 : nullptr;
    r.dum0_stride = dstride;
    r.e = c.e;
    r.e32 = nullptr;
    r.i2 = c.i2; r.j2 = c.j2; r.periodic = c.periodic; r.nb = nb; r.ieb = c.ie[nb - 1];
    r.nrhs = nrhs;
    return r;
  };
  {
    const ResidArgs r = resid_args(c.nb0);
    if (c.prof) c.prof->start(SEVP_KCLASS_RESID);
    HIPC(launch_resid(r, st));
    if (c.prof) c.prof->stop(SEVP_KCLASS_RESID);
  }
  for (int nbs = 1; nbs <= c.nb0; ++nbs) {
    const int nb = c.nb0 - nbs + 1;
    const int ieb = c.ie[nb - 1];
    js = (nb != 1) ? c.ie[nb - 2] : 1;
    CHK(corr(c.e, c.rinv + (long)(nb - 1) * ww, EP_BWD, js + 1, nullptr, c.hfree,
             nb != 1 ? c.g : nullptr, js - 1));
    if (nb != 1)
      CHK(corr(c.g, c.rinv1 + (long)(nb - 2) * ww, EP_CPL, 0, nullptr, c.hcpl, nullptr, 0));
    MarchArgs m{};
    m.ax = c.ax; m.ay = c.ay; m.bb = c.bb; m.cx = c.cx; m.cy = c.cy;
    m.cld = c.i2;
    m.x = x; m.x_stride = x_stride;
    m.hcpl = (nb != 1) ? c.hcpl : nullptr;
    m.hfree = c.hfree;
    m.i2 = c.i2; m.j2 = c.j2; m.periodic = c.periodic;
    m.js = js; m.jf = ieb - 2; m.ieb = ieb;
    m.w = c.i2;
    m.map = ColMap{1, 0, 0};
    m.nrhs = nrhs;
    ResidArgs rb{};
    if (nb != 1) rb = resid_args(nb - 1);
    if (c.prof) c.prof->start(SEVP_KCLASS_MARCH);
    HIPC(launch_hom_resid(m, nb != 1 ? &rb : nullptr, st));
    if (c.prof) c.prof->stop(SEVP_KCLASS_MARCH);
  }
  if (c.periodic) HIPC(launch_halo_fix(x, x_stride, c.i2, c.j2, nrhs, st));
  return SEVP_OK;
}

// ------------------------------------------------------------- residual
// part[(r*nchunk + chunk)*2 + {0,1}] = partial sums of (s - A x)^2 and s^2.
namespace sevp {
template <int U>
__device__ __forceinline__ void resid_norm_body(const float *al, const float *ab, const float *ac,
                                                const float *ar, const float *at, const float *s,
                                                const double *s64, long s_stride, const double *x,
                                                long x_stride, double *rout, int i2, int j2,
                                                int nchunk, double *part, int r, int ch) {
  __shared__ double red[2][256];
  const int t = threadIdx.x;
  const long n = (long)i2 * j2;
  // chunk ch takes whole rows [j_lo, j_hi) (1-based j = row + 1); each
  // thread strides over the columns of each row, so no index division
  const int j_lo = (int)((long)j2 * ch / nchunk), j_hi = (int)((long)j2 * (ch + 1) / nchunk);
  const double *xr = x + (long)r * x_stride;
  const int i0 = i2 + 2;
  const float *sr = s ? s + (long)r * s_stride : nullptr;
  const double *s64r = s64 ? s64 + (long)r * s_stride : nullptr;
  double *ro = rout ? rout + (long)r * n : nullptr;
  double a0 = 0.0, a1 = 0.0;
  // the operands of U columns (t, t + 256, ...) are loaded before any of
  // them is used, so their loads are in flight together; the columns are
  // then summed in the same order as one at a time
  auto rows = [&](auto wide) {
    constexpr bool W64 = decltype(wide)::value;
    for (int jj = j_lo; jj < j_hi; ++jj) {
      const int j = jj + 1;
      const double *xs = xr + f2(i0, 1, j), *xc = xr + f2(i0, 1, j + 1), *xn = xr + f2(i0, 1, j + 2);
      for (int i00 = t; i00 < i2; i00 += 256 * U) {
        float cl[U], cb[U], cc[U], cr[U], ct[U];
        double x0[U], x1[U], x2[U], x3[U], x4[U], sv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int ii = min(i00 + 256 * u, i2 - 1);  // a column past I2 repeats the last
          const long q = (long)jj * i2 + ii;
          cb[u] = ab[q]; cl[u] = al[q]; cc[u] = ac[q]; cr[u] = ar[q]; ct[u] = at[q];
          x0[u] = xs[ii + 1]; x1[u] = xc[ii]; x2[u] = xc[ii + 1]; x3[u] = xc[ii + 2];
          x4[u] = xn[ii + 1];
          if constexpr (W64) sv[u] = s64r[q];
          else sv[u] = (double)sr[q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int ii = i00 + 256 * u;
          if (ii < i2) {
            const double ax = (double)cb[u] * x0[u] + (double)cl[u] * x1[u] +
                              (double)cc[u] * x2[u] + (double)cr[u] * x3[u] +
                              (double)ct[u] * x4[u];
            const double rv = sv[u] - ax;
            if (ro) ro[(long)jj * i2 + ii] = rv;
            a0 += rv * rv;
            a1 += sv[u] * sv[u];
          }
        }
      }
    }
  };
  if (s64r) rows(std::true_type{});
  else rows(std::false_type{});
  red[0][t] = a0;
  red[1][t] = a1;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      red[0][t] += red[0][t + o];
      red[1][t] += red[1][t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[((long)r * nchunk + ch) * 2] = red[0][0];
    part[((long)r * nchunk + ch) * 2 + 1] = red[1][0];
  }
}

// U: columns per thread loaded together (chosen by launch_resid_norm_cols).
template <int U>
__global__ __launch_bounds__(256) void resid_norm_kernel(const float *al, const float *ab,
                                                         const float *ac, const float *ar,
                                                         const float *at, const float *s,
                                                         const double *s64, long s_stride,
                                                         const double *x, long x_stride,
                                                         double *rout, int i2, int j2, int nchunk,
                                                         double *part, const int *gate) {
  if (gate && gate[1]) return;  // captured refinement: converged, the residual stays
  resid_norm_body<U>(al, ab, ac, ar, at, s, s64, s_stride, x, x_stride, rout, i2, j2, nchunk, part,
                  blockIdx.y, blockIdx.x);
}

// The residual norms of many problems in one launch: entry e runs the
// nchunk x nrhs workgroups of its own resid_norm_kernel launch.
struct NormEnt {
  const float *al, *ab, *ac, *ar, *at, *s;
  const double *x;
  long s_stride, x_stride;
  double *rout, *part;
  int i2, j2, nrhs, nchunk;
};

__global__ __launch_bounds__(256) void batch_resid_norm_kernel(const NormEnt *__restrict__ tab,
                                                               const int *__restrict__ start, int n) {
  const int b = blockIdx.x;
  const int e = group_entry(start, n, b);
  const NormEnt &q = tab[e];
  const int l = b - start[e];
  resid_norm_body<1>(q.al, q.ab, q.ac, q.ar, q.at, q.s, nullptr, q.s_stride, q.x, q.x_stride, q.rout,
                  q.i2, q.j2, q.nchunk, q.part, l / q.nchunk, l % q.nchunk);
}

__global__ void zero_kernel(unsigned int *p, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0u;
}

__global__ void axpy_interior_kernel(double *x, long x_stride, const double *d, long d_stride,
                                     int i2, int j2, int nrhs) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)i2 * j2;
  if (t >= n * nrhs) return;
  const int r = (int)(t / n);
  const long q = t % n;
  const int i = (int)(q % i2) + 2, j = (int)(q / i2) + 2;
  const long o = f2(i2 + 2, i, j);
  x[(long)r * x_stride + o] += d[(long)r * d_stride + o];
}

// Captured refinement (one graph for all passes).  ctl[0] = passes run,
// ctl[1] = done, ctl[2] = max_iter; tol in *tolp.  The update x += dx with
// the periodic halo columns refreshed from the new interior (what
// axpy_interior_kernel and halo_fix_kernel do), skipped once done.
__global__ void refine_axpy_kernel(double *x, long x_stride, const double *d, long d_stride,
                                   int i2, int j2, int nrhs, int periodic, const int *ctl) {
  if (ctl[1]) return;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)i2 * j2;
  if (t >= n * nrhs) return;
  const int r = (int)(t / n);
  const long q = t % n;
  const int i = (int)(q % i2) + 2, j = (int)(q / i2) + 2;
  const int i0 = i2 + 2;
  const long o = f2(i0, i, j);
  double *xr = x + (long)r * x_stride;
  const double v = xr[o] + d[(long)r * d_stride + o];
  xr[o] = v;
  if (periodic) {
    if (i == i0 - 1) xr[f2(i0, 1, j)] = v;
    if (i == 2) xr[f2(i0, i0, j)] = v;
  }
}

// Relative residuals from the residual kernel's partial sums, summed in the
// order the host sums them, and the stop test of the refinement loop:
// done = !(worst > tol) || passes >= max_iter, worst the running maximum in
// RHS order with the host's std::max semantics.  first: the residual of the
// initial guess (passes = 0); otherwise a pass was just added (unless done).
__global__ void refine_finish_kernel(const double *part, int nrhs, int nchunk, double *rel, int *ctl,
                                     const double *tolp, int first) {
  if (!first && ctl[1]) return;
  for (int r = threadIdx.x; r < nrhs; r += blockDim.x) {
    double a = 0.0, b = 0.0;
    for (int c = 0; c < nchunk; ++c) {
      a += part[((long)r * nchunk + c) * 2];
      b += part[((long)r * nchunk + c) * 2 + 1];
    }
    rel[r] = b > 0 ? sqrt(a / b) : sqrt(a);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double worst = 0.0;
    for (int k = 0; k < nrhs; ++k) worst = (worst < rel[k]) ? rel[k] : worst;
    const int it = first ? 0 : ctl[0] + 1;
    ctl[0] = it;
    ctl[1] = (!(worst > *tolp) || it >= ctl[2]) ? 1 : 0;
    rel[nrhs] = worst;
  }
}
}  // namespace sevp

static constexpr int NCHUNK = 64;

// Clear `bytes` (a multiple of 4) of device memory on a stream with a kernel.
static int zero_async(void *ptr, size_t bytes, hipStream_t st) {
  const long n = (long)(bytes / 4);
  if (n == 0) return SEVP_OK;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     (unsigned int *)ptr, n);
  HIPC(hipGetLastError());
  return SEVP_OK;
}

// The residual-norm launch with U = the columns per thread of a row (at most
// 3) loaded together: a kernel's registers are those of its widest path, so
// narrow grids keep the one-column kernel.
static void launch_resid_norm_cols(sevp_plan p, const float *s, const double *s64, long s_stride,
                                   const double *x, long x_stride, double *rout, int nrhs,
                                   const int *gate, hipStream_t st) {
  const int i2 = p->d.i2, j2 = p->d.j2;
  const dim3 grid(NCHUNK, nrhs);
  if (i2 > 512)
    hipLaunchKernelGGL(resid_norm_kernel<3>, grid, dim3(256), 0, st, p->al(), p->ab(), p->ac(),
                       p->ar(), p->at(), s, s64, s_stride, x, x_stride, rout, i2, j2, NCHUNK,
                       p->part.p, gate);
  else if (i2 > 256)
    hipLaunchKernelGGL(resid_norm_kernel<2>, grid, dim3(256), 0, st, p->al(), p->ab(), p->ac(),
                       p->ar(), p->at(), s, s64, s_stride, x, x_stride, rout, i2, j2, NCHUNK,
                       p->part.p, gate);
  else
    hipLaunchKernelGGL(resid_norm_kernel<1>, grid, dim3(256), 0, st, p->al(), p->ab(), p->ac(),
                       p->ar(), p->at(), s, s64, s_stride, x, x_stride, rout, i2, j2, NCHUNK,
                       p->part.p, gate);
}

static int residual_norms(sevp_plan p, const float *s, const double *s64, long s_stride,
                          const double *x, long x_stride, double *rout, int nrhs,
                          std::vector<double> &rel) {
  p->prof.start(SEVP_KCLASS_RESID);
  launch_resid_norm_cols(p, s, s64, s_stride, x, x_stride, rout, nrhs, nullptr, p->stream);
  HIPC(hipGetLastError());
  p->prof.stop(SEVP_KCLASS_RESID);
  std::vector<double> h((size_t)nrhs * NCHUNK * 2);
  HIPC(hipMemcpyAsync(h.data(), p->part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                      p->stream));
  HIPC(hipStreamSynchronize(p->stream));
  p->prof.harvest();
  rel.assign(nrhs, 0.0);
  for (int r = 0; r < nrhs; ++r) {
    double a = 0, b = 0;
    for (int c = 0; c < NCHUNK; ++c) {
      a += h[((size_t)r * NCHUNK + c) * 2];
      b += h[((size_t)r * NCHUNK + c) * 2 + 1];
    }
    rel[r] = b > 0 ? std::sqrt(a / b) : std::sqrt(a);
  }
  return SEVP_OK;
}

static int alloc_mixed(sevp_plan p);
static int make_low_precision_copies(sevp_plan p);
static bool fused_eligible(sevp_plan p);
static int alloc_operands(sevp_plan p);

// A block after the first that is one row high: its backward closure
// residual reads the row below it as REP1's forward correction left it.
static bool has_unit_block(sevp_plan p) {
  for (int b = 1; b < p->d.nb0; ++b)
    if (p->ie[b] - p->ie[b - 1] == 1) return true;
  return false;
}

// Whether pass kernels `kind` can run corrections of `precision` on this
// plan's block rows: the fused pass takes blocks of 1 to 4 rows; the split-K
// pass keeps no corrected rows for the backward pass, so no one-row block may
// follow the first; BF16 runs in the fused pass only.
static int check_pass(sevp_plan p, int kind, int precision) {
  if (precision == SEVP_PREC_FP64 || p->d.form == SEVP_FORM_BIR) return SEVP_OK;
  if (kind == SEVP_PASS_FUSED && !fused_eligible(p))
    return fail(SEVP_E_ARG, "the fused pass needs block heights <= 4 and I2 divisible by 4 (>= 16)");
  if (kind == SEVP_PASS_SPLITK && precision == SEVP_PREC_BF16)
    return fail(SEVP_E_ARG, "BF16 corrections run in the fused pass only");
  if (kind == SEVP_PASS_SPLITK && has_unit_block(p))
    return fail(SEVP_E_ARG, "the split-K pass needs blocks of at least 2 rows after the first");
  return SEVP_OK;
}

// ------------------------------------------------------------------ C ABI
extern "C" {

const char *sevp_last_error(void) { return g_err.c_str(); }
const char *sevp_version(void) { return "sevp-mi355x 0.1 (gfx950)"; }

int sevp_plan_create(const sevp_desc *desc, const float *al, const float *ab, const float *ac,
                     const float *ar, const float *at, sevp_plan *out) {
  if (!desc || !out || !al || !ab || !ac || !ar || !at) return fail(SEVP_E_ARG, "null argument");
  *out = nullptr;
  const sevp_desc &d = *desc;
  if (d.i2 < 2 || d.j2 < 2) return fail(SEVP_E_ARG, "grid must be at least 2x2");
  if (d.nrhs_max < 1) return fail(SEVP_E_ARG, "nrhs_max must be >= 1");
  if (d.form != SEVP_FORM_BIR && d.form != SEVP_FORM_FULL_PERIODIC && d.form != SEVP_FORM_FULL)
    return fail(SEVP_E_ARG, "unknown form");
  if (d.precision < SEVP_PREC_FP64 || d.precision > SEVP_PREC_FP64_MFMA)
    return fail(SEVP_E_ARG, "unknown precision");
  CHK(check_ie(d.ie, d.nb0, d.j2));
  auto *p = new sevp_plan_s();
  p->d = d;
  p->ie.assign(d.ie, d.ie + d.nb0);
  p->d.ie = p->ie.data();
  if (d.form == SEVP_FORM_BIR) {
    if (d.nbir < 2 || (2 * d.i2) % d.nbir != 0) {
      delete p;
      return fail(SEVP_E_ARG, "BIR form needs NBIR >= 2 dividing 2*I2");
    }
    p->ibir = 2 * d.i2 / d.nbir;
    if (p->ibir % 2 != 0 || p->ibir > 64) {
      delete p;
      return fail(SEVP_E_ARG, "BIR strip width IBIR must be even and <= 64");
    }
    if (d.nb0 > BIR_MAX_BLOCKS) {
      delete p;
      return fail(SEVP_E_ARG, "BIR form supports at most 4096 blocks (NB0)");
    }
    if (d.precision != SEVP_PREC_FP64) {
      delete p;
      return fail(SEVP_E_ARG, "BIR form runs the plain FP64 arithmetic only");
    }
    if (d.nsweep < 1) {
      delete p;
      return fail(SEVP_E_ARG, "nsweep must be >= 1");
    }
    p->w = p->ibir;
    p->nmat = d.nbir;
  } else {
    p->ibir = d.i2;
    p->w = d.i2;
    p->nmat = 1;
  }
  p->pass_kind = fused_eligible(p) ? SEVP_PASS_FUSED : SEVP_PASS_SPLITK;
  int rc = SEVP_OK;
  auto bail = [&](int code) {
    sevp_plan_destroy(p);
    return code;
  };
  // a blocking stream: work on it is ordered with the null stream, which is
  // where callers such as PyTorch's default stream enqueue by default
  if (hipStreamCreateWithFlags(&p->own_stream, hipStreamDefault) != hipSuccess)
    return bail(fail(SEVP_E_HIP, "hipStreamCreate failed"));
  p->stream = p->own_stream;
  if (hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SEVP_E_HIP, "hipStreamCreate failed"));
  const size_t nc = (size_t)d.i2 * d.j2;
  if ((rc = p->coef.alloc(5 * nc))) return bail(rc);
  const float *src[5] = {al, ab, ac, ar, at};
  for (int k = 0; k < 5; ++k)
    if (hipMemcpy(p->coef.p + k * nc, src[k], nc * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(SEVP_E_HIP, "coefficient upload failed"));
  if ((rc = p->ie_d.alloc(d.nb0))) return bail(rc);
  if (hipMemcpy(p->ie_d.p, p->ie.data(), d.nb0 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(SEVP_E_HIP, "IE upload failed"));
  if (d.form == SEVP_FORM_BIR) {
    std::vector<int> rt;
    bir_record_table(p->ie.data(), d.nb0, rt, p->brec_fwd);
    if ((rc = p->brec.alloc(rt.size()))) return bail(rc);
    if (hipMemcpy(p->brec.p, rt.data(), rt.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(SEVP_E_HIP, "record table upload failed"));
  }
  if ((rc = p->zp.alloc(1))) return bail(rc);
  if (!matinv_uses_lds(p->w))
    if ((rc = p->b1s.alloc(matinv_blocked_scratch(p->w, p->nmat)))) return bail(rc);
  const int nr = d.nrhs_max;
  const int nb1 = d.nb0 > 1 ? d.nb0 - 1 : 1;
  const size_t nx = (size_t)(d.i2 + 2) * (d.j2 + 2);
  if ((rc = p->part.alloc((size_t)nr * NCHUNK * 2))) return bail(rc);
  if (d.form == SEVP_FORM_BIR) {
    if ((rc = p->xx.alloc((size_t)nr * (p->ibir + 2) * (d.j2 + 2)))) return bail(rc);
    if ((rc = p->bdum0.alloc((size_t)nr * p->ibir * nb1))) return bail(rc);
    if (hipMemset(p->xx.p, 0, p->xx.bytes()) != hipSuccess) return bail(fail(SEVP_E_HIP, "memset"));
    if (hipMemset(p->bdum0.p, 0, p->bdum0.bytes()) != hipSuccess) return bail(fail(SEVP_E_HIP, "memset"));
  } else {
    if ((rc = p->e.alloc((size_t)nr * d.i2))) return bail(rc);
    if ((rc = p->dum0.alloc((size_t)nr * d.i2 * nb1))) return bail(rc);
    if ((rc = p->hfree.alloc((size_t)nr * d.i2))) return bail(rc);
    if ((rc = p->hcpl.alloc((size_t)nr * d.i2))) return bail(rc);
    if ((rc = p->g.alloc((size_t)nr * d.i2))) return bail(rc);
    if ((rc = p->r64.alloc((size_t)nr * nc))) return bail(rc);
    if ((rc = p->dx.alloc((size_t)nr * nx))) return bail(rc);
    if ((rc = check_pass(p, p->pass_kind, d.precision))) return bail(rc);
    if (d.precision != SEVP_PREC_FP64)
      if ((rc = alloc_mixed(p))) return bail(rc);
  }
  *out = p;
  return SEVP_OK;
}

static void drop_graphs(sevp_plan p) {
  for (auto &g : p->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto &e : g.events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }
  p->graphs.clear();
  for (auto &g : p->rgraphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto &e : g.events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }
  p->rgraphs.clear();
  p->prof.pending.clear();
}

int sevp_plan_destroy(sevp_plan p) {
  if (!p) return SEVP_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  drop_graphs(p);
  for (auto st : p->gstreams) (void)hipStreamDestroy(st);
  for (auto ev : p->gevents) (void)hipEventDestroy(ev);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
  if (p->ctl_h) (void)hipHostFree(p->ctl_h);
  if (p->part_h) (void)hipHostFree(p->part_h);
  delete p;
  return SEVP_OK;
}

int sevp_set_stream(sevp_plan p, void *stream) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  HIPC(hipStreamSynchronize(p->stream));
  p->stream = stream ? (hipStream_t)stream : p->own_stream;
  p->prof.st = p->stream;
  return SEVP_OK;
}

size_t sevp_plan_device_bytes(sevp_plan p) {
  if (!p) return 0;
  return p->coef.bytes() + p->ie_d.bytes() + p->brec.bytes() + p->rinv.bytes() + p->rinv1.bytes() +
         p->rinv32.bytes() + p->rinv1_32.bytes() + p->w32.bytes() + p->e.bytes() + p->dum0.bytes() +
         p->hfree.bytes() + p->hcpl.bytes() + p->g.bytes() + p->e32.bytes() +
         p->kpart.bytes() + p->xx.bytes() + p->rinv64t.bytes() + p->rinv1_64t.bytes() +
         p->w64t.bytes() + p->kpart64.bytes() +
         p->bdum0.bytes() + p->s14.bytes() + p->r64.bytes() + p->dx.bytes() + p->part.bytes() + p->b1s.bytes() +
         p->zp.bytes() + p->rinv16t.bytes() + p->rinv1_16t.bytes() + p->w16t.bytes() + p->fe.bytes() + p->fh.bytes() +
         p->fhc.bytes() + p->fe32.bytes() + p->fd.bytes() + p->ctl.bytes() + p->ctl_d.bytes() +
         p->s_stage.bytes();
}

// FP64 influence matrices of an SEVP_STORE_FP64 plan, allocated on first use,
// with RINV_PAD spare doubles: the strip kernels load a whole register column
// of 32 or 64 entries per lane, past the last column's end.
constexpr size_t RINV_PAD = 64;
static int ensure_inverses(sevp_plan p) {
  if (!p->rinv.p) CHK(p->rinv.alloc(p->rinv_count() + RINV_PAD));
  if (!p->rinv1.p) CHK(p->rinv1.alloc(p->rinv1_count() + RINV_PAD));
  return SEVP_OK;
}

// SEVP_STORE_PRODUCTS: the influence matrices of one block at a time in FP64,
// turned into the FP32 product operands (RINV_b, RINV1_b and the coupling
// operator W_b in lane tiles) as soon as they are complete, so the plan never
// holds the FP64 set (186 GB at 3600 columns and NB0 = 900).  Block b needs
// only RINV1_(b-1) of the block before it (PRE1's coupling row and W_b), kept
// in a two-slot ring.  With h_rinv / h_rinv1 (host, reference layout) the
// blocks are uploaded instead of built.  The FP64 arithmetic of each block is
// the one run_pre and make_low_precision_copies perform, so the FP32 operands
// are bitwise those of an SEVP_STORE_FP64 plan.
static int build_products(sevp_plan p, const double *h_rinv, const double *h_rinv1,
                          int *zero_pivots) {
  if (p->d.precision != SEVP_PREC_FP32) return fail(SEVP_E_ARG, "products-only plans are FP32");
  CHK(alloc_mixed(p));
  CHK(alloc_operands(p));
  const int i2 = p->d.i2, nb0 = p->d.nb0;
  const long ww = (long)i2 * i2;
  const long tw = lane_tile_elems(i2, i2);
  hipStream_t st = p->stream;
  DBuf<double> cur, tm, r1a, r1b, sc, wt;
  CHK(cur.alloc(ww));
  CHK(tm.alloc(ww));
  CHK(r1a.alloc(ww));
  CHK(r1b.alloc(ww));
  CHK(sc.alloc(ww));
  CHK(wt.alloc(ww));
  double *r1[2] = {r1a.p, r1b.p};
  CHK(zero_async(p->zp.p, sizeof(int), st));
  const int periodic = p->d.form == SEVP_FORM_FULL_PERIODIC;
  int jl = 1;
  for (int nb = 1; nb <= nb0; ++nb) {
    const int jh = p->ie[nb - 1];
    double *r1cur = r1[(nb - 1) & 1], *r1prev = r1[nb & 1];  // RINV1 of blocks nb and nb-1
    if (h_rinv) {
      HIPC(hipMemcpyAsync(cur.p, h_rinv + (long)(nb - 1) * ww, ww * sizeof(double),
                          hipMemcpyHostToDevice, st));
      if (nb != nb0)
        HIPC(hipMemcpyAsync(r1cur, h_rinv1 + (long)(nb - 1) * ww, ww * sizeof(double),
                            hipMemcpyHostToDevice, st));
    } else {
      MarchArgs m{};
      m.ax = p->al(); m.ay = p->ab(); m.bb = p->ac(); m.cx = p->ar(); m.cy = p->at();
      m.cld = i2;
      m.i2 = i2; m.j2 = p->d.j2;
      m.periodic = periodic;
      m.js = jl; m.jf = jh - 2; m.ieb = jh;
      m.w = i2;
      m.map = ColMap{1, 0, 0};
      m.nrhs = 1;
      m.rinv_b = cur.p;
      m.tmat = (nb != nb0) ? tm.p : nullptr;
      m.rinv1_prev = (nb != 1) ? r1prev : nullptr;
      HIPC(launch_march(MODE_PRE, m, i2, st));
      HIPC(launch_matinv(cur.p, ww, i2, 1, p->zp.p, p->b1s.p, st));
      if (nb != nb0) {
        GemmArgs gm{};
        gm.a = cur.p; gm.a_sm = 1; gm.a_sk = i2;
        gm.b = tm.p; gm.b_sk = 1; gm.b_sn = i2;
        gm.M = i2; gm.N = i2; gm.K = i2;
        gm.ep = EP_NEGSUB;
        gm.out = r1cur;
        gm.out_ld = i2;
        HIPC(launch_gemm(gm, st));
      }
    }
    HIPC(launch_cvt_lane_tiles(cur.p, p->rinv32.p + (long)(nb - 1) * tw, i2, i2, 1, st));
    if (nb != nb0)
      HIPC(launch_cvt_lane_tiles(r1cur, p->rinv1_32.p + (long)(nb - 1) * tw, i2, i2, 1, st));
    if (nb >= 2) {  // W_nb = RINV_nb diag(CY(:, IE(nb-1)-1)) RINV1_(nb-1)
      HIPC(launch_scale_cols(cur.p, p->at(), jl - 1, i2, sc.p, st));
      GemmArgs gm{};
      gm.a = sc.p; gm.a_sm = 1; gm.a_sk = i2;
      gm.b = r1prev; gm.b_sk = 1; gm.b_sn = i2;
      gm.M = i2; gm.N = i2; gm.K = i2;
      gm.ep = EP_COLMAJ;
      gm.out = wt.p; gm.out_ld = i2;
      HIPC(launch_gemm(gm, st));
      HIPC(launch_cvt_lane_tiles(wt.p, p->w32.p + (long)(nb - 2) * tw, i2, i2, 1, st));
    }
    jl = jh;
  }
  int z = 0;
  HIPC(hipMemcpyAsync(&z, p->zp.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  if (zero_pivots) *zero_pivots = z;
  p->has32 = true;
  p->has64 = p->has16 = false;
  p->have_inv = true;
  return SEVP_OK;
}

// The operand copies of the current precision's products, in lane tiles:
// FP32 (SEVP_PREC_FP32), FP64 (SEVP_PREC_FP64_MFMA) or BF16 (SEVP_PREC_BF16)
// copies of RINV, RINV1 and the coupling operators W.
static int make_low_precision_copies(sevp_plan p) {
  if (p->d.precision == SEVP_PREC_FP64 || p->d.form == SEVP_FORM_BIR) return SEVP_OK;
  if (p->storage == SEVP_STORE_PRODUCTS) return SEVP_OK;  // made while the blocks were built
  const bool dbl = p->d.precision == SEVP_PREC_FP64_MFMA;
  const bool b16 = p->d.precision == SEVP_PREC_BF16;
  if (b16 && p->pass_kind != SEVP_PASS_FUSED)
    return fail(SEVP_E_ARG, "BF16 corrections run in the fused pass only");
  bool &have = b16 ? p->has16 : (dbl ? p->has64 : p->has32);
  if (have) return SEVP_OK;
  CHK(alloc_operands(p));
  const int i2 = p->d.i2, nb1 = p->d.nb0 > 1 ? p->d.nb0 - 1 : 1;
  const long ww = (long)i2 * i2;
  const long tw = lane_tile_elems(i2, i2);
  if (b16) {
    HIPC(launch_cvt_lane_tiles_bf16(p->rinv.p, p->rinv16t.p, i2, i2, p->d.nb0, p->stream));
    HIPC(launch_cvt_lane_tiles_bf16(p->rinv1.p, p->rinv1_16t.p, i2, i2, nb1, p->stream));
  } else if (dbl) {
    HIPC(launch_cvt_lane_tiles64(p->rinv.p, p->rinv64t.p, i2, i2, p->d.nb0, p->stream));
    HIPC(launch_cvt_lane_tiles64(p->rinv1.p, p->rinv1_64t.p, i2, i2, nb1, p->stream));
  } else {
    HIPC(launch_cvt_lane_tiles(p->rinv.p, p->rinv32.p, i2, i2, p->d.nb0, p->stream));
    HIPC(launch_cvt_lane_tiles(p->rinv1.p, p->rinv1_32.p, i2, i2, nb1, p->stream));
  }
  // W_b = RINV_b diag(CY(:, IE(b-1)-1)) RINV1_(b-1): the coupling row of the
  // backward pass as one product with the error row (FP64, then the
  // precision of the products)
  if (p->d.nb0 > 1) {
    DBuf<double> t, w;
    CHK(t.alloc(ww));
    CHK(w.alloc(ww));
    for (int nb = 2; nb <= p->d.nb0; ++nb) {
      const int js = p->ie[nb - 2];
      HIPC(launch_scale_cols(p->rinv.p + (long)(nb - 1) * ww, p->at(), js - 1, i2, t.p, p->stream));
      GemmArgs gm{};
      gm.a = t.p; gm.a_sm = 1; gm.a_sk = i2;
      gm.b = p->rinv1.p + (long)(nb - 2) * ww; gm.b_sk = 1; gm.b_sn = i2;
      gm.M = i2; gm.N = i2; gm.K = i2;
      gm.ep = EP_COLMAJ;
      gm.out = w.p;
      gm.out_ld = i2;
      HIPC(launch_gemm(gm, p->stream));
      const long off = (long)(nb - 2) * tw;
      if (b16) HIPC(launch_cvt_lane_tiles_bf16(w.p, p->w16t.p + off, i2, i2, 1, p->stream));
      else if (dbl) HIPC(launch_cvt_lane_tiles64(w.p, p->w64t.p + off, i2, i2, 1, p->stream));
      else HIPC(launch_cvt_lane_tiles(w.p, p->w32.p + off, i2, i2, 1, p->stream));
    }
    HIPC(hipStreamSynchronize(p->stream));
  }
  have = true;
  return SEVP_OK;
}

int sevp_build_influence(sevp_plan p, int *zero_pivots) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  if (p->storage == SEVP_STORE_PRODUCTS) return build_products(p, nullptr, nullptr, zero_pivots);
  CHK(ensure_inverses(p));
  CHK(zero_async(p->zp.p, sizeof(int), p->stream));
  CHK(zero_async(p->rinv1.p, p->rinv1.bytes(), p->stream));
  PreCtx c{};
  c.ax = p->al(); c.ay = p->ab(); c.bb = p->ac(); c.cx = p->ar(); c.cy = p->at();
  c.cld = p->d.i2;
  c.i2 = p->d.i2; c.j2 = p->d.j2;
  c.w = p->w;
  c.nmat = p->nmat;
  c.rinv = p->rinv.p; c.rinv1 = p->rinv1.p;
  c.ie = p->ie.data();
  c.nb0 = p->d.nb0;
  c.zp = p->zp.p;
  c.b1s = p->b1s.p;
  if (p->d.form == SEVP_FORM_BIR) {
    c.periodic = 0;
    c.map = ColMap{1, p->ibir / 2, 1};
  } else {
    c.periodic = (p->d.form == SEVP_FORM_FULL_PERIODIC);
    c.map = ColMap{1, 0, 0};
  }
  CHK(run_pre(c, p->stream));
  p->has32 = p->has64 = p->has16 = false;
  CHK(make_low_precision_copies(p));
  int z = 0;
  HIPC(hipMemcpyAsync(&z, p->zp.p, sizeof(int), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipStreamSynchronize(p->stream));
  if (zero_pivots) *zero_pivots = z;
  p->have_inv = true;
  return SEVP_OK;
}

int sevp_load_influence(sevp_plan p, const double *rinv, const double *rinv1) {
  if (!p || !rinv || !rinv1) return fail(SEVP_E_ARG, "null argument");
  if (p->storage == SEVP_STORE_PRODUCTS) return build_products(p, rinv, rinv1, nullptr);
  CHK(ensure_inverses(p));
  HIPC(hipMemcpy(p->rinv.p, rinv, p->rinv_count() * sizeof(double), hipMemcpyHostToDevice));
  if (p->d.nb0 > 1)
    HIPC(hipMemcpy(p->rinv1.p, rinv1, p->rinv1_count() * sizeof(double), hipMemcpyHostToDevice));
  p->has32 = p->has64 = p->has16 = false;
  CHK(make_low_precision_copies(p));
  HIPC(hipStreamSynchronize(p->stream));
  p->have_inv = true;
  return SEVP_OK;
}

int sevp_get_influence(sevp_plan p, double *rinv, double *rinv1) {
  if (!p || !rinv || !rinv1) return fail(SEVP_E_ARG, "null argument");
  if (!p->have_inv) return fail(SEVP_E_STATE, "influence matrices not built or loaded");
  if (p->storage == SEVP_STORE_PRODUCTS)
    return fail(SEVP_E_STATE, "the plan keeps only the FP32 product operands (SEVP_STORE_PRODUCTS)");
  HIPC(hipStreamSynchronize(p->stream));
  HIPC(hipMemcpy(rinv, p->rinv.p, p->rinv_count() * sizeof(double), hipMemcpyDeviceToHost));
  if (p->d.nb0 > 1)
    HIPC(hipMemcpy(rinv1, p->rinv1.p, p->rinv1_count() * sizeof(double), hipMemcpyDeviceToHost));
  return SEVP_OK;
}

int sevp_reset_seam(sevp_plan p) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  if (p->xx.p) CHK(zero_async(p->xx.p, p->xx.bytes(), p->stream));
  return SEVP_OK;
}

static FullCtx full_ctx(sevp_plan p) {
  const bool mixed = p->d.precision == SEVP_PREC_FP32;
  const bool mfma64 = p->d.precision == SEVP_PREC_FP64_MFMA;
  FullCtx c{};
  c.ax = p->al(); c.ay = p->ab(); c.bb = p->ac(); c.cx = p->ar(); c.cy = p->at();
  c.i2 = p->d.i2; c.j2 = p->d.j2;
  c.periodic = (p->d.form == SEVP_FORM_FULL_PERIODIC);
  c.nb0 = p->d.nb0;
  c.ie = p->ie.data();
  c.rinv = p->rinv.p; c.rinv1 = p->rinv1.p;
  if (mixed) {
    c.rinv32 = p->rinv32.p;
    c.rinv1_32 = p->rinv1_32.p;
    c.w32 = p->w32.p;
  }
  c.e = p->e.p; c.dum0 = p->dum0.p; c.hfree = p->hfree.p; c.hcpl = p->hcpl.p; c.g = p->g.p;
  c.e32 = mixed ? p->e32.p : nullptr;
  c.kpart = mixed ? p->kpart.p : nullptr;
  if (mfma64) {
    c.rinv64t = p->rinv64t.p;
    c.rinv1_64t = p->rinv1_64t.p;
    c.w64t = p->w64t.p;
    c.kpart64 = p->kpart64.p;
  }
  const bool bf16 = p->d.precision == SEVP_PREC_BF16;
  c.fused = (mixed || mfma64 || bf16) && p->pass_kind == SEVP_PASS_FUSED;
  c.farith = mfma64 ? FUSED_F64 : (bf16 ? FUSED_B16 : FUSED_F32);
  c.fwaves = p->fused_waves;
  c.frows = p->fused_rows;
  if (bf16) {
    c.rinv16t = p->rinv16t.p;
    c.rinv1_16t = p->rinv1_16t.p;
    c.w16t = p->w16t.p;
  }
  c.fe = p->fe.p; c.fh = p->fh.p; c.fhc = p->fhc.p; c.fe32 = p->fe32.p;
  c.fd = p->fd.p;
  c.fhalf = (long)p->d.nrhs_max * p->d.i2;
  c.fe_half = p->fe_half;
  c.prof = &p->prof;
  return c;
}

// The context of RHS r0.. of a batch split in groups of ng: every per-RHS
// scratch row offset by r0 right-hand sides (the split-K partials by r0
// times their per-RHS size); the fused pass's lane-tile error rows take a
// region of ceil(ng / 16) row tiles per group.
static FullCtx group_ctx(const FullCtx &c0, int r0, int ng) {
  FullCtx c = c0;
  const long i2 = c.i2;
  const long dstride = i2 * (c.nb0 > 1 ? c.nb0 - 1 : 1);
  c.e += r0 * i2;
  c.hfree += r0 * i2;
  c.hcpl += r0 * i2;
  c.g += r0 * i2;
  c.dum0 += r0 * dstride;
  if (c.e32) c.e32 += r0 * i2;
  const long part_rhs = (long)gemm_part_slices((int)i2, (int)i2, nullptr) * 2 * ((i2 + 15) / 16 * 16);
  if (c.kpart) c.kpart += r0 * part_rhs;
  if (c.kpart64) c.kpart64 += r0 * part_rhs;
  const long lt_off = (long)(r0 / ng) * lane_tile_elems(ng, (int)i2);
  if (c.fe) c.fe += lt_off;
  if (c.fh) c.fh += r0 * i2;
  if (c.fhc) c.fhc += r0 * i2;
  if (c.fe32) c.fe32 += lt_off;
  if (c.fd) c.fd += r0 * i2;
  return c;
}

// Buffers of the split-K products of the plan's precision (FP32 or FP64
// MFMA), allocated on first use.
static int alloc_mixed(sevp_plan p) {
  const bool dbl = p->d.precision == SEVP_PREC_FP64_MFMA;
  if (dbl ? p->kpart64.p != nullptr : p->e32.p != nullptr) return SEVP_OK;
  const int nr = p->d.nrhs_max;
  if (nr > 64) return fail(SEVP_E_ARG, "MFMA corrections solve at most 64 RHS per call");
  if (p->d.i2 % 4 != 0) return fail(SEVP_E_ARG, "MFMA corrections need I2 divisible by 4");
  const int i2 = p->d.i2, np = (i2 + 15) / 16 * 16;
  const size_t n1 = (size_t)gemm_part_slices(i2, i2, nullptr) * i2;
  const size_t n2 = (size_t)gemm_part_slices(2 * np, i2, nullptr) * 2 * np;
  const size_t nf = 2 * (size_t)nr * p->d.i2;  // two parities of (nrhs, I2)
  if (!p->fe.p) {
    // lane-tile error rows: a region of ceil(ng / 16) row tiles per RHS
    // group of ng, for any split into 1 to 16 groups
    long half = 0;
    for (int g = 1; g <= 16; ++g)
      half = std::max(half, g * lane_tile_elems((nr + g - 1) / g, p->d.i2));
    p->fe_half = half;
    CHK(p->fe.alloc(2 * (size_t)half));
    CHK(p->fe32.alloc(2 * (size_t)half));
    CHK(zero_async(p->fe.p, p->fe.bytes(), p->stream));
    CHK(zero_async(p->fe32.p, p->fe32.bytes(), p->stream));
    CHK(p->fh.alloc(nf));
    CHK(p->fhc.alloc(nf));
    if (has_unit_block(p))
      CHK(p->fd.alloc((size_t)(p->d.nb0 - 1) * nr * p->d.i2));
  }
  if (dbl) {
    CHK(p->kpart64.alloc(std::max(n1, n2) * nr));
    return SEVP_OK;
  }
  CHK(p->e32.alloc((size_t)nr * p->d.i2));
  CHK(p->kpart.alloc(std::max(n1, n2) * nr));
  return SEVP_OK;
}

// The product operands of the plan's precision, in lane tiles (both pass
// kinds read them), allocated on first use.
static int alloc_operands(sevp_plan p) {
  const size_t nb1 = p->d.nb0 > 1 ? p->d.nb0 - 1 : 1;
  const size_t tw = (size_t)lane_tile_elems(p->d.i2, p->d.i2);
  if (p->d.precision == SEVP_PREC_BF16) {
    if (!p->rinv16t.p) {
      CHK(p->rinv16t.alloc(tw * p->d.nb0));
      CHK(p->rinv1_16t.alloc(tw * nb1));
      CHK(p->w16t.alloc(tw * nb1));
    }
  } else if (p->d.precision == SEVP_PREC_FP64_MFMA) {
    if (!p->rinv64t.p) {
      CHK(p->rinv64t.alloc(tw * p->d.nb0));
      CHK(p->rinv1_64t.alloc(tw * nb1));
      CHK(p->w64t.alloc(tw * nb1));
    }
  } else if (!p->rinv32.p) {
    CHK(p->rinv32.alloc(tw * p->d.nb0));
    CHK(p->rinv1_32.alloc(tw * nb1));
    CHK(p->w32.alloc(tw * nb1));
  }
  return SEVP_OK;
}

// The fused pass marches a block's rows inside a 16-column product tile and
// keeps the 16 - 2 (h - 1) columns that stay exact: blocks up to 4 rows high.
static bool fused_eligible(sevp_plan p) {
  if (p->d.form == SEVP_FORM_BIR || p->d.i2 % 4 != 0 || p->d.i2 < 16) return false;
  int prev = 1;
  for (int b = 0; b < p->d.nb0; ++b) {
    if (p->ie[b] - prev > 4) return false;
    prev = p->ie[b];
  }
  return true;
}

int sevp_set_pass(sevp_plan p, int kind) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  if (kind != SEVP_PASS_SPLITK && kind != SEVP_PASS_FUSED) return fail(SEVP_E_ARG, "unknown pass kind");
  if (kind == p->pass_kind) return SEVP_OK;
  if (p->d.form == SEVP_FORM_BIR) return fail(SEVP_E_ARG, "the BIR form has no correction products to fuse");
  if (kind == SEVP_PASS_FUSED && !fused_eligible(p))
    return fail(SEVP_E_ARG, "the fused pass needs block heights <= 4 and I2 divisible by 4 (>= 16)");
  CHK(check_pass(p, kind, p->d.precision));
  HIPC(hipStreamSynchronize(p->stream));
  drop_graphs(p);
  const int old = p->pass_kind;
  p->pass_kind = kind;
  if (p->have_inv && p->d.precision != SEVP_PREC_FP64) {
    if (int rc = make_low_precision_copies(p)) {
      p->pass_kind = old;
      return rc;
    }
  }
  return SEVP_OK;
}

int sevp_get_pass(sevp_plan p, int *kind) {
  if (!p || !kind) return fail(SEVP_E_ARG, "null argument");
  *kind = p->pass_kind;
  return SEVP_OK;
}

int sevp_get_tuning(sevp_plan p, sevp_tuning *t) {
  if (!p || !t) return fail(SEVP_E_ARG, "null argument");
  t->pass_kind = p->pass_kind;
  t->waves = p->fused_waves;
  t->rhs_groups = p->rhs_groups;
  t->rows = p->fused_rows;
  return SEVP_OK;
}

int sevp_set_tuning(sevp_plan p, const sevp_tuning *t) {
  if (!p || !t) return fail(SEVP_E_ARG, "null argument");
  if (t->waves != 8 && t->waves != 16) return fail(SEVP_E_ARG, "waves per workgroup must be 8 or 16");
  if (t->rhs_groups < 1 || t->rhs_groups > 16) return fail(SEVP_E_ARG, "RHS groups must be 1..16");
  if (t->rows < 0 || t->rows > 2) return fail(SEVP_E_ARG, "row tiles per workgroup must be 0, 1 or 2");
  CHK(sevp_set_pass(p, t->pass_kind));
  HIPC(hipStreamSynchronize(p->stream));
  if (p->fused_waves != t->waves || p->fused_rows != t->rows) drop_graphs(p);
  p->fused_waves = t->waves;
  p->fused_rows = t->rows;
  p->rhs_groups = t->rhs_groups;
  return SEVP_OK;
}

// Time refined solves of the caller's problem over the pass kernels, the
// waves and row tiles per fused workgroup and the RHS groups, and keep the
// fastest.
int sevp_autotune(sevp_plan p, const float *s, double *x, int nrhs, double tol, int max_iter,
                  sevp_tuning *best, double *best_ms) {
  if (!p || !s || !x) return fail(SEVP_E_ARG, "null argument");
  if (p->d.form == SEVP_FORM_BIR || p->d.precision == SEVP_PREC_FP64)
    return fail(SEVP_E_ARG, "the autotuner chooses among MFMA pass kernels: full forms, MFMA precision");
  if (nrhs < 1 || nrhs > p->d.nrhs_max) return fail(SEVP_E_ARG, "nrhs out of range");
  if (!p->have_inv) return fail(SEVP_E_STATE, "influence matrices not built or loaded");
  const long nx = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
  std::vector<sevp_tuning> cand;
  for (int kind : {SEVP_PASS_FUSED, SEVP_PASS_SPLITK}) {
    if (check_pass(p, kind, p->d.precision) != SEVP_OK) continue;
    for (int waves : {16, 8})
      for (int rows : {1, 2}) {
        // waves and row tiles are not split-K parameters
        if (kind == SEVP_PASS_SPLITK && (waves != FUSED_WAVES || rows != 1)) continue;
        for (int g : {1, 2, 4})
          if (g <= nrhs) cand.push_back(sevp_tuning{kind, waves, g, kind == SEVP_PASS_SPLITK ? 0 : rows});
      }
  }
  if (cand.empty()) return fail(SEVP_E_ARG, "no pass kernels apply to this plan");
  sevp_tuning keep{};
  CHK(sevp_get_tuning(p, &keep));
  double bms = 1e300;
  sevp_tuning b = keep;
  for (const auto &t : cand) {
    CHK(sevp_set_tuning(p, &t));
    double ms = 1e300;
    for (int rep = 0; rep < 3; ++rep) {  // the first run captures the graphs
      CHK(zero_async(x, sizeof(double) * nx * nrhs, p->stream));
      HIPC(hipStreamSynchronize(p->stream));
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = sevp_solve_refined_dev(p, s, x, nrhs, tol, max_iter, nullptr, nullptr);
      if (rc != SEVP_OK && rc != SEVP_E_NOTCONV) return rc;
      const double dt =
          1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (rep > 0) ms = std::min(ms, dt);
    }
    if (ms < bms) {
      bms = ms;
      b = t;
    }
  }
  CHK(sevp_set_tuning(p, &b));
  if (best) *best = b;
  if (best_ms) *best_ms = bms;
  return SEVP_OK;
}

int sevp_set_storage(sevp_plan p, int storage) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  if (storage != SEVP_STORE_FP64 && storage != SEVP_STORE_PRODUCTS)
    return fail(SEVP_E_ARG, "unknown storage");
  if (storage == p->storage) return SEVP_OK;
  if (storage == SEVP_STORE_PRODUCTS) {
    if (p->d.form == SEVP_FORM_BIR) return fail(SEVP_E_ARG, "products-only storage needs a full-width form");
    if (p->d.precision != SEVP_PREC_FP32)
      return fail(SEVP_E_ARG, "products-only storage needs SEVP_PREC_FP32");
  }
  HIPC(hipStreamSynchronize(p->stream));
  drop_graphs(p);
  p->rinv.release();
  p->rinv1.release();
  p->has32 = p->has64 = p->has16 = false;
  p->have_inv = false;
  p->storage = storage;
  return SEVP_OK;
}

int sevp_set_precision(sevp_plan p, int precision) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  if (precision < SEVP_PREC_FP64 || precision > SEVP_PREC_FP64_MFMA)
    return fail(SEVP_E_ARG, "unknown precision");
  if (p->d.form == SEVP_FORM_BIR && precision != SEVP_PREC_FP64)
    return fail(SEVP_E_ARG, "BIR form runs the plain FP64 arithmetic only");
  CHK(check_pass(p, p->pass_kind, precision));
  if (p->storage == SEVP_STORE_PRODUCTS && precision != SEVP_PREC_FP32)
    return fail(SEVP_E_ARG, "a plan that keeps only FP32 product operands runs SEVP_PREC_FP32 only");
  const int old = p->d.precision;
  p->d.precision = precision;
  if (precision != SEVP_PREC_FP64) {
    if (int rc = alloc_mixed(p)) {
      p->d.precision = old;
      return rc;
    }
    if (p->have_inv) CHK(make_low_precision_copies(p));
    HIPC(hipStreamSynchronize(p->stream));
  }
  return SEVP_OK;
}

__global__ void span_init_kernel(unsigned long long *t, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = 0ull;
}

int sevp_profile_begin(sevp_plan p, int kclass) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  if (kclass < SEVP_KCLASS_CORR || kclass > SEVP_KCLASS_RESID) return fail(SEVP_E_ARG, "bad kernel class");
  HIPC(hipStreamSynchronize(p->stream));
  p->prof.harvest();
  p->prof.harvest_spans();
  KProf &k = p->prof;
  if (!k.spans) HIPC(hipMalloc(&k.spans, KProf::kSpanWords * sizeof(unsigned long long)));
  const long n = (long)KProf::kSpanWords;
  hipLaunchKernelGGL(span_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream,
                     k.spans, n);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(p->stream));
  k.kclass = kclass;
  k.total_ms = 0.0;
  k.launches = 0;
  k.span_ms = 0.0;
  k.span_launches = 0;
  k.span_runs.clear();
  k.span_next = 0;
  k.st = p->stream;
  return SEVP_OK;
}

int sevp_profile_spans(sevp_plan p, double *total_ms, long *launches) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  HIPC(hipStreamSynchronize(p->stream));
  for (auto st : p->gstreams) HIPC(hipStreamSynchronize(st));
  p->prof.harvest_spans();
  if (total_ms) *total_ms = p->prof.span_ms;
  if (launches) *launches = p->prof.span_launches;
  return SEVP_OK;
}

__global__ void empty_kernel() {}

int sevp_profile_end(sevp_plan p, double *total_ms, long *launches) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  HIPC(hipStreamSynchronize(p->stream));
  p->prof.harvest();
  if (total_ms) *total_ms = p->prof.total_ms;
  if (launches) *launches = p->prof.launches;
  p->prof.kclass = -1;
  return SEVP_OK;
}

int sevp_solve_dev(sevp_plan p, const float *s, double *x, int nrhs, void *stream) {
  if (!p || !s || !x) return fail(SEVP_E_ARG, "null argument");
  if (nrhs < 1 || nrhs > p->d.nrhs_max) return fail(SEVP_E_ARG, "nrhs out of range");
  if (!p->have_inv) return fail(SEVP_E_STATE, "influence matrices not built or loaded");
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  const long s_stride = (long)p->d.i2 * p->d.j2;
  const long x_stride = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
  if (p->d.form == SEVP_FORM_BIR) {
    BirArgs a{};
    a.al = p->al(); a.ab = p->ab(); a.ac = p->ac(); a.ar = p->ar(); a.at = p->at();
    a.s = s; a.rinv = p->rinv.p; a.rinv1 = p->rinv1.p;
    a.x = x; a.xx = p->xx.p; a.dum0 = p->bdum0.p; a.ie = p->ie_d.p;
    a.recs = p->brec.p; a.fwd_records = p->brec_fwd;
    a.i2 = p->d.i2; a.j2 = p->d.j2; a.nbir = p->d.nbir; a.ibir = p->ibir; a.nb0 = p->d.nb0;
    a.nsweep = p->d.nsweep; a.periodic = p->d.periodic; a.seam_fix = p->d.seam_fix;
    a.s_stride = s_stride; a.x_stride = x_stride;
    a.xx_stride = (long)(p->ibir + 2) * (p->d.j2 + 2);
    a.dum0_stride = (long)p->ibir * (p->d.nb0 > 1 ? p->d.nb0 - 1 : 1);
    HIPC(launch_bir_sweeps(a, nrhs, st));
    return SEVP_OK;
  }
  FullCtx c = full_ctx(p);
  return run_full_pass(c, s, nullptr, s_stride, x, x_stride, nrhs, st);
}

int sevp_solve(sevp_plan p, const float *s, double *x, int nrhs) {
  if (!p || !s || !x) return fail(SEVP_E_ARG, "null argument");
  if (nrhs < 1 || nrhs > p->d.nrhs_max) return fail(SEVP_E_ARG, "nrhs out of range");
  const size_t ns = (size_t)p->d.i2 * p->d.j2 * nrhs;
  const size_t nx = (size_t)(p->d.i2 + 2) * (p->d.j2 + 2) * nrhs;
  DBuf<float> sd;
  DBuf<double> xd;
  CHK(sd.alloc(ns));
  CHK(xd.alloc(nx));
  HIPC(hipMemcpyAsync(sd.p, s, ns * sizeof(float), hipMemcpyHostToDevice, p->stream));
  HIPC(hipMemcpyAsync(xd.p, x, nx * sizeof(double), hipMemcpyHostToDevice, p->stream));
  CHK(sevp_solve_dev(p, sd.p, xd.p, nrhs, nullptr));
  HIPC(hipMemcpyAsync(x, xd.p, nx * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipStreamSynchronize(p->stream));
  return SEVP_OK;
}

int sevp_residual_dev(sevp_plan p, const float *s, const double *x, int nrhs, double *rel) {
  if (!p || !s || !x || !rel) return fail(SEVP_E_ARG, "null argument");
  if (nrhs < 1 || nrhs > p->d.nrhs_max) return fail(SEVP_E_ARG, "nrhs out of range");
  std::vector<double> v;
  CHK(residual_norms(p, s, nullptr, (long)p->d.i2 * p->d.j2, x, (long)(p->d.i2 + 2) * (p->d.j2 + 2),
                     nullptr, nrhs, v));
  std::memcpy(rel, v.data(), sizeof(double) * nrhs);
  return SEVP_OK;
}

// Streams (default priority, non-blocking) and join events for `groups` RHS
// groups, created on first use.
static int ensure_group_streams(sevp_plan p, int groups) {
  while ((int)p->gstreams.size() < groups - 1) {
    hipStream_t st;
    HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    p->gstreams.push_back(st);
  }
  while ((int)p->gevents.size() < groups) {
    hipEvent_t ev;
    HIPC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    p->gevents.push_back(ev);
  }
  return SEVP_OK;
}

// Put an event-record node before and after each of `knodes` in `graph`: the
// first waits for the kernel's dependencies and the kernel waits for it, the
// second follows the kernel and its dependents wait for it.
static int add_timing_nodes(hipGraph_t graph, const std::vector<hipGraphNode_t> &knodes,
                            std::vector<std::pair<hipEvent_t, hipEvent_t>> &events) {
  std::vector<hipGraphNode_t> deps;
  for (hipGraphNode_t k : knodes) {
    hipEvent_t a, b;
    HIPC(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return fail(SEVP_E_HIP, "hipEventCreate failed");
    }
    events.push_back({a, b});
    size_t n = 0;
    HIPC(hipGraphNodeGetDependencies(k, nullptr, &n));
    deps.resize(n);
    if (n) HIPC(hipGraphNodeGetDependencies(k, deps.data(), &n));
    hipGraphNode_t ra, rb;
    HIPC(hipGraphAddEventRecordNode(&ra, graph, deps.data(), n, a));
    for (size_t i = 0; i < n; ++i) HIPC(hipGraphRemoveDependencies(graph, &deps[i], &k, 1));
    HIPC(hipGraphAddDependencies(graph, &ra, &k, 1));
    n = 0;
    HIPC(hipGraphNodeGetDependentNodes(k, nullptr, &n));
    deps.resize(n);
    if (n) HIPC(hipGraphNodeGetDependentNodes(k, deps.data(), &n));
    HIPC(hipGraphAddEventRecordNode(&rb, graph, &k, 1, b));
    for (size_t i = 0; i < n; ++i) {
      HIPC(hipGraphRemoveDependencies(graph, &k, &deps[i], 1));
      HIPC(hipGraphAddDependencies(graph, &rb, &deps[i], 1));
    }
  }
  return SEVP_OK;
}

// The captured pass for (nrhs, precision, profiling), captured on first use.
static int get_pass_graph(sevp_plan p, const FullCtx &c0, int nrhs, int r0, int count, int ng,
                          PassGraph **out) {
  const bool prof = p->prof.kclass >= 0;
  for (auto &g : p->graphs)
    if (g.nrhs == nrhs && g.r0 == r0 && g.count == count && g.ng == ng && g.prec == p->d.precision &&
        g.kind == p->pass_kind && g.waves == p->fused_waves && g.rows == p->fused_rows &&
        g.prof == prof) {
      *out = &g;
      return SEVP_OK;
    }
  const long ns = (long)p->d.i2 * p->d.j2, nx = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
  const FullCtx c = group_ctx(c0, r0, ng);
  PassGraph g;
  g.nrhs = nrhs;
  g.r0 = r0;
  g.count = count;
  g.ng = ng;
  g.prec = p->d.precision;
  g.kind = p->pass_kind;
  g.waves = p->fused_waves;
  g.rows = p->fused_rows;
  g.prof = prof;
  hipStream_t cs = p->cap_stream;
  HIPC(hipStreamSynchronize(p->stream));
  HIPC(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  hipStream_t saved = p->prof.st;
  p->prof.st = cs;
  std::vector<hipGraphNode_t> knodes;
  if (prof) p->prof.capture_into = &knodes;
  int rc = zero_async(p->dx.p + r0 * nx, sizeof(double) * nx * count, cs);
  if (rc == SEVP_OK)
    rc = run_full_pass(c, nullptr, p->r64.p + r0 * ns, ns, p->dx.p + r0 * nx, nx, count, cs);
  p->prof.capture_into = nullptr;
  p->prof.st = saved;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(cs, &graph);
  if (rc != SEVP_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != SEVP_OK) return rc;
    HIPC(e);
  }
  if (prof) {
    rc = add_timing_nodes(graph, knodes, g.events);
    if (rc != SEVP_OK) {
      (void)hipGraphDestroy(graph);
      for (auto &ev : g.events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
      }
      return rc;
    }
  }
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    for (auto &ev : g.events) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    HIPC(ei);
  }
  p->graphs.push_back(std::move(g));
  *out = &p->graphs.back();
  return SEVP_OK;
}

int sevp_step14_dev(sevp_plan p, const sevp_step14_desc *d, const sevp_step14_grid *g,
                    const sevp_step14_state *st, int *itmask, float *tmp) {
  if (!p || !d || !g || !st) return fail(SEVP_E_ARG, "null argument");
  if (d->i2 != p->d.i2 || d->j2 != p->d.j2) return fail(SEVP_E_ARG, "step 14 grid differs from the plan's");
  if (d->k1 < 1 || d->mxmask < 1) return fail(SEVP_E_ARG, "K1 and MXMASK must be >= 1");
  if (!p->have_inv) return fail(SEVP_E_STATE, "influence matrices not built or loaded");
  if (!g->kb || !g->iu0 || !g->iv0 || !g->iu || !g->iv || !g->odx || !g->ody || !g->ocs ||
      !g->csv || !g->odyv || !g->odz || !st->u || !st->v || !st->w || !st->p0 || !st->scr ||
      !st->s || !st->x)
    return fail(SEVP_E_ARG, "null array");
  const long np = step14_partials(d->i2, d->j2);
  if (!p->s14.p) CHK(p->s14.alloc((size_t)np + 1));
  Step14Args a{};
  a.i2 = d->i2; a.j2 = d->j2; a.k1 = d->k1;
  a.mask = 1; a.lopenw = d->lopenw; a.harness = d->harness;
  a.free_surface = d->free_surface; a.periodic_we = d->periodic_we;
  a.odt = d->odt;
  a.kb = g->kb; a.iu0 = g->iu0; a.iv0 = g->iv0; a.iu = g->iu; a.iv = g->iv;
  a.odx = g->odx; a.ody = g->ody; a.ocs = g->ocs; a.csv = g->csv; a.odyv = g->odyv; a.odz = g->odz;
  a.u = st->u; a.v = st->v; a.w = st->w; a.p0 = st->p0; a.scr = st->scr; a.s = st->s; a.x = st->x;
  a.tmp_part = p->s14.p;
  hipStream_t s = p->stream;
  HIPC(launch_step14_zero_scr(st->scr, 4L * (d->i2 + 2) * (d->j2 + 2), s));
  int it = 0;
  float t = 0.f;
  for (;;) {
    ++it;
    HIPC(launch_step14_prepare(a, s));
    CHK(sevp_solve_dev(p, st->s, st->x, 1, s));
    HIPC(launch_step14_correct(a, p->s14.p + np, s));
    HIPC(hipMemcpyAsync(&t, p->s14.p + np, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (t <= d->tolerance || it >= d->mxmask) break;
  }
  if (itmask) *itmask = it;
  if (tmp) *tmp = t;
  return SEVP_OK;
}

int sevp_set_rhs_groups(sevp_plan p, int groups) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  if (groups < 1 || groups > 16) return fail(SEVP_E_ARG, "RHS groups must be 1..16");
  HIPC(hipStreamSynchronize(p->stream));
  for (auto st : p->gstreams) HIPC(hipStreamSynchronize(st));
  p->rhs_groups = groups;
  return SEVP_OK;
}

int sevp_set_graphs(sevp_plan p, int on) {
  if (!p) return fail(SEVP_E_ARG, "null plan");
  HIPC(hipStreamSynchronize(p->stream));
  p->use_graphs = on != 0;
  if (!p->use_graphs) drop_graphs(p);
  return SEVP_OK;
}

// Buffers of the captured refinement, allocated on first use.
static int ensure_refine_ctl(sevp_plan p) {
  if (p->ctl.p) return SEVP_OK;
  CHK(p->ctl.alloc(4));
  CHK(p->ctl_d.alloc((size_t)p->d.nrhs_max + 2));
  HIPC(hipHostMalloc((void **)&p->ctl_h, 16 * sizeof(int), hipHostMallocDefault));
  HIPC(hipHostMalloc((void **)&p->part_h, sizeof(double) * 2 * NCHUNK * p->d.nrhs_max,
                     hipHostMallocDefault));
  CHK(p->s_stage.alloc((size_t)p->d.nrhs_max * p->d.i2 * p->d.j2));
  return SEVP_OK;
}

// The passes of RHS groups (ng right-hand sides each) on their streams,
// forked from and joined back into `st` (which may be capturing).
static int run_group_passes(sevp_plan p, const FullCtx &c, int nrhs, int ngroups, int ng,
                            hipStream_t st, bool graph_groups) {
  const long ns = (long)p->d.i2 * p->d.j2, nx = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
  // a profiled eager pass runs the groups one after another on st, so each
  // kernel's timed duration is its own and not shared with another stream's
  // (captured and graph-replayed passes keep them concurrent)
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  HIPC(hipStreamIsCapturing(st, &cap));
  const bool concurrent = ngroups > 1 && (graph_groups || p->prof.kclass < 0 ||
                                          cap == hipStreamCaptureStatusActive);
  if (concurrent) HIPC(hipEventRecord(p->gevents[0], st));
  for (int g = 0; g < ngroups; ++g) {
    const int r0 = g * ng;
    if (r0 >= nrhs) break;
    const int count = std::min(ng, nrhs - r0);
    hipStream_t gs = (g == 0 || !concurrent) ? st : p->gstreams[g - 1];
    if (g > 0 && concurrent) HIPC(hipStreamWaitEvent(gs, p->gevents[0], 0));
    PassGraph *pg = nullptr;
    if (graph_groups) CHK(get_pass_graph(p, c, nrhs, r0, count, ng, &pg));
    if (pg) {
      HIPC(hipGraphLaunch(pg->exec, gs));
      if (pg->prof) p->prof.pending.push_back(&pg->events);
    } else {
      const FullCtx cg = group_ctx(c, r0, ng);
      hipStream_t saved = p->prof.st;
      p->prof.st = gs;
      int rc = zero_async(p->dx.p + r0 * nx, sizeof(double) * nx * count, gs);
      if (rc == SEVP_OK)
        rc = run_full_pass(cg, nullptr, p->r64.p + r0 * ns, ns, p->dx.p + r0 * nx, nx, count, gs);
      p->prof.st = saved;
      CHK(rc);
    }
    if (g > 0 && concurrent) {
      HIPC(hipEventRecord(p->gevents[g], gs));
      HIPC(hipStreamWaitEvent(st, p->gevents[g], 0));
    }
  }
  return SEVP_OK;
}

static void launch_resid_norm(sevp_plan p, const float *s, const double *x, int nrhs, const int *gate,
                              hipStream_t st) {
  const int i2 = p->d.i2, j2 = p->d.j2;
  launch_resid_norm_cols(p, s, nullptr, (long)i2 * j2, x, (long)(i2 + 2) * (j2 + 2), p->r64.p,
                         nrhs, gate, st);
}

// Capture the whole refinement for `passes` passes (see RefineGraph).  It
// reads the right-hand sides from the plan's staging copy (so a caller that
// passes a new s each time step reuses the graph) and updates the caller's x.
static int get_refine_graph(sevp_plan p, const FullCtx &c, double *x, int nrhs, int ngroups,
                            int ng, int passes, RefineGraph **out) {
  const bool prof = p->prof.kclass >= 0;
  CHK(ensure_refine_ctl(p));
  const float *s = p->s_stage.p;
  for (auto &g : p->rgraphs)
    if (g.nrhs == nrhs && g.prec == p->d.precision && g.kind == p->pass_kind &&
        g.groups == ngroups && g.passes == passes && g.prof == prof && g.x == x &&
        g.waves == p->fused_waves && g.rows == p->fused_rows) {
      *out = &g;
      return SEVP_OK;
    }
  if (p->rgraphs.size() >= 8) {  // the oldest capture goes
    RefineGraph &o = p->rgraphs.front();
    HIPC(hipStreamSynchronize(p->stream));
    if (o.exec) (void)hipGraphExecDestroy(o.exec);
    for (auto &e : o.events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    p->rgraphs.pop_front();
  }
  const int i2 = p->d.i2, j2 = p->d.j2;
  const long nx = (long)(i2 + 2) * (j2 + 2);
  RefineGraph g;
  g.nrhs = nrhs; g.prec = p->d.precision; g.kind = p->pass_kind; g.groups = ngroups;
  g.passes = passes; g.prof = prof; g.s = s; g.x = x; g.waves = p->fused_waves;
  g.rows = p->fused_rows;
  hipStream_t cs = p->cap_stream;
  HIPC(hipStreamSynchronize(p->stream));
  HIPC(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  hipStream_t saved = p->prof.st;
  p->prof.st = cs;
  std::vector<hipGraphNode_t> knodes;
  if (prof) p->prof.capture_into = &knodes;
  int rc = SEVP_OK;
  const long n = (long)i2 * j2 * nrhs;
  launch_resid_norm(p, s, x, nrhs, nullptr, cs);
  hipLaunchKernelGGL(refine_finish_kernel, dim3(1), dim3(256), 0, cs, (const double *)p->part.p,
                     nrhs, NCHUNK, p->ctl_d.p + 1, p->ctl.p, (const double *)p->ctl_d.p, 1);
  for (int it = 0; it < passes && rc == SEVP_OK; ++it) {
    rc = run_group_passes(p, c, nrhs, ngroups, ng, cs, false);
    if (rc != SEVP_OK) break;
    hipLaunchKernelGGL(refine_axpy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cs, x, nx,
                       (const double *)p->dx.p, nx, i2, j2, nrhs, c.periodic, (const int *)p->ctl.p);
    launch_resid_norm(p, s, x, nrhs, p->ctl.p, cs);
    hipLaunchKernelGGL(refine_finish_kernel, dim3(1), dim3(256), 0, cs, (const double *)p->part.p,
                       nrhs, NCHUNK, p->ctl_d.p + 1, p->ctl.p, (const double *)p->ctl_d.p, 0);
  }
  if (rc == SEVP_OK && hipGetLastError() != hipSuccess) rc = fail(SEVP_E_HIP, "launch failed in capture");
  p->prof.capture_into = nullptr;
  p->prof.st = saved;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(cs, &graph);
  if (rc != SEVP_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != SEVP_OK) return rc;
    HIPC(e);
  }
  if (prof) {
    rc = add_timing_nodes(graph, knodes, g.events);
    if (rc != SEVP_OK) {
      (void)hipGraphDestroy(graph);
      for (auto &ev : g.events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
      }
      return rc;
    }
  }
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    for (auto &ev : g.events) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    HIPC(ei);
  }
  p->rgraphs.push_back(std::move(g));
  *out = &p->rgraphs.back();
  return SEVP_OK;
}

int sevp_solve_refined_dev(sevp_plan p, const float *s, double *x, int nrhs, double tol,
                           int max_iter, double *rel_out, sevp_stats *stats) {
  if (!p || !s || !x) return fail(SEVP_E_ARG, "null argument");
  if (p->d.form == SEVP_FORM_BIR) return fail(SEVP_E_ARG, "refinement needs a full-width form");
  if (nrhs < 1 || nrhs > p->d.nrhs_max) return fail(SEVP_E_ARG, "nrhs out of range");
  if (!p->have_inv) return fail(SEVP_E_STATE, "influence matrices not built or loaded");
  auto t0 = std::chrono::steady_clock::now();
  const int i2 = p->d.i2, j2 = p->d.j2;
  const long ns = (long)i2 * j2, nx = (long)(i2 + 2) * (j2 + 2);
  FullCtx c = full_ctx(p);
  // RHS groups: ng right-hand sides each (the last may be smaller)
  const int ngroups = std::min(p->rhs_groups, nrhs);
  const int ng = (nrhs + ngroups - 1) / ngroups;
  CHK(ensure_group_streams(p, ngroups));
  std::vector<double> rel;
  int it = 0;
  double worst = -1.0;
  bool done = false;
  int &expect = p->expect_passes[p->d.precision & 3][p->pass_kind & 1];
  if (p->use_graphs && expect > 0 && max_iter > 0) {
    // the passes the last solve took, as one captured graph; the stop flag
    // lives on the device and the host reads it once
    RefineGraph *rg = nullptr;
    CHK(get_refine_graph(p, c, x, nrhs, ngroups, ng, std::min(expect, max_iter), &rg));
    HIPC(hipMemcpyAsync(p->s_stage.p, s, sizeof(float) * ns * nrhs, hipMemcpyDeviceToDevice,
                        p->stream));
    p->ctl_h[0] = 0;
    p->ctl_h[1] = 0;
    p->ctl_h[2] = max_iter;
    std::memcpy(p->ctl_h + 8, &tol, sizeof(double));
    HIPC(hipMemcpyAsync(p->ctl.p, p->ctl_h, 3 * sizeof(int), hipMemcpyHostToDevice, p->stream));
    HIPC(hipMemcpyAsync(p->ctl_d.p, p->ctl_h + 8, sizeof(double), hipMemcpyHostToDevice, p->stream));
    HIPC(hipGraphLaunch(rg->exec, p->stream));
    if (rg->prof) p->prof.pending.push_back(&rg->events);
    HIPC(hipMemcpyAsync(p->ctl_h + 4, p->ctl.p, 2 * sizeof(int), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipMemcpyAsync(p->part_h, p->part.p, sizeof(double) * 2 * NCHUNK * nrhs,
                        hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    p->prof.harvest();
    it = p->ctl_h[4];
    rel.assign(nrhs, 0.0);
    worst = 0.0;
    for (int r = 0; r < nrhs; ++r) {  // as residual_norms sums them
      double a = 0, b = 0;
      for (int k = 0; k < NCHUNK; ++k) {
        a += p->part_h[((size_t)r * NCHUNK + k) * 2];
        b += p->part_h[((size_t)r * NCHUNK + k) * 2 + 1];
      }
      rel[r] = b > 0 ? std::sqrt(a / b) : std::sqrt(a);
      worst = std::max(worst, rel[r]);
    }
    done = !(worst > tol) || it >= max_iter;
  }
  if (!done) {
    for (;;) {
      CHK(residual_norms(p, s, nullptr, ns, x, nx, p->r64.p, nrhs, rel));
      worst = 0.0;
      for (double v : rel) worst = std::max(worst, v);
      if (!(worst > tol) || it >= max_iter) break;
      CHK(run_group_passes(p, c, nrhs, ngroups, ng, p->stream, p->use_graphs));
      const long n = ns * nrhs;
      hipLaunchKernelGGL(axpy_interior_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                         p->stream, x, nx, p->dx.p, nx, i2, j2, nrhs);
      HIPC(hipGetLastError());
      if (c.periodic) HIPC(launch_halo_fix(x, nx, i2, j2, nrhs, p->stream));
      ++it;
    }
  }
  if (!(worst > tol)) expect = it;
  if (rel_out) std::memcpy(rel_out, rel.data(), sizeof(double) * nrhs);
  if (stats) {
    stats->iterations = it;
    stats->max_rel_resid = worst;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (worst > tol) return fail(SEVP_E_NOTCONV, "refinement did not reach the tolerance");
  return SEVP_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batched problems
// Many independent full-width problems (regional basins, ensemble members on
// different operators) refined together.  Every step of the mixed pass is one
// grouped launch over all problems: forward step t marches block t of every
// problem with at least t blocks, then forms their products; backward step t
// takes block NB0 - t + 1 of each.  A problem's entries carry exactly the
// arguments of its own launches (mixed_pass_ops), so its results are bitwise
// those of sevp_solve_refined_dev on its plan.  The pass is one captured
// graph of 4 * max(NB0) + 2 launches; each refinement step reads every
// problem's residual norms with one host synchronisation, and a problem
// that has converged keeps its X (its correction is computed and not added).
enum { BL_FILL, BL_MARCH, BL_PROD, BL_PROD64, BL_RESID, BL_HOM, BL_HALO };
struct BatchLaunch {
  int kind;
  size_t tab, start;  // byte offsets of the entry table and workgroup starts
  int n;
  long nwg;
};

struct sevp_batch_s {
  std::vector<sevp_plan> plans;
  hipStream_t stream = nullptr, cap = nullptr;
  int ks = 0;                 // the product slice length every problem uses
  std::vector<int> key;       // nrhs per problem of the captured pass
  int key_prec = -1;          // the plans' correction precision of the captured pass
  int kmarch = 1, mb = 1;     // lane width of the grouped marches, 16-row blocks of products
  std::vector<BatchLaunch> launches;
  DBuf<unsigned char> tables;
  hipGraphExec_t exec = nullptr;
  DBuf<double> norm_part;     // residual-norm partials, NCHUNK * 2 per RHS per problem
  DBuf<unsigned char> xtab;   // per call: X += DX entries (all, then periodic) and starts
  DBuf<int> act;              // per problem: still refining
  int *act_h = nullptr;       // pinned
  double *norm_h = nullptr;   // pinned
  size_t norm_h_n = 0;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
    key_prec = -1;
  }
  ~sevp_batch_s() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    if (act_h) (void)hipHostFree(act_h);
    if (norm_h) (void)hipHostFree(norm_h);
    if (cap) (void)hipStreamDestroy(cap);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Host image of the batch's tables: entry arrays and their starts, each at a
// 256-byte boundary.
struct TableImage {
  std::vector<unsigned char> bytes;
  template <class T>
  size_t put(const T *v, size_t n) {
    const size_t off = (bytes.size() + 255) / 256 * 256;
    bytes.resize(off + n * sizeof(T));
    std::memcpy(bytes.data() + off, v, n * sizeof(T));
    return off;
  }
};

template <class T>
static void add_launch(TableImage &img, std::vector<BatchLaunch> &out, int kind,
                       const std::vector<T> &ents, const std::vector<long> &nwg) {
  if (ents.empty()) return;
  std::vector<int> start(ents.size());
  long total = 0;
  for (size_t i = 0; i < ents.size(); ++i) {
    start[i] = (int)total;
    total += nwg[i];
  }
  BatchLaunch l;
  l.kind = kind;
  l.tab = img.put(ents.data(), ents.size());
  l.start = img.put(start.data(), start.size());
  l.n = (int)ents.size();
  l.nwg = total;
  out.push_back(l);
}

static long cdiv(long a, long b) { return (a + b - 1) / b; }

static int batch_launch(sevp_batch b, const BatchLaunch &l, hipStream_t st) {
  const unsigned char *base = b->tables.p;
  const int *start = reinterpret_cast<const int *>(base + l.start);
  switch (l.kind) {
    case BL_FILL:
      HIPC(launch_batch_fill(reinterpret_cast<const FillEnt *>(base + l.tab), start, l.n, l.nwg, st));
      break;
    case BL_MARCH:
      HIPC(launch_batch_march(b->kmarch, reinterpret_cast<const MarchEnt *>(base + l.tab), start, l.n,
                              l.nwg, st));
      break;
    case BL_PROD:
      HIPC(launch_batch_gemm_part(b->ks, b->mb, reinterpret_cast<const GemmPartArgs *>(base + l.tab),
                                  start, l.n, l.nwg, st));
      break;
    case BL_PROD64:
      HIPC(launch_batch_gemm_part64(b->ks, b->mb, reinterpret_cast<const GemmPart64Args *>(base + l.tab),
                                    start, l.n, l.nwg, st));
      break;
    case BL_RESID:
      HIPC(launch_batch_resid(reinterpret_cast<const ResidArgs *>(base + l.tab), start, l.n, l.nwg, st));
      break;
    case BL_HOM:
      HIPC(launch_batch_hom_resid(b->kmarch, reinterpret_cast<const HomResidEnt *>(base + l.tab), start,
                                  l.n, l.nwg, st));
      break;
    case BL_HALO:
      HIPC(launch_batch_axpy(reinterpret_cast<const AxpyEnt *>(base + l.tab), start, l.n, l.nwg,
                             nullptr, true, st));
      break;
    default:
      return fail(SEVP_E_STATE, "unknown batch launch");
  }
  return SEVP_OK;
}

// The grouped launches of one pass for nrhs[i] right-hand sides of problem i,
// and the graph that runs them.
static int batch_build(sevp_batch b, const int *nrhs) {
  b->drop_graph();
  b->launches.clear();
  const int np = (int)b->plans.size();
  std::vector<MixedPassOps> ops(np);
  int maxnb = 0, maxsteps = 0, maxm = 1;
  bool dbl = false;
  for (int i = 0; i < np; ++i) {
    sevp_plan p = b->plans[i];
    const long ns = (long)p->d.i2 * p->d.j2, nx = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
    mixed_pass_ops(full_ctx(p), nullptr, p->r64.p, ns, p->dx.p, nx, nrhs[i], ops[i]);
    dbl = ops[i].dbl;
    maxnb = std::max(maxnb, p->d.nb0);
    maxm = std::max(maxm, nrhs[i]);
    for (const auto &m : ops[i].fmarch) maxsteps = std::max(maxsteps, m.jf - m.js + 1);
    for (const auto &m : ops[i].bmarch) maxsteps = std::max(maxsteps, m.jf - m.js + 1);
  }
  // the narrowest lane width whose tile is positive for every block
  const int halo = std::max(maxsteps, 0) + 1;
  b->kmarch = (64 - 2 * halo > 0) ? 1 : (128 - 2 * halo > 0) ? 2 : (256 - 2 * halo > 0) ? 4 : -1;
  if (b->kmarch < 0) return fail(SEVP_E_ARG, "a block is too tall for the grouped march (more than 126 rows)");
  b->mb = maxm <= 16 ? 1 : 2;  // 32-row groups beyond 32 rows, as in a single launch
  TableImage img;
  {
    std::vector<FillEnt> f;
    std::vector<long> w;
    for (int i = 0; i < np; ++i) {
      sevp_plan p = b->plans[i];
      const long nx = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
      f.push_back(FillEnt{reinterpret_cast<unsigned int *>(p->dx.p), 2 * nx * nrhs[i]});
      w.push_back(cdiv(f.back().n, 256));
    }
    add_launch(img, b->launches, BL_FILL, f, w);
  }
  for (int t = 1; t <= maxnb; ++t) {
    std::vector<MarchEnt> me;
    std::vector<long> mw;
    std::vector<GemmPartArgs> g32;
    std::vector<GemmPart64Args> g64;
    std::vector<long> gw;
    for (int i = 0; i < np; ++i) {
      const int nb0 = b->plans[i]->d.nb0;
      if (t > nb0) continue;
      MarchEnt e;
      e.m = ops[i].fmarch[t - 1];
      e.ntiles = march_tiles(e.m, b->kmarch);
      if (e.ntiles <= 0) return fail(SEVP_E_STATE, "grouped march geometry");
      me.push_back(e);
      mw.push_back((long)e.ntiles * nrhs[i]);
      if (t == nb0) continue;
      if (dbl) {
        g64.push_back(ops[i].fprod64[t - 1]);
        gw.push_back(batch_gemm_part64_workgroups(g64.back()));
      } else {
        g32.push_back(ops[i].fprod[t - 1]);
        gw.push_back(batch_gemm_part_workgroups(g32.back()));
      }
    }
    add_launch(img, b->launches, BL_MARCH, me, mw);
    if (dbl) add_launch(img, b->launches, BL_PROD64, g64, gw);
    else add_launch(img, b->launches, BL_PROD, g32, gw);
  }
  {
    std::vector<ResidArgs> r;
    std::vector<long> w;
    for (int i = 0; i < np; ++i) {
      r.push_back(ops[i].mid);
      w.push_back(cdiv((long)nrhs[i] * b->plans[i]->d.i2, 256));
    }
    add_launch(img, b->launches, BL_RESID, r, w);
  }
  for (int t = 1; t <= maxnb; ++t) {
    std::vector<HomResidEnt> he;
    std::vector<long> hw;
    std::vector<GemmPartArgs> g32;
    std::vector<GemmPart64Args> g64;
    std::vector<long> gw;
    for (int i = 0; i < np; ++i) {
      const int nb0 = b->plans[i]->d.nb0;
      if (t > nb0) continue;
      const int nb = nb0 - t + 1;
      if (dbl) {
        g64.push_back(ops[i].bprod64[nb - 1]);
        gw.push_back(batch_gemm_part64_workgroups(g64.back()));
      } else {
        g32.push_back(ops[i].bprod[nb - 1]);
        gw.push_back(batch_gemm_part_workgroups(g32.back()));
      }
      HomResidEnt e;
      const long n = hom_resid_prepare(ops[i].bmarch[nb - 1], &ops[i].bresid[nb - 1], b->kmarch, e);
      if (n < 0) return fail(SEVP_E_STATE, "grouped march geometry");
      if (n == 0) continue;
      he.push_back(e);
      hw.push_back(n);
    }
    if (dbl) add_launch(img, b->launches, BL_PROD64, g64, gw);
    else add_launch(img, b->launches, BL_PROD, g32, gw);
    add_launch(img, b->launches, BL_HOM, he, hw);
  }
  {
    std::vector<AxpyEnt> h;
    std::vector<long> w;
    for (int i = 0; i < np; ++i) {
      sevp_plan p = b->plans[i];
      if (p->d.form != SEVP_FORM_FULL_PERIODIC) continue;
      const long nx = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
      h.push_back(AxpyEnt{p->dx.p, nx, nullptr, 0, p->d.i2, p->d.j2, nrhs[i], i});
      w.push_back(cdiv((long)nrhs[i] * p->d.j2, 256));
    }
    add_launch(img, b->launches, BL_HALO, h, w);
  }
  CHK(b->tables.alloc(img.bytes.size()));
  HIPC(hipMemcpy(b->tables.p, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice));
  // capture the pass
  HIPC(hipStreamSynchronize(b->stream));
  HIPC(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  int rc = SEVP_OK;
  for (const auto &l : b->launches)
    if ((rc = batch_launch(b, l, b->cap)) != SEVP_OK) break;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cap, &graph);
  if (rc != SEVP_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != SEVP_OK) return rc;
    HIPC(e);
  }
  const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPC(ei);
  b->key.assign(nrhs, nrhs + np);
  b->key_prec = b->plans[0]->d.precision;
  return SEVP_OK;
}

extern "C" {

int sevp_batch_create(const sevp_plan *plans, int nplans, sevp_batch *out) {
  if (!plans || !out || nplans < 1) return fail(SEVP_E_ARG, "null argument or no plans");
  *out = nullptr;
  int ks = 0, prec = -1;
  for (int i = 0; i < nplans; ++i) {
    sevp_plan p = plans[i];
    if (!p) return fail(SEVP_E_ARG, "null plan");
    if (p->d.form == SEVP_FORM_BIR) return fail(SEVP_E_ARG, "batched solves need full-width forms");
    if (p->d.precision != SEVP_PREC_FP32 && p->d.precision != SEVP_PREC_FP64_MFMA)
      return fail(SEVP_E_ARG, "batched solves need FP32 or FP64_MFMA corrections");
    if (prec >= 0 && p->d.precision != prec) return fail(SEVP_E_ARG, "plans of a batch must share the precision");
    prec = p->d.precision;
    if (!p->have_inv) return fail(SEVP_E_STATE, "influence matrices not built or loaded");
    if (p->d.i2 % 4 != 0) return fail(SEVP_E_ARG, "batched solves need I2 divisible by 4");
    // the grouped launches are the split-K pass's; the plans switch to it so
    // that their own solves give the batch's results
    if (p->pass_kind != SEVP_PASS_SPLITK) CHK(sevp_set_pass(p, SEVP_PASS_SPLITK));
    int k = 0;
    gemm_part_slices(p->d.i2, p->d.i2, &k);
    if (ks && k != ks)
      return fail(SEVP_E_ARG, "plans of a batch must share the product slice length "
                              "(all I2 <= 1536 or all I2 > 1536)");
    ks = k;
    for (int j = 0; j < i; ++j)
      if (plans[j] == p) return fail(SEVP_E_ARG, "a plan appears twice in the batch");
  }
  auto *b = new sevp_batch_s();
  b->plans.assign(plans, plans + nplans);
  b->ks = ks;
  auto bail = [&](int code) {
    delete b;
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamDefault) != hipSuccess ||
      hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SEVP_E_HIP, "hipStreamCreate failed"));
  if (int rc = b->act.alloc(nplans)) return bail(rc);
  if (hipHostMalloc((void **)&b->act_h, sizeof(int) * nplans, hipHostMallocDefault) != hipSuccess) {
    b->act_h = nullptr;
    return bail(fail(SEVP_E_NOMEM, "hipHostMalloc failed"));
  }
  *out = b;
  return SEVP_OK;
}

int sevp_batch_destroy(sevp_batch b) {
  delete b;
  return SEVP_OK;
}

int sevp_batch_solve_refined_dev(sevp_batch b, const float *const *s, double *const *x,
                                 const int *nrhs, double tol, int max_iter, double *rel_out,
                                 int *passes, sevp_stats *stats) {
  if (!b || !s || !x || !nrhs) return fail(SEVP_E_ARG, "null argument");
  const int np = (int)b->plans.size();
  size_t total_rhs = 0;
  for (int i = 0; i < np; ++i) {
    sevp_plan p = b->plans[i];
    if (!s[i] || !x[i]) return fail(SEVP_E_ARG, "null right-hand side or solution array");
    if (nrhs[i] < 1 || nrhs[i] > p->d.nrhs_max || nrhs[i] > 64) return fail(SEVP_E_ARG, "nrhs out of range");
    if (!p->have_inv) return fail(SEVP_E_STATE, "influence matrices not built or loaded");
    if (p->d.precision != b->plans[0]->d.precision)
      return fail(SEVP_E_STATE, "the plans of a batch no longer share one precision");
    if (p->d.precision != SEVP_PREC_FP32 && p->d.precision != SEVP_PREC_FP64_MFMA)
      return fail(SEVP_E_STATE, "batched solves need FP32 or FP64_MFMA corrections");
    total_rhs += nrhs[i];
  }
  auto t0 = std::chrono::steady_clock::now();
  // the plans' own queued work (their scratch arrays are shared with the batch)
  for (sevp_plan p : b->plans) {
    HIPC(hipStreamSynchronize(p->stream));
    for (auto st : p->gstreams) HIPC(hipStreamSynchronize(st));
  }
  // the pass is rebuilt when the RHS counts or the plans' (common) precision change
  if (b->key.size() != (size_t)np || !std::equal(b->key.begin(), b->key.end(), nrhs) ||
      b->key_prec != b->plans[0]->d.precision)
    CHK(batch_build(b, nrhs));
  // per call: X += DX entries of every problem, then the periodic ones
  const size_t nn = total_rhs * NCHUNK * 2;
  if (b->norm_part.n < nn) CHK(b->norm_part.alloc(nn));
  std::vector<AxpyEnt> ax, hx;
  std::vector<NormEnt> nx_;
  std::vector<int> as, hs, ns_;
  long anwg = 0, hnwg = 0, nnwg = 0;
  size_t poff = 0;
  for (int i = 0; i < np; ++i) {
    sevp_plan p = b->plans[i];
    const long nx = (long)(p->d.i2 + 2) * (p->d.j2 + 2);
    // every problem's residual and its norms (sevp_solve_refined_dev's
    // kernel, chunks and summation order)
    NormEnt q{p->al(), p->ab(), p->ac(), p->ar(), p->at(), s[i], x[i], (long)p->d.i2 * p->d.j2, nx,
              p->r64.p, b->norm_part.p + poff, p->d.i2, p->d.j2, nrhs[i], NCHUNK};
    nx_.push_back(q);
    ns_.push_back((int)nnwg);
    nnwg += (long)NCHUNK * nrhs[i];
    poff += (size_t)nrhs[i] * NCHUNK * 2;
    const AxpyEnt e{x[i], nx, p->dx.p, nx, p->d.i2, p->d.j2, nrhs[i], i};
    ax.push_back(e);
    as.push_back((int)anwg);
    anwg += cdiv((long)nrhs[i] * p->d.i2 * p->d.j2, 256);
    if (p->d.form == SEVP_FORM_FULL_PERIODIC) {
      hx.push_back(e);
      hs.push_back((int)hnwg);
      hnwg += cdiv((long)nrhs[i] * p->d.j2, 256);
    }
  }
  if (anwg > 0x7fffffffL || nnwg > 0x7fffffffL) return fail(SEVP_E_ARG, "batch too large");
  TableImage img;
  const size_t o_ax = img.put(ax.data(), ax.size()), o_as = img.put(as.data(), as.size());
  const size_t o_nx = img.put(nx_.data(), nx_.size()), o_ns = img.put(ns_.data(), ns_.size());
  size_t o_hx = 0, o_hs = 0;
  if (!hx.empty()) {
    o_hx = img.put(hx.data(), hx.size());
    o_hs = img.put(hs.data(), hs.size());
  }
  if (b->xtab.n < img.bytes.size()) CHK(b->xtab.alloc(img.bytes.size()));
  HIPC(hipMemcpy(b->xtab.p, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice));
  if (b->norm_h_n < nn) {
    if (b->norm_h) (void)hipHostFree(b->norm_h);
    b->norm_h = nullptr;
    b->norm_h_n = 0;
    if (hipHostMalloc((void **)&b->norm_h, nn * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      b->norm_h = nullptr;
      return fail(SEVP_E_NOMEM, "hipHostMalloc failed");
    }
    b->norm_h_n = nn;
  }
  std::vector<int> it(np, 0);
  std::vector<double> rel(total_rhs, 0.0), worst(np, 0.0);
  for (;;) {
    // every problem's residual and norm partials in one launch, one synchronisation
    hipLaunchKernelGGL(batch_resid_norm_kernel, dim3((unsigned)nnwg), dim3(256), 0, b->stream,
                       reinterpret_cast<const NormEnt *>(b->xtab.p + o_nx),
                       reinterpret_cast<const int *>(b->xtab.p + o_ns), np);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(b->norm_h, b->norm_part.p, nn * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
    bool any = false;
    size_t r0 = 0, off = 0;
    for (int i = 0; i < np; ++i) {
      worst[i] = 0.0;
      for (int r = 0; r < nrhs[i]; ++r) {
        double a = 0, c = 0;
        for (int k = 0; k < NCHUNK; ++k) {
          a += b->norm_h[off + ((size_t)r * NCHUNK + k) * 2];
          c += b->norm_h[off + ((size_t)r * NCHUNK + k) * 2 + 1];
        }
        rel[r0 + r] = c > 0 ? std::sqrt(a / c) : std::sqrt(a);
        worst[i] = std::max(worst[i], rel[r0 + r]);
      }
      r0 += nrhs[i];
      off += (size_t)nrhs[i] * NCHUNK * 2;
      b->act_h[i] = (worst[i] > tol && it[i] < max_iter) ? 1 : 0;
      any = any || b->act_h[i];
    }
    if (!any) break;
    HIPC(hipMemcpyAsync(b->act.p, b->act_h, sizeof(int) * np, hipMemcpyHostToDevice, b->stream));
    HIPC(hipGraphLaunch(b->exec, b->stream));
    const int *act = b->act.p;
    HIPC(launch_batch_axpy(reinterpret_cast<const AxpyEnt *>(b->xtab.p + o_ax),
                           reinterpret_cast<const int *>(b->xtab.p + o_as), (int)ax.size(), anwg, act,
                           false, b->stream));
    if (!hx.empty())
      HIPC(launch_batch_axpy(reinterpret_cast<const AxpyEnt *>(b->xtab.p + o_hx),
                             reinterpret_cast<const int *>(b->xtab.p + o_hs), (int)hx.size(), hnwg,
                             act, true, b->stream));
    for (int i = 0; i < np; ++i) it[i] += b->act_h[i];
  }
  double w = 0.0;
  int itmax = 0;
  for (int i = 0; i < np; ++i) {
    w = std::max(w, worst[i]);
    itmax = std::max(itmax, it[i]);
    if (passes) passes[i] = it[i];
  }
  if (rel_out) std::memcpy(rel_out, rel.data(), sizeof(double) * total_rhs);
  if (stats) {
    stats->iterations = itmax;
    stats->max_rel_resid = w;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  for (int i = 0; i < np; ++i)
    if (worst[i] > tol) return fail(SEVP_E_NOTCONV, "refinement did not reach the tolerance");
  return SEVP_OK;
}

}  // extern "C"

// --------------------------------------------- one-to-one reference entry points

// Copy rows 1..nrows of `ncols` columns from a column-major array of leading
// dimension ld into a compact device array of leading dimension ncols.
template <class T>
static int up2d(DBuf<T> &dst, const T *src, int ld, int ncols, int nrows) {
  CHK(dst.alloc((size_t)ncols * nrows));
  HIPC(hipMemcpy2D(dst.p, ncols * sizeof(T), src, (size_t)ld * sizeof(T), ncols * sizeof(T), nrows,
                   hipMemcpyHostToDevice));
  return SEVP_OK;
}

static hipStream_t aux_stream() { return nullptr; }

extern "C" int sevp_repbir(const float *ax, const float *ay, const float *bb, const float *cx,
                const float *cy, const double *rinv, const double *rinv1, double *dum0,
                double *dum1, double *dum2, const float *f, double *h, double *x, const int *ie,
                int i0, int i2, int m0, int m2, int nblk) {
  (void)dum0; (void)dum1; (void)dum2; (void)h; (void)m0;
  if (!ax || !ay || !bb || !cx || !cy || !rinv || !f || !x || !ie) return fail(SEVP_E_ARG, "null argument");
  if (m2 < 2 || m2 > 64 || nblk < 1 || i2 < m2 || i0 < m2 + 2) return fail(SEVP_E_ARG, "bad REPBIR shape");
  if (nblk > BIR_MAX_BLOCKS) return fail(SEVP_E_ARG, "REPBIR supports at most 4096 blocks (NBLK)");
  if (nblk > 1 && !rinv1) return fail(SEVP_E_ARG, "null RINV1");
  for (int b = 0; b < nblk; ++b)
    if (ie[b] < 2 || (b && ie[b] <= ie[b - 1])) return fail(SEVP_E_ARG, "IE must be increasing from 2");
  const int jr = ie[nblk - 1] - 1;  // last coefficient row used
  const int jx = ie[nblk - 1] + 1;  // last X row used
  DBuf<float> c[5], fd;
  const float *cs[5] = {ax, ay, bb, cx, cy};
  for (int k = 0; k < 5; ++k) CHK(up2d(c[k], cs[k], i2, m2, jr));
  CHK(up2d(fd, f, m2, m2, jr));
  DBuf<double> xd, rv, rv1, d0;
  CHK(up2d(xd, x, i0, m2 + 2, jx));
  const size_t nrv = (size_t)m2 * m2 * nblk, nrv1 = (size_t)m2 * m2 * (nblk > 1 ? nblk - 1 : 1);
  CHK(rv.alloc(nrv + RINV_PAD));
  HIPC(hipMemcpy(rv.p, rinv, nrv * sizeof(double), hipMemcpyHostToDevice));
  CHK(rv1.alloc(nrv1 + RINV_PAD));
  if (nblk > 1) HIPC(hipMemcpy(rv1.p, rinv1, nrv1 * sizeof(double), hipMemcpyHostToDevice));
  CHK(d0.alloc((size_t)m2 * (nblk > 1 ? nblk - 1 : 1)));
  DBuf<int> ied, recd;
  CHK(ied.alloc(nblk));
  HIPC(hipMemcpy(ied.p, ie, nblk * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> rt;
  int rt_fwd = 0;
  bir_record_table(ie, nblk, rt, rt_fwd);
  CHK(recd.alloc(rt.size()));
  HIPC(hipMemcpy(recd.p, rt.data(), rt.size() * sizeof(int), hipMemcpyHostToDevice));
  RepbirArgs a{};
  a.ax = c[0].p; a.ay = c[1].p; a.bb = c[2].p; a.cx = c[3].p; a.cy = c[4].p;
  a.f = fd.p; a.x = xd.p; a.rinv = rv.p; a.rinv1 = rv1.p; a.dum0 = d0.p; a.ie = ied.p;
  a.recs = recd.p; a.fwd_records = rt_fwd;
  a.i0 = m2 + 2; a.i2 = m2; a.m2 = m2; a.nblk = nblk;
  HIPC(launch_repbir(a, 1, aux_stream()));
  HIPC(hipMemcpy2D(x, (size_t)i0 * sizeof(double), xd.p, (m2 + 2) * sizeof(double),
                   (m2 + 2) * sizeof(double), jx, hipMemcpyDeviceToHost));
  return SEVP_OK;
}

static int full_once(const float *ax, const float *ay, const float *bb, const float *cx,
                     const float *cy, const double *rinv, const double *rinv1, const float *f,
                     double *x, const int *ie, int i0, int i2, int nblk, int periodic) {
  if (!ax || !ay || !bb || !cx || !cy || !rinv || !f || !x || !ie) return fail(SEVP_E_ARG, "null argument");
  if (i0 != i2 + 2 || i2 < 2 || nblk < 1) return fail(SEVP_E_ARG, "bad shape (I0 must be I2+2)");
  if (nblk > 1 && !rinv1) return fail(SEVP_E_ARG, "null RINV1");
  for (int b = 0; b < nblk; ++b)
    if (ie[b] < 2 || (b && ie[b] <= ie[b - 1])) return fail(SEVP_E_ARG, "IE must be increasing from 2");
  const int j2 = ie[nblk - 1] - 1;  // rows used: coefficients 1..J2, X 1..J2+2
  const size_t nc = (size_t)i2 * j2;
  DBuf<float> c[5], fd;
  const float *cs[5] = {ax, ay, bb, cx, cy};
  for (int k = 0; k < 5; ++k) {
    CHK(c[k].alloc(nc));
    HIPC(hipMemcpy(c[k].p, cs[k], nc * sizeof(float), hipMemcpyHostToDevice));
  }
  CHK(fd.alloc(nc));
  HIPC(hipMemcpy(fd.p, f, nc * sizeof(float), hipMemcpyHostToDevice));
  const size_t nx = (size_t)i0 * (j2 + 2);
  const int nb1 = nblk > 1 ? nblk - 1 : 1;
  DBuf<double> xd, rv, rv1, e, d0, hf, hc, g;
  CHK(xd.alloc(nx));
  HIPC(hipMemcpy(xd.p, x, nx * sizeof(double), hipMemcpyHostToDevice));
  CHK(rv.alloc((size_t)i2 * i2 * nblk));
  HIPC(hipMemcpy(rv.p, rinv, rv.bytes(), hipMemcpyHostToDevice));
  CHK(rv1.alloc((size_t)i2 * i2 * nb1));
  if (nblk > 1) HIPC(hipMemcpy(rv1.p, rinv1, rv1.bytes(), hipMemcpyHostToDevice));
  CHK(e.alloc(i2));
  CHK(d0.alloc((size_t)i2 * nb1));
  CHK(hf.alloc(i2));
  CHK(hc.alloc(i2));
  CHK(g.alloc(i2));
  FullCtx fc{};
  fc.ax = c[0].p; fc.ay = c[1].p; fc.bb = c[2].p; fc.cx = c[3].p; fc.cy = c[4].p;
  fc.i2 = i2; fc.j2 = j2; fc.periodic = periodic; fc.nb0 = nblk; fc.ie = ie;
  fc.rinv = rv.p; fc.rinv1 = rv1.p;
  fc.e = e.p; fc.dum0 = d0.p; fc.hfree = hf.p; fc.hcpl = hc.p; fc.g = g.p;
  CHK(run_full_pass(fc, fd.p, nullptr, (long)nc, xd.p, (long)nx, 1, aux_stream()));
  HIPC(hipMemcpy(x, xd.p, nx * sizeof(double), hipMemcpyDeviceToHost));
  return SEVP_OK;
}

extern "C" int sevp_rep1(const float *ax, const float *ay, const float *bb, const float *cx,
              const float *cy, const double *rinv, const double *rinv1, double *dum0,
              double *dum1, double *dum2, const float *f, double *h, double *x, const int *ie,
              int i0, int i2, int nblk) {
  (void)dum0; (void)dum1; (void)dum2; (void)h;
  return full_once(ax, ay, bb, cx, cy, rinv, rinv1, f, x, ie, i0, i2, nblk, 1);
}

extern "C" int sevp_rep(const float *ax, const float *ay, const float *bb, const float *cx,
             const float *cy, const double *rinv, const double *rinv1, double *dum0,
             double *dum1, double *dum2, const float *f, double *h, double *x, const int *ie,
             int i0, int i2, int nblk) {
  (void)dum0; (void)dum1; (void)dum2; (void)h;
  return full_once(ax, ay, bb, cx, cy, rinv, rinv1, f, x, ie, i0, i2, nblk, 0);
}

extern "C" int sevp_repbir_period(const float *ax, const float *ay, const float *bb, const float *cx,
                       const float *cy, const double *rinv, const double *rinv1, double *dum0,
                       double *dum1, double *dum2, const float *f, double *h, double *x,
                       const int *ie, int i0, int i2, int m0, int m2, int nblk) {
  if (m0 != i0 || m2 != i2)
    return fail(SEVP_E_ARG, "REPBIR_PERIOD is defined for M0=I0, M2=I2 only");
  return sevp_rep1(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, i2, nblk);
}

static int pre_once(const float *ax, const float *ay, const float *bb, const float *cx,
                    const float *cy, double *rinv, double *rinv1, const int *ie, int cld, int w,
                    int nblk, int periodic) {
  if (!ax || !ay || !bb || !cx || !cy || !rinv || !ie) return fail(SEVP_E_ARG, "null argument");
  if (w < 1 || nblk < 1 || cld < w) return fail(SEVP_E_ARG, "bad PRE shape");
  if (nblk > 1 && !rinv1) return fail(SEVP_E_ARG, "null RINV1");
  for (int b = 0; b < nblk; ++b)
    if (ie[b] < 2 || (b && ie[b] <= ie[b - 1])) return fail(SEVP_E_ARG, "IE must be increasing from 2");
  const int jr = ie[nblk - 1] - 1;
  DBuf<float> c[5];
  const float *cs[5] = {ax, ay, bb, cx, cy};
  for (int k = 0; k < 5; ++k) CHK(up2d(c[k], cs[k], cld, w, jr));
  const int nb1 = nblk > 1 ? nblk - 1 : 1;
  DBuf<double> rv, rv1, b1;
  DBuf<int> zp;
  CHK(rv.alloc((size_t)w * w * nblk));
  CHK(rv1.alloc((size_t)w * w * nb1));
  if (!matinv_uses_lds(w)) CHK(b1.alloc(matinv_blocked_scratch(w, 1)));
  CHK(zp.alloc(1));
  HIPC(hipMemset(zp.p, 0, sizeof(int)));
  HIPC(hipMemset(rv1.p, 0, rv1.bytes()));
  PreCtx pc{};
  pc.ax = c[0].p; pc.ay = c[1].p; pc.bb = c[2].p; pc.cx = c[3].p; pc.cy = c[4].p;
  pc.cld = w; pc.i2 = w; pc.j2 = jr; pc.w = w; pc.nmat = 1; pc.periodic = periodic;
  pc.map = ColMap{1, 0, 0};
  pc.rinv = rv.p; pc.rinv1 = rv1.p; pc.ie = ie; pc.nb0 = nblk; pc.zp = zp.p; pc.b1s = b1.p;
  CHK(run_pre(pc, aux_stream()));
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(rinv, rv.p, rv.bytes(), hipMemcpyDeviceToHost));
  if (nblk > 1) HIPC(hipMemcpy(rinv1, rv1.p, rv1.bytes(), hipMemcpyDeviceToHost));
  return SEVP_OK;
}

extern "C" int sevp_pre(const float *ax, const float *ay, const float *bb, const float *cx,
             const float *cy, double *rinv, double *rinv1, double *h, const int *ie, int i2,
             int m0, int m2, int nblk) {
  (void)h; (void)m0;
  return pre_once(ax, ay, bb, cx, cy, rinv, rinv1, ie, i2, m2, nblk, 0);
}

extern "C" int sevp_pre1(const float *ax, const float *ay, const float *bb, const float *cx,
              const float *cy, double *rinv, double *rinv1, double *h, const int *ie, int i0,
              int i2, int nblk) {
  (void)h;
  if (i0 != i2 + 2) return fail(SEVP_E_ARG, "I0 must be I2+2");
  return pre_once(ax, ay, bb, cx, cy, rinv, rinv1, ie, i2, i2, nblk, 1);
}

extern "C" int sevp_matinv(double *b, int n) {
  if (!b || n < 1) return fail(SEVP_E_ARG, "bad MATINV argument");
  DBuf<double> bd, b1;
  DBuf<int> zp;
  CHK(bd.alloc((size_t)n * n));
  if (!matinv_uses_lds(n)) CHK(b1.alloc(matinv_blocked_scratch(n, 1)));
  CHK(zp.alloc(1));
  HIPC(hipMemset(zp.p, 0, sizeof(int)));
  HIPC(hipMemcpy(bd.p, b, bd.bytes(), hipMemcpyHostToDevice));
  HIPC(launch_matinv(bd.p, (long)n * n, n, 1, zp.p, b1.p, aux_stream()));
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(b, bd.p, bd.bytes(), hipMemcpyDeviceToHost));
  return SEVP_OK;
}



// ------------------------------------------- MATINV with no allocation
// A device-pointer MATINV that allocates nothing: the blocked inversion's
// scratch (panel rows and multipliers, 2 * 32 * n doubles) and the zero-pivot
// counter live in `scratch`, a file-scope device array of a fixed size, one
// copy per GPU for the life of the process.  Host code never passes the
// array's name to a kernel (it is a host-side stand-in there): its device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and that address is what the kernels get.  Calls on one
// GPU share the array: calls that overlap in time on one GPU race.
namespace {
constexpr int kNoAllocMaxN = 4096;
constexpr size_t kNoAllocCounter = 2 * 32 * (size_t)kNoAllocMaxN;  // counter slot (in doubles)
constexpr size_t kNoAllocDoubles = kNoAllocCounter + 32;           // 2 MiB + 256 B
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kNoAllocDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

__global__ void copy_int_kernel(const int *src, int *dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = src[0];
}

// The calling thread's current device's copy of `scratch`.
static int noalloc_scratch(double **out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return fail(SEVP_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(SEVP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[dev] = (double *)ptr;
  }
  *out = g_noalloc_ptr[dev];
  return SEVP_OK;
}

extern "C" int sevp_matinv_dev_noalloc(double *b_dev, int n, int *zero_pivots_dev, void *hip_stream) {
  if (!b_dev || n < 1) return fail(SEVP_E_ARG, "bad MATINV argument");
  if (n > kNoAllocMaxN)
    return fail(SEVP_E_ARG, "sevp_matinv_dev_noalloc: order " + std::to_string(n) +
                                " exceeds its fixed scratch (order <= 4096)");
  double *base = nullptr;
  CHK(noalloc_scratch(&base));
  hipStream_t st = (hipStream_t)hip_stream;
  int *zp = (int *)(base + kNoAllocCounter);
  CHK(zero_async(zp, sizeof(int), st));  // a kernel, not a memset node
  HIPC(launch_matinv(b_dev, (long)n * n, n, 1, zp, matinv_uses_lds(n) ? nullptr : base, st));
  if (zero_pivots_dev) {
    hipLaunchKernelGGL(copy_int_kernel, dim3(1), dim3(64), 0, st, (const int *)zp, zero_pivots_dev);
    HIPC(hipGetLastError());
  }
  return SEVP_OK;
}

// ------------------------------------------------------------ Krylov solvers
// Drivers of the device BiCGSTAB, P-CSI and Lanczos (sevp_krylov.hip) with
// the control flow of MAIN/template/GLOBAL/src/sevp.f90:443-694; the
// scalar recurrences of BiCGSTAB and P-CSI run on the device, the host
// checks the stop flag between batches of iterations.

static double kry_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int kry_check_opts(const sevp_krylov_opts &o) {
  if (o.max_iter < 0 || o.restart < 1) return fail(SEVP_E_ARG, "max_iter must be >= 0, restart >= 1");
  if (!(o.tol >= 0.0) || !(o.rtol >= 0.0 && o.rtol < 1.0))
    return fail(SEVP_E_ARG, "tol must be >= 0 and rtol in [0, 1)");
  return SEVP_OK;
}

static int run_bicgstab(const KryOp &op, KryWork &wk, const float *b, double *x,
                        const sevp_krylov_opts &o, sevp_krylov_stats *stats, hipStream_t s) {
  CHK(kry_check_opts(o));
  const double t0 = kry_now();
  const long n = (long)(op.i2 + 2) * (op.j2 + 2);
  CHK(wk.ensure(op.i2, op.j2));
  double *base = wk.vec.p;
  KryVecs q{};
  q.b = b;
  q.x = x;
  q.r = base;
  q.rh = base + n;
  q.p[0] = base + 2 * n;
  q.p[1] = base + 3 * n;
  q.v[0] = base + 4 * n;
  q.v[1] = base + 5 * n;
  q.s = base + 6 * n;
  q.t = base + 7 * n;
  HIPC(kry_fill(base, 8 * n, 0.0, s));
  KryState h{};
  h.tp = 1.0;
  h.tol = o.tol;
  h.rtol = o.rtol;
  h.restart = o.restart;
  h.max_iter = o.max_iter;
  HIPC(hipMemcpyAsync(wk.st.p, &h, sizeof h, hipMemcpyHostToDevice, s));
  if (op.periodic) HIPC(kry_bnd(x, op.i2, op.j2, s));
  const int fast = !o.ref_order;
  bool early = false;
  while (std::sqrt(h.tp) > o.tol && h.iter < o.max_iter) {
    HIPC(kry_bicg_restart(op, q, wk.st.p, wk.part.p, fast, s));
    // the device stops by the next multiple of `restart` at the latest
    const int limit = o.restart - h.iter % o.restart;
    int launched = 0;
    do {
      if (launched >= limit)
        return fail(SEVP_E_HIP, "BiCGSTAB: the device did not reach its stop condition");
      int batch = limit - launched;
      if (batch > 32) batch = 32;
      for (int k = 0; k < batch; ++k) HIPC(kry_bicg_iter(op, q, wk.st.p, wk.part.p, fast, s));
      launched += batch;
      HIPC(hipMemcpyAsync(&h, wk.st.p, sizeof h, hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
    } while (!h.stop);
    if (h.stop == 2) {  // converged at the restart: the reference returns at once
      early = true;
      break;
    }
  }
  if (!early && op.periodic) HIPC(kry_bnd(x, op.i2, op.j2, s));
  HIPC(hipStreamSynchronize(s));
  if (stats) {
    stats->iterations = h.iter;
    stats->residual = std::sqrt(h.tp);
    stats->converged = stats->residual <= o.tol;
    stats->seconds = kry_now() - t0;
  }
  return SEVP_OK;
}

static int run_pcsi(const KryOp &op, KryWork &wk, const float *b, double *x, double mineig,
                    double maxeig, const sevp_krylov_opts &o, sevp_krylov_stats *stats,
                    hipStream_t s) {
  CHK(kry_check_opts(o));
  if (!(maxeig != mineig)) return fail(SEVP_E_ARG, "P-CSI needs mineig != maxeig");
  const double t0 = kry_now();
  const long n = (long)(op.i2 + 2) * (op.j2 + 2);
  CHK(wk.ensure(op.i2, op.j2));
  // the OMEGA recurrence depends on no data: evaluated here, in the
  // reference's order (p_csi, :580-601)
  const double alpha = 2.0 / (maxeig - mineig);
  const double beta = (maxeig + mineig) / (maxeig - mineig);
  const double csy = beta / alpha;
  double omega = 2.0 / csy;
  const int mi = o.max_iter;
  std::vector<double> tab(2 * (size_t)(mi > 0 ? mi : 1));
  for (int k = 0; k < mi; ++k) {
    omega = 1.0 / (csy - omega / (4.0 * alpha * alpha));
    tab[k] = omega;
    tab[mi + k] = csy * omega - 1.0;
  }
  const double c0 = 1.0 / csy;
  if (wk.omega.n < tab.size()) CHK(wk.omega.alloc(tab.size()));
  HIPC(hipMemcpyAsync(wk.omega.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s));
  double *r = wk.vec.p, *q = wk.vec.p + n;
  HIPC(kry_fill(wk.vec.p, 2 * n, 0.0, s));
  KryState h{};
  h.tp = 1.0;
  h.tol = o.tol;
  h.max_iter = mi;
  h.restart = 1;
  h.stop = (std::sqrt(h.tp) > o.tol && 0 < mi) ? 0 : 1;
  HIPC(hipMemcpyAsync(wk.st.p, &h, sizeof h, hipMemcpyHostToDevice, s));
  const int fast = !o.ref_order;
  if (op.periodic) HIPC(kry_bnd(x, op.i2, op.j2, s));
  HIPC(kry_pcsi_r(op, b, x, r, wk.st.p, wk.part.p, 0, fast, s));
  HIPC(kry_pcsi_qx(op.i2, op.j2, op.periodic, r, q, x, wk.omega.p, wk.omega.p + mi, c0, 1, wk.st.p, s));
  HIPC(kry_pcsi_r(op, b, x, r, wk.st.p, wk.part.p, 0, fast, s));
  int launched = 0;
  while (!h.stop) {
    if (launched >= mi) return fail(SEVP_E_HIP, "P-CSI: the device did not reach its stop condition");
    int batch = mi - launched;
    if (batch > 32) batch = 32;
    launched += batch;
    for (int k = 0; k < batch; ++k) {
      HIPC(kry_pcsi_qx(op.i2, op.j2, op.periodic, r, q, x, wk.omega.p, wk.omega.p + mi, c0, 0,
                       wk.st.p, s));
      HIPC(kry_pcsi_r(op, b, x, r, wk.st.p, wk.part.p, 1, fast, s));
    }
    HIPC(hipMemcpyAsync(&h, wk.st.p, sizeof h, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
  }
  if (op.periodic) HIPC(kry_bnd(x, op.i2, op.j2, s));
  HIPC(hipStreamSynchronize(s));
  if (stats) {
    stats->iterations = h.iter;
    stats->residual = std::sqrt(h.tp);
    stats->converged = stats->residual <= o.tol;
    stats->seconds = kry_now() - t0;
  }
  return SEVP_OK;
}

static int kry_read_dot(KryWork &wk, int i2, int j2, const double *a, const double *b, int fast,
                        hipStream_t s, double &out) {
  HIPC(kry_dot(i2, j2, a, b, wk.dot.p, wk.part.p, fast, s));
  HIPC(hipMemcpyAsync(&out, wk.dot.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  return SEVP_OK;
}

static int run_lanczos(const KryOp &op, KryWork &wk, int m, int fast, double *mineig,
                       double *maxeig, double *alpha_out, double *sigma_out, hipStream_t s) {
  if (m < 1) return fail(SEVP_E_ARG, "Lanczos needs m >= 1");
  const int i2 = op.i2, j2 = op.j2;
  const long n = (long)(i2 + 2) * (j2 + 2);
  CHK(wk.ensure(i2, j2));
  double *v = wk.vec.p, *vp = v + n, *w = v + 2 * n, *wp = v + 3 * n, *sv = v + 4 * n, *qv = v + 5 * n;
  HIPC(kry_fill(wk.vec.p, 6 * n, 0.0, s));
  KryState h{};
  HIPC(hipMemcpyAsync(wk.st.p, &h, sizeof h, hipMemcpyHostToDevice, s));
  std::vector<double> alpha(m, 0.0), beta(m + 1, 0.0), sigma(m + 1, 0.0);
  double tp = 0.0;
  HIPC(kry_fill(v, n, 1.0, s));
  CHK(kry_read_dot(wk, i2, j2, v, v, fast, s, tp));
  const double c = 1.0 / std::sqrt(tp);
  HIPC(kry_fill(v, n, c * 1.0, s));
  HIPC(kry_fill(w, n, c * 1.0, s));
  for (int it = 1; it <= m; ++it) {
    HIPC(kry_spmv(op, v, sv, 0, s));
    CHK(kry_read_dot(wk, i2, j2, sv, w, fast, s, tp));
    alpha[it - 1] = tp;
    HIPC(kry_lincomb(vp, sv, alpha[it - 1], v, beta[it - 1], vp, n, s));
    std::swap(v, vp);
    HIPC(kry_spmv(op, w, qv, 1, s));
    HIPC(kry_lincomb(wp, qv, alpha[it - 1], w, sigma[it - 1], wp, n, s));
    std::swap(w, wp);
    CHK(kry_read_dot(wk, i2, j2, v, w, fast, s, tp));
    sigma[it] = std::sqrt(std::fabs(tp));
    beta[it] = std::copysign(sigma[it], tp);
    HIPC(kry_scale(v, sigma[it], w, beta[it], n, s));
  }
  HIPC(kry_spmv(op, v, sv, 0, s));
  CHK(kry_read_dot(wk, i2, j2, sv, w, fast, s, tp));
  alpha[m - 1] = tp;
  // FIGI's symmetric tridiagonal: D = ALPHA, E(I) = SQRT(SIGMA(I)*SIGMA(I));
  // E(1), which FIGI leaves unset and the bound below reads, is zero
  std::vector<double> d(m), e(m);
  for (int i = 0; i < m; ++i) {
    d[i] = alpha[i];
    e[i] = i > 0 ? std::sqrt(sigma[i] * sigma[i]) : 0.0;
  }
  double mx = d[0] + e[0];
  for (int i = 1; i < m; ++i) {
    const double g = d[i] + e[i] + e[i - 1];
    mx = g > mx ? g : mx;
  }
  if (maxeig) *maxeig = mx;
  if (mineig) *mineig = kry_tridiag_min_eig(d.data(), e.data(), m);
  if (alpha_out) std::memcpy(alpha_out, alpha.data(), sizeof(double) * m);
  if (sigma_out) std::memcpy(sigma_out, sigma.data(), sizeof(double) * (m + 1));
  return SEVP_OK;
}

static KryOp plan_kry_op(sevp_plan p, int periodic) {
  return KryOp{p->al(), p->ab(), p->ac(), p->ar(), p->at(), p->d.i2, p->d.j2, periodic ? 1 : 0};
}

static int kry_plan_work(sevp_plan p) {
  if (!p->kry) p->kry.reset(new KryWork());
  return p->kry->ensure(p->d.i2, p->d.j2);
}

// A temporary device copy of host coefficients for the one-to-one calls.
struct KryHostOp {
  DBuf<float> coef;
  KryOp op{};
  int load(const float *al, const float *ab, const float *ac, const float *ar, const float *at,
           int i2, int j2, int periodic) {
    if (!al || !ab || !ac || !ar || !at || i2 < 2 || j2 < 2) return fail(SEVP_E_ARG, "bad operator");
    const size_t m = (size_t)i2 * j2;
    CHK(coef.alloc(5 * m));
    const float *src[5] = {al, ab, ac, ar, at};
    for (int k = 0; k < 5; ++k)
      HIPC(hipMemcpy(coef.p + k * m, src[k], m * sizeof(float), hipMemcpyHostToDevice));
    op = KryOp{coef.p, coef.p + m, coef.p + 2 * m, coef.p + 3 * m, coef.p + 4 * m, i2, j2, periodic ? 1 : 0};
    return SEVP_OK;
  }
};

extern "C" {

int sevp_krylov_defaults(sevp_krylov_opts *o, int solver) {
  if (!o) return fail(SEVP_E_ARG, "null argument");
  o->periodic = 1;
  o->ref_order = 0;
  o->max_iter = 1000;
  o->restart = 100;
  o->tol = solver == SEVP_KRYLOV_PCSI ? (double)1e-7f : (double)1e-8f;
  o->rtol = (double)1e-8f;
  return SEVP_OK;
}

int sevp_bicgstab_dev(sevp_plan p, const float *b, double *x, const sevp_krylov_opts *opts,
                      sevp_krylov_stats *stats) {
  if (!p || !b || !x) return fail(SEVP_E_ARG, "null argument");
  sevp_krylov_opts o;
  sevp_krylov_defaults(&o, SEVP_KRYLOV_BICGSTAB);
  if (opts) o = *opts;
  CHK(kry_plan_work(p));
  return run_bicgstab(plan_kry_op(p, o.periodic), *p->kry, b, x, o, stats, p->stream);
}

int sevp_pcsi_dev(sevp_plan p, double mineig, double maxeig, const float *b, double *x,
                  const sevp_krylov_opts *opts, sevp_krylov_stats *stats) {
  if (!p || !b || !x) return fail(SEVP_E_ARG, "null argument");
  sevp_krylov_opts o;
  sevp_krylov_defaults(&o, SEVP_KRYLOV_PCSI);
  if (opts) o = *opts;
  CHK(kry_plan_work(p));
  return run_pcsi(plan_kry_op(p, o.periodic), *p->kry, b, x, mineig, maxeig, o, stats, p->stream);
}

int sevp_lanczos_dev(sevp_plan p, int m, const sevp_krylov_opts *opts, double *mineig,
                     double *maxeig, double *alpha, double *sigma) {
  if (!p) return fail(SEVP_E_ARG, "null argument");
  CHK(kry_plan_work(p));
  return run_lanczos(plan_kry_op(p, 0), *p->kry, m, opts ? !opts->ref_order : 1, mineig, maxeig,
                     alpha, sigma, p->stream);
}

int sevp_p_bicgstab_le(const float *al, const float *ab, const float *ac, const float *ar,
                       const float *at, const float *b, double *x, int i2, int j2,
                       const sevp_krylov_opts *opts, sevp_krylov_stats *stats) {
  if (!b || !x) return fail(SEVP_E_ARG, "null argument");
  sevp_krylov_opts o;
  sevp_krylov_defaults(&o, SEVP_KRYLOV_BICGSTAB);
  if (opts) o = *opts;
  KryHostOp h;
  CHK(h.load(al, ab, ac, ar, at, i2, j2, o.periodic));
  KryWork wk;
  DBuf<float> bd;
  DBuf<double> xd;
  const size_t nx = (size_t)(i2 + 2) * (j2 + 2);
  CHK(bd.alloc((size_t)i2 * j2));
  CHK(xd.alloc(nx));
  HIPC(hipMemcpy(bd.p, b, bd.bytes(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(xd.p, x, xd.bytes(), hipMemcpyHostToDevice));
  CHK(run_bicgstab(h.op, wk, bd.p, xd.p, o, stats, aux_stream()));
  HIPC(hipMemcpy(x, xd.p, xd.bytes(), hipMemcpyDeviceToHost));
  return SEVP_OK;
}

int sevp_p_csi(const float *al, const float *ab, const float *ac, const float *ar, const float *at,
               double mineig, double maxeig, const float *b, double *x, int i2, int j2,
               const sevp_krylov_opts *opts, sevp_krylov_stats *stats) {
  if (!b || !x) return fail(SEVP_E_ARG, "null argument");
  sevp_krylov_opts o;
  sevp_krylov_defaults(&o, SEVP_KRYLOV_PCSI);
  if (opts) o = *opts;
  KryHostOp h;
  CHK(h.load(al, ab, ac, ar, at, i2, j2, o.periodic));
  KryWork wk;
  DBuf<float> bd;
  DBuf<double> xd;
  const size_t nx = (size_t)(i2 + 2) * (j2 + 2);
  CHK(bd.alloc((size_t)i2 * j2));
  CHK(xd.alloc(nx));
  HIPC(hipMemcpy(bd.p, b, bd.bytes(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(xd.p, x, xd.bytes(), hipMemcpyHostToDevice));
  CHK(run_pcsi(h.op, wk, bd.p, xd.p, mineig, maxeig, o, stats, aux_stream()));
  HIPC(hipMemcpy(x, xd.p, xd.bytes(), hipMemcpyDeviceToHost));
  return SEVP_OK;
}

int sevp_lanczos(const float *al, const float *ab, const float *ac, const float *ar,
                 const float *at, int i2, int j2, int m, const sevp_krylov_opts *opts,
                 double *mineig, double *maxeig, double *alpha, double *sigma) {
  KryHostOp h;
  CHK(h.load(al, ab, ac, ar, at, i2, j2, 0));
  KryWork wk;
  return run_lanczos(h.op, wk, m, opts ? !opts->ref_order : 1, mineig, maxeig, alpha, sigma,
                     aux_stream());
}

}  // extern "C"
