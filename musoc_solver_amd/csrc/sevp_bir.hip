// Plain FP64 strip SEVP (REPBIR) and the BIR multiplicative strip sweeps.
//
// One workgroup owns one right-hand side and runs the whole chain the
// reference runs on one CPU core: nsweep sweeps x NBIR overlapping strips x
// (forward + backward block march).  The chain is a true data dependency
// (each strip reads the halo columns and interior guess written by the one
// before it, and each block's march starts from the corrected last row of the
// block below), so the only parallelism inside one solve is the strip's
// columns: lane l of the marching wavefront marches strip column l+1.
// Independent right-hand sides run as independent workgroups.
//
// A row step is a dependent FP64 chain (two wave shifts, four products and
// the ordered differences, a correctly rounded division), so the workgroup's
// four wavefronts split the work to keep memory off that chain:
//  * wave 0, the marcher, reads every row operand from LDS, keeps the live
//    rows in registers, exchanges the stencil's east/west neighbours with DPP
//    wave shifts and writes its results to LDS.  Its only global accesses are
//    the influence-matrix columns (column n of RINV/RINV1 for lane n), loaded
//    into registers a whole block before the closure that uses them.
//  * waves 1-3, the stagers, stream each pass's operands into LDS with
//    global_load_lds DMAs (no registers, no waits while they fly) and store
//    the march's results back to X and DUM0, each taking every third record.
// A pass is a sequence of records, one per march row or closure, in the order
// the marcher consumes them (the host-built record table, sevp_records.h).
// Records travel in chunks of R through three input buffers: while the
// marcher works on chunk c the stagers fill chunk c+1 and drain chunk c-1,
// and one workgroup barrier ends each chunk on all four wavefronts, so they
// agree on the barrier count by construction (1 + ceil(records / R) per
// pass, plus a pass-end barrier).
//
// A record is 8 slots of WMAX dwords; lane l's value of slot s is dword
// s*WMAX + l.  A double per lane takes two slots: low words, then high words.
//   forward header       slots 0-1 X(., 1), 2-3 X(., 2)
//   march row, closure   slots 0-5 AX, AY, BB, CX, CY, F of the row (REAL(4));
//                        slots 6-7 the halo of a forward march row (lane 0
//                        X(1, J+1), lane W-1 X(W+2, J+1)), the forward
//                        closure's north row X(., IE+1), the backward
//                        closure's X(., JR) or the backward march row's old
//                        X(., J+2)
//   backward closure B   slots 0-1 the restored row X(., IE), 2-3 X(., JS+1),
//                        4 CY(., JS-1), 6-7 X(., IE+1) of the top block
// plus 4 dwords of halo for the closure records (X(1, row) and X(W+2, row)).
//
// Arithmetic follows MAIN/template/GLOBAL/src/sevp.f90:218-328 operation for
// operation (see sevp_dev.h), so results are bitwise equal to the reference.
#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

// Stager wavefronts per workgroup: a single wavefront's stream of DMAs and
// stores is bound by how many vector memory operations it may have in flight,
// so the records of each chunk are dealt round-robin to NSTG stagers.
constexpr int NSTG = 3;
constexpr int BIR_THREADS = 64 * (1 + NSTG);

template <int WMAX>
struct alignas(16) Stage {
  static constexpr int R = WMAX == 32 ? 32 : 16;  // records per chunk
  static constexpr int NIN = 3;                     // input buffers
  static constexpr int REC = 8 * WMAX;              // dwords per record
  static constexpr int NF = 6 * WMAX / 64;          // DMAs of slots 0-5
  static constexpr int ND = 2 * WMAX / 64;          // DMAs of one double per lane
  uint32_t in[NIN][R][REC];
  uint32_t halo[NIN][R][4];
  double out[2][R][2][WMAX];  // up to two results per record
  double e[WMAX];             // the marcher's residual row, read back as broadcasts
  int desc[4][R][4];          // record-table entries of four chunks
};

// One strip: strip column c (0-based) of the coefficient arrays is physical
// column col(c); X(i, j) = xv[(i-1) + (j-1)*ldx] with the halo columns at
// i = 1 and i = w+2.
struct Strip {
  const float *arr[6];  // AX, AY, BB, CX, CY, F
  int ld[6];
  int il, inc, seam;
  double *xv;
  int ldx;
  const double *rinv, *rinv1;
  double *dum0;  // (w, nblk-1)
  const int *ie;
  int nblk, w;
  __device__ int col(int c) const {
    c = c < w ? c : 0;
    return (seam && c >= inc) ? c - inc : il - 1 + c;
  }
  __device__ double *xrow(int j) const { return xv + 1 + (long)(j - 1) * ldx; }  // &X(2, j)
};

// IE(k+1) from the LDS copy, as a wave-uniform (scalar) value.
__device__ __forceinline__ int ie_at(const int *ie, int k) { return __builtin_amdgcn_readfirstlane(ie[k]); }

__device__ __forceinline__ int fwd_jlast(const int *ie, int nblk) {
  const int ieb = ie_at(ie, nblk - 1), js = nblk > 1 ? ie_at(ie, nblk - 2) : 1;
  return ieb - 2 >= js ? ieb - 2 : js - 1;
}

// Chunk hand-over.  The marcher's outstanding global loads are influence
// columns for a later closure, so its barrier waits for LDS only; the
// stager's also waits for its DMAs and stores.
__device__ __forceinline__ void bar_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
__device__ __forceinline__ void bar_all() {
  asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
__device__ __forceinline__ void wait_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Lane t's dword of a 64-dword run at dst (wave-uniform) comes from src.
__device__ __forceinline__ void dma(const void *src, uint32_t *dst) {
  __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)(src),
                                   (__attribute__((address_space(3))) void *)(dst), 4, 0, 0);
}

// Column n of a w x w matrix into registers from one base: WX entries when
// the kernel is specialised for strips of width WX, else WMAX (the entries
// past w belong to the next column or the allocation's padding and are never
// used).
template <int WMAX, int WX>
__device__ __forceinline__ void load_col(double (&r)[WMAX], const double *mat, int w, int n) {
  const double *c = mat + (long)n * w;
#pragma unroll
  for (int m = 0; m < (WX > 0 ? WX : WMAX); ++m) r[m] = c[m];
}

// sum_{m=1..w} e(m) * R(m, n), sequential from zero (lane n holds column n).
// e goes through LDS (eb, the marcher's own buffer) and comes back as
// broadcast reads of two entries.  Specialised for strips of width WX, the
// reads are all issued first and the sum is one straight chain; otherwise
// they go a group of eight entries at a time (unconditional: eb holds WMAX
// entries) and each term is guarded by m < w.
template <int WMAX, int WX>
__device__ __forceinline__ double dot_col(double e, const double (&r)[WMAX], int w, double *eb, int lane) {
  eb[lane] = e;
  const double2 *e2 = reinterpret_cast<const double2 *>(eb);
  double s = 0.0;
  if constexpr (WX > 0) {
    double2 em[(WX + 1) / 2];
#pragma unroll
    for (int i = 0; i < (WX + 1) / 2; ++i) em[i] = e2[i];
#pragma unroll
    for (int m = 0; m < WX; ++m) s = s + ((m & 1) ? em[m >> 1].y : em[m >> 1].x) * r[m];
  } else {
#pragma unroll
    for (int g = 0; g < WMAX / 8; ++g) {
      if (8 * g >= w) break;
      double2 cur[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) cur[i] = e2[4 * g + i];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = 8 * g + 2 * i;
        if (m < w) {
          s = s + cur[i].x * r[m];
          if (m + 1 < w) s = s + cur[i].y * r[m + 1];
        }
      }
    }
  }
  return s;
}

// ------------------------------------------------------------ record order
// One record of a pass (an entry of bir_record_table): its kind, block nb,
// row j and the block's IE and JS, all wave-uniform.
struct RecInfo {
  int kind, j, nb, js, ieb;
};

// ------------------------------------------------------------------ stager
template <int WMAX>
struct StagerLane {
  const float *fp[Stage<WMAX>::NF];  // this lane's source at row 1, per slot DMA
  int fld[Stage<WMAX>::NF];
  int doff[Stage<WMAX>::ND];  // byte offset of this lane's dword in a row of doubles
  int hoff[Stage<WMAX>::ND];  // the same for the halo pair: X(1, .) or, at lane w-1, X(w+2, .)
  __device__ void init(const Strip &s, int t) {
#pragma unroll
    for (int d = 0; d < Stage<WMAX>::NF; ++d) {
      const int u = 64 * d + t, slot = u / WMAX, c = u % WMAX;
      const float *a = s.arr[0];
      int l = s.ld[0];
#pragma unroll
      for (int k = 1; k < 6; ++k)
        if (slot == k) { a = s.arr[k]; l = s.ld[k]; }
      fp[d] = a + s.col(c);
      fld[d] = l;
    }
#pragma unroll
    for (int e = 0; e < Stage<WMAX>::ND; ++e) {
      const int u = 64 * e + t, half = u / WMAX, k = min(u % WMAX, s.w - 1);
      doff[e] = (2 * k + half) * 4;
      hoff[e] = (k == s.w - 1 ? (s.w + 1) * 8 : 0) + half * 4;
    }
  }
};

template <int WMAX>
__device__ __forceinline__ void fill_record(Stage<WMAX> &S, const Strip &s, const StagerLane<WMAX> &L,
                                            int b, int p, const RecInfo &k, int t) {
  using G = Stage<WMAX>;
  uint32_t *rec = S.in[b][p];
  auto frows = [&](int j) {
#pragma unroll
    for (int d = 0; d < G::NF; ++d) dma(L.fp[d] + (long)(j - 1) * L.fld[d], rec + 64 * d);
  };
  auto drow = [&](uint32_t *dst, const double *row) {
#pragma unroll
    for (int e = 0; e < G::ND; ++e) dma((const char *)row + L.doff[e], dst + 64 * e);
  };
  auto halo = [&](int j) {
    if (t < 4) dma((const char *)(s.xv + (long)(j - 1) * s.ldx + (t < 2 ? 0 : s.w + 1)) + 4 * (t & 1),
                   S.halo[b][p]);
  };
  const double *xr = (k.nb != s.nblk) ? s.dum0 + (long)(k.nb - 1) * s.w : s.xrow(k.ieb);
  switch (k.kind) {
    case BIR_HEAD:
      drow(rec, s.xrow(1));
      drow(rec + 2 * WMAX, s.xrow(2));
      break;
    case BIR_ROW: {
      const char *h = (const char *)(s.xv + (long)k.j * s.ldx);  // &X(1, j+1)
      frows(k.j);
#pragma unroll
      for (int e = 0; e < G::ND; ++e) dma(h + L.hoff[e], rec + 6 * WMAX + 64 * e);
      break;
    }
    case BIR_CLOS:
      frows(k.j);
      halo(k.j + 1);
      drow(rec + 6 * WMAX, s.xrow(k.ieb + 1));
      break;
    case BIR_A:
      frows(k.ieb - 1);
      halo(k.ieb);
      drow(rec + 6 * WMAX, s.xrow(k.ieb - 1));
      break;
    case BIR_B: {
      constexpr int d4 = 4 * WMAX / 64;  // the DMA holding slot 4
      drow(rec, xr);
      // with a one-row block X(., JS+1) is row IE, read after its restore
      drow(rec + 2 * WMAX, k.js + 1 == k.ieb ? xr : s.xrow(k.js + 1));
      if (k.nb != 1) dma(L.fp[d4] + (long)(k.js - 2) * L.fld[d4], rec + 64 * d4);
      if (k.nb == s.nblk) drow(rec + 6 * WMAX, s.xrow(k.ieb + 1));
      break;
    }
    default:  // BIR_BROW: the last row reads row IE after its restore
      frows(k.j);
      drow(rec + 6 * WMAX, k.j + 2 == k.ieb ? xr : s.xrow(k.j + 2));
      break;
  }
}

// The record's results to memory.  The destination rows are wave-uniform and
// chosen first; only the stores themselves are limited to the strip's lanes.
template <int WMAX>
__device__ __forceinline__ void drain_record(Stage<WMAX> &S, const Strip &s, int b, int ob, int p,
                                             const RecInfo &k, int t, double o0) {
  double *d0 = nullptr, *d1 = nullptr, *dr = nullptr;
  if (k.kind == BIR_ROW) {  // row IE(nb) of a closing block waits for its correction
    if (!(k.nb < s.nblk && k.j == k.ieb - 2)) d0 = s.xrow(k.j + 2);
  } else if (k.kind == BIR_CLOS) {
    d0 = s.xrow(k.ieb);
    d1 = s.dum0 + (long)(k.nb - 1) * s.w;
  } else if (k.kind == BIR_B) {
    // restore row IE from DUM0 (still in the record); a block with march
    // rows overwrites it with its last row, so only a one-row block keeps it
    if (k.nb != s.nblk && k.ieb - 2 < k.js) dr = s.xrow(k.ieb);
    d0 = s.xrow(k.js + 1);
  } else if (k.kind == BIR_BROW) {
    d0 = s.xrow(k.j + 2);
  }
  const bool act = t < s.w;
  if (dr && act) dr[t] = __hiloint2double((int)S.in[b][p][WMAX + t], (int)S.in[b][p][t]);
  if (d0 && act) d0[t] = o0;
  if (d1 && act) d1[t] = S.out[ob][p][1][t];
}

// One pass for stager `sid`: it fills and drains the records p of each chunk
// with p % NSTG == sid.  The record-table entries of each chunk come into an
// LDS ring by DMA two chunks ahead: stager 0 issues them, except the first two
// chunks' entries, which every stager loads for itself.  Records of one pass read only rows
// that no record of the same pass writes, and write distinct rows, so the
// stagers need no ordering among themselves inside a pass; the barrier after
// the last drain makes every result visible to the next pass's DMAs.
template <int WMAX>
__device__ void stage_pass(Stage<WMAX> &S, const Strip &s, const StagerLane<WMAX> &L, int t, int sid,
                           bool fwd, const int *recs) {
  constexpr int R = Stage<WMAX>::R;
  const int T = fwd ? 1 + fwd_jlast(s.ie, s.nblk) : s.nblk + ie_at(s.ie, s.nblk - 1) - 1;
  const int C = (T + R - 1) / R;
  auto load_desc = [&](int c) {
#pragma unroll
    for (int k = 0; k < R * 4 / 64; ++k)
      dma(recs + (long)c * R * 4 + 64 * k + t, reinterpret_cast<uint32_t *>(&S.desc[c & 3][0][0]) + 64 * k);
  };
  // This stager's entries of chunk c, one per lane (lane i: record sid +
  // i*NSTG), read at once; the i-th is then taken with lane reads.
  auto descs = [&](int c, int n) {
    const int p = sid + NSTG * t;
    int4 d = make_int4(0, 0, 0, 0);
    if (p < n) d = *reinterpret_cast<const int4 *>(S.desc[c & 3][p]);
    return d;
  };
  auto cur_at = [&](const int4 &d, int i) {
    RecInfo k;
    const int kn = __builtin_amdgcn_readlane(d.x, i);
    k.kind = kn & 255;
    k.nb = kn >> 8;
    k.j = __builtin_amdgcn_readlane(d.y, i);
    k.ieb = __builtin_amdgcn_readlane(d.z, i);
    k.js = __builtin_amdgcn_readlane(d.w, i);
    return k;
  };
  auto fill = [&](int c) {
    const int n = min(R, T - c * R);
    const int4 d = descs(c, n);
    for (int i = 0, p = sid; p < n; ++i, p += NSTG) fill_record<WMAX>(S, s, L, c % 3, p, cur_at(d, i), t);
  };
  // Each result is read from LDS a record ahead of its store.
  auto drain = [&](int c) {
    const int n = min(R, T - c * R);
    const int4 d = descs(c, n);
    const int tl = t & (WMAX - 1);
    double on = sid < n ? S.out[c & 1][sid][0][tl] : 0.0;
    for (int i = 0, p = sid; p < n; ++i, p += NSTG) {
      const double o0 = on;
      if (p + NSTG < n) on = S.out[c & 1][p + NSTG][0][tl];
      drain_record<WMAX>(S, s, c % 3, c & 1, p, cur_at(d, i), t, o0);
    }
  };
  load_desc(0);
  if (C > 1) load_desc(1);
  wait_vm();
  fill(0);
  bar_all();
  for (int c = 0; c < C; ++c) {
    if (sid == 0 && c + 2 < C) load_desc(c + 2);
    if (c + 1 < C) fill(c + 1);
    if (c >= 1) drain(c - 1);
    bar_all();
  }
  drain(C - 1);
  bar_all();  // pass end
}

// ----------------------------------------------------------------- marcher
template <int WMAX>
struct Rec {
  uint32_t v[8];
  __device__ double fl(int s) const { return (double)__uint_as_float(v[s]); }
  __device__ double dbl(int s) const { return __hiloint2double((int)v[s + 1], (int)v[s]); }
};

// The marcher's walk through the records.  Byte offsets of the current
// record (input, halo, output) advance by one record per step and wrap with
// the buffers at chunk ends.
template <int WMAX>
struct MarchIO {
  using G = Stage<WMAX>;
  static constexpr int R = G::R;
  static constexpr int RB = G::REC * 4;    // bytes per input record
  static constexpr int OB = 2 * WMAX * 8;  // bytes per output record
  G &S;
  int l, qo = 0, ho = 0, oo = 0, p = 0;
  Rec<WMAX> c0, c1;  // c0 is the current record between calls
  __device__ MarchIO(G &s, int lane) : S(s), l(lane) {}
  template <int K>
  __device__ Rec<WMAX> &set() { return K ? c1 : c0; }
  // Record at input offset q (past the third buffer's end it reads the halo
  // array that follows: a discarded prefetch).
  template <int K>
  __device__ void read(int q) {
    Rec<WMAX> &c = set<K>();
    const uint32_t *r = reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(&S) + q) + l;
#pragma unroll
    for (int k = 0; k < 8; ++k) c.v[k] = r[k * WMAX];
  }
  __device__ void start() {
    bar_lds();
    qo = 0; ho = 0; oo = 0; p = 0;
    read<0>(0);
  }
  // Record set K is current: return it and request the next record into the
  // other set at once, so its LDS latency hides behind this row's chain.
  // Rows alternate the two sets, so a row's record is never copied; other
  // records go through take(), which moves the prefetched record back into
  // set 0.
  template <int K>
  __device__ const Rec<WMAX> &take_row() {
    read<1 - K>(qo + RB);
    return set<K>();
  }
  __device__ Rec<WMAX> take() {
    const Rec<WMAX> r = c0;
    read<0>(qo + RB);
    return r;
  }
  // The current closure record's halo: X(1, row) and X(W+2, row).
  __device__ void halo(double &hw, double &he) const {
    const uint4 h = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(&S.halo[0][0][0]) + ho);
    hw = __hiloint2double((int)h.y, (int)h.x);
    he = __hiloint2double((int)h.w, (int)h.z);
  }
  __device__ void put(int k, double v) {
    *reinterpret_cast<double *>(reinterpret_cast<char *>(&S.out[0][0][0][0]) + oo + k * WMAX * 8 + l * 8) = v;
  }
  // After the record: at the chunk's end its barrier, then the new chunk's
  // first record into set K (the set that becomes current).
  template <int K = 0>
  __device__ void next() {
    qo += RB;
    ho += 16;
    oo += OB;
    if (++p == R) {
      bar_lds();
      p = 0;
      if (qo == G::NIN * R * RB) { qo = 0; ho = 0; }
      if (oo == 2 * R * OB) oo = 0;
      read<K>(qo);
    }
  }
  // The last chunk's barrier, then the pass-end barrier that waits for the
  // stagers' last results to reach memory.
  __device__ void finish() {
    if (p != 0) bar_lds();
    bar_lds();
  }
};

// REPBIR on one strip (MAIN/template/GLOBAL/src/sevp.f90:218-328), consuming
// the forward and backward record sequences.  WX > 0: the strip width is WX.
template <int WMAX, int WX>
__device__ void march_strip(Stage<WMAX> &S, const Strip &s, int lane) {
  const int w = s.w, nblk = s.nblk;
  const int n = lane < w ? lane : 0;
  const long ww = (long)w * w;
  double rc[WMAX], rc2[WMAX];
  MarchIO<WMAX> io(S, lane);

  // ---------------- forward ----------------
  if (nblk > 1) load_col<WMAX, WX>(rc, s.rinv1, w, n);
  else load_col<WMAX, WX>(rc, s.rinv, w, n);
  io.start();
  const int jlast = fwd_jlast(s.ie, nblk);
  double xs, xc;
  {
    const Rec<WMAX> h = io.take();
    xs = h.dbl(0);
    xc = h.dbl(2);
    io.next();
  }
  int js = 1, ieb = ie_at(s.ie, 0);
  for (int nb = 1;; ++nb) {
    const int jend = nb < nblk ? ieb - 2 : jlast;
    const int ieb_next = nb < nblk ? ie_at(s.ie, nb) : 0;  // requested a block ahead
    auto row = [&](const Rec<WMAX> &r) {
      const double a = r.fl(0), b = r.fl(1), cc = r.fl(2), d = r.fl(3), e = r.fl(4), ff = r.fl(5);
      const double h = r.dbl(6);  // lane 0: X(1, J+1); lane w-1: X(w+2, J+1)
      double xw = wave_shr1_zero(xc);
      if (lane == 0) xw = h;
      double xe = wave_shl1_zero(xc);
      if (lane == w - 1) xe = h;
      const double xn = (ff - a * xw - b * xs - cc * xc - d * xe) / e;
      io.put(0, xn);
      xs = xc;
      xc = xn;
    };
    for (int j = js; j <= jend; j += 2) {  // rows in pairs: record sets 0, 1
      row(io.template take_row<0>());
      io.template next<1>();
      if (j == jend) {
        io.c0 = io.c1;
        break;
      }
      row(io.template take_row<1>());
      io.template next<0>();
    }
    if (nb == nblk) break;
    {  // closure residual of block nb at row IE-1, correction of row IE
      const Rec<WMAX> r = io.take();
      const double a = r.fl(0), b = r.fl(1), cc = r.fl(2), d = r.fl(3), e = r.fl(4), ff = r.fl(5);
      const double xnn = r.dbl(6);
      double hw, he;
      io.halo(hw, he);
      double xw = wave_shr1_zero(xc);
      if (lane == 0) xw = hw;
      double xe = wave_shl1_zero(xc);
      if (lane == w - 1) xe = he;
      const double res = ff - a * xw - b * xs - cc * xc - d * xe - e * xnn;
      const double dd = dot_col<WMAX, WX>(res, rc, w, S.e, lane);
      const double xcor = xc - dd;
      io.put(0, xcor);
      io.put(1, xc);  // DUM0
      io.next();
      xs = xcor;
      xc = xnn;
      if (nb + 1 < nblk) {
        load_col<WMAX, WX>(rc, s.rinv1 + nb * ww, w, n);
      } else {  // no forward closure left: the backward pass's first block
        load_col<WMAX, WX>(rc, s.rinv + (nblk - 1) * ww, w, n);
        load_col<WMAX, WX>(rc2, s.rinv1 + (nblk - 2) * ww, w, n);
      }
    }
    js = ieb;
    ieb = ieb_next;
  }
  io.finish();

  // ---------------- backward ----------------
  io.start();
  double xnn2 = 0.0;  // X(., IE+1): the block above's corrected row JS+1
  int iebk = ie_at(s.ie, nblk - 1), jsb = nblk != 1 ? ie_at(s.ie, nblk - 2) : 1;
  for (int nbk = nblk; nbk >= 1; --nbk) {
    const int jsb_next = nbk > 2 ? ie_at(s.ie, nbk - 3) : 1;  // block nbk-1's JS, a block ahead
    const int jf = iebk - 2;
    const Rec<WMAX> A = io.take();
    double hw, he;
    io.halo(hw, he);  // record A's
    io.next();
    const Rec<WMAX> B = io.take();
    const double xrow = B.dbl(0), xf = B.dbl(2);
    if (nbk == nblk) xnn2 = B.dbl(6);
    double xw = wave_shr1_zero(xrow);
    if (lane == 0) xw = hw;
    double xe = wave_shl1_zero(xrow);
    if (lane == w - 1) xe = he;
    const double res = A.fl(5) - A.fl(0) * xw - A.fl(1) * A.dbl(6) - A.fl(2) * xrow -
                       A.fl(3) * xe - A.fl(4) * xnn2;
    const double cv = dot_col<WMAX, WX>(res, rc, w, S.e, lane);
    const double xnew = xf + cv;
    io.put(0, xnew);
    double hs = 0.0;
    if (nbk != 1) {
      const double g = cv * B.fl(4);
      hs = dot_col<WMAX, WX>(g, rc2, w, S.e, lane);
    }
    io.next();
    if (nbk > 1) {  // the next block's matrices, a block ahead
      load_col<WMAX, WX>(rc, s.rinv + (nbk - 2) * ww, w, n);
      if (nbk > 2) load_col<WMAX, WX>(rc2, s.rinv1 + (nbk - 3) * ww, w, n);
    }
    xnn2 = xnew;
    double hc = cv;
    auto row = [&](const Rec<WMAX> &m) {
      const double a = m.fl(0), b = m.fl(1), cc = m.fl(2), d = m.fl(3), e = m.fl(4), xo = m.dbl(6);
      const double hwv = wave_shr1_zero(hc);
      double hev = wave_shl1_zero(hc);
      if (lane == w - 1) hev = 0.0;
      const double hn = (-(a * hwv) - b * hs - cc * hc - d * hev) / e;
      io.put(0, xo + hn);
      hs = hc;
      hc = hn;
    };
    for (int j = jsb; j <= jf; j += 2) {
      row(io.template take_row<0>());
      io.template next<1>();
      if (j == jf) {
        io.c0 = io.c1;
        break;
      }
      row(io.template take_row<1>());
      io.template next<0>();
    }
    iebk = jsb;
    jsb = jsb_next;
  }
  io.finish();
}

// ----------------------------------------------------------------- kernels
// The block rows IE live in LDS for the whole launch: every stager copies
// them (the same values) before its first chunk, so neither the marcher, which
// reads them only after the first chunk barrier, nor the stagers wait on
// global memory for them.
__device__ inline void load_ie(int *dst, const int *src, int nblk, int t) {
  for (int b = t; b < nblk; b += SEVP_WAVE) dst[b] = src[b];
  wait_vm();
}

__device__ inline Strip bir_strip(const BirArgs &p, const int *ie, int r, int n) {
  Strip s;
  const int w = p.ibir, inc = p.ibir / 2;
  const int i0 = p.i2 + 2;
  const int il = (n - 1) * inc + 1;
  s.arr[0] = p.al; s.arr[1] = p.ab; s.arr[2] = p.ac; s.arr[3] = p.ar; s.arr[4] = p.at;
  s.arr[5] = p.s + (long)r * p.s_stride;
  for (int k = 0; k < 6; ++k) s.ld[k] = p.i2;
  s.il = il;
  s.inc = inc;
  s.seam = n == p.nbir;
  if (!s.seam) {
    s.xv = p.x + (long)r * p.x_stride + f2(i0, il, 1);
    s.ldx = i0;
  } else {
    s.xv = p.xx + (long)r * p.xx_stride;
    s.ldx = w + 2;
  }
  const long ibr2 = (long)w * w;
  s.rinv = p.rinv + (long)(n - 1) * ibr2 * p.nb0;
  s.rinv1 = p.rinv1 + (long)(n - 1) * ibr2 * (p.nb0 - 1);
  s.dum0 = p.dum0 + (long)r * p.dum0_stride;
  s.ie = ie;
  s.nblk = p.nb0;
  s.w = w;
  return s;
}

// Whole FS_SOLVER BIR loop (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1902-1966)
// for one RHS per workgroup.  The seam strip's halo gather and write-back
// (and the periodic halo refresh) are done by every stager in full: the
// stagers write identical values, and each reads only what it has itself
// written before its next pass's DMAs.
template <int WMAX, int WX>
__global__ __launch_bounds__(BIR_THREADS) void bir_sweeps_kernel(BirArgs p) {
  __shared__ Stage<WMAX> S;
  __shared__ int ie[BIR_MAX_BLOCKS];
  const int r = blockIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = threadIdx.x & 63;
  if (wave == 0) {
    if (t < WMAX)
      for (int it = 0; it < p.nsweep; ++it)
        for (int n = 1; n <= p.nbir; ++n) march_strip<WMAX, WX>(S, bir_strip(p, ie, r, n), t);
    return;
  }
  const int sid = wave - 1;
  load_ie(ie, p.ie, p.nb0, t);
  const int w = p.ibir, inc = p.ibir / 2;
  const int i0 = p.i2 + 2, m0 = w + 2;
  const int j1 = p.j2 + 1;
  double *x = p.x + (long)r * p.x_stride;
  double *xx = p.xx + (long)r * p.xx_stride;
  for (int it = 0; it < p.nsweep; ++it) {
    for (int n = 1; n <= p.nbir; ++n) {
      const Strip s = bir_strip(p, ie, r, n);
      const int il = s.il;
      if (s.seam) {
        // seam strip: halo columns of XX for rows 1..J2 as the reference does
        // (row J1 keeps its old value), or 1..J1 with seam_fix
        const int jg = p.seam_fix ? j1 : p.j2;
        for (int j = 1 + t; j <= jg; j += SEVP_WAVE) {
          xx[f2(m0, 1, j)] = x[f2(i0, il, j)];
          xx[f2(m0, w + 2, j)] = x[f2(i0, inc + 2, j)];
        }
        wait_vm();
      }
      StagerLane<WMAX> L;
      L.init(s, t);
      stage_pass<WMAX>(S, s, L, t, sid, true, p.recs);
      stage_pass<WMAX>(S, s, L, t, sid, false, p.recs + 4L * p.fwd_records);
      if (s.seam) {
        for (long q = t; q < (long)w * (j1 - 1); q += SEVP_WAVE) {
          const int i = (int)(q % w) + 1, j = (int)(q / w) + 2;
          const double v = xx[f2(m0, i + 1, j)];
          if (i <= inc) x[f2(i0, il + i, j)] = v;
          else x[f2(i0, i - inc + 1, j)] = v;
        }
        wait_vm();
        if (p.periodic) {
          for (int j = 2 + t; j <= j1; j += SEVP_WAVE) {
            x[f2(i0, 1, j)] = x[f2(i0, i0 - 1, j)];
            x[f2(i0, i0, j)] = x[f2(i0, 2, j)];
          }
          wait_vm();
        }
      }
    }
  }
}

// One REPBIR call with the reference's argument conventions, batched over
// independent right-hand sides (F and X strided by f_stride / x_stride).
template <int WMAX, int WX>
__global__ __launch_bounds__(BIR_THREADS) void repbir_kernel(RepbirArgs p) {
  __shared__ Stage<WMAX> S;
  __shared__ int ie[BIR_MAX_BLOCKS];
  const int r = blockIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = threadIdx.x & 63;
  Strip s;
  s.arr[0] = p.ax; s.arr[1] = p.ay; s.arr[2] = p.bb; s.arr[3] = p.cx; s.arr[4] = p.cy;
  s.arr[5] = p.f + (long)r * p.f_stride;
  for (int k = 0; k < 5; ++k) s.ld[k] = p.i2;
  s.ld[5] = p.m2;
  s.il = 1;
  s.inc = 0;
  s.seam = 0;
  s.xv = p.x + (long)r * p.x_stride;
  s.ldx = p.i0;
  s.rinv = p.rinv;
  s.rinv1 = p.rinv1;
  s.dum0 = p.dum0 + (long)r * p.dum0_stride;
  s.ie = ie;
  s.nblk = p.nblk;
  s.w = p.m2;
  if (wave == 0) {
    if (t < WMAX) march_strip<WMAX, WX>(S, s, t);
    return;
  }
  load_ie(ie, p.ie, p.nblk, t);
  StagerLane<WMAX> L;
  L.init(s, t);
  stage_pass<WMAX>(S, s, L, t, wave - 1, true, p.recs);
  stage_pass<WMAX>(S, s, L, t, wave - 1, false, p.recs + 4L * p.fwd_records);
}

// Strips of width 24 and 30 (the 0.5, 0.25 and 0.1 degree configurations)
// run kernels specialised for their width.
hipError_t launch_bir_sweeps(const BirArgs &a, int nrhs, hipStream_t st) {
  const dim3 g(nrhs), b(BIR_THREADS);
  if (a.ibir == 24) hipLaunchKernelGGL((bir_sweeps_kernel<32, 24>), g, b, 0, st, a);
  else if (a.ibir == 30) hipLaunchKernelGGL((bir_sweeps_kernel<32, 30>), g, b, 0, st, a);
  else if (a.ibir <= 32) hipLaunchKernelGGL((bir_sweeps_kernel<32, 0>), g, b, 0, st, a);
  else hipLaunchKernelGGL((bir_sweeps_kernel<64, 0>), g, b, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_repbir(const RepbirArgs &a, int nrhs, hipStream_t st) {
  const dim3 g(nrhs), b(BIR_THREADS);
  if (a.m2 == 24) hipLaunchKernelGGL((repbir_kernel<32, 24>), g, b, 0, st, a);
  else if (a.m2 == 30) hipLaunchKernelGGL((repbir_kernel<32, 30>), g, b, 0, st, a);
  else if (a.m2 <= 32) hipLaunchKernelGGL((repbir_kernel<32, 0>), g, b, 0, st, a);
  else hipLaunchKernelGGL((repbir_kernel<64, 0>), g, b, 0, st, a);
  return hipGetLastError();
}

}  // namespace sevp
