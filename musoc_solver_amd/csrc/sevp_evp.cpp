// The EVP file that PREP writes and MAIN reads: one Fortran unformatted
// sequential record, big-endian (OPEN(..., CONVERT='BIG_ENDIAN',
// FORM='unformatted')).
//
//   PREP  WRITE(FNO_EVP) AL,AR,AB,AT,AC,RINV,RINV1,IE
//         (PREP/template/GLOBAL/src/prep_GLOBAL.f90:320)
//   MAIN  READ(FN_EVP) AL,AR,AB,AT,AC,OAC,RINV,RINV1,IE
//         (MAIN/template/GLOBAL/src/input_GLOBAL.f90:442)
//
// The two lists differ by OAC, so a record written by PREP is shorter than
// what MAIN reads.  The reader accepts either layout (with_oac = -1 picks
// the one that matches the record length) and the writer produces either.
//
// Record framing: each record is one or more subrecords, each preceded and
// followed by a 4-byte signed length; a record longer than the largest
// subrecord (2147483639 bytes) is split, the leading length of every
// subrecord but the last and the trailing length of every subrecord but the
// first being negated (gfortran's convention for records over 2 GiB; a
// record that fits in one subrecord carries two equal positive lengths, as
// every Fortran compiler writes it).
#include <sys/types.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sevp.h"
#include "sevp_internal.h"

namespace {

int evp_fail(const std::string &m) { return sevp::set_error(SEVP_E_ARG, m); }
int io_fail(const std::string &m) { return sevp::set_error(SEVP_E_IO, m); }

constexpr long kMaxSub = 2147483639L;
constexpr size_t kChunk = 1 << 20;  // elements converted per buffer fill

// One array of the record's item list.
struct Item {
  const void *src;  // writer: the values
  void *dst;        // reader: where they go (null: skip them)
  size_t n, size;   // element count and size in bytes
};

void swap_copy(unsigned char *d, const unsigned char *s, size_t n, size_t size) {
  for (size_t i = 0; i < n; ++i)
    for (size_t k = 0; k < size; ++k) d[i * size + k] = s[i * size + size - 1 - k];
}

bool put_marker(FILE *f, int32_t v) {
  unsigned char b[4];
  const uint32_t u = (uint32_t)v;
  b[0] = (unsigned char)(u >> 24); b[1] = (unsigned char)(u >> 16);
  b[2] = (unsigned char)(u >> 8); b[3] = (unsigned char)u;
  return std::fwrite(b, 1, 4, f) == 4;
}

bool get_marker(FILE *f, int32_t &v) {
  unsigned char b[4];
  if (std::fread(b, 1, 4, f) != 4) return false;
  v = (int32_t)(((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3]);
  return true;
}

// Writes the payload of one record of known length n, opening and closing
// subrecords at their boundaries as bytes arrive.
struct RecordWriter {
  FILE *f;
  long n, maxsub, done = 0, len = 0, left = 0;
  bool ok = true;
  RecordWriter(FILE *f_, long n_, long maxsub_) : f(f_), n(n_), maxsub(maxsub_) { open_sub(); }
  void open_sub() {
    len = n - done < maxsub ? n - done : maxsub;
    left = len;
    ok = ok && put_marker(f, done + len == n ? (int32_t)len : -(int32_t)len);
  }
  void close_sub() { ok = ok && put_marker(f, done == len ? (int32_t)len : -(int32_t)len); }
  void put(const unsigned char *b, long k) {
    while (ok && k > 0) {
      if (left == 0) {
        close_sub();
        open_sub();
      }
      const long m = k < left ? k : left;
      ok = ok && std::fwrite(b, 1, (size_t)m, f) == (size_t)m;
      b += m, k -= m, left -= m, done += m;
    }
  }
  // The trailing length of the last subrecord (positive only if it is also
  // the first).
  bool finish() {
    if (ok && (done != n || left != 0)) ok = false;
    if (ok) ok = put_marker(f, done == len ? (int32_t)len : -(int32_t)len);
    return ok;
  }
};

struct Layout {
  long ncoef, nrinv, nrinv1, nie;
  long bytes(bool oac) const {
    return 4 * ncoef * (oac ? 6 : 5) + 8 * (nrinv + nrinv1) + 4 * nie;
  }
};

int layout_of(const sevp_desc *d, Layout &l) {
  if (!d || d->i2 < 1 || d->j2 < 1 || d->nb0 < 1) return evp_fail("bad descriptor");
  long w = d->i2, nmat = 1;
  if (d->form == SEVP_FORM_BIR) {
    if (d->nbir < 1 || (2 * d->i2) % d->nbir) return evp_fail("NBIR must divide 2*I2");
    w = 2L * d->i2 / d->nbir;
    nmat = d->nbir;
  }
  l.ncoef = (long)d->i2 * d->j2;
  l.nrinv = w * w * d->nb0 * nmat;
  l.nrinv1 = w * w * (d->nb0 - 1) * nmat;
  l.nie = d->nb0;
  return SEVP_OK;
}

// Walks the subrecords of the file's first record without reading their
// payload: checks that every leading length has its matching trailing
// length and sign, and returns the payload length (or -1).  limit bounds
// the total so that a corrupt length is refused (-2).
long scan_record(FILE *f, long limit) {
  long n = 0;
  for (bool first = true;; first = false) {
    int32_t m1 = 0, m2 = 0;
    if (!get_marker(f, m1)) return -1;
    const long len = m1 < 0 ? -(long)m1 : (long)m1;
    if (n + len > limit) return -2;
    if (fseeko(f, (off_t)len, SEEK_CUR) != 0 || !get_marker(f, m2)) return -1;
    const long len2 = m2 < 0 ? -(long)m2 : (long)m2;
    if (len2 != len || (m2 < 0) == first) return -1;
    n += len;
    if (m1 >= 0) return n;
  }
}

// Reads the payload of the record scan_record accepted into the items, in
// order, skipping the subrecord markers.
bool read_items(FILE *f, const std::vector<Item> &items) {
  std::vector<unsigned char> buf;
  long left = 0;         // payload bytes left in the current subrecord
  bool started = false;  // past the first leading length
  auto take = [&](unsigned char *d, long k) {
    while (k > 0) {
      if (left == 0) {
        int32_t m = 0;
        if (started && !get_marker(f, m)) return false;  // trailing length
        if (!get_marker(f, m)) return false;              // next leading length
        started = true;
        left = m < 0 ? -(long)m : (long)m;
        continue;
      }
      const long c = k < left ? k : left;
      if (std::fread(d, 1, (size_t)c, f) != (size_t)c) return false;
      d += c, k -= c, left -= c;
    }
    return true;
  };
  for (const Item &it : items) {
    for (size_t i = 0; i < it.n; i += kChunk) {
      const size_t m = it.n - i < kChunk ? it.n - i : kChunk;
      buf.resize(m * it.size);
      if (!take(buf.data(), (long)(m * it.size))) return false;
      if (it.dst)
        swap_copy(static_cast<unsigned char *>(it.dst) + i * it.size, buf.data(), m, it.size);
    }
  }
  return true;
}

}  // namespace

extern "C" {

int sevp_evp_write(const char *path, const sevp_desc *d, const float *al, const float *ab,
                   const float *ac, const float *ar, const float *at, const double *rinv,
                   const double *rinv1, const float *oac, long max_subrecord) {
  Layout l{};
  if (int rc = layout_of(d, l)) return rc;
  if (!path || !al || !ab || !ac || !ar || !at || !rinv || (!rinv1 && l.nrinv1) || !d->ie)
    return evp_fail("null argument");
  const long maxsub = max_subrecord > 0 ? max_subrecord : kMaxSub;
  if (maxsub > kMaxSub) return evp_fail("subrecords are at most 2147483639 bytes");
  static_assert(sizeof(int) == 4, "IE is INTEGER(4)");
  std::vector<Item> items;
  for (const float *c : {al, ar, ab, at, ac}) items.push_back({c, nullptr, (size_t)l.ncoef, 4});
  if (oac) items.push_back({oac, nullptr, (size_t)l.ncoef, 4});
  items.push_back({rinv, nullptr, (size_t)l.nrinv, 8});
  items.push_back({rinv1, nullptr, (size_t)l.nrinv1, 8});
  items.push_back({d->ie, nullptr, (size_t)l.nie, 4});
  FILE *f = std::fopen(path, "wb");
  if (!f) return io_fail(std::string("cannot open ") + path + " for writing");
  RecordWriter w(f, l.bytes(oac != nullptr), maxsub);
  std::vector<unsigned char> buf;
  for (const Item &it : items)
    for (size_t i = 0; i < it.n && w.ok; i += kChunk) {
      const size_t m = it.n - i < kChunk ? it.n - i : kChunk;
      buf.resize(m * it.size);
      swap_copy(buf.data(), static_cast<const unsigned char *>(it.src) + i * it.size, m, it.size);
      w.put(buf.data(), (long)(m * it.size));
    }
  bool ok = w.finish();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SEVP_OK : io_fail(std::string("write to ") + path + " failed");
}

int sevp_evp_read(const char *path, const sevp_desc *d, float *al, float *ab, float *ac,
                  float *ar, float *at, double *rinv, double *rinv1, int *ie, float *oac,
                  int with_oac, int *had_oac) {
  Layout l{};
  if (int rc = layout_of(d, l)) return rc;
  if (!path || !al || !ab || !ac || !ar || !at || !rinv || (!rinv1 && l.nrinv1) || !ie)
    return evp_fail("null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return io_fail(std::string("cannot open ") + path);
  // No valid record is longer than the larger layout.
  const long limit = l.bytes(with_oac != 0);
  const long n = scan_record(f, limit);
  if (n < 0) {
    std::fclose(f);
    if (n == -2)
      return io_fail(std::string(path) + ": record longer than the " + std::to_string(limit) +
                     " bytes this grid's layout holds");
    return io_fail(std::string(path) + ": not one well-formed unformatted record");
  }
  bool use_oac;
  if (with_oac < 0) {
    use_oac = n == l.bytes(true);
    if (!use_oac && n != l.bytes(false)) {
      std::fclose(f);
      return io_fail("record of " + std::to_string(n) + " bytes matches neither layout (" +
                     std::to_string(l.bytes(false)) + " without OAC, " +
                     std::to_string(l.bytes(true)) + " with)");
    }
  } else {
    use_oac = with_oac != 0;
    if (n != l.bytes(use_oac)) {
      std::fclose(f);
      return io_fail("record of " + std::to_string(n) + " bytes, expected " +
                     std::to_string(l.bytes(use_oac)) + (use_oac ? " with OAC" : " without OAC"));
    }
  }
  std::vector<Item> items;
  for (float *c : {al, ar, ab, at, ac}) items.push_back({nullptr, c, (size_t)l.ncoef, 4});
  if (use_oac) items.push_back({nullptr, oac, (size_t)l.ncoef, 4});
  items.push_back({nullptr, rinv, (size_t)l.nrinv, 8});
  items.push_back({nullptr, rinv1, (size_t)l.nrinv1, 8});
  items.push_back({nullptr, ie, (size_t)l.nie, 4});
  std::rewind(f);
  const bool ok = read_items(f, items);
  std::fclose(f);
  if (!ok) return io_fail(std::string(path) + ": read failed");
  if (had_oac) *had_oac = use_oac ? 1 : 0;
  return SEVP_OK;
}

}  // extern "C"

// ------------------------------------------------------------ plan cache

namespace {

uint64_t mix64(uint64_t h) {  // splitmix64 finalizer
  h ^= h >> 30;
  h *= 0xbf58476d1ce4e5b9ULL;
  h ^= h >> 27;
  h *= 0x94d049bb133111ebULL;
  return h ^ (h >> 31);
}

struct Hasher {
  uint64_t h = 0x5345565045565031ULL;  // format tag: change it to invalidate every cache
  void word(uint64_t w) { h = mix64(h ^ w) + 0x9e3779b97f4a7c15ULL; }
  void bytes(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, b + i, 8);
      word(w);
    }
    uint64_t w = 0;
    std::memcpy(&w, b + i, n - i);
    word(w ^ ((uint64_t)(n - i) << 56));
  }
};

}  // namespace

extern "C" {

int sevp_operator_hash(const sevp_desc *d, const float *al, const float *ab, const float *ac,
                       const float *ar, const float *at, unsigned long long *out) {
  Layout l{};
  if (int rc = layout_of(d, l)) return rc;
  if (!al || !ab || !ac || !ar || !at || !d->ie || !out) return evp_fail("null argument");
  Hasher h;
  h.word((uint64_t)d->i2);
  h.word((uint64_t)d->j2);
  h.word((uint64_t)d->nb0);
  h.word((uint64_t)d->form);
  h.word((uint64_t)(d->form == SEVP_FORM_BIR ? d->nbir : 0));
  h.bytes(d->ie, sizeof(int) * (size_t)d->nb0);
  for (const float *c : {al, ab, ac, ar, at}) h.bytes(c, sizeof(float) * (size_t)l.ncoef);
  *out = h.h;
  return SEVP_OK;
}

int sevp_plan_create_cached(const sevp_desc *d, const float *al, const float *ab, const float *ac,
                            const float *ar, const float *at, const char *cache_dir,
                            sevp_plan *out, int *hit) {
  if (!cache_dir || !out) return evp_fail("null argument");
  unsigned long long key = 0;
  if (int rc = sevp_operator_hash(d, al, ab, ac, ar, at, &key)) return rc;
  Layout l{};
  layout_of(d, l);
  sevp_plan plan = nullptr;
  if (int rc = sevp_plan_create(d, al, ab, ac, ar, at, &plan)) return rc;
  char name[32];
  std::snprintf(name, sizeof name, "sevp_%016llx.evp", key);
  const std::string path = std::string(cache_dir) + "/" + name;
  std::vector<float> c[5];
  for (auto &v : c) v.resize((size_t)l.ncoef);
  std::vector<double> rinv((size_t)l.nrinv), rinv1((size_t)(l.nrinv1 > 0 ? l.nrinv1 : 1));
  std::vector<int> ie((size_t)l.nie);
  // A hit needs the stored operator and IE to equal the caller's bit for bit.
  if (sevp_evp_read(path.c_str(), d, c[0].data(), c[1].data(), c[2].data(), c[3].data(),
                    c[4].data(), rinv.data(), rinv1.data(), ie.data(), nullptr, -1,
                    nullptr) == SEVP_OK) {
    bool same = std::memcmp(ie.data(), d->ie, sizeof(int) * (size_t)l.nie) == 0;
    const float *mine[5] = {al, ab, ac, ar, at};
    for (int k = 0; k < 5 && same; ++k)
      same = std::memcmp(c[k].data(), mine[k], sizeof(float) * (size_t)l.ncoef) == 0;
    if (same) {
      if (int rc = sevp_load_influence(plan, rinv.data(), rinv1.data())) {
        sevp_plan_destroy(plan);
        return rc;
      }
      if (hit) *hit = 1;
      *out = plan;
      return SEVP_OK;
    }
  }
  int zp = 0;
  int rc = sevp_build_influence(plan, &zp);
  if (rc == SEVP_OK) rc = sevp_get_influence(plan, rinv.data(), rinv1.data());
  if (rc != SEVP_OK) {
    sevp_plan_destroy(plan);
    return rc;
  }
  // Written under a private name and renamed, so that a reader never sees
  // a partial file and concurrent writers of one key leave a whole one.
  char tmp_suffix[48];
  std::snprintf(tmp_suffix, sizeof tmp_suffix, ".tmp%ld.%p", (long)getpid(), (void *)plan);
  const std::string tmp = path + tmp_suffix;
  int stored = sevp_evp_write(tmp.c_str(), d, al, ab, ac, ar, at, rinv.data(), rinv1.data(),
                              nullptr, 0) == SEVP_OK &&
               std::rename(tmp.c_str(), path.c_str()) == 0;
  if (!stored) std::remove(tmp.c_str());
  if (hit) *hit = stored ? 0 : -1;
  *out = plan;
  return SEVP_OK;
}

}  // extern "C"
