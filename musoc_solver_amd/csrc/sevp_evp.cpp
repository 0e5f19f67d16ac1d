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

template <class T>
void put_be(std::vector<unsigned char> &buf, const T *src, size_t n) {
  const size_t at = buf.size();
  buf.resize(at + n * sizeof(T));
  unsigned char *d = buf.data() + at;
  for (size_t i = 0; i < n; ++i) {
    unsigned char b[sizeof(T)];
    std::memcpy(b, src + i, sizeof(T));
    for (size_t k = 0; k < sizeof(T); ++k) d[i * sizeof(T) + k] = b[sizeof(T) - 1 - k];
  }
}

template <class T>
void get_be(const unsigned char *s, T *dst, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    unsigned char b[sizeof(T)];
    for (size_t k = 0; k < sizeof(T); ++k) b[k] = s[i * sizeof(T) + sizeof(T) - 1 - k];
    std::memcpy(dst + i, b, sizeof(T));
  }
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
  std::vector<unsigned char> rec;
  rec.reserve((size_t)l.bytes(oac != nullptr));
  for (const float *c : {al, ar, ab, at, ac}) put_be(rec, c, (size_t)l.ncoef);
  if (oac) put_be(rec, oac, (size_t)l.ncoef);
  put_be(rec, rinv, (size_t)l.nrinv);
  if (l.nrinv1) put_be(rec, rinv1, (size_t)l.nrinv1);
  std::vector<int32_t> ie(d->ie, d->ie + d->nb0);
  put_be(rec, ie.data(), ie.size());
  FILE *f = std::fopen(path, "wb");
  if (!f) return io_fail(std::string("cannot open ") + path + " for writing");
  bool ok = true;
  const long n = (long)rec.size();
  long at_ = 0;
  do {
    const long len = n - at_ < maxsub ? n - at_ : maxsub;
    const bool first = at_ == 0, last = at_ + len == n;
    ok = ok && put_marker(f, last ? (int32_t)len : -(int32_t)len);
    ok = ok && std::fwrite(rec.data() + at_, 1, (size_t)len, f) == (size_t)len;
    ok = ok && put_marker(f, first ? (int32_t)len : -(int32_t)len);
    at_ += len;
  } while (ok && at_ < n);
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
  // No valid record is longer than the larger layout: a length beyond it is
  // refused before anything is allocated for it.
  const long limit = l.bytes(with_oac != 0);
  std::vector<unsigned char> rec;
  bool ok = true, first = true, too_long = false;
  for (;;) {
    int32_t m1 = 0, m2 = 0;
    if (!get_marker(f, m1)) { ok = false; break; }
    const long len = m1 < 0 ? -(long)m1 : (long)m1;
    if ((long)rec.size() + len > limit) { ok = false; too_long = true; break; }
    const size_t at_ = rec.size();
    rec.resize(at_ + (size_t)len);
    if (std::fread(rec.data() + at_, 1, (size_t)len, f) != (size_t)len || !get_marker(f, m2)) {
      ok = false;
      break;
    }
    const long len2 = m2 < 0 ? -(long)m2 : (long)m2;
    if (len2 != len || (m2 < 0) == first) { ok = false; break; }
    first = false;
    if (m1 >= 0) break;
  }
  std::fclose(f);
  if (too_long)
    return io_fail(std::string(path) + ": record longer than the " + std::to_string(limit) +
                   " bytes this grid's layout holds");
  if (!ok) return io_fail(std::string(path) + ": not one well-formed unformatted record");
  const long n = (long)rec.size();
  bool use_oac;
  if (with_oac < 0) {
    if (n == l.bytes(false)) use_oac = false;
    else if (n == l.bytes(true)) use_oac = true;
    else
      return io_fail("record of " + std::to_string(n) + " bytes matches neither layout (" +
                      std::to_string(l.bytes(false)) + " without OAC, " +
                      std::to_string(l.bytes(true)) + " with)");
  } else {
    use_oac = with_oac != 0;
    if (n != l.bytes(use_oac))
      return io_fail("record of " + std::to_string(n) + " bytes, expected " +
                      std::to_string(l.bytes(use_oac)) + (use_oac ? " with OAC" : " without OAC"));
  }
  const unsigned char *p = rec.data();
  for (float *c : {al, ar, ab, at, ac}) {
    get_be(p, c, (size_t)l.ncoef);
    p += 4 * l.ncoef;
  }
  if (use_oac) {
    if (oac) get_be(p, oac, (size_t)l.ncoef);
    p += 4 * l.ncoef;
  }
  get_be(p, rinv, (size_t)l.nrinv);
  p += 8 * l.nrinv;
  if (l.nrinv1) get_be(p, rinv1, (size_t)l.nrinv1);
  p += 8 * l.nrinv1;
  std::vector<int32_t> iev((size_t)l.nie);
  get_be(p, iev.data(), iev.size());
  for (long i = 0; i < l.nie; ++i) ie[i] = iev[(size_t)i];
  if (had_oac) *had_oac = use_oac ? 1 : 0;
  return SEVP_OK;
}

}  // extern "C"
