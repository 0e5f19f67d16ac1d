// Record order of the strip kernels (REPBIR, BIR sweeps), host side.  Plain
// C++ with no HIP dependency, so host tools and tests can build it alone.
#pragma once
#include <vector>

namespace sevp {

// Record sequence of one strip's march, the order in which the strip kernels
// stage operands and consume them (the same for every strip of a plan).  One
// 4-int entry per record: (kind | block << 8, row, IE(block), JS(block)).
// Forward: the header, then rows 1..jlast, row IE(nb)-1 of every block but
// the last being its closure.  Backward, blocks NBLK..1: closure records A and
// B, then the H-march rows JS..IE-2.  The backward table follows the forward
// one; each is padded with empty entries to a whole number of 32-record
// chunks plus one spare chunk.
enum BirKind : int { BIR_HEAD = 0, BIR_ROW, BIR_CLOS, BIR_A, BIR_B, BIR_BROW };

inline void bir_record_table(const int *ie, int nblk, std::vector<int> &out, int &fwd_records) {
  out.clear();
  auto push = [&](int kind, int nb, int j, int ieb, int js) {
    out.push_back(kind | (nb << 8));
    out.push_back(j);
    out.push_back(ieb);
    out.push_back(js);
  };
  auto pad = [&]() {
    const size_t n = out.size() / 4, np = (n + 31) / 32 * 32 + 32;
    out.resize(4 * np, 0);
  };
  const int ieb_last = ie[nblk - 1], js_last = nblk > 1 ? ie[nblk - 2] : 1;
  const int jlast = ieb_last - 2 >= js_last ? ieb_last - 2 : js_last - 1;
  int nb = 1, js = 1, ieb = ie[0];
  push(BIR_HEAD, 1, 0, ieb, js);
  for (int j = 1; j <= jlast; ++j) {
    while (nb < nblk && j > ieb - 1) {
      ++nb;
      js = ieb;
      ieb = ie[nb - 1];
    }
    push((nb < nblk && j == ieb - 1) ? BIR_CLOS : BIR_ROW, nb, j, ieb, js);
  }
  pad();
  fwd_records = (int)(out.size() / 4);
  for (nb = nblk; nb >= 1; --nb) {
    ieb = ie[nb - 1];
    js = nb != 1 ? ie[nb - 2] : 1;
    push(BIR_A, nb, 0, ieb, js);
    push(BIR_B, nb, 0, ieb, js);
    for (int j = js; j <= ieb - 2; ++j) push(BIR_BROW, nb, j, ieb, js);
  }
  pad();
}

}  // namespace sevp
