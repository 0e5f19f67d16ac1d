"""The strip kernels' record order (csrc/sevp_records.h) against the order in
which the marching wavefront consumes records.

The HIP strip kernels (csrc/sevp_bir.hip) run one marching wavefront and
three stager wavefronts that meet at one barrier per chunk of records.  The
stagers follow the host-built record table; the marcher follows the block
loops of `march_strip`, restated below (MAIN/template/GLOBAL/src/sevp.f90:
218-328: forward rows 1..JLAST with a closure at row IE(nb)-1 of every block
but the last; backward, per block from the top, the closure and the H-march
rows JS..IE-2).  The two must agree record for record, and the table's
lengths must equal the kernels' closed-form counts (1 + JLAST forward,
NBLK + IE(NBLK) - 1 backward), or the barriers would not pair up.  The table
is built from the header alone with g++, so this runs without a GPU.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "musoc_solver_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include "sevp_records.h"
int main() {
  int nblk;
  while (std::scanf("%d", &nblk) == 1) {
    std::vector<int> ie(nblk), t;
    for (int &v : ie) std::scanf("%d", &v);
    int fwd = 0;
    sevp::bir_record_table(ie.data(), nblk, t, fwd);
    std::printf("%d %zu\n", fwd, t.size() / 4);
    for (size_t k = 0; k < t.size(); k += 4) std::printf("%d %d %d %d\n", t[k], t[k + 1], t[k + 2], t[k + 3]);
  }
  return 0;
}
"""

KIND = {0: "H", 1: "R", 2: "C", 3: "A", 4: "B", 5: "BR"}


def marcher_order(ie):
    """The records march_strip consumes, as (kind, block, row)."""
    nblk = len(ie)
    js_last = ie[-2] if nblk > 1 else 1
    jlast = ie[-1] - 2 if ie[-1] - 2 >= js_last else js_last - 1
    fwd = [("H", 1, 0)]
    js, ieb, nb = 1, ie[0], 1
    while True:
        jend = ieb - 2 if nb < nblk else jlast
        fwd += [("R", nb, j) for j in range(js, jend + 1)]
        if nb == nblk:
            break
        fwd.append(("C", nb, ieb - 1))
        js, ieb, nb = ieb, ie[nb], nb + 1
    bwd = []
    for nb in range(nblk, 0, -1):
        ieb, js = ie[nb - 1], (ie[nb - 2] if nb != 1 else 1)
        bwd += [("A", nb, 0), ("B", nb, 0)] + [("BR", nb, j) for j in range(js, ieb - 1)]
    return fwd, bwd


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("records")
    src = d / "records.cpp"
    src.write_text(HARNESS)
    exe = d / "records"
    subprocess.run([cxx, "-O1", "-std=c++17", f"-I{CSRC}", str(src), "-o", str(exe)], check=True)
    return str(exe)


def random_ie(rng):
    nblk = int(rng.integers(1, 14))
    return list(np.cumsum(rng.choice([1, 1, 2, 3, 5, 24], size=nblk)) + 1)


def test_record_table_matches_the_marcher(harness):
    rng = np.random.default_rng(7)
    cases = [random_ie(rng) for _ in range(400)]
    cases += [[2], [3], [41], list(range(2, 42)), [2, 3], [40, 41]]
    inp = "".join(f"{len(ie)} {' '.join(map(str, ie))}\n" for ie in cases)
    out = subprocess.run([harness], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    k = 0
    for ie in cases:
        fwd_len, total = map(int, out[k].split())
        rows = [tuple(map(int, out[k + 1 + i].split())) for i in range(total)]
        k += 1 + total
        fwd_m, bwd_m = marcher_order(ie)
        nblk = len(ie)
        js_last = ie[-2] if nblk > 1 else 1
        jlast = ie[-1] - 2 if ie[-1] - 2 >= js_last else js_last - 1
        assert len(fwd_m) == 1 + jlast and len(bwd_m) == nblk + ie[-1] - 1
        # padded to whole 32-record chunks plus a spare chunk
        assert fwd_len == (len(fwd_m) + 31) // 32 * 32 + 32
        assert total - fwd_len == (len(bwd_m) + 31) // 32 * 32 + 32

        def decode(r):
            kind = KIND[r[0] & 255]
            return (kind, r[0] >> 8, r[1] if kind in ("R", "C", "BR") else 0)
        assert [decode(r) for r in rows[:len(fwd_m)]] == fwd_m, ie
        assert [decode(r) for r in rows[fwd_len:fwd_len + len(bwd_m)]] == bwd_m, ie
        # block bounds carried with every record
        for r in rows[:len(fwd_m)] + rows[fwd_len:fwd_len + len(bwd_m)]:
            nb = r[0] >> 8
            assert r[2] == ie[nb - 1] and r[3] == (ie[nb - 2] if nb > 1 else 1)
        # padding is empty
        assert all(r == (0, 0, 0, 0) for r in rows[len(fwd_m):fwd_len] + rows[fwd_len + len(bwd_m):])
