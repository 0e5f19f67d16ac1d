"""Write the golden vectors in tests/golden/*.npz.  Run in the build
container, where oracle/_ref/libsevpref.so is built from the reference
sources (make -C oracle ref):  python tests/golden/make_golden.py

Every fixture holds inputs and the outputs the compiled reference routines
(REPBIR, REP1, REP, REPBIR_PERIOD from MAIN/template/GLOBAL/src/sevp.f90)
produce on them.  The influence matrices among the inputs come from the
oracle's PRE/PRE1 restatement: the reference PREP program is not buildable
here, so RINV/RINV1 are inputs of these vectors, not outputs.

Cases
  bir_harness_192x96_h24, _h6
                       the reference's own manufactured test
                       (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1814-1848):
                       ideal operator with Neumann north/south edges
                       (PREP/template/GLOBAL/src/prep_GLOBAL.f90:246-261),
                       S = A*0 = 0, X = 1 on [3, I2-1] x [3, J2-1], then the
                       5-sweep BIR strip loop, 16 strips of width 24, blocks
                       of 24 rows (the 192x96, 8x4-block layout; the march
                       overflows and the reference returns NaN everywhere in
                       the interior) and of 6 rows (finite)
  bir_synth_48x24      synthetic operator, Mercator metric, NBIR 8, blocks of 6
  rep1_synth_48x24     REP1 on the synthetic operator, blocks of 4
  rep_synth_48x24      REP on the synthetic operator, blocks of 4
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from musoc_solver_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes(order="F")).hexdigest()


def save(name, meta, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
    print(name, os.path.getsize(path), "bytes")


def harness_bir(n0):
    i2, j2, nbir = 192, 96, 16
    coefs = synth.ideal_operator(i2, j2)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, n0))
    rinv, rinv1, zp = O.bir_prep(coefs, ie, nbir)
    x = np.zeros((i2 + 2, j2 + 2), order="F")
    s = np.asfortranarray(O.sqprod(coefs, x)[1:-1, 1:-1].astype(np.float32))  # S = A*0
    x[2:i2 - 1, 2:j2 - 1] = 1.0  # X(3:I2-1, 3:J2-1) = 1
    x0 = x.copy(order="F")
    xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
    O.bir_solve_ref(coefs, s, rinv, rinv1, x, xx, ie, nbir, nsweep=5, periodic=True)
    # the influence matrices are large and incompressible: keep their hash,
    # the tests rebuild them with the oracle's PRE and check it
    save(f"bir_harness_192x96_h{n0}",
         dict(i2=i2, j2=j2, nbir=nbir, n0=n0, nsweep=5, periodic=True, operator="ideal",
              zero_pivots=zp, rinv_sha256=sha(rinv), rinv1_sha256=sha(rinv1[:, :len(ie) - 1])),
         ie=ie, s=s, x0=x0, x=x, xx=xx)


def synth_bir():
    i2, j2, nbir, n0 = 48, 24, 8, 6
    case = synth.make_case(i2, j2, 101, metric="mercator", h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, n0))
    rinv, rinv1, zp = O.bir_prep(case.coefs, ie, nbir)
    s = synth.make_rhs(case.mask, 102, 1)[:, :, 0].copy(order="F")
    x0 = np.asfortranarray(np.random.default_rng(103).standard_normal((i2 + 2, j2 + 2)))
    x = x0.copy(order="F")
    xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
    O.bir_solve_ref(case.coefs, s, rinv, rinv1, x, xx, ie, nbir, nsweep=5, periodic=True)
    al, ab, ac, ar, at = case.coefs
    save("bir_synth_48x24", dict(i2=i2, j2=j2, nbir=nbir, n0=n0, nsweep=5, periodic=True,
                                 operator="synthetic mercator seed 101", zero_pivots=zp),
         al=al, ab=ab, ac=ac, ar=ar, at=at, ie=ie, s=s, rinv=rinv, rinv1=rinv1, x0=x0, x=x, xx=xx)


def synth_full(periodic):
    i2, j2, n0 = 48, 24, 4
    case = synth.make_case(i2, j2, 201, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, n0))
    if periodic:
        rinv, rinv1, zp = O.pre1(case.coefs, ie)
    else:
        rinv, rinv1, zp = O.pre(case.coefs, ie)
    s = synth.make_rhs(case.mask, 202, 1)[:, :, 0].copy(order="F")
    x0 = np.asfortranarray(np.random.default_rng(203).standard_normal((i2 + 2, j2 + 2)))
    x = x0.copy(order="F")
    if periodic:
        O.rep1(case.coefs, rinv, rinv1, s, x, ie, use_ref=True)
        xp = x0.copy(order="F")
        d0 = np.zeros((i2, len(ie)), order="F")
        d1 = np.zeros(i2)
        h = np.zeros((i2 + 2, j2 + 2), order="F")
        cs = [(np.asfortranarray(c), 0) for c in case.coefs]
        O.repbir_call(cs, rinv, rinv1, d0, d1, d1.copy(), s, h, (xp, 0),
                      np.ascontiguousarray(ie, np.int32), i2 + 2, i2, i2 + 2, i2, len(ie),
                      use_ref=True, period=True)
        assert np.array_equal(x, xp), "REPBIR_PERIOD(M0=I0, M2=I2) differs from REP1"
    else:
        O.rep(case.coefs, rinv, rinv1, s, x, ie, use_ref=True)
    al, ab, ac, ar, at = case.coefs
    name = "rep1_synth_48x24" if periodic else "rep_synth_48x24"
    save(name, dict(i2=i2, j2=j2, n0=n0, periodic=periodic,
                    operator="synthetic latlon +-80 seed 201", zero_pivots=zp),
         al=al, ab=ab, ac=ac, ar=ar, at=at, ie=ie, s=s, rinv=rinv, rinv1=rinv1, x0=x0, x=x)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libsevpref.so missing: run make -C oracle ref first")
    O.build()
    harness_bir(24)
    harness_bir(6)
    synth_bir()
    synth_full(True)
    synth_full(False)
