"""The C oracle against the reference routines compiled from
MAIN/template/GLOBAL/src/sevp.f90 (oracle/_ref), bitwise, on random
operators, strips and offsets.  Skipped where oracle/_ref was not built."""
import os

import numpy as np
import pytest

from musoc_solver_amd import synth
from tests.conftest import fields_equal, have_ref

pytestmark = [pytest.mark.ref,
              pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (make -C oracle ref)")]


def case(i2, j2, n0, seed, metric="mercator"):
    c = synth.make_case(i2, j2, seed, metric=metric, lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, n0))
    return c, ie


@pytest.mark.parametrize("i2,j2,n0,seed", [(24, 12, 3, 1), (40, 20, 5, 2), (64, 32, 4, 3)])
@pytest.mark.parametrize("fn", ["rep1", "rep"])
def test_full_width_matches_reference(oracle, i2, j2, n0, seed, fn):
    c, ie = case(i2, j2, n0, seed, "latlon")
    rinv, rinv1, _ = (oracle.pre1 if fn == "rep1" else oracle.pre)(c.coefs, ie)
    s = synth.make_rhs(c.mask, seed, 1)[:, :, 0].copy(order="F")
    x0 = np.asfortranarray(np.random.default_rng(seed).standard_normal((i2 + 2, j2 + 2)))
    xa, xb = x0.copy(order="F"), x0.copy(order="F")
    getattr(oracle, fn)(c.coefs, rinv, rinv1, s, xa, ie)
    getattr(oracle, fn)(c.coefs, rinv, rinv1, s, xb, ie, use_ref=True)
    assert fields_equal(xa, xb)


@pytest.mark.parametrize("il", [1, 7, 13])
def test_repbir_strip_with_offsets_matches_reference(oracle, il):
    i2, j2, n0, ibir = 36, 20, 4, 12
    c, ie = case(i2, j2, n0, 5)
    off = il - 1
    nblk = len(ie)
    from oracle.oracle import P, _ci
    rv = np.zeros((ibir, ibir, nblk), order="F")
    rv1 = np.zeros((ibir, ibir, nblk), order="F")
    h = np.zeros((ibir + 2, j2 + 2), order="F")
    oracle.lib().orc_pre(*[P(a, off) for a in c.coefs], P(rv), P(rv1), P(h),
                         P(np.ascontiguousarray(ie, np.int32)), _ci(i2), _ci(ibir + 2), _ci(ibir),
                         _ci(nblk))
    src = np.asfortranarray(synth.make_rhs(c.mask, 6, 1)[il - 1:il - 1 + ibir, :, 0])
    x0 = np.asfortranarray(np.random.default_rng(il).standard_normal((i2 + 2, j2 + 2)))
    out = []
    for use_ref in (False, True):
        x = x0.copy(order="F")
        d0 = np.zeros((ibir, nblk), order="F")
        d1, d2 = np.zeros(ibir), np.zeros(ibir)
        oracle.repbir_call([(a, off) for a in c.coefs], rv, rv1, d0, d1, d2, src,
                           np.zeros((ibir + 2, j2 + 2), order="F"), (x, off),
                           np.ascontiguousarray(ie, np.int32), i2 + 2, i2, ibir + 2, ibir, nblk,
                           use_ref=use_ref)
        out.append(x)
    assert fields_equal(out[0], out[1])


@pytest.mark.parametrize("i2,j2,nbir,n0,nsweep", [(48, 24, 8, 6, 5), (60, 30, 10, 5, 3),
                                                  (96, 48, 8, 4, 2)])
def test_bir_driver_matches_reference_driver(oracle, i2, j2, nbir, n0, nsweep):
    """orc_bir_solve (C) against the FS_SOLVER strip loop driven from Python
    around the compiled reference REPBIR, two calls in a row (the seam
    scratch carries over)."""
    c, ie = case(i2, j2, n0, 9)
    rinv, rinv1, _ = oracle.bir_prep(c.coefs, ie, nbir)
    s = synth.make_rhs(c.mask, 10, 1)[:, :, 0].copy(order="F")
    x0 = np.asfortranarray(np.random.default_rng(11).standard_normal((i2 + 2, j2 + 2)))
    xa, xb = x0.copy(order="F"), x0.copy(order="F")
    w = 2 * i2 // nbir
    xxa = np.zeros((w + 2, j2 + 2), order="F")
    xxb = np.zeros((w + 2, j2 + 2), order="F")
    for _ in range(2):
        oracle.bir_solve(c.coefs, s, rinv, rinv1, xa, xxa, ie, nbir, nsweep)
        oracle.bir_solve_ref(c.coefs, s, rinv, rinv1, xb, xxb, ie, nbir, nsweep)
        assert fields_equal(xa, xb) and fields_equal(xxa, xxb)


def test_repbir_period_equals_rep1_in_the_reference(oracle):
    """REPBIR_PERIOD's one call shape (M0=I0, M2=I2) is REP1, statement for
    statement; the compiled reference agrees bitwise."""
    i2, j2 = 32, 16
    c, ie = case(i2, j2, 4, 12, "latlon")
    rinv, rinv1, _ = oracle.pre1(c.coefs, ie)
    s = synth.make_rhs(c.mask, 13, 1)[:, :, 0].copy(order="F")
    x0 = np.asfortranarray(np.random.default_rng(14).standard_normal((i2 + 2, j2 + 2)))
    xa, xb = x0.copy(order="F"), x0.copy(order="F")
    oracle.rep1(c.coefs, rinv, rinv1, s, xa, ie, use_ref=True)
    nblk = len(ie)
    oracle.repbir_call([(a, 0) for a in c.coefs], rinv, rinv1, np.zeros((i2, nblk), order="F"),
                       np.zeros(i2), np.zeros(i2), s, np.zeros((i2 + 2, j2 + 2), order="F"),
                       (xb, 0), np.ascontiguousarray(ie, np.int32), i2 + 2, i2, i2 + 2, i2, nblk,
                       use_ref=True, period=True)
    assert fields_equal(xa, xb)


def test_planted_mistake_is_caught(oracle, tmp_path):
    """A plausible slip planted in a copy of the oracle (RINV indexed
    transposed in the backward correction) must break bitwise agreement with
    the compiled reference and must also break the solve's accuracy: the
    pins have teeth."""
    import ctypes
    import subprocess
    from oracle.oracle import HERE, P, _ci
    so = str(tmp_path / "liboracle_planted.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-std=c99",
                    "-DORC_PLANT_TRANSPOSED_RINV", os.path.join(HERE, "sevp_oracle.c"), "-o", so],
                   check=True)
    bad = ctypes.CDLL(so)
    i2, j2 = 24, 12
    c, ie = case(i2, j2, 3, 15, "latlon")
    rinv, rinv1, _ = oracle.pre(c.coefs, ie)
    s = synth.make_rhs(c.mask, 16, 1)[:, :, 0].copy(order="F")
    x_ref = np.zeros((i2 + 2, j2 + 2), order="F")
    oracle.rep(c.coefs, rinv, rinv1, s, x_ref, ie, use_ref=True)
    x_bad = np.zeros((i2 + 2, j2 + 2), order="F")
    nblk = len(ie)
    d0 = np.zeros((i2, nblk), order="F")
    d1, d2 = np.zeros(i2), np.zeros(i2)
    h = np.zeros((i2 + 2, j2 + 2), order="F")
    ie32 = np.ascontiguousarray(ie, np.int32)
    bad.orc_rep(*[P(a) for a in c.coefs], P(rinv), P(rinv1), P(d0), P(d1), P(d2), P(s), P(h),
                P(x_bad), P(ie32), _ci(i2 + 2), _ci(i2), _ci(nblk))
    assert not fields_equal(x_bad, x_ref)
    assert oracle.rel_residual(c.coefs, s, x_ref) < 1e-9
    assert oracle.rel_residual(c.coefs, s, x_bad) > 1e-6


def test_reference_rep1_is_reentrant_across_threads(oracle):
    """The CPU baseline times the compiled reference REP1 on several host
    threads at once; concurrent calls on independent arrays must give
    bitwise the sequential results (the routines keep no static state)."""
    import threading
    i2, j2, nrhs = 48, 24, 8
    c, ie = case(i2, j2, 3, 9, "latlon")
    rinv, rinv1, _ = oracle.pre1(c.coefs, ie)
    s = synth.make_rhs(c.mask, 10, nrhs)
    seq = []
    for r in range(nrhs):
        x = np.zeros((i2 + 2, j2 + 2), order="F")
        oracle.rep1(c.coefs, rinv, rinv1, np.asfortranarray(s[:, :, r]), x, ie, use_ref=True)
        seq.append(x)
    out = [np.zeros((i2 + 2, j2 + 2), order="F") for _ in range(nrhs)]

    def work(r):
        for _ in range(10):
            out[r][:] = 0.0
            oracle.rep1(c.coefs, rinv, rinv1, np.asfortranarray(s[:, :, r]), out[r], ie, use_ref=True)

    threads = [threading.Thread(target=work, args=(r,)) for r in range(nrhs)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for r in range(nrhs):
        assert np.array_equal(out[r], seq[r]), r
