"""ctypes front end of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the C restatement in sevp_oracle.c) and, where it
was built here, oracle/_ref/libsevpref.so (the reference's own REP1,
REPBIR_PERIOD, REPBIR and REP compiled from MAIN/template/GLOBAL/src/
sevp.f90).  Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
leg import this module; the product path never does.

All arrays are numpy arrays in Fortran order with the reference's shapes:
coefficients and RHS float32 (I2, J2); solution X float64 (I2+2, J2+2);
strip influence matrices RINV float64 (IBIR*IBIR, NB0, NBIR) and RINV1
(IBIR*IBIR, NB0-1, NBIR); full-width ones (I2, I2, NB0) and (I2, I2, NB0-1).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsevpref.so")
REF_FAST_PATH = os.path.join(HERE, "_ref", "libsevpref_fast.so")  # -O3, FMA: timing only

_lib = None
_ref = None
_ref_fast = None


def build(ref: bool = False) -> None:
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for name in ("orc_matinv", "orc_pre", "orc_pre1", "orc_bir_prep", "orc_bir_prep_strips"):
            getattr(_lib, name).restype = ctypes.c_int
        _lib.orc_step14_correct.restype = ctypes.c_float
        for name in ("orc_inprod", "orc_tridiag_min_eig"):
            getattr(_lib, name).restype = ctypes.c_double
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_PATH)
    return _ref


def ref_fast_available() -> bool:
    return os.path.exists(REF_FAST_PATH)


def ref_fast():
    """The reference routines built with -O3 and multiply-add contraction:
    for timing the CPU path, not bitwise parity."""
    global _ref_fast
    if _ref_fast is None:
        _ref_fast = ctypes.CDLL(REF_FAST_PATH)
    return _ref_fast


def P(a: np.ndarray, off: int = 0):
    """Pointer to element ``off`` (in elements) of a contiguous array."""
    assert a.flags.f_contiguous or a.flags.c_contiguous
    return ctypes.c_void_p(a.ctypes.data + off * a.itemsize)


def _ci(v):
    return ctypes.c_int(int(v))


def _ri(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def f32(a):
    return np.asfortranarray(a, dtype=np.float32)


def f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def strip_geometry(i2: int, nbir: int):
    ibir = 2 * i2 // nbir
    if ibir * nbir != 2 * i2:
        raise ValueError("2*I2 must be divisible by NBIR")
    return ibir, ibir // 2


# ---------------------------------------------------------------- C oracle

def matinv(b: np.ndarray) -> tuple[np.ndarray, int]:
    b = f64(b).copy(order="F")
    zp = lib().orc_matinv(P(b), _ci(b.shape[0]))
    return b, zp


def bir_prep(coefs, ie, nbir):
    """Influence matrices of every BIR strip (PREP INITFS strip loop)."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    nb0 = len(ie)
    ibir, _ = strip_geometry(i2, nbir)
    rinv = np.zeros((ibir * ibir, nb0, nbir), order="F")
    rinv1 = np.zeros((ibir * ibir, max(nb0 - 1, 1), nbir), order="F")
    h = np.zeros((ibir + 2, j2 + 2), order="F")
    ie = np.ascontiguousarray(ie, np.int32)
    zp = lib().orc_bir_prep(P(al), P(ab), P(ac), P(ar), P(at), P(rinv), P(rinv1),
                            P(h), P(ie), _ci(i2), _ci(j2), _ci(nbir), _ci(nb0))
    return rinv, rinv1, zp


def bir_prep_strips(coefs, ie, nbir, strips):
    """bir_prep for the listed strips only (0-based); the other strips' slices
    are zero.  Each strip is computed exactly as in the full loop."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    nb0 = len(ie)
    ibir, _ = strip_geometry(i2, nbir)
    rinv = np.zeros((ibir * ibir, nb0, nbir), order="F")
    rinv1 = np.zeros((ibir * ibir, max(nb0 - 1, 1), nbir), order="F")
    h = np.zeros((ibir + 2, j2 + 2), order="F")
    ie = np.ascontiguousarray(ie, np.int32)
    zp = 0
    for n in strips:
        zp += lib().orc_bir_prep_strips(P(al), P(ab), P(ac), P(ar), P(at), P(rinv), P(rinv1),
                                        P(h), P(ie), _ci(i2), _ci(j2), _ci(nbir), _ci(nb0),
                                        _ci(n + 1), _ci(n + 1))
    return rinv, rinv1, zp


def bir_solve(coefs, s, rinv, rinv1, x, xx, ie, nbir, nsweep=5, periodic=True, seam_fix=False):
    """FS_SOLVER BIR sweeps on one RHS; x and xx are updated in place."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    s = f32(s)
    assert x.flags.f_contiguous and x.dtype == np.float64 and x.shape == (i2 + 2, j2 + 2)
    assert xx.flags.f_contiguous and xx.dtype == np.float64
    ie = np.ascontiguousarray(ie, np.int32)
    lib().orc_bir_solve(P(al), P(ab), P(ac), P(ar), P(at), P(s), P(f64(rinv)),
                        P(f64(rinv1)), P(x), P(xx), P(ie), _ci(i2), _ci(j2),
                        _ci(nbir), _ci(len(ie)), _ci(nsweep), _ci(1 if periodic else 0),
                        _ci(1 if seam_fix else 0))


def pre1(coefs, ie):
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    nb0 = len(ie)
    rinv = np.zeros((i2, i2, nb0), order="F")
    rinv1 = np.zeros((i2, i2, max(nb0 - 1, 1)), order="F")
    h = np.zeros((i2 + 2, j2 + 2), order="F")
    ie = np.ascontiguousarray(ie, np.int32)
    zp = lib().orc_pre1(P(al), P(ab), P(ac), P(ar), P(at), P(rinv), P(rinv1), P(h),
                        P(ie), _ci(i2 + 2), _ci(i2), _ci(nb0))
    return rinv, rinv1, zp


def pre(coefs, ie):
    """Non-periodic full-width influence matrices (PRE with M2=I2)."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    nb0 = len(ie)
    rinv = np.zeros((i2, i2, nb0), order="F")
    rinv1 = np.zeros((i2, i2, max(nb0 - 1, 1)), order="F")
    h = np.zeros((i2 + 2, j2 + 2), order="F")
    ie = np.ascontiguousarray(ie, np.int32)
    zp = lib().orc_pre(P(al), P(ab), P(ac), P(ar), P(at), P(rinv), P(rinv1), P(h),
                       P(ie), _ci(i2), _ci(i2 + 2), _ci(i2), _ci(nb0))
    return rinv, rinv1, zp


def _full_width(fn_c, fn_ref, coefs, rinv, rinv1, s, x, ie, use_ref):
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    nb0 = len(ie)
    s = f32(s)
    assert x.flags.f_contiguous and x.dtype == np.float64
    ie = np.ascontiguousarray(ie, np.int32)
    dum0 = np.zeros((i2, max(nb0 - 1, 1)), order="F")
    dum1 = np.zeros(i2)
    dum2 = np.zeros(i2)
    h = np.zeros((i2 + 2, j2 + 2), order="F")
    rinv, rinv1 = f64(rinv), f64(rinv1)
    ptrs = [P(al), P(ab), P(ac), P(ar), P(at), P(rinv), P(rinv1), P(dum0), P(dum1),
            P(dum2), P(s), P(h), P(x), P(ie)]
    if use_ref == "fast":
        getattr(ref_fast(), fn_ref)(*ptrs, _ri(i2 + 2), _ri(i2), _ri(nb0))
    elif use_ref:
        getattr(ref(), fn_ref)(*ptrs, _ri(i2 + 2), _ri(i2), _ri(nb0))
    else:
        getattr(lib(), fn_c)(*ptrs, _ci(i2 + 2), _ci(i2), _ci(nb0))


def rep1(coefs, rinv, rinv1, s, x, ie, use_ref=False):
    _full_width("orc_rep1", "rep1_", coefs, rinv, rinv1, s, x, ie, use_ref)


def rep(coefs, rinv, rinv1, s, x, ie, use_ref=False):
    _full_width("orc_rep", "rep_", coefs, rinv, rinv1, s, x, ie, use_ref)


def repbir_call(coefs_off, rinv, rinv1, dum0, dum1, dum2, f, h, x_off, ie,
                i0, i2, m0, m2, nblk, use_ref=False, period=False):
    """Direct call with the reference's argument list.  ``coefs_off`` is a list
    of (array, element offset) pairs, ``x_off`` an (array, offset) pair."""
    ptrs = [P(a, o) for a, o in coefs_off]
    ptrs += [P(rinv), P(rinv1), P(dum0), P(dum1), P(dum2), P(f), P(h), P(x_off[0], x_off[1]), P(ie)]
    if use_ref:
        fn = ref().repbir_period_ if period else ref().repbir_
        fn(*ptrs, _ri(i0), _ri(i2), _ri(m0), _ri(m2), _ri(nblk))
    else:
        fn = lib().orc_repbir_period if period else lib().orc_repbir
        fn(*ptrs, _ci(i0), _ci(i2), _ci(m0), _ci(m2), _ci(nblk))


def bir_solve_ref(coefs, s, rinv, rinv1, x, xx, ie, nbir, nsweep=5, periodic=True):
    """The FS_SOLVER strip loop (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:
    1902-1966) driven from Python around the compiled reference REPBIR.  An
    independent second statement of the driver, used to pin orc_bir_solve."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    i0, j0 = i2 + 2, j2 + 2
    nb0 = len(ie)
    ibir, inc = strip_geometry(i2, nbir)
    ie = np.ascontiguousarray(ie, np.int32)
    s = f32(s)
    h = np.zeros((ibir + 2, j0), order="F")
    dum0 = np.zeros((ibir, max(nb0 - 1, 1)), order="F")
    dum1 = np.zeros(ibir)
    dum2 = np.zeros(ibir)
    rinv, rinv1 = f64(rinv), f64(rinv1)
    seam_cols = np.concatenate([np.arange(i2 - inc, i2), np.arange(0, inc)])
    for _ in range(nsweep):
        for n in range(1, nbir + 1):
            il = (n - 1) * inc + 1
            rv = np.asfortranarray(rinv[:, :, n - 1])
            rv1 = np.asfortranarray(rinv1[:, :, n - 1])
            if n != nbir:
                src = f32(s[il - 1:il - 1 + ibir, :])
                off = (il - 1)
                repbir_call([(c, off) for c in (al, ab, ac, ar, at)], rv, rv1, dum0,
                            dum1, dum2, src, h, (x, il - 1), ie, i0, i2, ibir + 2,
                            ibir, nb0, use_ref=True)
            else:
                xx[0, :j2] = x[il - 1, :j2]
                xx[ibir + 1, :j2] = x[inc + 1, :j2]
                cs = [f32(c[seam_cols, :]) for c in (al, ab, ac, ar, at)]
                src = f32(s[seam_cols, :])
                repbir_call([(c, 0) for c in cs], rv, rv1, dum0, dum1, dum2, src, h,
                            (xx, 0), ie, ibir + 2, ibir, ibir + 2, ibir, nb0, use_ref=True)
                x[il:il + inc, 1:j2 + 1] = xx[1:inc + 1, 1:j2 + 1]
                x[1:inc + 1, 1:j2 + 1] = xx[inc + 1:ibir + 1, 1:j2 + 1]
                if periodic:
                    x[0, 1:j2 + 1] = x[i2, 1:j2 + 1]
                    x[i2 + 1, 1:j2 + 1] = x[1, 1:j2 + 1]


def sqprod(coefs, x):
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    r = np.zeros((i2 + 2, j2 + 2), order="F")
    lib().orc_sqprod(P(al), P(ab), P(ac), P(ar), P(at), P(f64(x)), P(r), _ci(i2), _ci(j2))
    return r


def rel_residual(coefs, s, x):
    """||s - A x||_2 / ||s||_2 over the interior, in float64."""
    r = sqprod(coefs, x)[1:-1, 1:-1]
    s = np.asarray(s, np.float64)
    return float(np.linalg.norm(s - r) / np.linalg.norm(s))


def _i16(a):
    assert a.dtype == np.int16 and a.flags.f_contiguous
    return a


def _f32w(a):
    assert a.dtype == np.float32 and a.flags.f_contiguous, "REAL(4) Fortran-ordered array expected"
    return a


def step14_prepare(oc, u, v, w, s, mask=True, lopenw=0, harness=False, free_surface=True):
    """First half of one step-14 ITMASK iteration (land masking, optional
    harness override, W integration, RHS S); u, v, w, s updated in place."""
    lib().orc_step14_prepare(
        _ci(oc.i2), _ci(oc.j2), _ci(oc.k1), _ci(int(mask)), _ci(lopenw), _ci(int(harness)),
        _ci(int(free_surface)), P(_i16(oc.kb)), P(_i16(oc.iu0)), P(_i16(oc.iv0)), P(oc.odx),
        P(oc.ody), P(oc.ocs), P(oc.csv), P(oc.odz), P(_f32w(u)), P(_f32w(v)), P(_f32w(w)),
        P(_f32w(s)))


def step14_correct(oc, x, p0, scr, u, v, odt, periodic_we=False):
    """Second half: P0, U, V and SCR updated in place; returns the largest
    top-layer speed on land faces."""
    assert x.dtype == np.float64 and x.flags.f_contiguous
    return float(lib().orc_step14_correct(
        _ci(oc.i2), _ci(oc.j2), _ci(oc.k1), _ci(int(periodic_we)), ctypes.c_float(odt),
        P(_i16(oc.iu)), P(_i16(oc.iv)), P(_i16(oc.iu0)), P(_i16(oc.iv0)), P(oc.odx), P(oc.odyv),
        P(x), P(_f32w(p0)), P(_f32w(scr)), P(_f32w(u)), P(_f32w(v))))


# ------------------------------------------------------------------ Krylov
# The reference's stop thresholds are REAL(4) literals compared in double.
BICG_TOL = float(np.float32(1e-8))
PCSI_TOL = float(np.float32(1e-7))


def sqprod_t(coefs, x):
    """p_sqprod_T (MAIN/template/GLOBAL/src/sevp.f90:714-740)."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    r = np.zeros((i2 + 2, j2 + 2), order="F")
    lib().orc_sqprod_t(P(al), P(ab), P(ac), P(ar), P(at), P(f64(x)), P(r), _ci(i2), _ci(j2))
    return r


def inprod(r, p):
    """p_inprod (:741-755): sequential interior sum."""
    i2, j2 = r.shape[0] - 2, r.shape[1] - 2
    return lib().orc_inprod(P(f64(r)), P(f64(p)), _ci(i2), _ci(j2))


def bicgstab(coefs, b, x, periodic=True, max_iter=1000, restart=100, tol=BICG_TOL,
             rtol=BICG_TOL):
    """p_bicgstab_le (:443-552) on X in place; returns (iterations, ||r||)."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    res = ctypes.c_double(0.0)
    it = lib().orc_bicgstab(P(al), P(ab), P(ac), P(ar), P(at), P(f32(b)), P(f64(x)), _ci(i2),
                            _ci(j2), _ci(1 if periodic else 0), _ci(max_iter), _ci(restart),
                            ctypes.c_double(tol), ctypes.c_double(rtol), ctypes.byref(res))
    return it, res.value


def pcsi(coefs, mineig, maxeig, b, x, periodic=True, max_iter=1000, tol=PCSI_TOL):
    """p_csi (:553-616) on X in place; returns (iterations, ||r||)."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    res = ctypes.c_double(0.0)
    it = lib().orc_pcsi(P(al), P(ab), P(ac), P(ar), P(at), ctypes.c_double(mineig),
                        ctypes.c_double(maxeig), P(f32(b)), P(f64(x)), _ci(i2), _ci(j2),
                        _ci(1 if periodic else 0), _ci(max_iter), ctypes.c_double(tol),
                        ctypes.byref(res))
    return it, res.value


def lanczos(coefs, m=10):
    """lanczos (:618-694): returns (mineig, maxeig, alpha[m], beta[m+1], sigma[m+1])."""
    al, ab, ac, ar, at = [f32(c) for c in coefs]
    i2, j2 = al.shape
    alpha = np.zeros(m)
    beta = np.zeros(m + 1)
    sigma = np.zeros(m + 1)
    mn, mx = ctypes.c_double(0.0), ctypes.c_double(0.0)
    lib().orc_lanczos(P(al), P(ab), P(ac), P(ar), P(at), _ci(i2), _ci(j2), _ci(m), P(alpha),
                      P(beta), P(sigma), ctypes.byref(mn), ctypes.byref(mx))
    return mn.value, mx.value, alpha, beta, sigma


def tridiag_min_eig(d, e):
    d = np.ascontiguousarray(d, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    return lib().orc_tridiag_min_eig(P(d), P(e), _ci(len(d)))
