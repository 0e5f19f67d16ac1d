"""Python front end of the SEVP C ABI (include/sevp.h).

Mirrors the reference's solver interface:

  repbir / rep1 / rep / repbir_period   MAIN/template/GLOBAL/src/sevp.f90
  pre / pre1 / matinv                   PREP/template/GLOBAL/src/prep_GLOBAL.f90
  SevpPlan.solve                        FS_SOLVER's dispatch
                                        (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1850-1966)

with the same argument meaning: numpy arrays in Fortran order, coefficients
and right-hand sides float32, solutions and influence matrices float64,
updated in place where the reference updates its arguments.  Unlike the
reference (which has no error path), a failure raises SevpError with the
library's message.

The library is the compiled HIP code in musoc_solver_amd/lib/libsevp_hip.so;
there is no CPU fallback: if the library cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libsevp_hip.so")

FORM_FULL_PERIODIC = 0
FORM_BIR = 1
FORM_FULL = 2
PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2
PREC_FP64_MFMA = 3
STORE_FP64 = 0
STORE_PRODUCTS = 1
PASS_SPLITK = 0
PASS_FUSED = 1

SEVP_OK = 0
SEVP_E_NOTCONV = -5
SEVP_E_IO = -6
KCLASS_CORR = 0
KCLASS_MARCH = 1
KCLASS_RESID = 2


class SevpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sevp error {code}: {msg}")
        self.code = code


class _Desc(ctypes.Structure):
    _fields_ = [("i2", ctypes.c_int), ("j2", ctypes.c_int), ("nb0", ctypes.c_int),
                ("ie", ctypes.POINTER(ctypes.c_int)), ("form", ctypes.c_int),
                ("nbir", ctypes.c_int), ("nsweep", ctypes.c_int), ("periodic", ctypes.c_int),
                ("nrhs_max", ctypes.c_int), ("precision", ctypes.c_int),
                ("seam_fix", ctypes.c_int)]


class SevpTuning(ctypes.Structure):
    _fields_ = [("pass_kind", ctypes.c_int), ("waves", ctypes.c_int), ("rhs_groups", ctypes.c_int),
                ("rows", ctypes.c_int)]

    def as_dict(self):
        return {"pass_kernels": "fused" if self.pass_kind == PASS_FUSED else "splitk",
                "waves": self.waves, "rhs_groups": self.rhs_groups, "rows": self.rows}


class SevpStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("max_rel_resid", ctypes.c_double),
                ("seconds", ctypes.c_double)]


_lib = None
VP = ctypes.c_void_p
CI = ctypes.c_int


def lib():
    """Load libsevp_hip.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SevpError(-2, f"{LIB_PATH} is missing; run python -m musoc_solver_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        L.sevp_last_error.restype = ctypes.c_char_p
        L.sevp_version.restype = ctypes.c_char_p
        L.sevp_plan_device_bytes.restype = ctypes.c_size_t
        L.sevp_plan_device_bytes.argtypes = [VP]
        L.sevp_plan_create.argtypes = [ctypes.POINTER(_Desc), VP, VP, VP, VP, VP, ctypes.POINTER(VP)]
        L.sevp_plan_destroy.argtypes = [VP]
        L.sevp_build_influence.argtypes = [VP, ctypes.POINTER(CI)]
        L.sevp_load_influence.argtypes = [VP, VP, VP]
        L.sevp_get_influence.argtypes = [VP, VP, VP]
        L.sevp_reset_seam.argtypes = [VP]
        L.sevp_solve.argtypes = [VP, VP, VP, CI]
        L.sevp_solve_dev.argtypes = [VP, VP, VP, CI, VP]
        L.sevp_solve_refined_dev.argtypes = [VP, VP, VP, CI, ctypes.c_double, CI, VP,
                                             ctypes.POINTER(SevpStats)]
        L.sevp_residual_dev.argtypes = [VP, VP, VP, CI, VP]
        L.sevp_repbir.argtypes = [VP] * 14 + [CI] * 5
        L.sevp_repbir_period.argtypes = [VP] * 14 + [CI] * 5
        L.sevp_rep1.argtypes = [VP] * 14 + [CI] * 3
        L.sevp_rep.argtypes = [VP] * 14 + [CI] * 3
        L.sevp_pre.argtypes = [VP] * 9 + [CI] * 4
        L.sevp_pre1.argtypes = [VP] * 9 + [CI] * 3
        L.sevp_matinv.argtypes = [VP, CI]
        L.sevp_matinv_dev_noalloc.argtypes = [VP, CI, VP, VP]
        L.sevp_set_precision.argtypes = [VP, CI]
        L.sevp_set_storage.argtypes = [VP, CI]
        L.sevp_set_pass.argtypes = [VP, CI]
        TP = ctypes.POINTER(SevpTuning)
        L.sevp_get_tuning.argtypes = [VP, TP]
        L.sevp_set_tuning.argtypes = [VP, TP]
        L.sevp_autotune.argtypes = [VP, VP, VP, CI, ctypes.c_double, CI, TP,
                                    ctypes.POINTER(ctypes.c_double)]
        L.sevp_get_pass.argtypes = [VP, ctypes.POINTER(CI)]
        L.sevp_set_stream.argtypes = [VP, VP]
        L.sevp_set_graphs.argtypes = [VP, CI]
        L.sevp_set_rhs_groups.argtypes = [VP, CI]
        L.sevp_step14_dev.argtypes = [VP, ctypes.POINTER(Step14Desc), ctypes.POINTER(Step14Grid),
                                      ctypes.POINTER(Step14State), ctypes.POINTER(CI),
                                      ctypes.POINTER(ctypes.c_float)]
        L.sevp_profile_begin.argtypes = [VP, CI]
        L.sevp_operator_hash.argtypes = ([ctypes.POINTER(_Desc)] + [VP] * 5
                                         + [ctypes.POINTER(ctypes.c_ulonglong)])
        L.sevp_plan_create_cached.argtypes = ([ctypes.POINTER(_Desc)] + [VP] * 5
                                              + [ctypes.c_char_p, ctypes.POINTER(VP),
                                                 ctypes.POINTER(CI)])
        KO, KS = ctypes.POINTER(KrylovOpts), ctypes.POINTER(KrylovStats)
        DP = ctypes.POINTER(ctypes.c_double)
        CD = ctypes.c_double
        L.sevp_krylov_defaults.argtypes = [KO, CI]
        L.sevp_bicgstab_dev.argtypes = [VP, VP, VP, KO, KS]
        L.sevp_pcsi_dev.argtypes = [VP, CD, CD, VP, VP, KO, KS]
        L.sevp_lanczos_dev.argtypes = [VP, CI, KO, DP, DP, VP, VP]
        L.sevp_p_bicgstab_le.argtypes = [VP] * 7 + [CI, CI, KO, KS]
        L.sevp_p_csi.argtypes = [VP] * 5 + [CD, CD, VP, VP, CI, CI, KO, KS]
        L.sevp_lanczos.argtypes = [VP] * 5 + [CI, CI, CI, KO, DP, DP, VP, VP]
        L.sevp_evp_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Desc)] + [VP] * 8 + [ctypes.c_long]
        L.sevp_evp_read.argtypes = ([ctypes.c_char_p, ctypes.POINTER(_Desc)] + [VP] * 9
                                    + [CI, ctypes.POINTER(CI)])
        L.sevp_profile_end.argtypes = [VP, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_long)]
        L.sevp_profile_spans.argtypes = [VP, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_long)]
        L.sevp_batch_create.argtypes = [VP, CI, ctypes.POINTER(VP)]
        L.sevp_batch_destroy.argtypes = [VP]
        L.sevp_batch_solve_refined_dev.argtypes = [VP, VP, VP, VP, ctypes.c_double, CI, VP, VP,
                                                   ctypes.POINTER(SevpStats)]
        _lib = L
    return _lib


def _check(rc):
    if rc != SEVP_OK:
        raise SevpError(rc, lib().sevp_last_error().decode())


def _p(a, off=0):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not (a.flags.f_contiguous or a.flags.c_contiguous):
            raise ValueError("array must be contiguous")
        return ctypes.c_void_p(a.ctypes.data + off * a.itemsize)
    # torch tensor on the device
    return ctypes.c_void_p(a.data_ptr() + off * a.element_size())


def _f32(a):
    if a.dtype != np.float32 or not a.flags.f_contiguous:
        raise TypeError("expected a Fortran-ordered float32 array")
    return a


def _f64(a):
    if a.dtype != np.float64 or not a.flags.f_contiguous:
        raise TypeError("expected a Fortran-ordered float64 array")
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def version() -> str:
    return lib().sevp_version().decode()


# ------------------------------------------------------------------ plans

class Step14Desc(ctypes.Structure):
    """sevp_step14_desc (include/sevp.h)."""
    _fields_ = [("i2", ctypes.c_int), ("j2", ctypes.c_int), ("k1", ctypes.c_int),
                ("free_surface", ctypes.c_int), ("periodic_we", ctypes.c_int),
                ("lopenw", ctypes.c_int), ("harness", ctypes.c_int), ("mxmask", ctypes.c_int),
                ("tolerance", ctypes.c_float), ("odt", ctypes.c_float)]


class Step14Grid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("kb", "iu0", "iv0", "iu", "iv", "odx", "ody", "ocs", "csv", "odyv", "odz")]


class Step14State(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("u", "v", "w", "p0", "scr", "s", "x")]


class KrylovOpts(ctypes.Structure):
    """sevp_krylov_opts (include/sevp.h)."""
    _fields_ = [("periodic", ctypes.c_int), ("ref_order", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("restart", ctypes.c_int), ("tol", ctypes.c_double), ("rtol", ctypes.c_double)]


class KrylovStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("residual", ctypes.c_double), ("seconds", ctypes.c_double)]


KRYLOV_BICGSTAB = 0
KRYLOV_PCSI = 1


def krylov_opts(solver=KRYLOV_BICGSTAB, **kw) -> KrylovOpts:
    """The reference's settings for ``solver``, with fields overridden by
    keyword (periodic, ref_order, max_iter, restart, tol, rtol)."""
    o = KrylovOpts()
    _check(lib().sevp_krylov_defaults(ctypes.byref(o), solver))
    for k, v in kw.items():
        if v is None:
            continue
        setattr(o, k, int(v) if k in ("periodic", "ref_order", "max_iter", "restart") else float(v))
    return o


class SevpPlan:
    """A device-resident SEVP problem: operator, IE, influence matrices and
    scratch for up to ``nrhs_max`` simultaneous right-hand sides."""

    def __init__(self, coefs, ie, form=FORM_BIR, nbir=None, nsweep=5, periodic=True,
                 nrhs_max=1, precision=PREC_FP64, seam_fix=False, cache_dir=None,
                 storage=STORE_FP64):
        """With ``cache_dir`` the influence matrices come from the plan cache
        (sevp_plan_create_cached): ``cache_hit`` is 1 when they were loaded,
        0 when built and stored, -1 when built but not stored.  ``storage``
        STORE_PRODUCTS keeps only the FP32 product operands (sevp_set_storage)."""
        al, ab, ac, ar, at = [_f32(np.asfortranarray(c, dtype=np.float32)) for c in coefs]
        self.coefs = (al, ab, ac, ar, at)
        self.i2, self.j2 = al.shape
        self.ie = _i32(ie)
        self.form, self.precision, self.nrhs_max = form, precision, nrhs_max
        self.nbir = nbir if nbir is not None else 0
        self.ibir = 2 * self.i2 // self.nbir if form == FORM_BIR else self.i2
        d = _Desc(self.i2, self.j2, len(self.ie), self.ie.ctypes.data_as(ctypes.POINTER(CI)),
                  form, self.nbir, nsweep, 1 if periodic else 0, nrhs_max, precision,
                  1 if seam_fix else 0)
        h = VP()
        self.cache_hit = None
        if cache_dir is None:
            _check(lib().sevp_plan_create(ctypes.byref(d), _p(al), _p(ab), _p(ac), _p(ar),
                                          _p(at), ctypes.byref(h)))
        else:
            hit = CI(0)
            _check(lib().sevp_plan_create_cached(ctypes.byref(d), _p(al), _p(ab), _p(ac), _p(ar),
                                                 _p(at), os.fsencode(str(cache_dir)),
                                                 ctypes.byref(h), ctypes.byref(hit)))
            self.cache_hit = hit.value
        self._h = h
        self._stream = None
        if storage != STORE_FP64:
            self.set_storage(storage)

    def close(self):
        # at interpreter exit the module's globals may already be gone; the
        # process exit releases the device memory then
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.sevp_plan_destroy(h)

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def nb0(self):
        return len(self.ie)

    @property
    def device_bytes(self):
        return int(lib().sevp_plan_device_bytes(self._h))

    def influence_shapes(self):
        nb1 = max(self.nb0 - 1, 1)
        if self.form == FORM_BIR:
            w2 = self.ibir * self.ibir
            return (w2, self.nb0, self.nbir), (w2, nb1, self.nbir)
        return (self.i2, self.i2, self.nb0), (self.i2, self.i2, nb1)

    def set_storage(self, storage):
        """STORE_FP64 (default) or STORE_PRODUCTS: keep only the FP32 product
        operands (set before build_influence / load_influence)."""
        _check(lib().sevp_set_storage(self._h, storage))

    def set_pass(self, kind):
        """PASS_SPLITK or PASS_FUSED kernels for the MFMA corrections."""
        _check(lib().sevp_set_pass(self._h, kind))

    @property
    def tuning(self) -> SevpTuning:
        t = SevpTuning()
        _check(lib().sevp_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, pass_kind, waves, rhs_groups, rows=0):
        t = SevpTuning(pass_kind, waves, rhs_groups, rows)
        _check(lib().sevp_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, s, x, tol=1e-10, max_iter=20, nrhs=None):
        """Time refined solves of (s, x) (torch device tensors; x is
        overwritten) over the pass kernels, waves and row tiles per fused
        workgroup and RHS groups, keep the fastest; returns (SevpTuning, ms per
        solve call)."""
        nrhs = nrhs or s.shape[0]
        self._follow_torch_stream()
        t = SevpTuning()
        ms = ctypes.c_double(0.0)
        _check(lib().sevp_autotune(self._h, _p(s), _p(x), nrhs, tol, max_iter, ctypes.byref(t),
                                   ctypes.byref(ms)))
        return t, ms.value

    @property
    def pass_kind(self) -> int:
        k = CI(0)
        _check(lib().sevp_get_pass(self._h, ctypes.byref(k)))
        return k.value

    def build_influence(self) -> int:
        z = CI(0)
        _check(lib().sevp_build_influence(self._h, ctypes.byref(z)))
        return z.value

    def load_influence(self, rinv, rinv1):
        _check(lib().sevp_load_influence(self._h, _p(_f64(rinv)), _p(_f64(rinv1))))

    def get_influence(self):
        s0, s1 = self.influence_shapes()
        rinv = np.zeros(s0, order="F")
        rinv1 = np.zeros(s1, order="F")
        _check(lib().sevp_get_influence(self._h, _p(rinv), _p(rinv1)))
        return rinv, rinv1

    def write_evp(self, path, oac=None, max_subrecord=0):
        """Write the plan's operator, influence matrices and IE as the EVP
        record PREP writes (PREP/template/GLOBAL/src/prep_GLOBAL.f90:320);
        with ``oac`` the record carries OAC after AC, as MAIN reads it
        (MAIN/template/GLOBAL/src/input_GLOBAL.f90:442)."""
        rinv, rinv1 = self.get_influence()
        evp_write(path, self.coefs, rinv, rinv1, self.ie, self.form, self.nbir, oac=oac,
                  max_subrecord=max_subrecord)

    @classmethod
    def from_evp(cls, path, i2, j2, nb0, form=FORM_BIR, nbir=None, with_oac=None, **kw):
        """A plan loaded from an EVP file, as MAIN starts from PREP's output
        (MAIN/template/GLOBAL/src/input_GLOBAL.f90:438-443): the operator and IE
        come from the record, the influence matrices are uploaded, not rebuilt."""
        rec = evp_read(path, i2, j2, nb0, form, nbir or 0, with_oac=with_oac)
        plan = cls(rec["coefs"], rec["ie"], form=form, nbir=nbir, **kw)
        plan.load_influence(rec["rinv"], rec["rinv1"])
        return plan

    # --------------------------------------------------- Krylov solvers
    def bicgstab_dev(self, b, x, **opts) -> KrylovStats:
        """p_bicgstab_le on device arrays: b (J2, I2) float32 and x
        (J2+2, I2+2) float64 torch tensors (the Fortran arrays, transposed
        views); x is updated in place."""
        self._follow_torch_stream()
        st = KrylovStats()
        o = krylov_opts(KRYLOV_BICGSTAB, **opts)
        _check(lib().sevp_bicgstab_dev(self._h, _p(b), _p(x), ctypes.byref(o), ctypes.byref(st)))
        return st

    def pcsi_dev(self, mineig, maxeig, b, x, **opts) -> KrylovStats:
        """p_csi on device arrays (as bicgstab_dev)."""
        self._follow_torch_stream()
        st = KrylovStats()
        o = krylov_opts(KRYLOV_PCSI, **opts)
        _check(lib().sevp_pcsi_dev(self._h, float(mineig), float(maxeig), _p(b), _p(x),
                                   ctypes.byref(o), ctypes.byref(st)))
        return st

    def lanczos_dev(self, m=10, ref_order=False):
        """lanczos: (mineig, maxeig, alpha[m], sigma[m+1])."""
        self._follow_torch_stream()
        o = krylov_opts(ref_order=ref_order)
        mn, mx = ctypes.c_double(), ctypes.c_double()
        alpha, sigma = np.zeros(m), np.zeros(m + 1)
        _check(lib().sevp_lanczos_dev(self._h, m, ctypes.byref(o), ctypes.byref(mn),
                                      ctypes.byref(mx), _p(alpha), _p(sigma)))
        return mn.value, mx.value, alpha, sigma

    def set_precision(self, precision):
        _check(lib().sevp_set_precision(self._h, precision))
        self.precision = precision

    def set_graphs(self, on: bool):
        _check(lib().sevp_set_graphs(self._h, 1 if on else 0))

    def set_rhs_groups(self, groups: int):
        """Run the refinement passes of `groups` RHS groups concurrently."""
        _check(lib().sevp_set_rhs_groups(self._h, groups))

    def profile_spans(self):
        """(total ms, launches) of the self-timed correction products since
        profile_begin, from the kernels' own clock."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        _check(lib().sevp_profile_spans(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def profile_begin(self, kclass):
        _check(lib().sevp_profile_begin(self._h, kclass))

    def profile_end(self):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_long(0)
        _check(lib().sevp_profile_end(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def reset_seam(self):
        _check(lib().sevp_reset_seam(self._h))

    def solve(self, s, x):
        """Host arrays: s (I2, J2[, nrhs]) float32, x (I2+2, J2+2[, nrhs])
        float64 updated in place."""
        nrhs = s.shape[2] if s.ndim == 3 else 1
        _check(lib().sevp_solve(self._h, _p(_f32(s)), _p(_f64(x)), nrhs))
        return x

    def _follow_torch_stream(self):
        """Enqueue on torch's current stream so the solve is ordered after
        the torch work that produced its inputs."""
        import torch
        # handle 0 is the null stream: the plan's own (blocking) stream is
        # ordered with it; any other stream is used directly
        st = torch.cuda.current_stream().cuda_stream
        if st != self._stream:
            _check(lib().sevp_set_stream(self._h, ctypes.c_void_p(st) if st else None))
            self._stream = st

    def solve_dev(self, s, x, nrhs=None):
        """Device tensors (torch, cuda): s float32 (nrhs, J2, I2) contiguous
        (i.e. Fortran (I2, J2) per RHS), x float64 (nrhs, J2+2, I2+2)."""
        nrhs = nrhs or s.shape[0]
        self._follow_torch_stream()
        _check(lib().sevp_solve_dev(self._h, _p(s), _p(x), nrhs, None))

    def step14_dev(self, desc: Step14Desc, grid: dict, state: dict):
        """FS_GLOBAL step 14 on the device (sevp_step14_dev).  grid and state
        map the sevp_step14_grid / sevp_step14_state field names to torch
        device tensors laid out in the reference's Fortran shapes (a Fortran
        (A, B, C) array is a contiguous tensor of shape (C, B, A)).  Returns
        (iterations, final land speed)."""
        self._follow_torch_stream()
        g = Step14Grid(**{k: _p(grid[k]).value for k, _ in Step14Grid._fields_})
        st = Step14State(**{k: _p(state[k]).value for k, _ in Step14State._fields_})
        it, tmp = ctypes.c_int(), ctypes.c_float()
        _check(lib().sevp_step14_dev(self._h, ctypes.byref(desc), ctypes.byref(g),
                                     ctypes.byref(st), ctypes.byref(it), ctypes.byref(tmp)))
        return it.value, tmp.value

    def solve_refined_dev(self, s, x, tol=1e-10, max_iter=20, nrhs=None):
        nrhs = nrhs or s.shape[0]
        self._follow_torch_stream()
        rel = np.zeros(nrhs)
        st = SevpStats()
        rc = lib().sevp_solve_refined_dev(self._h, _p(s), _p(x), nrhs, tol, max_iter, _p(rel),
                                          ctypes.byref(st))
        if rc not in (SEVP_OK, SEVP_E_NOTCONV):
            _check(rc)
        return rel, st, rc == SEVP_OK

    def residual_dev(self, s, x, nrhs=None):
        nrhs = nrhs or s.shape[0]
        self._follow_torch_stream()
        rel = np.zeros(nrhs)
        _check(lib().sevp_residual_dev(self._h, _p(s), _p(x), nrhs, _p(rel)))
        return rel


class SevpBatch:
    """Independent full-width plans (regional basins) refined together
    (sevp_batch_*): each step of a pass is one launch over every problem,
    and each problem's result is bitwise its plan's solve_refined_dev.  The
    plans must stay open while the batch is used."""

    def __init__(self, plans):
        self.plans = list(plans)
        arr = (VP * len(self.plans))(*[p._h.value if isinstance(p._h, VP) else p._h
                                       for p in self.plans])
        h = VP()
        _check(lib().sevp_batch_create(arr, len(self.plans), ctypes.byref(h)))
        self._h = h

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.sevp_batch_destroy(h)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def solve_refined_dev(self, s, x, tol=1e-10, max_iter=20):
        """s[i] float32 (nrhs_i, J2, I2), x[i] float64 (nrhs_i, J2+2, I2+2),
        torch device tensors of problem i.  Returns (per-problem relative
        residuals, passes per problem, SevpStats, converged)."""
        import torch
        n = len(self.plans)
        if len(s) != n or len(x) != n:
            raise SevpError(SEVP_E_ARG, "one right-hand side and one solution array per plan")
        for p, si, xi in zip(self.plans, s, x):
            ok = (si.is_cuda and xi.is_cuda and si.dtype == torch.float32
                  and xi.dtype == torch.float64 and si.is_contiguous() and xi.is_contiguous()
                  and tuple(si.shape[1:]) == (p.j2, p.i2)
                  and tuple(xi.shape) == (si.shape[0], p.j2 + 2, p.i2 + 2))
            if not ok:
                raise SevpError(SEVP_E_ARG, f"arrays of the {p.i2}x{p.j2} problem: s float32 "
                                "(nrhs, J2, I2) and x float64 (nrhs, J2+2, I2+2), contiguous")
        torch.cuda.synchronize()  # the batch's stream follows the null stream only
        nrhs = (CI * n)(*[int(t.shape[0]) for t in s])
        sp = (VP * n)(*[t.data_ptr() for t in s])
        xp = (VP * n)(*[t.data_ptr() for t in x])
        rel = np.zeros(sum(nrhs))
        passes = np.zeros(n, np.int32)
        st = SevpStats()
        rc = lib().sevp_batch_solve_refined_dev(self._h, sp, xp, nrhs, tol, max_iter, _p(rel),
                                                _p(passes), ctypes.byref(st))
        if rc not in (SEVP_OK, SEVP_E_NOTCONV):
            _check(rc)
        out, o = [], 0
        for k in nrhs:
            out.append(rel[o:o + k])
            o += k
        return out, passes, st, rc == SEVP_OK


# ------------------------------------------------- one-to-one reference calls

def repbir(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, i2, m0, m2, nblk,
           coef_off=0, x_off=0):
    """REPBIR (MAIN/template/GLOBAL/src/sevp.f90:218-328).  ``coef_off`` and
    ``x_off`` are element offsets into the coefficient arrays and X, the
    equivalent of the Fortran caller passing AL(IL,1) and X(IL,1)."""
    cs = [_p(_f32(c), coef_off) for c in (ax, ay, bb, cx, cy)]
    ie32 = _i32(ie)  # kept alive across the call
    _check(lib().sevp_repbir(*cs, _p(rinv), _p(rinv1), _p(dum0), _p(dum1), _p(dum2), _p(_f32(f)),
                             _p(h), _p(_f64(x), x_off), _p(ie32), i0, i2, m0, m2, nblk))


def repbir_period(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, i2, m0, m2,
                  nblk):
    """REPBIR_PERIOD (MAIN/template/GLOBAL/src/sevp.f90:109-217), M0=I0, M2=I2."""
    cs = [_p(_f32(c)) for c in (ax, ay, bb, cx, cy)]
    ie32 = _i32(ie)  # kept alive across the call
    _check(lib().sevp_repbir_period(*cs, _p(rinv), _p(rinv1), _p(dum0), _p(dum1), _p(dum2),
                                    _p(_f32(f)), _p(h), _p(_f64(x)), _p(ie32), i0, i2, m0, m2,
                                    nblk))


def _full(fn, ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, i2, nblk):
    cs = [_p(_f32(c)) for c in (ax, ay, bb, cx, cy)]
    ie32 = _i32(ie)  # kept alive across the call
    _check(fn(*cs, _p(rinv), _p(rinv1), _p(dum0), _p(dum1), _p(dum2), _p(_f32(f)), _p(h),
              _p(_f64(x)), _p(ie32), i0, i2, nblk))


def rep1(*args):
    """REP1 (MAIN/template/GLOBAL/src/sevp.f90:11-107)."""
    _full(lib().sevp_rep1, *args)


def rep(*args):
    """REP (MAIN/template/GLOBAL/src/sevp.f90:335-438)."""
    _full(lib().sevp_rep, *args)


def pre(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i2, m0, m2, nblk, coef_off=0):
    """PRE (PREP/template/GLOBAL/src/prep_GLOBAL.f90:457-513); RINV, RINV1 out."""
    cs = [_p(_f32(c), coef_off) for c in (ax, ay, bb, cx, cy)]
    ie32 = _i32(ie)  # kept alive across the call
    _check(lib().sevp_pre(*cs, _p(_f64(rinv)), _p(_f64(rinv1)), _p(h), _p(ie32), i2, m0, m2,
                          nblk))


def pre1(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i0, i2, nblk):
    """PRE1 (PREP/template/GLOBAL/src/prep_GLOBAL.f90:327-386)."""
    cs = [_p(_f32(c)) for c in (ax, ay, bb, cx, cy)]
    ie32 = _i32(ie)  # kept alive across the call
    _check(lib().sevp_pre1(*cs, _p(_f64(rinv)), _p(_f64(rinv1)), _p(h), _p(ie32), i0, i2,
                           nblk))


def matinv(b):
    """MATINV (PREP/template/GLOBAL/src/prep_GLOBAL.f90:516-568), in place."""
    _check(lib().sevp_matinv(_p(_f64(b)), b.shape[0]))
    return b


def matinv_dev_noalloc(b_dev, n, zero_pivots_dev=None, stream=None):
    """MATINV of the n x n column-major FP64 matrix at device tensor b_dev, in
    place, without allocating (sevp_matinv_dev_noalloc); enqueued on stream
    (a raw HIP stream handle, None: the null stream), not waited for."""
    _check(lib().sevp_matinv_dev_noalloc(_p(b_dev), n, _p(zero_pivots_dev), stream))


# ---------------------------------------------------------------- EVP file

def _evp_desc(i2, j2, ie, form, nbir):
    return _Desc(i2, j2, len(ie), ie.ctypes.data_as(ctypes.POINTER(CI)), form, nbir or 0, 0, 0,
                 1, PREC_FP64, 0)


def evp_shapes(i2, nb0, form, nbir=0):
    """RINV, RINV1 shapes of the EVP record: (IBIR*IBIR, NB0, NBIR) and
    (IBIR*IBIR, NB0-1, NBIR) for BIR (PREP/template/GLOBAL/src/grid_var_GLOBAL.f90:431),
    (I2, I2, NB0) and (I2, I2, NB0-1) otherwise (:436)."""
    if form == FORM_BIR:
        w = 2 * i2 // nbir
        return (w * w, nb0, nbir), (w * w, nb0 - 1, nbir)
    return (i2, i2, nb0), (i2, i2, nb0 - 1)


def evp_write(path, coefs, rinv, rinv1, ie, form=FORM_BIR, nbir=0, oac=None, max_subrecord=0):
    """Write the EVP record: AL,AR,AB,AT,AC,[OAC,]RINV,RINV1,IE, big-endian,
    Fortran unformatted sequential.  ``coefs`` is (AL, AB, AC, AR, AT) as
    everywhere in this package; the file order is the reference's."""
    al, ab, ac, ar, at = [_f32(np.asfortranarray(c, dtype=np.float32)) for c in coefs]
    i2, j2 = al.shape
    ie = _i32(ie)
    s0, s1 = evp_shapes(i2, len(ie), form, nbir)
    rinv = _f64(np.asfortranarray(rinv))
    rinv1 = _f64(np.asfortranarray(rinv1))
    if rinv.size != int(np.prod(s0)) or rinv1.size < int(np.prod(s1)):
        raise ValueError(f"RINV {rinv.shape} / RINV1 {rinv1.shape} do not match {s0} / {s1}")
    if oac is not None:
        oac = _f32(np.asfortranarray(oac, dtype=np.float32))
        if oac.shape != (i2, j2):
            raise ValueError("OAC must be (I2, J2)")
    d = _evp_desc(i2, j2, ie, form, nbir)
    _check(lib().sevp_evp_write(os.fsencode(path), ctypes.byref(d), _p(al), _p(ab), _p(ac),
                                    _p(ar), _p(at), _p(rinv), _p(rinv1), _p(oac), max_subrecord))


def evp_read(path, i2, j2, nb0, form=FORM_BIR, nbir=0, with_oac=None):
    """Read an EVP record written by PREP (no OAC) or in MAIN's layout (with
    OAC); ``with_oac=None`` accepts whichever the record length matches.
    Returns a dict with coefs (AL, AB, AC, AR, AT), rinv, rinv1, ie and oac
    (None when the record has none)."""
    s0, s1 = evp_shapes(i2, nb0, form, nbir)
    coefs = [np.zeros((i2, j2), np.float32, order="F") for _ in range(5)]
    oac = np.zeros((i2, j2), np.float32, order="F")
    rinv = np.zeros(s0, order="F")
    rinv1 = np.zeros(s1, order="F")
    ie = np.zeros(nb0, np.int32)
    had = CI(0)
    mode = -1 if with_oac is None else (1 if with_oac else 0)
    d = _evp_desc(i2, j2, ie, form, nbir)
    _check(lib().sevp_evp_read(os.fsencode(path), ctypes.byref(d), *[_p(c) for c in coefs],
                                   _p(rinv), _p(rinv1) if rinv1.size else None, _p(ie), _p(oac),
                                   mode, ctypes.byref(had)))
    return {"coefs": tuple(coefs), "rinv": rinv, "rinv1": rinv1, "ie": ie,
            "oac": oac if had.value else None}


def operator_hash(coefs, ie, form=FORM_BIR, nbir=0) -> int:
    """The plan cache's key: a 64-bit hash of the shape, IE, form, NBIR and
    the coefficient bits (sevp_operator_hash; host only)."""
    al, ab, ac, ar, at = [_f32(np.asfortranarray(c, dtype=np.float32)) for c in coefs]
    ie = _i32(ie)
    d = _evp_desc(al.shape[0], al.shape[1], ie, form, nbir)
    out = ctypes.c_ulonglong(0)
    _check(lib().sevp_operator_hash(ctypes.byref(d), _p(al), _p(ab), _p(ac), _p(ar), _p(at),
                                    ctypes.byref(out)))
    return out.value


# ------------------------------------------------------------ Krylov solvers

def p_bicgstab_le(coefs, b, x, **opts) -> KrylovStats:
    """P_BICGSTAB_LE (MAIN/template/GLOBAL/src/sevp.f90:443-552) on host
    arrays: b (I2, J2) float32, x (I2+2, J2+2) float64 updated in place."""
    al, ab, ac, ar, at = [_f32(np.asfortranarray(c, dtype=np.float32)) for c in coefs]
    i2, j2 = al.shape
    st = KrylovStats()
    o = krylov_opts(KRYLOV_BICGSTAB, **opts)
    _check(lib().sevp_p_bicgstab_le(_p(al), _p(ab), _p(ac), _p(ar), _p(at), _p(_f32(b)),
                                    _p(_f64(x)), i2, j2, ctypes.byref(o), ctypes.byref(st)))
    return st


def p_csi(coefs, mineig, maxeig, b, x, **opts) -> KrylovStats:
    """P_CSI (MAIN/template/GLOBAL/src/sevp.f90:553-616) on host arrays."""
    al, ab, ac, ar, at = [_f32(np.asfortranarray(c, dtype=np.float32)) for c in coefs]
    i2, j2 = al.shape
    st = KrylovStats()
    o = krylov_opts(KRYLOV_PCSI, **opts)
    _check(lib().sevp_p_csi(_p(al), _p(ab), _p(ac), _p(ar), _p(at), float(mineig), float(maxeig),
                            _p(_f32(b)), _p(_f64(x)), i2, j2, ctypes.byref(o), ctypes.byref(st)))
    return st


def lanczos(coefs, m=10, ref_order=False):
    """LANCZOS (MAIN/template/GLOBAL/src/sevp.f90:618-694): (mineig, maxeig,
    alpha[m], sigma[m+1])."""
    al, ab, ac, ar, at = [_f32(np.asfortranarray(c, dtype=np.float32)) for c in coefs]
    i2, j2 = al.shape
    o = krylov_opts(ref_order=ref_order)
    mn, mx = ctypes.c_double(), ctypes.c_double()
    alpha, sigma = np.zeros(m), np.zeros(m + 1)
    _check(lib().sevp_lanczos(_p(al), _p(ab), _p(ac), _p(ar), _p(at), i2, j2, m, ctypes.byref(o),
                              ctypes.byref(mn), ctypes.byref(mx), _p(alpha), _p(sigma)))
    return mn.value, mx.value, alpha, sigma
