"""Why the corrections run in FP32 or FP64 and not in BF16 (CPU model).

Iterative refinement x <- x + P(s - A x) where P is the mixed-precision
pass as the GPU runs it: REP1's marches and closure residuals in FP64
(MAIN/template/GLOBAL/src/sevp.f90:11-107), and its three corrections per
block -- e RINV1_b forward, e RINV_b and e W_b backward, with
W_b = RINV_b diag(CY(:, IE(b-1)-1)) RINV1_(b-1) -- as products whose BOTH
operands, the error row e and the influence matrix, are rounded to the
trial precision and whose sums are accumulated in FP32 (the matrix cores'
accumulator).  The matrices come from the oracle's PRE1; the residual from
the oracle's p_sqprod.  The residual contracts per pass by roughly the
operands' rounding times the block march's amplification:

  - FP32 (2^-24): 1e-5 .. 4e-5 per pass on these grids (the GPU measures
    about 4e-4 at 1440 x 720), so 1e-10 takes two or three passes;
  - BF16 (2^-8): no better than 0.2 per pass, and at 192 x 96 the first pass
    increases the residual, so BF16 operands cannot reach 1e-10 in a useful
    number of passes.  Rounding only the matrices gives nearly the same
    contraction: the matrices' rounding dominates.
"""
import numpy as np
import pytest

from musoc_solver_amd import synth


def round_bf16(a):
    """Round FP64 values to BF16 (via FP32, round to nearest even), back to FP64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def round_fp32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def mixed_pass(coefs, rinv, rinv1, w, f, ie, rnd, acc=np.float32):
    """One REP1 pass from x = 0 on the FP64 right-hand side f (I2, J2), with
    the corrections of the GPU's mixed pass: rnd(e) @ M with sums in `acc`
    (M already rounded).  Returns X (I2+2, J2+2) with periodic halo columns."""
    ax, ay, bb, cx, cy = [np.asarray(c, np.float64) for c in coefs]
    i2, j2 = ax.shape
    nb0 = len(ie)
    X = np.zeros((i2, j2 + 3))  # X[:, r]: interior of Fortran row r = 1 .. J2+2

    def prod(e, m):
        return (rnd(e).astype(acc) @ m.astype(acc)).astype(np.float64)

    def row(j, south, c, north=None):
        # equation row j: west/east of the centre row by periodic wrap
        t = (f[:, j - 1] - ax[:, j - 1] * np.roll(c, 1) - ay[:, j - 1] * south
             - bb[:, j - 1] * c - cx[:, j - 1] * np.roll(c, -1))
        return t / cy[:, j - 1] if north is None else t - cy[:, j - 1] * north

    d, e_top = None, None
    for b in range(1, nb0 + 1):  # forward
        js, ieb = (1 if b == 1 else ie[b - 2]), ie[b - 1]
        south = X[:, js] - (d if d is not None else 0.0)  # the corrected row IE(b-1)
        for j in range(js, ieb - 1):
            X[:, j + 2] = row(j, south, X[:, j + 1])
            south = X[:, j + 1]
        e = row(ieb - 1, south, X[:, ieb], X[:, ieb + 1])
        if b < nb0:
            d = prod(e, rinv1[:, :, b - 1])
        else:
            e_top = e
    for b in range(nb0, 0, -1):  # backward (X(IE(b)) was never overwritten)
        js, ieb = (1 if b == 1 else ie[b - 2]), ie[b - 1]
        e = e_top if b == nb0 else row(ieb - 1, X[:, ieb - 1], X[:, ieb], X[:, ieb + 1])
        c = prod(e, rinv[:, :, b - 1])
        hs = prod(e, w[b - 2]) if b > 1 else np.zeros(i2)
        X[:, js + 1] += c
        hc = c
        for j in range(js, ieb - 1):
            hn = (-(ax[:, j - 1] * np.roll(hc, 1)) - ay[:, j - 1] * hs - bb[:, j - 1] * hc
                  - cx[:, j - 1] * np.roll(hc, -1)) / cy[:, j - 1]
            X[:, j + 2] += hn
            hs, hc = hc, hn
    out = np.zeros((i2 + 2, j2 + 2), order="F")
    out[1:-1, 1:-1] = X[:, 2:j2 + 2]
    out[0, :], out[-1, :] = out[-2, :], out[1, :]
    return out


def couplings(rinv, rinv1, coefs, ie):
    """W_b = RINV_b diag(CY(:, IE(b-1)-1)) RINV1_(b-1), b = 2 .. NB0 (FP64)."""
    cy = np.asarray(coefs[4], np.float64)
    return [rinv[:, :, b - 1] @ np.diag(cy[:, ie[b - 2] - 2]) @ rinv1[:, :, b - 2]
            for b in range(2, len(ie) + 1)]


def residual_history(oracle, i2, j2, rnd_matrix, rnd_error, passes=4):
    case = synth.make_case(i2, j2, 3, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    rinv, rinv1, _ = oracle.pre1(case.coefs, ie)
    w = couplings(rinv, rinv1, case.coefs, ie)
    rv, rv1, ww = rnd_matrix(rinv), rnd_matrix(rinv1), [rnd_matrix(m) for m in w]
    s = synth.make_rhs(case.mask, 4, 1)[:, :, 0].astype(np.float64)
    x = np.zeros((i2 + 2, j2 + 2), order="F")
    hist = []
    for k in range(passes + 1):
        res = s - oracle.sqprod(case.coefs, x)[1:-1, 1:-1]
        hist.append(np.linalg.norm(res) / np.linalg.norm(s))
        if k < passes:
            x += mixed_pass(case.coefs, rv, rv1, ww, res, ie, rnd_error)
    return np.array(hist)


def test_model_pass_is_rep1_in_fp64(oracle):
    """With FP64 operands and FP64 sums the model pass is REP1 (the
    oracle's) to rounding, so its structure (the corrected row never stored,
    the coupling as one product with W) is the reference's algorithm; a
    planted slip (W built from RINV1 of the wrong block) breaks it."""
    i2, j2 = 48, 24
    case = synth.make_case(i2, j2, 3, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    rinv, rinv1, _ = oracle.pre1(case.coefs, ie)
    w = couplings(rinv, rinv1, case.coefs, ie)
    s = synth.make_rhs(case.mask, 4, 1)[:, :, 0]
    xo = np.zeros((i2 + 2, j2 + 2), order="F")
    oracle.rep1(case.coefs, rinv, rinv1, np.asfortranarray(s), xo, ie)
    ident = lambda a: np.asarray(a, np.float64)
    x = mixed_pass(case.coefs, rinv, rinv1, w, s.astype(np.float64), ie, ident, np.float64)
    scale = np.abs(xo[1:-1, 1:-1]).max()
    assert np.abs(x[1:-1, 1:-1] - xo[1:-1, 1:-1]).max() <= 1e-11 * scale
    bad = [w[0]] + w[:-1]
    xb = mixed_pass(case.coefs, rinv, rinv1, bad, s.astype(np.float64), ie, ident, np.float64)
    assert np.abs(xb[1:-1, 1:-1] - xo[1:-1, 1:-1]).max() > 1e-6 * scale


def test_bf16_rounding_is_exact_on_representable_values():
    v = np.array([1.0, -2.5, 0.15625, 3.0 * 2.0 ** -20])
    assert np.array_equal(round_bf16(v), v)
    # BF16 keeps 8 significant bits (ulp 2^-7 at 1): 1 + 2^-8 is a tie and goes
    # to the even 1.0; 1 + 3 * 2^-8 is a tie between 1 + 2^-7 (odd) and 1 + 2^-6
    assert round_bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert round_bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert round_bf16(np.array([1.0 + 2.0 ** -9]))[0] == 1.0


@pytest.mark.parametrize("i2,j2", [(96, 48), (192, 96)])
def test_contraction_per_pass_by_operand_precision(oracle, i2, j2):
    h32 = residual_history(oracle, i2, j2, round_fp32, round_fp32)
    h16 = residual_history(oracle, i2, j2, round_bf16, round_bf16)
    h16m = residual_history(oracle, i2, j2, round_bf16, round_fp32)  # BF16 matrices only
    assert h32[1] < 1e-4 and h32[2] < 1e-9     # FP32: 1e-10 within two or three passes
    c16 = h16[1:] / h16[:-1]
    assert c16.max() > 0.1                     # BF16: at best ~0.2 per pass
    assert h16[-1] > 1e-6                      # far from 1e-10 after four passes
    # the matrices' rounding dominates: the error rows' adds little
    assert np.all(np.abs(np.log(h16[1:] / h16m[1:])) < 0.2)
    if i2 == 192:
        assert h16[1] > 1.0                    # the first pass does not reduce the residual
