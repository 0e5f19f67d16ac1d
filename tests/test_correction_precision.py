"""Why the corrections run in FP32 or FP64 and not in BF16 (CPU, oracle).

Iterative refinement x <- x + REP1(s - A x) with the influence matrices
RINV / RINV1 rounded before use (the arithmetic stays FP64, as in the
oracle's REP1 restatement, MAIN/template/GLOBAL/src/sevp.f90:11-107).  The
residual contracts per pass by roughly the rounding of the matrices times the
block march's amplification:

  - exact FP64 matrices: one pass reaches the FP64 floor;
  - FP32 (2^-24): 1e-6 .. 1e-5 per pass on these grids, and about 4e-4 at
    1440 x 720, which is why the FP32 leg needs three passes to reach 1e-10;
  - BF16 (2^-8): no better than 0.2 per pass, and the first pass does not
    reduce the residual at 192 x 96 at all, so BF16 operands cannot reach
    1e-10 in a useful number of passes (and diverge on larger grids, where
    the FP32 contraction is 25x worse).
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


def residual_history(oracle, i2, j2, rounding, passes=4):
    case = synth.make_case(i2, j2, 3, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    rinv, rinv1, _ = oracle.pre1(case.coefs, ie)
    rv, rv1 = np.asfortranarray(rounding(rinv)), np.asfortranarray(rounding(rinv1))
    s = synth.make_rhs(case.mask, 4, 1)[:, :, 0].astype(np.float64)
    x = np.zeros((i2 + 2, j2 + 2), order="F")
    hist = []
    for _ in range(passes + 1):
        res = s - oracle.sqprod(case.coefs, x)[1:-1, 1:-1]
        hist.append(np.linalg.norm(res) / np.linalg.norm(s))
        d = np.zeros((i2 + 2, j2 + 2), order="F")
        oracle.rep1(case.coefs, rv, rv1, np.asfortranarray(res.astype(np.float32)), d, ie)
        x += d
    return np.array(hist)


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
    h64 = residual_history(oracle, i2, j2, lambda a: a, passes=1)
    h32 = residual_history(oracle, i2, j2, round_fp32)
    h16 = residual_history(oracle, i2, j2, round_bf16)
    assert h64[1] < 1e-12                      # exact matrices: one pass
    assert h32[1] < 1e-4 and h32[2] < 1e-9     # FP32: 1e-10 within two or three passes
    c16 = h16[1:] / h16[:-1]
    assert c16.max() > 0.1                     # BF16: at best ~0.2 per pass
    assert h16[-1] > 1e-6                      # far from 1e-10 after four passes
    if i2 == 192:
        assert h16[1] > 1.0                    # the first pass does not reduce the residual
