"""The plan cache's key (sevp_operator_hash, host only).

The influence matrices depend on the grid shape, IE, the form, NBIR (BIR
strips) and every bit of AL..AT, and on nothing else; the key must follow
exactly that: equal for equal inputs, different when any one of them moves
(a single flipped low bit of one coefficient included), unchanged by the
solve options.  The device half (a hit loads the same matrices, a collision
is caught by comparing the stored operator) is in tests/test_gpu_evp.py.
"""
import numpy as np
import pytest

from musoc_solver_amd import sevp as S
from musoc_solver_amd import synth


@pytest.fixture(scope="module")
def op():
    case = synth.make_case(40, 24, 3, metric="latlon", lat_max=70.0)
    ie = synth.reference_ie(24, synth.nb0_for_height(24, 2))
    return [np.asfortranarray(c, dtype=np.float32) for c in case.coefs], np.asarray(ie, np.int32)


def flip_low_bit(a, idx):
    b = a.copy(order="F")
    v = b.reshape(-1, order="F").view(np.uint32)
    v[idx] ^= 1
    return b


def test_deterministic_and_copy_invariant(op):
    coefs, ie = op
    h = S.operator_hash(coefs, ie, S.FORM_BIR, 4)
    assert h == S.operator_hash([c.copy(order="F") for c in coefs], ie.copy(), S.FORM_BIR, 4)


def test_every_input_moves_the_key(op):
    coefs, ie = op
    base = S.operator_hash(coefs, ie, S.FORM_BIR, 4)
    seen = {base}
    rng = np.random.default_rng(0)
    pairs = {(k, int(i)) for k in range(5) for i in rng.integers(0, coefs[k].size, 40)}
    for k, idx in pairs:
        c = list(coefs)
        c[k] = flip_low_bit(coefs[k], idx)
        seen.add(S.operator_hash(c, ie, S.FORM_BIR, 4))
    assert len(seen) == 1 + len(pairs)  # every single-bit change gives its own key
    swapped = [coefs[3], coefs[1], coefs[2], coefs[0], coefs[4]]  # AL <-> AR
    ie2 = ie.copy()
    ie2[0] += 1
    others = [S.operator_hash(swapped, ie, S.FORM_BIR, 4), S.operator_hash(coefs, ie2, S.FORM_BIR, 4),
              S.operator_hash(coefs, ie, S.FORM_BIR, 8), S.operator_hash(coefs, ie, S.FORM_FULL, 0),
              S.operator_hash(coefs, ie, S.FORM_FULL_PERIODIC, 0)]
    assert len(set(others) | seen) == len(seen) + len(others)


def test_nbir_ignored_outside_bir(op):
    coefs, ie = op
    assert S.operator_hash(coefs, ie, S.FORM_FULL, 0) == S.operator_hash(coefs, ie, S.FORM_FULL, 8)


def test_rejects_bad_descriptor(op):
    coefs, ie = op
    with pytest.raises(S.SevpError, match="NBIR"):
        S.operator_hash(coefs, ie, S.FORM_BIR, 7)
