"""PREP then MAIN through the EVP file, on the device.

PREP's role (build the influence matrices, write the record,
PREP/template/GLOBAL/src/prep_GLOBAL.f90:264-320) is SevpPlan.build_influence
plus write_evp; MAIN's (read the record, solve with it,
MAIN/template/GLOBAL/src/input_GLOBAL.f90:438-443) is SevpPlan.from_evp plus
solve.  The solve from the file must equal the oracle's solve with the
oracle's own matrices bit for bit, in both of PREP's and MAIN's layouts.
"""
import numpy as np
import pytest

from musoc_solver_amd import synth
from musoc_solver_amd import sevp as S
from tests.conftest import fields_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_oac", [False, True], ids=["prep-layout", "main-layout"])
def test_bir_prep_file_main(oracle, device, tmp_path, with_oac):
    i2, j2, nbir = 96, 48, 8
    print(f"[evp-bir] {i2}x{j2} nbir={nbir} oac={with_oac}", flush=True)
    case = synth.make_case(i2, j2, 31, metric="latlon", lat_max=75.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    path = str(tmp_path / "evp.bin")
    with S.SevpPlan(case.coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=5) as prep:
        assert prep.build_influence() == 0
        oac = (1.0 / case.coefs[2]).astype(np.float32) if with_oac else None
        prep.write_evp(path, oac=oac)
    r_o, r1_o, _ = oracle.bir_prep(case.coefs, ie, nbir)
    rec = S.evp_read(path, i2, j2, len(ie), S.FORM_BIR, nbir)
    assert (rec["oac"] is not None) == with_oac
    assert fields_equal(rec["rinv"], r_o) and fields_equal(rec["rinv1"], r1_o)
    s = synth.make_rhs(case.mask, 32, 1)
    with S.SevpPlan.from_evp(path, i2, j2, len(ie), form=S.FORM_BIR, nbir=nbir, nsweep=5) as main:
        xg = np.zeros((i2 + 2, j2 + 2, 1), order="F")
        main.solve(s, xg)
    xo = np.zeros((i2 + 2, j2 + 2), order="F")
    xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
    oracle.bir_solve(case.coefs, s[:, :, 0], r_o, r1_o, xo, xx, ie, nbir)
    assert fields_equal(xg[:, :, 0], xo)


def test_full_periodic_prep_file_main(oracle, device, tmp_path):
    i2, j2 = 64, 40
    print(f"[evp-rep1] {i2}x{j2}", flush=True)
    case = synth.make_case(i2, j2, 33, metric="latlon", lat_max=70.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    path = str(tmp_path / "evp.bin")
    with S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC) as prep:
        prep.build_influence()
        prep.write_evp(path)
    s = synth.make_rhs(case.mask, 34, 2)
    with S.SevpPlan.from_evp(path, i2, j2, len(ie), form=S.FORM_FULL_PERIODIC, nrhs_max=2) as main:
        xg = np.zeros((i2 + 2, j2 + 2, 2), order="F")
        main.solve(s, xg)
    r_o, r1_o, _ = oracle.pre1(case.coefs, ie)
    for r in range(2):
        xo = np.zeros((i2 + 2, j2 + 2), order="F")
        oracle.rep1(case.coefs, r_o, r1_o, s[:, :, r], xo, ie)
        assert fields_equal(xg[:, :, r], xo), f"rhs {r}"


def test_plan_cache_miss_hit_and_collision(oracle, device, tmp_path):
    """sevp_plan_create_cached: a miss builds and stores, a hit loads the same
    matrices (bitwise equal solves), a different operator misses, and a file
    under an operator's key holding another operator (a collision) is caught
    by the stored-operator comparison, rebuilt and replaced."""
    import os
    import shutil
    i2, j2, nbir = 96, 48, 8
    print(f"[plan-cache] {i2}x{j2} nbir={nbir}", flush=True)
    case = synth.make_case(i2, j2, 41, metric="latlon", lat_max=75.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    s = synth.make_rhs(case.mask, 42, 1)
    kw = dict(form=S.FORM_BIR, nbir=nbir, nsweep=5, cache_dir=str(tmp_path))

    def solve(plan):
        x = np.zeros((i2 + 2, j2 + 2, 1), order="F")
        plan.solve(s, x)
        return x

    with S.SevpPlan(case.coefs, ie, **kw) as p1:
        assert p1.cache_hit == 0
        x1 = solve(p1)
    key = S.operator_hash(case.coefs, ie, S.FORM_BIR, nbir)
    path = tmp_path / f"sevp_{key:016x}.evp"
    assert path.exists() and os.listdir(tmp_path) == [path.name]
    with S.SevpPlan(case.coefs, ie, **kw) as p2:
        assert p2.cache_hit == 1
        assert fields_equal(solve(p2), x1)
    r_o, r1_o, _ = oracle.bir_prep(case.coefs, ie, nbir)
    xo = np.zeros((i2 + 2, j2 + 2), order="F")
    xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
    oracle.bir_solve(case.coefs, s[:, :, 0], r_o, r1_o, xo, xx, ie, nbir)
    assert fields_equal(x1[:, :, 0], xo)
    # another operator (one coefficient one ulp away) misses
    coefs2 = [np.asfortranarray(c, dtype=np.float32).copy(order="F") for c in case.coefs]
    coefs2[0].reshape(-1, order="F").view(np.uint32)[123] ^= 1
    key2 = S.operator_hash(coefs2, ie, S.FORM_BIR, nbir)
    assert key2 != key
    # plant operator 1's record under operator 2's key: a collision
    path2 = tmp_path / f"sevp_{key2:016x}.evp"
    shutil.copy(path, path2)
    with S.SevpPlan(coefs2, ie, **kw) as p3:
        assert p3.cache_hit == 0
        x3 = solve(p3)
    got = S.evp_read(str(path2), i2, j2, len(ie), S.FORM_BIR, nbir)
    assert fields_equal(got["coefs"][0], coefs2[0])
    with S.SevpPlan(coefs2, ie, **kw) as p4:
        assert p4.cache_hit == 1
        assert fields_equal(solve(p4), x3)
    # an unwritable cache directory: built, not stored, still solves
    with S.SevpPlan(case.coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=5,
                    cache_dir=str(tmp_path / "missing" / "dir")) as p5:
        assert p5.cache_hit == -1
        assert fields_equal(solve(p5), x1)
