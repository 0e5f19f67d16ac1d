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
