"""The Fortran drop-in layer: a Fortran program that calls PRE, REPBIR, PRE1
and REP1 the way the reference's PREP and FS_SOLVER do (array-element
offsets, seam gathers, periodic halo refresh), and FS_SOLVER's Krylov
cross-checks P_BICGSTAB_LE, LANCZOS and P_CSI, gets the oracle's results
bitwise, through ISO_C_BINDING into the HIP library."""
import os
import subprocess

import numpy as np
import pytest

from musoc_solver_amd import synth
from tests.conftest import fields_equal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "dropin_driver.exe")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_dropin_driver_bitwise(oracle, device, tmp_path):
    i2, j2, nbir, n0, nsweep = 48, 24, 8, 4, 3
    case = synth.make_case(i2, j2, 61, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, n0)).astype(np.int32)
    s = synth.make_rhs(case.mask, 62, 1)[:, :, 0].copy(order="F")
    x0 = np.asfortranarray(np.random.default_rng(63).standard_normal((i2 + 2, j2 + 2)))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array([i2, j2, nbir, len(ie), nsweep], np.int32).tofile(fh)
        ie.tofile(fh)
        for a in (*case.coefs, s):
            np.asarray(a, np.float32).ravel(order="F").tofile(fh)
        x0.ravel(order="F").tofile(fh)
    print("[fortran] running the drop-in driver", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nb0, ibir = len(ie), 2 * i2 // nbir
    sizes = [ibir * ibir * nb0 * nbir, ibir * ibir * (nb0 - 1) * nbir, (i2 + 2) * (j2 + 2),
             i2 * i2 * nb0, i2 * i2 * (nb0 - 1), (i2 + 2) * (j2 + 2), (i2 + 2) * (j2 + 2), 2,
             (i2 + 2) * (j2 + 2)]
    out = np.fromfile(fout, np.float64)
    assert out.size == sum(sizes)
    parts = np.split(out, np.cumsum(sizes)[:-1])
    rinv = parts[0].reshape((ibir * ibir, nb0, nbir), order="F")
    rinv1 = parts[1].reshape((ibir * ibir, nb0 - 1, nbir), order="F")
    x = parts[2].reshape((i2 + 2, j2 + 2), order="F")
    rf = parts[3].reshape((i2, i2, nb0), order="F")
    rf1 = parts[4].reshape((i2, i2, nb0 - 1), order="F")
    xf = parts[5].reshape((i2 + 2, j2 + 2), order="F")
    xk = parts[6].reshape((i2 + 2, j2 + 2), order="F")
    emin, emax = parts[7]
    xc = parts[8].reshape((i2 + 2, j2 + 2), order="F")
    r_o, r1_o, _ = oracle.bir_prep(case.coefs, ie, nbir)
    assert fields_equal(rinv, r_o) and fields_equal(rinv1, r1_o[:, :nb0 - 1])
    xo = x0.copy(order="F")
    xx = np.zeros((ibir + 2, j2 + 2), order="F")
    oracle.bir_solve(case.coefs, s, r_o, r1_o, xo, xx, ie, nbir, nsweep=nsweep)
    assert fields_equal(x, xo)
    f_o, f1_o, _ = oracle.pre1(case.coefs, ie)
    assert fields_equal(rf, f_o) and fields_equal(rf1, f1_o[:, :, :nb0 - 1])
    xfo = np.zeros((i2 + 2, j2 + 2), order="F")
    oracle.rep1(case.coefs, f_o, f1_o, s, xfo, ie)
    assert fields_equal(xf, xfo)
    # FS_SOLVER's cross-checks: P_BICGSTAB_LE, LANCZOS, P_CSI (reference order)
    xko = np.zeros((i2 + 2, j2 + 2), order="F")
    oracle.bicgstab(case.coefs, s, xko, periodic=True)
    assert fields_equal(xk, xko)
    mn, mx, _, _, _ = oracle.lanczos(case.coefs, 10)
    assert fields_equal(np.array([emin, emax]), np.array([mn, mx]))
    xco = np.zeros((i2 + 2, j2 + 2), order="F")
    oracle.pcsi(case.coefs, mn, mx, s, xco, periodic=True)
    assert fields_equal(xc, xco)
