"""GPU step-14 caller (sevp_step14_dev) against the oracle's step-14 loop:
land masking, RHS assembly, the solve (FS_SOLVER: full-width REP1 / REP or
the BIR sweeps, warm-started from X), P0 / U / V / SCR corrections and the
land-speed stopping test, bitwise after every iteration's state."""
import numpy as np
import pytest

from musoc_solver_amd import ocean3d, synth
from musoc_solver_amd import sevp as S
from tests.conftest import fields_equal

pytestmark = pytest.mark.gpu

CASES = [
    # (form, free_surface, periodic_we, lopenw, harness, mxmask, tolerance)
    ("rep1", True, True, 1, False, 3, 0.0),
    ("rep1", False, True, 1, False, 2, 0.0),
    ("rep", True, False, 0, False, 3, 0.0),
    ("rep", True, False, 0, True, 2, 0.0),
    ("bir", False, True, 1, False, 2, 0.0),
    ("rep1", True, True, 1, False, 18, 1e9),  # stops after the first iteration
]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda()


def host(t, like):
    return t.cpu().numpy().reshape(like.shape, order="F")


@pytest.mark.parametrize("form,fs,pwe,lopenw,harness,mxmask,tol", CASES)
def test_step14_matches_oracle(oracle, device, form, fs, pwe, lopenw, harness, mxmask, tol):
    import torch
    i2, j2, k1 = 48, 32, 4
    nbir = 8 if form == "bir" else 1
    print(f"[step14] {form} fs={fs} pwe={pwe} harness={harness} mxmask={mxmask}", flush=True)
    oc = ocean3d.make_ocean3d(i2, j2, k1, 21, free_surface=fs, periodic=(form != "rep"), nbir=nbir)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    fmap = {"rep1": S.FORM_FULL_PERIODIC, "rep": S.FORM_FULL, "bir": S.FORM_BIR}
    kw = dict(nbir=nbir, nsweep=5) if form == "bir" else {}
    plan = S.SevpPlan(oc.coefs, ie, form=fmap[form], nrhs_max=1, **kw)
    plan.build_influence()
    rinv, rinv1 = plan.get_influence()
    u, v, w, p0 = ocean3d.make_state(oc, 22)
    x = np.asfortranarray(np.random.default_rng(23).standard_normal((i2 + 2, j2 + 2)) * 1e-3)
    s = np.zeros((i2, j2), np.float32, order="F")
    scr = np.zeros((i2 + 2, j2 + 2, 4), np.float32, order="F")
    odt = np.float32(1.0 / 1800.0)
    # device run
    grid = {k: dev(getattr(oc, k)) for k in ("kb", "iu0", "iv0", "iu", "iv", "odx", "ody", "ocs",
                                             "csv", "odyv", "odz")}
    state = {"u": dev(u), "v": dev(v), "w": dev(w), "p0": dev(p0), "scr": dev(scr), "s": dev(s),
             "x": dev(x)}
    desc = S.Step14Desc(i2=i2, j2=j2, k1=k1, free_surface=int(fs), periodic_we=int(pwe),
                        lopenw=lopenw, harness=int(harness), mxmask=mxmask, tolerance=tol,
                        odt=float(odt))
    it_g, tmp_g = plan.step14_dev(desc, grid, state)
    torch.cuda.synchronize()
    # the oracle's loop on the same inputs
    xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
    it_o, tmp_o = 0, 0.0
    while True:
        it_o += 1
        oracle.step14_prepare(oc, u, v, w, s, mask=True, lopenw=lopenw, harness=harness,
                              free_surface=fs)
        if form == "bir":
            oracle.bir_solve(oc.coefs, s, rinv, rinv1, x, xx, ie, nbir)
        elif form == "rep1":
            oracle.rep1(oc.coefs, rinv, rinv1, s, x, ie)
        else:
            oracle.rep(oc.coefs, rinv, rinv1, s, x, ie)
        tmp_o = oracle.step14_correct(oc, x, p0, scr, u, v, odt, periodic_we=pwe)
        if tmp_o <= tol or it_o >= mxmask:
            break
    assert it_g == it_o
    assert np.float32(tmp_g) == np.float32(tmp_o)
    for name, ref in (("u", u), ("v", v), ("w", w), ("p0", p0), ("scr", scr), ("s", s), ("x", x)):
        assert fields_equal(host(state[name], ref), ref), name
    plan.close()
