"""GPU parity at the benchmark configurations' full grid sizes.

The reference's default solver form (FL_EVP_STP = 1: multiplicative sweeps
over overlapping strips, REPBIR per strip) is checked bitwise against the
oracle at the sizes of the 0.5, 0.25 and 0.1 degree configurations, with the
configurations' strip widths (24, 24 and 30 columns) and block height 2 (the
24- and 30-row blocks overflow; the oracle tests pin that).  The 0.1 degree
grid's influence matrices are compared on a subset of strips, including the
seam strip, because the oracle needs 17 s for all 240; its solve runs on the
GPU-built matrices.  The 0.5 degree configuration's sequence of solves is
checked as the time stepping uses it: X and the seam scratch carry over from
one solve to the next while the forcing drifts.  Regional basins (the batched
uneven-domain configuration) use the non-periodic full-width form with block
rows placed by the autotuner: FP64 bitwise where the oracle is quick, and the
mixed-precision refinement to 1e-10 at every size.
"""
import numpy as np
import pytest

from musoc_solver_amd import autotune, synth
from musoc_solver_amd import sevp as S
from tests.conftest import fields_equal

pytestmark = pytest.mark.gpu


def bir_case(i2, j2, seed):
    case = synth.make_case(i2, j2, seed, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    return case, ie


@pytest.mark.parametrize("i2,j2,nbir", [(720, 360, 60), (1440, 720, 120)])
def test_bir_full_size_bitwise(oracle, device, i2, j2, nbir):
    print(f"[bir-full] {i2}x{j2} nbir={nbir}", flush=True)
    case, ie = bir_case(i2, j2, 4000 + i2)
    nrhs = 2
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=5, nrhs_max=nrhs)
    zp = plan.build_influence()
    rinv, rinv1 = plan.get_influence()
    r_o, r1_o, zp_o = oracle.bir_prep(case.coefs, ie, nbir)
    assert zp == zp_o == 0
    assert fields_equal(rinv, r_o)
    assert fields_equal(rinv1, r1_o)
    s = synth.make_rhs(case.mask, 5, nrhs)
    xg = np.zeros((i2 + 2, j2 + 2, nrhs), order="F")
    plan.solve(s, xg)
    for r in range(nrhs):
        xo = np.zeros((i2 + 2, j2 + 2), order="F")
        xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
        oracle.bir_solve(case.coefs, s[:, :, r], r_o, r1_o, xo, xx, ie, nbir)
        assert fields_equal(xg[:, :, r], xo), f"rhs {r}"
    plan.close()


def test_bir_time_sequence_bitwise(oracle, device):
    """Three successive solves with drifting forcing, each starting from the
    previous X, with the seam scratch carried over (0.5 degree grid)."""
    i2, j2, nbir = 720, 360, 60
    print(f"[bir-seq] {i2}x{j2}", flush=True)
    case, ie = bir_case(i2, j2, 4720)
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=5, nrhs_max=1)
    plan.build_influence()
    rinv, rinv1 = plan.get_influence()
    forcing = synth.make_rhs(case.mask, 9, 3, lowfreq=2.0)
    xg = np.zeros((i2 + 2, j2 + 2, 1), order="F")
    xo = np.zeros((i2 + 2, j2 + 2), order="F")
    xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
    for t in range(3):
        s = np.asfortranarray(forcing[:, :, t:t + 1])
        plan.solve(s, xg)
        oracle.bir_solve(case.coefs, s[:, :, 0], rinv, rinv1, xo, xx, ie, nbir)
        assert fields_equal(xg[:, :, 0], xo), f"step {t}"
    plan.close()


def test_bir_tenth_degree(oracle, device):
    """3600x1800 with 30-column strips (NBIR = 240, NB0 = 900)."""
    i2, j2, nbir = 3600, 1800, 240
    print(f"[bir-0.1] {i2}x{j2} nbir={nbir}", flush=True)
    case, ie = bir_case(i2, j2, 5000)
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=5, nrhs_max=1)
    assert plan.build_influence() == 0
    rinv, rinv1 = plan.get_influence()
    strips = [0, 1, 119, nbir - 2, nbir - 1]  # the last one is the seam strip
    r_o, r1_o, _ = oracle.bir_prep_strips(case.coefs, ie, nbir, strips)
    for n in strips:
        assert fields_equal(rinv[:, :, n], r_o[:, :, n]), f"strip {n}"
        assert fields_equal(rinv1[:, :, n], r1_o[:, :, n]), f"strip {n}"
    s = synth.make_rhs(case.mask, 11, 1)
    xg = np.zeros((i2 + 2, j2 + 2, 1), order="F")
    plan.solve(s, xg)
    xo = np.zeros((i2 + 2, j2 + 2), order="F")
    xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
    oracle.bir_solve(case.coefs, s[:, :, 0], rinv, rinv1, xo, xx, ie, nbir)
    assert fields_equal(xg[:, :, 0], xo)
    plan.close()


BASINS = [  # (i2, j2, seed, land fraction): sizes spanning 64x64 .. 512x384
    (64, 64, 1, 0.20),
    (136, 96, 2, 0.30),
    (200, 152, 3, 0.45),
    (512, 384, 4, 0.35),
]


@pytest.mark.parametrize("i2,j2,seed,land", BASINS)
def test_regional_basins(oracle, device, i2, j2, seed, land):
    import torch
    print(f"[basin] {i2}x{j2} land={land}", flush=True)
    case = synth.make_case(i2, j2, 6000 + seed, metric="latlon", lat_max=60.0,
                           land_frac=land, h_land=50.0)
    ie = autotune.adaptive_ie(case.coefs, max_growth=1e3, hmax=8, hmin=2, periodic=False)
    nrhs = 8
    s = synth.make_rhs(case.mask, 7000 + seed, nrhs)
    # FP64: bitwise against the oracle's PRE + REP where the oracle is quick
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL, nrhs_max=nrhs)
    plan.build_influence()
    if i2 <= 200:
        rinv, rinv1 = plan.get_influence()
        r_o, r1_o, _ = oracle.pre(case.coefs, ie)
        assert fields_equal(rinv, r_o)
        assert fields_equal(rinv1[:, :, :len(ie) - 1], r1_o[:, :, :len(ie) - 1])
        xg = np.zeros((i2 + 2, j2 + 2, nrhs), order="F")
        plan.solve(s, xg)
        for r in range(0, nrhs, 3):
            xo = np.zeros((i2 + 2, j2 + 2), order="F")
            oracle.rep(case.coefs, r_o, r1_o, s[:, :, r], xo, ie)
            assert fields_equal(xg[:, :, r], xo), f"rhs {r}"
    # mixed precision, refined to 1e-10
    plan.set_precision(S.PREC_FP32)
    s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
    x_d = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=10)
    torch.cuda.synchronize()
    assert ok and rel.max() <= 1e-10, (rel.max(), st.iterations)
    x = x_d.cpu().numpy().transpose(2, 1, 0)
    for r in (0, nrhs - 1):
        assert oracle.rel_residual(case.coefs, s[:, :, r], np.asfortranarray(x[:, :, r])) <= 1.5e-10
    plan.close()


def test_metgen_global_operator(oracle, device):
    """A global operator on METGEN's square-cell grid (1 degree, +-85 deg,
    layered ocean, periodic), block rows from the autotuner: FP64 bitwise
    against the oracle's PRE1 + REP1, and the mixed-precision refinement to
    1e-10."""
    import torch
    from musoc_solver_amd import metgen, ocean3d
    i2 = 360
    j2 = metgen.rows_to_latitude(60.0, -85.0, 85.0)
    print(f"[metgen] {i2}x{j2}", flush=True)
    oc = ocean3d.make_ocean3d(i2, j2, 5, 8100, metric="metgen", dxmnut=60.0, y0deg=-85.0,
                              periodic=True, nbir=2, land_frac=0.3)
    ie = autotune.adaptive_ie(oc.coefs, max_growth=1e3, hmax=8, hmin=2, periodic=True)
    nrhs = 4
    mask = np.ones((i2, j2), bool)
    s = synth.make_rhs(mask, 8101, nrhs)
    plan = S.SevpPlan(oc.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs)
    plan.build_influence()
    rinv, rinv1 = plan.get_influence()
    r_o, r1_o, _ = oracle.pre1(oc.coefs, ie)
    assert fields_equal(rinv, r_o)
    assert fields_equal(rinv1[:, :, :len(ie) - 1], r1_o[:, :, :len(ie) - 1])
    xg = np.zeros((i2 + 2, j2 + 2, nrhs), order="F")
    plan.solve(s, xg)
    for r in (0, nrhs - 1):
        xo = np.zeros((i2 + 2, j2 + 2), order="F")
        oracle.rep1(oc.coefs, r_o, r1_o, s[:, :, r], xo, ie)
        assert fields_equal(xg[:, :, r], xo), f"rhs {r}"
    plan.set_precision(S.PREC_FP32)
    s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
    x_d = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=10)
    torch.cuda.synchronize()
    assert ok and rel.max() <= 1e-10, (rel.max(), st.iterations)
    plan.close()


def test_quarter_degree_mixed_64rhs(oracle, device):
    """The 0.25 degree configuration as the benchmark runs it: 1440x720
    (+-80 deg), block height 2 (NB0 = 360), 64 right-hand sides in one call,
    FP32 MFMA corrections refined in FP64 to 1e-10.  Checked per RHS against
    the plain FP64 solve of the same plan (the reference's REP1 arithmetic,
    bitwise the oracle at every size the oracle can run) and, for two RHS,
    with the oracle's own p_sqprod residual."""
    import torch
    i2, j2, nrhs = 1440, 720, 64
    print(f"[cfg-0.25] {i2}x{j2} nrhs={nrhs}", flush=True)
    case = synth.make_case(i2, j2, 3000, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    s = synth.make_rhs(case.mask, 3001, nrhs)
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs,
                      precision=S.PREC_FP64)
    assert plan.build_influence() == 0
    s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
    x64 = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
    rel64, st64, ok64 = plan.solve_refined_dev(s_d, x64, tol=1e-10, max_iter=4)
    assert ok64, rel64.max()
    print(f"[cfg-0.25] plain FP64: {st64.iterations} pass(es), max rel resid {rel64.max():.2e}",
          flush=True)
    plan.set_precision(S.PREC_FP32)
    x32 = torch.zeros_like(x64)
    rel, st, ok = plan.solve_refined_dev(s_d, x32, tol=1e-10, max_iter=10)
    torch.cuda.synchronize()
    print(f"[cfg-0.25] mixed: {st.iterations} passes, max rel resid {rel.max():.2e}", flush=True)
    assert ok and rel.max() <= 1e-10, (rel.max(), st.iterations)
    d = (x32 - x64)[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
    n = x64[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
    err = (d / n).cpu().numpy()
    print(f"[cfg-0.25] mixed vs plain FP64: max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-9, err.max()
    x = x32.cpu().numpy().transpose(2, 1, 0)
    for r in (0, nrhs - 1):
        ro = oracle.rel_residual(case.coefs, s[:, :, r], np.asfortranarray(x[:, :, r]))
        assert ro <= 1.5e-10, (r, ro)
    plan.close()


def test_tenth_degree_mixed_16rhs(oracle, device):
    """The 0.1 degree configuration: 3600x1800 (+-80 deg), 16 right-hand
    sides refined to 1e-10 with FP32 MFMA corrections, on one GPU.  With
    2-row blocks (NB0 = 900) the FP64 influence matrices would take 186 GB,
    so the plan keeps only the FP32 product operands (140 GB), built one block
    at a time.  Checked with the oracle's p_sqprod residual on two RHS."""
    import time
    import torch
    i2, j2, nrhs = 3600, 1800, 16
    print(f"[cfg-0.1] {i2}x{j2} nrhs={nrhs}", flush=True)
    case = synth.make_case(i2, j2, 5100, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    s = synth.make_rhs(case.mask, 5101, nrhs)
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs,
                      precision=S.PREC_FP32, storage=S.STORE_PRODUCTS)
    t0 = time.perf_counter()
    assert plan.build_influence() == 0
    print(f"[cfg-0.1] FP32 operands of NB0={len(ie)} blocks built in "
          f"{time.perf_counter() - t0:.1f} s, {plan.device_bytes / 2**30:.1f} GiB", flush=True)
    s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
    x_d = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=10)
    torch.cuda.synchronize()
    print(f"[cfg-0.1] mixed: {st.iterations} passes, max rel resid {rel.max():.2e}, "
          f"{st.seconds:.2f} s", flush=True)
    assert ok and rel.max() <= 1e-10, (rel.max(), st.iterations)
    for r in (0, nrhs - 1):
        xr = np.asfortranarray(x_d[r].cpu().numpy().T)
        ro = oracle.rel_residual(case.coefs, s[:, :, r], xr)
        assert ro <= 1.5e-10, (r, ro)
    plan.close()
