"""GPU parity: the HIP library against the CPU oracle on the same inputs.

Every plain-FP64 path must be bitwise equal to the oracle (which is itself
pinned bitwise to the compiled reference routines, see test_oracle_ref.py).
Mixed-precision paths are checked by their contract: relative residual
<= 1e-10 after refinement, and the distance to the FP64 oracle solution.
"""
import numpy as np
import pytest

from musoc_solver_amd import sevp as S
from musoc_solver_amd import synth
from tests.conftest import all_finite, fields_equal

pytestmark = pytest.mark.gpu


def make(i2, j2, n0, seed=11, metric="mercator", lat_max=90.0, h_land=50.0, ideal=False,
         dydx=1.0):
    if ideal:
        coefs = synth.ideal_operator(i2, j2)
        mask = np.ones((i2, j2), bool)
    else:
        c = synth.make_case(i2, j2, seed, metric=metric, lat_max=lat_max, h_land=h_land,
                            dydx=dydx)
        coefs, mask = c.coefs, c.mask
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, n0))
    return coefs, ie, mask


def rand_x(i2, j2, nrhs, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((i2 + 2, j2 + 2, nrhs)))


BIR_CASES = [
    # (i2, j2, nbir, n0, kwargs)
    (48, 24, 8, 6, dict(metric="mercator")),
    (192, 96, 16, 8, dict(metric="mercator")),
    # the 192x96 layout with 24x24 blocks (IBIR = 24, 24-row blocks) on an
    # operator whose cells are ten times wider than tall, so that the
    # 23-row march grows slowly and the solution stays finite
    (192, 96, 16, 24, dict(metric="mercator", dydx=0.1)),
    (120, 60, 8, 5, dict(metric="latlon", lat_max=80.0)),  # IBIR = 30
    (96, 48, 4, 4, dict(metric="latlon", lat_max=70.0)),  # IBIR = 48: the 64-lane kernel
    (16, 12, 16, 3, dict(metric="latlon", lat_max=60.0)),  # IBIR = 2, the narrowest strip
]


@pytest.mark.parametrize("i2,j2,nbir,n0,kw", BIR_CASES)
def test_bir_prep_and_solve_bitwise(oracle, device, i2, j2, nbir, n0, kw):
    _bir_case(oracle, i2, j2, nbir, n0, kw, finite=True)


def test_bir_24_row_blocks_overflow_pattern(oracle, device):
    """The reference's own test operator with the literal 24x24 blocks
    overflows in the march (the compiled reference returns NaN over the
    interior too): an overflow-pattern check only.  The influence matrices
    are finite and bitwise equal; the solutions have NaN at the same places
    and are bitwise equal elsewhere."""
    _bir_case(oracle, 192, 96, 16, 24, dict(ideal=True), finite=False)


def _bir_case(oracle, i2, j2, nbir, n0, kw, finite):
    coefs, ie, mask = make(i2, j2, n0, **kw)
    nrhs = 3
    print(f"[bir] i2={i2} j2={j2} nbir={nbir} n0={n0} {kw}", flush=True)
    plan = S.SevpPlan(coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=5, nrhs_max=nrhs)
    zp = plan.build_influence()
    rinv, rinv1 = plan.get_influence()
    r_o, r1_o, zp_o = oracle.bir_prep(coefs, ie, nbir)
    assert zp == zp_o
    assert all_finite(r_o) and all_finite(r1_o[:, :len(ie) - 1])
    assert fields_equal(rinv, r_o)
    assert fields_equal(rinv1[:, :len(ie) - 1], r1_o[:, :len(ie) - 1])
    s = synth.make_rhs(mask, 5, nrhs)
    x0 = rand_x(i2, j2, nrhs, 6)
    # two successive calls: the seam scratch XX persists between them
    xg = x0.copy(order="F")
    plan.solve(s, xg)
    xo = x0.copy(order="F")
    xxs = [np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F") for _ in range(nrhs)]
    for r in range(nrhs):
        xr = np.asfortranarray(xo[:, :, r])
        oracle.bir_solve(coefs, s[:, :, r], r_o, r1_o, xr, xxs[r], ie, nbir)
        xo[:, :, r] = xr
    if finite:
        assert all_finite(xo)
    else:
        assert np.isnan(xo[1:-1, 1:-1]).mean() > 0.5  # the overflow this case pins
    assert fields_equal(xg, xo)
    plan.solve(s, xg)
    for r in range(nrhs):
        xr = np.asfortranarray(xo[:, :, r])
        oracle.bir_solve(coefs, s[:, :, r], r_o, r1_o, xr, xxs[r], ie, nbir)
        xo[:, :, r] = xr
    assert fields_equal(xg, xo)
    plan.close()


FULL_CASES = [
    (96, 48, 3, S.FORM_FULL_PERIODIC),
    (96, 48, 4, S.FORM_FULL),
    (40, 24, 6, S.FORM_FULL_PERIODIC),
    (64, 30, 5, S.FORM_FULL_PERIODIC),  # trapezoid tiles wrap around the periodic width
    (192, 40, 3, S.FORM_FULL_PERIODIC),  # order-192 influence matrices: blocked MATINV
]


@pytest.mark.parametrize("i2,j2,n0,form", FULL_CASES)
def test_full_width_bitwise(oracle, device, i2, j2, n0, form):
    import torch
    coefs, ie, mask = make(i2, j2, n0, metric="latlon", lat_max=80.0)
    nrhs = 4
    print(f"[full] i2={i2} j2={j2} n0={n0} form={form}", flush=True)
    plan = S.SevpPlan(coefs, ie, form=form, nrhs_max=nrhs)
    zp = plan.build_influence()
    rinv, rinv1 = plan.get_influence()
    if form == S.FORM_FULL_PERIODIC:
        r_o, r1_o, zp_o = oracle.pre1(coefs, ie)
    else:
        r_o, r1_o, zp_o = oracle.pre(coefs, ie)
    assert zp == zp_o
    assert fields_equal(rinv, r_o)
    assert fields_equal(rinv1[:, :, :len(ie) - 1], r1_o[:, :, :len(ie) - 1])
    s = synth.make_rhs(mask, 7, nrhs)
    x0 = rand_x(i2, j2, nrhs, 8)
    xg = x0.copy(order="F")
    plan.solve(s, xg)
    for r in range(nrhs):
        xr = np.asfortranarray(x0[:, :, r])
        fn = oracle.rep1 if form == S.FORM_FULL_PERIODIC else oracle.rep
        fn(coefs, r_o, r1_o, s[:, :, r], xr, ie)
        assert all_finite(xr)
        assert fields_equal(xg[:, :, r], xr), f"rhs {r}"
    plan.close()


@pytest.mark.parametrize("i2,j2,n0,nrhs", [(96, 48, 3, 40), (192, 40, 3, 20)])
def test_full_width_many_rhs_bitwise(oracle, device, i2, j2, n0, nrhs):
    """The plain pass with several 16-row tiles of right-hand sides per product
    (3 and 2; the tiles of one column tile share an XCD) and K past one
    64-term chunk (the grouped sums and the chunk tail): the first, middle
    and last right-hand sides bit for bit against the oracle's REP1."""
    coefs, ie, mask = make(i2, j2, n0, metric="latlon", lat_max=80.0)
    print(f"[full-many] i2={i2} j2={j2} n0={n0} nrhs={nrhs}", flush=True)
    plan = S.SevpPlan(coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs)
    plan.build_influence()
    r_o, r1_o, _ = oracle.pre1(coefs, ie)
    s = synth.make_rhs(mask, 11, nrhs)
    x0 = rand_x(i2, j2, nrhs, 12)
    xg = x0.copy(order="F")
    plan.solve(s, xg)
    for r in (0, nrhs // 2, nrhs - 1):
        xr = np.asfortranarray(x0[:, :, r])
        oracle.rep1(coefs, r_o, r1_o, s[:, :, r], xr, ie)
        assert all_finite(xr)
        assert fields_equal(xg[:, :, r], xr), f"rhs {r}"
    plan.close()


def test_entry_points_bitwise(oracle, device):
    i2, j2, n0 = 48, 24, 4
    coefs, ie, mask = make(i2, j2, n0, metric="mercator")
    al, ab, ac, ar, at = coefs
    nblk = len(ie)
    s = synth.make_rhs(mask, 9, 1)[:, :, 0].copy(order="F")
    print("[entry] matinv", flush=True)
    rng = np.random.default_rng(3)
    for n in (24, 96):
        b = np.asfortranarray(rng.standard_normal((n, n)) + n * np.eye(n))
        bo, _ = oracle.matinv(b)
        bg = b.copy(order="F")
        S.matinv(bg)
        assert fields_equal(bg, bo), f"matinv n={n}"
    print("[entry] pre1/rep1", flush=True)
    rinv = np.zeros((i2, i2, nblk), order="F")
    rinv1 = np.zeros((i2, i2, nblk - 1), order="F")
    h = np.zeros((i2 + 2, j2 + 2), order="F")
    S.pre1(al, ab, ac, ar, at, rinv, rinv1, h, ie, i2 + 2, i2, nblk)
    r_o, r1_o, _ = oracle.pre1(coefs, ie)
    assert fields_equal(rinv, r_o) and fields_equal(rinv1, r1_o[:, :, :nblk - 1])
    d0 = np.zeros((i2, nblk), order="F")
    d1 = np.zeros(i2)
    x0 = rand_x(i2, j2, 1, 4)[:, :, 0].copy(order="F")
    for fn_g, fn_o, period in ((S.rep1, oracle.rep1, False), (S.repbir_period, oracle.rep1, True)):
        xg = x0.copy(order="F")
        if period:
            fn_g(al, ab, ac, ar, at, rinv, rinv1, d0, d1, d1, s, h, xg, ie, i2 + 2, i2, i2 + 2, i2, nblk)
        else:
            fn_g(al, ab, ac, ar, at, rinv, rinv1, d0, d1, d1, s, h, xg, ie, i2 + 2, i2, nblk)
        xo = x0.copy(order="F")
        fn_o(coefs, r_o, r1_o, s, xo, ie)
        assert fields_equal(xg, xo)
    print("[entry] pre/rep", flush=True)
    S.pre(al, ab, ac, ar, at, rinv, rinv1, h, ie, i2, i2 + 2, i2, nblk)
    r_o, r1_o, _ = oracle.pre(coefs, ie)
    assert fields_equal(rinv, r_o) and fields_equal(rinv1, r1_o[:, :, :nblk - 1])
    xg = x0.copy(order="F")
    S.rep(al, ab, ac, ar, at, rinv, rinv1, d0, d1, d1, s, h, xg, ie, i2 + 2, i2, nblk)
    xo = x0.copy(order="F")
    oracle.rep(coefs, r_o, r1_o, s, xo, ie)
    assert fields_equal(xg, xo)
    print("[entry] strip pre/repbir", flush=True)
    # one strip of width 12 starting at column il, as FS_SOLVER calls it
    nbir, ibir = 8, 12
    il = 13
    off = il - 1
    rv = np.zeros((ibir, ibir, nblk), order="F")
    rv1 = np.zeros((ibir, ibir, nblk - 1), order="F")
    S.pre(al, ab, ac, ar, at, rv, rv1, h, ie, i2, ibir + 2, ibir, nblk, coef_off=off)
    rvo = np.zeros((ibir, ibir, nblk), order="F")
    rv1o = np.zeros((ibir, ibir, nblk), order="F")
    ho = np.zeros((ibir + 2, j2 + 2), order="F")
    from oracle.oracle import P, _ci
    zpo = oracle.lib().orc_pre(*[P(c, off) for c in coefs], P(rvo), P(rv1o), P(ho),
                                P(np.ascontiguousarray(ie, np.int32)), _ci(i2), _ci(ibir + 2),
                                _ci(ibir), _ci(nblk))
    assert fields_equal(rv, rvo) and fields_equal(rv1, rv1o[:, :, :nblk - 1])
    src = np.asfortranarray(s[il - 1:il - 1 + ibir, :])
    xg = x0.copy(order="F")
    dd0 = np.zeros((ibir, nblk), order="F")
    dd1 = np.zeros(ibir)
    S.repbir(al, ab, ac, ar, at, rv, rv1, dd0, dd1, dd1, src, h, xg, ie, i2 + 2, i2, ibir + 2, ibir,
             nblk, coef_off=off, x_off=off)
    xo = x0.copy(order="F")
    oracle.repbir_call([(c, off) for c in coefs], rvo, rv1o, dd0.copy(order="F"), dd1.copy(),
                       dd1.copy(), src, ho, (xo, off), np.ascontiguousarray(ie, np.int32), i2 + 2,
                       i2, ibir + 2, ibir, nblk)
    assert fields_equal(xg, xo)


# Block rows the reference's IE rule never produces: one-row blocks among
# taller ones, a one-row first and last block, every block one row high, and
# a single block.  They exercise every record kind of the strip kernels'
# record order (closures right after closures, backward blocks with no march
# rows, restores that no march row overwrites).
IRREGULAR_IE = [
    (40, [2, 3, 5, 6, 9, 14, 15, 22, 30, 31, 41]),
    (40, list(range(4, 41, 3)) + [41]),
    (40, list(range(2, 42))),
    (12, [13]),
]


@pytest.mark.parametrize("m2", [10, 24, 30, 40])
@pytest.mark.parametrize("case", range(len(IRREGULAR_IE)))
def test_repbir_irregular_blocks_bitwise(oracle, device, m2, case):
    """REPBIR through the C ABI on one strip of width M2 (the generic and the
    width-specialised strip kernels, 32 and 64 lanes) with irregular block
    rows, bitwise against the oracle, influence matrices from the strip PRE."""
    from oracle.oracle import P, _ci
    j2, ie = IRREGULAR_IE[case]
    i2, off = 48, 4
    print(f"[repbir-irregular] m2={m2} ie={ie}", flush=True)
    coefs, _, mask = make(i2, j2, 2, seed=31 + case, metric="latlon", lat_max=70.0)
    al, ab, ac, ar, at = coefs
    ie = np.asarray(ie, np.int32)
    nblk = len(ie)
    h = np.zeros((i2 + 2, j2 + 2), order="F")
    rv = np.zeros((m2, m2, nblk), order="F")
    rv1 = np.zeros((m2, m2, max(nblk - 1, 1)), order="F")
    S.pre(al, ab, ac, ar, at, rv, rv1, h, ie, i2, m2 + 2, m2, nblk, coef_off=off)
    rvo = np.zeros((m2, m2, nblk), order="F")
    rv1o = np.zeros((m2, m2, nblk), order="F")
    ho = np.zeros((m2 + 2, j2 + 2), order="F")
    oracle.lib().orc_pre(*[P(c, off) for c in coefs], P(rvo), P(rv1o), P(ho),
                         P(np.ascontiguousarray(ie, np.int32)), _ci(i2), _ci(m2 + 2), _ci(m2),
                         _ci(nblk))
    assert fields_equal(rv, rvo)
    if nblk > 1:
        assert fields_equal(rv1, rv1o[:, :, :nblk - 1])
    s = synth.make_rhs(mask, 41 + case, 1)[:, :, 0]
    src = np.asfortranarray(s[off:off + m2, :])
    x0 = rand_x(i2, j2, 1, 51 + case)[:, :, 0].copy(order="F")
    xg = x0.copy(order="F")
    dd0 = np.zeros((m2, nblk), order="F")
    dd1 = np.zeros(m2)
    S.repbir(al, ab, ac, ar, at, rv, rv1, dd0, dd1, dd1, src, h, xg, ie, i2 + 2, i2, m2 + 2, m2,
             nblk, coef_off=off, x_off=off)
    xo = x0.copy(order="F")
    oracle.repbir_call([(c, off) for c in coefs], rvo, rv1o, dd0.copy(order="F"), dd1.copy(),
                       dd1.copy(), src, ho, (xo, off), ie, i2 + 2, i2, m2 + 2, m2, nblk)
    assert all_finite(xo)
    assert fields_equal(xg, xo)


@pytest.mark.parametrize("n", [89, 96, 150, 300, 1024, 1440])
def test_matinv_blocked_bitwise(oracle, device, n):
    rng = np.random.default_rng(n)
    print(f"[matinv] n={n}", flush=True)
    b = np.asfortranarray(rng.standard_normal((n, n)) + 0.5 * n * np.eye(n))
    bo, zo = oracle.matinv(b)
    bg = b.copy(order="F")
    S.matinv(bg)
    assert all_finite(bo)
    assert fields_equal(bg, bo)
    # the inverse it computes is an inverse
    assert np.abs(bg @ b - np.eye(n)).max() < 1e-10


def _device_case(i2, j2, n0, nrhs, precision, seed=21):
    import torch
    coefs, ie, mask = make(i2, j2, n0, seed=seed, metric="latlon", lat_max=80.0)
    plan = S.SevpPlan(coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs, precision=precision)
    plan.build_influence()
    s = synth.make_rhs(mask, seed + 1, nrhs)
    s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
    x_d = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
    return plan, coefs, ie, s, s_d, x_d


@pytest.mark.parametrize("precision", [S.PREC_FP64, S.PREC_FP32, S.PREC_FP64_MFMA])
def test_refined_solve_reaches_tolerance(oracle, device, precision):
    import torch
    i2, j2, n0, nrhs = 96, 48, 3, 8
    print(f"[refine] precision={precision}", flush=True)
    plan, coefs, ie, s, s_d, x_d = _device_case(i2, j2, n0, nrhs, precision)
    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=8)
    torch.cuda.synchronize()
    assert ok, (rel, st.iterations)
    assert rel.max() <= 1e-10
    # independent check of the residual with the oracle's p_sqprod
    x = x_d.cpu().numpy().transpose(2, 1, 0)
    r_o, r1_o, _ = oracle.pre1(coefs, ie)
    for r in range(nrhs):
        xr = np.asfortranarray(x[:, :, r])
        assert oracle.rel_residual(coefs, s[:, :, r], xr) <= 1.5e-10
        # error against the FP64 oracle solution (one REP1 pass, h=3)
        xo = np.zeros((i2 + 2, j2 + 2), order="F")
        oracle.rep1(coefs, r_o, r1_o, s[:, :, r], xo, ie)
        err = np.linalg.norm(xr[1:-1, 1:-1] - xo[1:-1, 1:-1]) / np.linalg.norm(xo[1:-1, 1:-1])
        assert err < 1e-8
    plan.close()


@pytest.mark.parametrize("i2", [64, 300, 600])
def test_residual_dev_matches_oracle(oracle, device, i2):
    """The FP64 residual norms against the oracle's p_sqprod, at widths whose
    rows give a thread one, two and three columns (the kernels that load
    that many columns' operands together)."""
    import torch
    j2 = 30
    print(f"[resid] i2={i2}", flush=True)
    plan, coefs, ie, s, s_d, x_d = _device_case(i2, j2, 3, 2, S.PREC_FP64)
    x = rand_x(i2, j2, 2, 1)
    x_d.copy_(torch.from_numpy(np.ascontiguousarray(x.transpose(2, 1, 0))))
    rel = plan.residual_dev(s_d, x_d)
    for r in range(2):
        ro = oracle.rel_residual(coefs, s[:, :, r], np.asfortranarray(x[:, :, r]))
        assert abs(rel[r] - ro) <= 1e-12 * ro
    plan.close()


GOLDEN_CASES = ["bir_harness_192x96_h24", "bir_harness_192x96_h6", "bir_synth_48x24",
                "rep1_synth_48x24", "rep_synth_48x24"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden_vectors_on_gpu(oracle, device, name):
    """The HIP library reproduces the outputs of the compiled reference
    routines stored in tests/golden, bitwise."""
    from tests.test_oracle_golden import coefs_of, load
    meta, d = load(name)
    coefs = coefs_of(meta, d)
    ie = d["ie"]
    print(f"[golden] {name}", flush=True)
    if name.startswith("bir"):
        plan = S.SevpPlan(coefs, ie, form=S.FORM_BIR, nbir=meta["nbir"], nsweep=meta["nsweep"],
                          periodic=meta["periodic"])
        plan.build_influence()
        x = np.asfortranarray(d["x0"]).copy(order="F")
        plan.solve(np.asfortranarray(d["s"]), x)
    else:
        form = S.FORM_FULL_PERIODIC if meta["periodic"] else S.FORM_FULL
        plan = S.SevpPlan(coefs, ie, form=form)
        plan.load_influence(np.asfortranarray(d["rinv"]), np.asfortranarray(d["rinv1"]))
        x = np.asfortranarray(d["x0"]).copy(order="F")
        plan.solve(np.asfortranarray(d["s"]), x)
    if name == "bir_harness_192x96_h24":
        # the compiled reference overflows with 24-row blocks on its own test
        # operator: an overflow-pattern check (NaN at the same places)
        assert np.isnan(d["x"][1:-1, 1:-1]).mean() > 0.5
    else:
        assert all_finite(d["x"])
    assert fields_equal(x, d["x"])
    plan.close()


def test_bir_seam_fix_bitwise(oracle, device):
    i2, j2, nbir, n0 = 96, 48, 8, 4
    coefs, ie, mask = make(i2, j2, n0, metric="latlon", lat_max=80.0)
    plan = S.SevpPlan(coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=3, nrhs_max=2, seam_fix=True)
    plan.build_influence()
    r_o, r1_o, _ = oracle.bir_prep(coefs, ie, nbir)
    s = synth.make_rhs(mask, 31, 2)
    xg = np.zeros((i2 + 2, j2 + 2, 2), order="F")
    plan.solve(s, xg)
    for r in range(2):
        xo = np.zeros((i2 + 2, j2 + 2), order="F")
        xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
        oracle.bir_solve(coefs, s[:, :, r], r_o, r1_o, xo, xx, ie, nbir, nsweep=3, seam_fix=True)
        assert fields_equal(xg[:, :, r], xo)
    plan.close()


@pytest.mark.parametrize("precision", [S.PREC_FP64, S.PREC_FP32, S.PREC_FP64_MFMA])
def test_graph_replays_match_eager_launches(oracle, device, precision):
    """The refinement pass replayed as a captured HIP graph gives bitwise the
    result of launching its kernels one by one, on every replay (each solve
    replays the graph several times with eager residual kernels between)."""
    import torch
    i2, j2, n0, nrhs = 96, 48, 3, 4
    print(f"[graph] precision={precision}", flush=True)
    plan, coefs, ie, s, s_d, x_d = _device_case(i2, j2, n0, nrhs, precision)
    plan.set_graphs(False)
    rel_e, st_e, ok_e = plan.solve_refined_dev(s_d, x_d, tol=1e-12, max_iter=6)
    x_eager = x_d.clone()
    plan.set_graphs(True)
    for rep in range(3):
        x_d.zero_()
        rel_g, st_g, ok_g = plan.solve_refined_dev(s_d, x_d, tol=1e-12, max_iter=6)
        torch.cuda.synchronize()
        assert st_g.iterations == st_e.iterations
        assert torch.equal(x_d, x_eager), f"replay {rep}"
        assert np.array_equal(rel_g, rel_e)
    # a profiled replay times its correction products through event nodes in
    # the graph and still gives the same result
    x_d.zero_()
    plan.profile_begin(S.KCLASS_CORR)
    rel_p, st_p, ok_p = plan.solve_refined_dev(s_d, x_d, tol=1e-12, max_iter=6)
    ms, n = plan.profile_end()
    torch.cuda.synchronize()
    assert torch.equal(x_d, x_eager)
    nb0 = len(ie)
    # forward: one product per closed block; backward: the free-row product of
    # every block plus, in FP64, a separate coupling product below the top
    per_pass = 3 * nb0 - 2 if precision == S.PREC_FP64 else 2 * nb0 - 1
    assert n == st_p.iterations * per_pass, (n, st_p.iterations)
    assert ms > 0.0
    plan.close()


@pytest.mark.parametrize("precision", [S.PREC_FP64, S.PREC_FP32, S.PREC_FP64_MFMA])
def test_rhs_groups_match_one_group(oracle, device, precision):
    """Refinement passes of RHS groups on concurrent streams give bitwise the
    result of one group: every right-hand side's pass is independent of the
    others, and the split-K slicing depends only on I2."""
    import torch
    i2, j2, n0, nrhs = 96, 48, 3, 5
    print(f"[groups] precision={precision}", flush=True)
    plan, coefs, ie, s, s_d, x_d = _device_case(i2, j2, n0, nrhs, precision)
    rel1, st1, ok1 = plan.solve_refined_dev(s_d, x_d, tol=1e-12, max_iter=6)
    torch.cuda.synchronize()
    x1 = x_d.clone()
    for groups in (2, 3):
        for graphs in (True, False):
            plan.set_graphs(graphs)
            plan.set_rhs_groups(groups)
            x_d.zero_()
            rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-12, max_iter=6)
            torch.cuda.synchronize()
            assert st.iterations == st1.iterations, (groups, graphs)
            assert torch.equal(x_d, x1), (groups, graphs)
            assert np.array_equal(rel, rel1)
    plan.close()


@pytest.mark.parametrize("precision", [S.PREC_FP64, S.PREC_FP32, S.PREC_FP64_MFMA])
def test_self_timed_products(oracle, device, precision):
    """Profiled eager solves time every correction product from the GPU clock
    and with HIP events; timing changes no result."""
    import torch
    i2, j2, n0, nrhs = 96, 48, 3, 4
    print(f"[spans] precision={precision}", flush=True)
    plan, coefs, ie, s, s_d, x_d = _device_case(i2, j2, n0, nrhs, precision)
    plan.set_graphs(False)
    rel0, st0, ok0 = plan.solve_refined_dev(s_d, x_d, tol=1e-12, max_iter=6)
    torch.cuda.synchronize()
    x0 = x_d.clone()
    plan.profile_begin(S.KCLASS_CORR)
    x_d.zero_()
    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-12, max_iter=6)
    span_ms, span_n = plan.profile_spans()
    ev_ms, ev_n = plan.profile_end()
    torch.cuda.synchronize()
    assert torch.equal(x_d, x0)
    assert span_n == ev_n > 0
    assert 0.0 < span_ms <= ev_ms
    plan.close()


def test_fp64_mfma_one_pass_and_precision_switching(oracle, device):
    """FP64 corrections on the matrix cores: one refinement pass reaches
    1e-10 (the FP32 products need three), the solution agrees with the plain
    FP64 (reference-order) solve to FP64 rounding, and switching a plan
    between the three arithmetics keeps each one's results."""
    import torch
    i2, j2, n0, nrhs = 192, 96, 2, 8
    print("[fp64-mfma] one pass", flush=True)
    plan, coefs, ie, s, s_d, x_d = _device_case(i2, j2, n0, nrhs, S.PREC_FP64)
    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=4)
    assert ok and st.iterations == 1
    x_plain = x_d.clone()
    results = {}
    for prec in (S.PREC_FP64_MFMA, S.PREC_FP32, S.PREC_FP64_MFMA, S.PREC_FP32):
        plan.set_precision(prec)
        x_d.zero_()
        rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=8)
        torch.cuda.synchronize()
        assert ok and rel.max() <= 1e-10, (prec, rel.max())
        if prec in results:
            assert torch.equal(results[prec][0], x_d) and results[prec][1] == st.iterations
        results[prec] = (x_d.clone(), st.iterations)
    x64, it64 = results[S.PREC_FP64_MFMA]
    assert it64 == 1 and results[S.PREC_FP32][1] >= 2
    d = (x64 - x_plain)[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
    n = x_plain[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
    assert float((d / n).max()) <= 1e-11, float((d / n).max())
    # and against the oracle's REP1 (FP64, the reference's order)
    r_o, r1_o, _ = oracle.pre1(coefs, ie)
    x = x64.cpu().numpy().transpose(2, 1, 0)
    for r in (0, nrhs - 1):
        xo = np.zeros((i2 + 2, j2 + 2), order="F")
        oracle.rep1(coefs, r_o, r1_o, s[:, :, r], xo, ie)
        err = np.linalg.norm(x[1:-1, 1:-1, r] - xo[1:-1, 1:-1]) / np.linalg.norm(xo[1:-1, 1:-1])
        assert err <= 1e-11, err
    plan.close()


def test_products_only_storage(oracle, device):
    """A plan that keeps only the FP32 product operands (built one block at a
    time, or uploaded block by block) refines bitwise like a plan that keeps
    the FP64 influence matrices, and refuses what needs them."""
    import torch
    i2, j2, n0, nrhs = 96, 48, 2, 4
    print("[products-only] build and load", flush=True)
    ref, coefs, ie, s, s_d, x_ref = _device_case(i2, j2, n0, nrhs, S.PREC_FP32)
    rel0, st0, ok0 = ref.solve_refined_dev(s_d, x_ref, tol=1e-10, max_iter=8)
    assert ok0
    rinv, rinv1 = ref.get_influence()
    for mode in ("build", "load"):
        plan = S.SevpPlan(coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs,
                          precision=S.PREC_FP32, storage=S.STORE_PRODUCTS)
        if mode == "build":
            assert plan.build_influence() == 0
        else:
            plan.load_influence(rinv, rinv1)
        x_d = torch.zeros_like(x_ref)
        rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=8)
        torch.cuda.synchronize()
        assert ok and st.iterations == st0.iterations
        assert torch.equal(x_d, x_ref), mode
        with pytest.raises(S.SevpError):
            plan.get_influence()
        with pytest.raises(S.SevpError):
            plan.set_precision(S.PREC_FP64)
        assert plan.device_bytes < ref.device_bytes
        if mode == "build":  # both pass kinds read the same lane-tile operands
            plan.set_pass(S.PASS_SPLITK)
            ref.set_pass(S.PASS_SPLITK)
            x_s, x_d = torch.zeros_like(x_ref), torch.zeros_like(x_ref)
            assert ref.solve_refined_dev(s_d, x_s, tol=1e-10, max_iter=8)[2]
            assert plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=8)[2]
            torch.cuda.synchronize()
            assert torch.equal(x_d, x_s)
            ref.set_pass(S.PASS_FUSED)
        plan.close()
    ref.close()


FUSED_CASES = [
    # (i2, j2, form, IE heights cycle, nrhs): Dirichlet and periodic, block
    # heights 1..4 (the fused tile keeps 16 - 2 (h - 1) columns), I2 neither a
    # multiple of 14 nor of 16, RHS counts across the 16-row blocks
    (200, 60, S.FORM_FULL_PERIODIC, (2,), 5),
    (200, 61, S.FORM_FULL, (2, 3, 1, 4), 21),
    (96, 48, S.FORM_FULL_PERIODIC, (3, 4, 2), 3),
    (64, 20, S.FORM_FULL, (4,), 16),
    (144, 40, S.FORM_FULL_PERIODIC, (2, 1), 40),  # three 16-row tiles: two per workgroup
]


def _ie_cycle(j2, heights):
    ie, row, k = [], 1, 0
    while row + heights[k % len(heights)] < j2 + 1:
        row += heights[k % len(heights)]
        ie.append(row)
        k += 1
    ie.append(j2 + 1)
    return np.array(ie, np.int32)


@pytest.mark.parametrize("i2,j2,form,heights,nrhs", FUSED_CASES)
def test_fused_pass_matches_reference_order(oracle, device, i2, j2, form, heights, nrhs):
    """One pass of the fused kernels with FP64 corrections (FP64 MFMA, whole
    dot products) against the plain FP64 pass of the same plan (the
    reference's order, bitwise the oracle's REP1/REP): the same FP64
    arithmetic summed in another order, so they agree to rounding; an index
    slip in the fused epilogue would show as O(1).  The FP32 fused pass and
    the split-K pass both refine to 1e-10."""
    import torch
    coefs, _, mask = make(i2, j2, 2, metric="latlon", lat_max=45.0)
    ie = _ie_cycle(j2, heights)
    unit = bool(np.any(np.diff(ie) == 1))  # a one-row block after the first
    print(f"[fused] {i2}x{j2} form={form} heights={heights} nrhs={nrhs} nb0={len(ie)}", flush=True)
    plan = S.SevpPlan(coefs, ie, form=form, nrhs_max=nrhs, precision=S.PREC_FP64)
    plan.build_influence()
    s = synth.make_rhs(mask, 41, nrhs)
    x0 = rand_x(i2, j2, nrhs, 42)
    x_plain = x0.copy(order="F")
    plan.solve(s, x_plain)
    plan.set_precision(S.PREC_FP64_MFMA)
    assert plan.pass_kind == S.PASS_FUSED
    x_f = x0.copy(order="F")
    plan.solve(s, x_f)
    assert all_finite(x_f)
    d = np.abs(x_f - x_plain).max() / np.abs(x_plain).max()
    # the rounding level of a correct FP64 pass in another summation order:
    # the split-K FP64 products (another code path) against the reference
    # order; the march and the influence matrices' conditioning amplify it
    # (up to ~1e-8 with 4-row blocks here), while an indexing slip shows as
    # 1e-6 or O(1)
    bound = 1e-9
    if not unit:
        plan.set_pass(S.PASS_SPLITK)
        x_k = x0.copy(order="F")
        plan.solve(s, x_k)
        d_k = np.abs(x_k - x_plain).max() / np.abs(x_plain).max()
        bound = max(1e-9, 100.0 * d_k)
        plan.set_pass(S.PASS_FUSED)
    print(f"[fused] FP64 fused vs reference order: {d:.2e} (bound {bound:.1e})", flush=True)
    assert d <= bound, (d, bound)
    s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
    sols = {}
    kinds = (S.PASS_FUSED,) if unit else (S.PASS_FUSED, S.PASS_SPLITK)
    if unit:  # the split-K pass refuses one-row blocks after the first
        with pytest.raises(S.SevpError, match="at least 2 rows"):
            plan.set_pass(S.PASS_SPLITK)
    for prec in (S.PREC_FP64_MFMA, S.PREC_FP32):
        plan.set_precision(prec)
        for kind in kinds:
            plan.set_pass(kind)
            x_d = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
            rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=10)
            torch.cuda.synchronize()
            print(f"[fused] prec={prec} pass={kind}: {st.iterations} passes, {rel.max():.2e}",
                  flush=True)
            assert ok and rel.max() <= 1e-10, (prec, kind, rel.max())
            sols[(prec, kind)] = x_d.clone()
    ref = sols[(S.PREC_FP64_MFMA, S.PASS_FUSED)]
    for k, v in sols.items():
        e = float((v - ref).abs().max() / ref.abs().max())
        assert e <= 1e-8, (k, e)
    x = sols[(S.PREC_FP32, S.PASS_FUSED)].cpu().numpy().transpose(2, 1, 0)
    for r in (0, nrhs - 1):
        assert oracle.rel_residual(coefs, s[:, :, r], np.asfortranarray(x[:, :, r])) <= 1.5e-10
    plan.close()


@pytest.mark.parametrize("i2,j2", [(96, 48), (192, 96)])
def test_bf16_trial_contraction(oracle, device, i2, j2):
    """BF16-input, FP32-accumulate corrections (the fused pass with BF16
    matrices, the FP32 error rows rounded to BF16 in registers): the residual
    history of one pass at a time, measured on the GPU, against the CPU model
    of the same pass (tests/test_correction_precision.py): at best about 0.2
    per pass, and at 192 x 96 the first pass does not reduce the residual.
    The refinement still reaches 1e-10 on these small grids, in many passes."""
    import torch
    from tests.test_correction_precision import residual_history, round_bf16
    case = synth.make_case(i2, j2, 3, metric="latlon", lat_max=80.0, h_land=50.0)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    print(f"[bf16] {i2}x{j2}", flush=True)
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=1,
                      precision=S.PREC_BF16)
    plan.build_influence()
    s = synth.make_rhs(case.mask, 4, 1)
    s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
    x_d = torch.zeros((1, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
    hist = [1.0]
    for _ in range(4):
        rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=0.0, max_iter=1)
        hist.append(float(rel[0]))
    model = residual_history(oracle, i2, j2, round_bf16, round_bf16, passes=4)
    print(f"[bf16] GPU {np.array(hist)} model {model}", flush=True)
    g = np.array(hist)
    assert np.all(np.abs(np.log(g[1:] / model[1:])) < np.log(3.0))  # same pass, other sums
    c = g[1:] / g[:-1]
    assert c.max() > 0.1  # the first pass: 0.2 at 96 x 48, above 1 at 192 x 96
    if i2 == 192:
        assert g[1] > 1.0
    x_d.zero_()
    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=60)
    print(f"[bf16] to 1e-10: {st.iterations} passes, {rel.max():.2e}", flush=True)
    assert ok and st.iterations > 8
    xr = np.asfortranarray(x_d[0].cpu().numpy().T)
    assert oracle.rel_residual(case.coefs, s[:, :, 0], xr) <= 1.5e-10
    plan.close()


def test_autotune_choices_and_cache(oracle, device, tmp_path):
    """The timing autotuner tries the pass kernels, waves and row tiles per
    fused workgroup and RHS groups, keeps the fastest, and every choice
    refines to the tolerance (row tiles and RHS groups bitwise alike); the
    cached choice is applied without timing."""
    import torch
    from musoc_solver_amd import autotune
    i2, j2, nrhs = 192, 96, 40
    print("[autotune]", flush=True)
    plan, coefs, ie, s, s_d, x_d = _device_case(i2, j2, 2, nrhs, S.PREC_FP32)
    for prec in (S.PREC_FP32, S.PREC_FP64_MFMA):
        plan.set_precision(prec)
        for kind in (S.PASS_FUSED, S.PASS_SPLITK):
            for waves in (8, 16):
                first = None
                for groups, rows in ((1, 1), (1, 2), (3, 2), (3, 1)):  # 3 groups: 14, 14, 12 RHS
                    if kind == S.PASS_SPLITK and rows == 2:
                        continue
                    plan.set_tuning(kind, waves, groups, rows)
                    x_d.zero_()
                    rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=10)
                    torch.cuda.synchronize()
                    assert ok and rel.max() <= 1e-10, (prec, kind, waves, groups, rows)
                    if first is None:
                        first = x_d.clone()
                    else:
                        assert torch.equal(x_d, first), (prec, kind, waves, groups, rows)
    plan.set_precision(S.PREC_FP32)
    r = autotune.tune(plan, s_d, x_d, cache_dir=tmp_path)
    print(f"[autotune] {r}", flush=True)
    assert not r["cached"] and r["ms"] > 0
    t = plan.tuning.as_dict()
    assert (t["pass_kernels"], t["waves"], t["rhs_groups"], t["rows"]) == (
        r["pass_kernels"], r["waves"], r["rhs_groups"], r["rows"])
    plan.set_tuning(S.PASS_SPLITK if r["pass_kernels"] == "fused" else S.PASS_FUSED, 8, 2)
    r2 = autotune.tune(plan, s_d, x_d, cache_dir=tmp_path)
    assert r2["cached"] and plan.tuning.as_dict() == t
    plan.close()
