"""Batched problems (sevp_batch_*): regional basins of different sizes, block
rows and right-hand-side counts refined together, one grouped launch per pass
step.  Each problem's iterate, residuals and pass count must equal its own
plan's sevp_solve_refined_dev bit for bit (the grouped launches run the same
arguments in the same order), for FP32 and FP64 MFMA corrections, periodic
and Dirichlet problems, a problem that starts converged (it must keep its X
and report no pass), and a second call whose RHS counts rebuild the pass.
"""
import numpy as np
import pytest

from musoc_solver_amd import autotune, synth
from musoc_solver_amd import sevp as S
from tests.conftest import fields_equal

pytestmark = pytest.mark.gpu

# (i2, j2, seed, land fraction, periodic, nrhs): I2 <= 1536 share the product slice
BASINS = [
    (64, 64, 1, 0.20, False, 8),
    (136, 96, 2, 0.30, False, 3),
    (200, 152, 3, 0.45, False, 40),  # more than 32 RHS: a second 32-row product group
    (96, 40, 4, 0.25, True, 5),
    (48, 120, 5, 0.35, False, 8),
    (512, 48, 6, 0.30, False, 4),
]


def make_basins(precision):
    out = []
    for i2, j2, seed, land, periodic, nrhs in BASINS:
        case = synth.make_case(i2, j2, 6100 + seed, metric="latlon", lat_max=60.0,
                               land_frac=land, h_land=50.0)
        ie = autotune.adaptive_ie(case.coefs, max_growth=1e3, hmax=8, hmin=2, periodic=periodic)
        form = S.FORM_FULL_PERIODIC if periodic else S.FORM_FULL
        plan = S.SevpPlan(case.coefs, ie, form=form, periodic=periodic, nrhs_max=64,
                          precision=precision)
        plan.set_pass(S.PASS_SPLITK)  # the pass kernels of the batch
        plan.build_influence()
        s = synth.make_rhs(case.mask, 7100 + seed, nrhs)
        out.append((plan, s))
    return out


def dev(plan, s, nrhs, x0=None):
    import torch
    s_d = torch.from_numpy(np.ascontiguousarray(s[:, :, :nrhs].transpose(2, 1, 0))).cuda()
    if x0 is None:
        x_d = torch.zeros((nrhs, plan.j2 + 2, plan.i2 + 2), dtype=torch.float64, device="cuda")
    else:
        x_d = x0.clone()
    return s_d, x_d


@pytest.mark.parametrize("precision", [S.PREC_FP32, S.PREC_FP64_MFMA])
def test_batch_equals_plan_solves(device, precision):
    import torch
    print(f"[batch] precision {precision}", flush=True)
    basins = make_basins(precision)
    try:
        for call, cut in enumerate((None, 2)):  # second call: fewer RHS, the pass is rebuilt
            ref, s_b, x_b = [], [], []
            for k, (plan, s) in enumerate(basins):
                nrhs = s.shape[2] if cut is None else min(cut, s.shape[2])
                s_d, x_d = dev(plan, s, nrhs)
                rel, st, ok = plan.solve_refined_dev(s_d, x_d, tol=1e-10, max_iter=12)
                torch.cuda.synchronize()
                assert ok, (k, rel.max())
                ref.append((x_d.cpu().numpy(), rel, st.iterations))
                s_b.append(s_d)
                # problem 0 starts from its converged solution: no pass, X kept
                x_b.append(x_d.clone() if k == 0 else torch.zeros_like(x_d))
            with S.SevpBatch([p for p, _ in basins]) as batch:
                rel_b, passes, st, ok = batch.solve_refined_dev(s_b, x_b, tol=1e-10, max_iter=12)
            torch.cuda.synchronize()
            print(f"[batch] call {call} passes {list(passes)} (plan solves "
                  f"{[r[2] for r in ref]}) worst {st.max_rel_resid:.2e}", flush=True)
            assert ok
            assert passes[0] == 0 and fields_equal(x_b[0].cpu().numpy(), ref[0][0])
            for k in range(1, len(basins)):
                assert passes[k] == ref[k][2], k
                assert fields_equal(rel_b[k], ref[k][1]), k
                assert fields_equal(x_b[k].cpu().numpy(), ref[k][0]), k
            assert st.iterations == max(r[2] for r in ref[1:])
    finally:
        for p, _ in basins:
            p.close()


def test_batch_follows_a_precision_switch(device):
    """All plans switched from FP32 to FP64 MFMA corrections after the batch
    ran: the next batched solve rebuilds its pass and matches the plans'
    FP64 MFMA solves; a batch whose plans disagree is refused."""
    import torch
    basins = make_basins(S.PREC_FP32)[:3]
    try:
        with S.SevpBatch([p for p, _ in basins]) as batch:
            s_b = [dev(p, s, 2)[0] for p, s in basins]
            x_b = [dev(p, s, 2)[1] for p, s in basins]
            batch.solve_refined_dev(s_b, x_b)
            for p, _ in basins:
                p.set_precision(S.PREC_FP64_MFMA)
            ref = []
            for (p, s), sd in zip(basins, s_b):
                x = torch.zeros((2, p.j2 + 2, p.i2 + 2), dtype=torch.float64, device="cuda")
                rel, st, ok = p.solve_refined_dev(sd, x, tol=1e-10, max_iter=12)
                assert ok
                ref.append((x.cpu().numpy(), st.iterations))
            x_b = [torch.zeros_like(x) for x in x_b]
            rel, passes, st, ok = batch.solve_refined_dev(s_b, x_b, tol=1e-10, max_iter=12)
            torch.cuda.synchronize()
            assert ok and list(passes) == [r[1] for r in ref]
            for x, r in zip(x_b, ref):
                assert fields_equal(x.cpu().numpy(), r[0])
            basins[0][0].set_precision(S.PREC_FP32)
            with pytest.raises(S.SevpError, match="no longer share"):
                batch.solve_refined_dev(s_b, x_b)
    finally:
        for p, _ in basins:
            p.close()


def test_batch_rejects_mismatched_plans(device):
    case = synth.make_case(64, 32, 6200, metric="latlon", lat_max=50.0, h_land=50.0)
    ie = synth.reference_ie(32, 16)
    wide = synth.make_case(1552, 32, 6201, metric="latlon", lat_max=50.0, h_land=50.0)
    plans = {
        "fp32": S.SevpPlan(case.coefs, ie, form=S.FORM_FULL, precision=S.PREC_FP32),
        "fp64": S.SevpPlan(case.coefs, ie, form=S.FORM_FULL, precision=S.PREC_FP64),
        "mfma": S.SevpPlan(case.coefs, ie, form=S.FORM_FULL, precision=S.PREC_FP64_MFMA),
        "wide": S.SevpPlan(wide.coefs, ie, form=S.FORM_FULL, precision=S.PREC_FP32),
        "bir": S.SevpPlan(case.coefs, ie, form=S.FORM_BIR, nbir=8, precision=S.PREC_FP64),
    }
    try:
        with pytest.raises(S.SevpError, match="influence"):
            S.SevpBatch([plans["fp32"]])
        for p in plans.values():
            p.build_influence()
        for bad, msg in ((["fp32", "mfma"], "share the precision"), (["fp64"], "FP32 or FP64_MFMA"),
                         (["bir"], "full-width"), (["fp32", "wide"], "slice length"),
                         (["fp32", "fp32"], "twice")):
            with pytest.raises(S.SevpError, match=msg):
                S.SevpBatch([plans[k] for k in bad])
    finally:
        for p in plans.values():
            p.close()
