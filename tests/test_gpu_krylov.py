"""The device Krylov solvers (sevp_bicgstab_dev, sevp_pcsi_dev,
sevp_lanczos_dev) against the oracle's restatement of
MAIN/template/GLOBAL/src/sevp.f90:443-694.

With ref_order the inner products are sequential sums in p_inprod's order,
and every iterate, iteration count and residual must equal the oracle's bit
for bit: BiCGSTAB through restarts, periodic and Dirichlet, P-CSI, and
Lanczos on a nonsymmetric operator whose halos take part (as the
reference's do).  The fast mode sums in a fixed parallel order: it must be
deterministic, converge to the oracle's solution, and, at the benchmark's
full size, cross-check the SEVP solution as FS_SOLVER uses it
(MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1853,1968).
"""
import numpy as np
import pytest

from musoc_solver_amd import synth
from musoc_solver_amd import sevp as S
from tests.conftest import fields_equal
from tests.test_krylov_oracle import case, hard_case, interior_matrix, symmetric_case

pytestmark = pytest.mark.gpu


def plan_for(coefs):
    i2, j2 = coefs[0].shape
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    return S.SevpPlan(coefs, ie, form=S.FORM_FULL)


def dev(b, x):
    import torch
    bt = torch.from_numpy(np.ascontiguousarray(np.asarray(b, np.float32).T)).cuda()
    xt = torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float64).T)).cuda()
    return bt, xt


def host(xt):
    return np.asfortranarray(xt.cpu().numpy().T)


BICG_CASES = [  # (name, periodic): synthetic operators converge in a few
    ("synth", True),  # iterations, the weakly dominant ones restart
    ("synth", False),
    ("hard", True),
    ("hard", False),
]


def bicg_case(name, periodic):
    if name == "synth":
        coefs, b, x0 = case(48, 24, 3, periodic)
    else:
        coefs, b, x0 = hard_case(32, 20, 5, periodic)
    return coefs, b, x0


@pytest.mark.parametrize("name,periodic", BICG_CASES)
def test_bicgstab_ref_order_bitwise(oracle, device, name, periodic):
    print(f"[bicg-ref] {name} periodic={periodic}", flush=True)
    coefs, b, x0 = bicg_case(name, periodic)
    xo = x0.copy(order="F")
    it_o, res_o = oracle.bicgstab(coefs, b, xo, periodic=periodic)
    with plan_for(coefs) as plan:
        bt, xt = dev(b, x0)
        st = plan.bicgstab_dev(bt, xt, periodic=periodic, ref_order=True)
        assert st.iterations == it_o and st.residual == res_o
        assert fields_equal(host(xt), xo)
    if name == "hard":
        assert it_o > 100  # the case restarts


@pytest.mark.parametrize("name,periodic", BICG_CASES)
def test_bicgstab_fast_converges_and_is_deterministic(oracle, device, name, periodic):
    print(f"[bicg-fast] {name} periodic={periodic}", flush=True)
    coefs, b, x0 = bicg_case(name, periodic)
    xo = x0.copy(order="F")
    it_o, _ = oracle.bicgstab(coefs, b, xo, periodic=periodic)
    with plan_for(coefs) as plan:
        runs = []
        for _ in range(2):
            bt, xt = dev(b, x0)
            st = plan.bicgstab_dev(bt, xt, periodic=periodic)
            runs.append((st.iterations, st.residual, host(xt)))
    (it1, r1, x1), (it2, r2, x2) = runs
    assert it1 == it2 and r1 == r2 and fields_equal(x1, x2)
    assert r1 <= oracle.BICG_TOL
    # the same solution to the accuracy the stop criterion gives
    scale = np.abs(xo[1:-1, 1:-1]).max()
    assert np.abs(x1 - xo).max() <= 1e-7 * scale, np.abs(x1 - xo).max() / scale
    assert abs(it1 - it_o) <= max(3, it_o // 10), (it1, it_o)


def test_bicgstab_returns_at_once_on_a_solution(oracle, device):
    coefs, b, x = case(32, 16, 7, True)
    x[0, 1:-1], x[-1, 1:-1] = x[-2, 1:-1], x[1, 1:-1]
    b = np.asfortranarray(oracle.sqprod(coefs, x)[1:-1, 1:-1].astype(np.float32))
    oracle.bicgstab(coefs, b, x)  # x now solves the REAL(4) system
    with plan_for(coefs) as plan:
        bt, xt = dev(b, x)
        st = plan.bicgstab_dev(bt, xt, ref_order=True)
        assert st.iterations == 0 and st.residual < oracle.BICG_TOL
        assert fields_equal(host(xt), x)


@pytest.mark.parametrize("periodic,kmax", [(False, 40), (True, 25)])
def test_pcsi_ref_order_bitwise(oracle, device, periodic, kmax):
    print(f"[pcsi-ref] periodic={periodic}", flush=True)
    if periodic:
        coefs, b, x0 = hard_case(24, 14, 9, True, gamma=0.5)
    else:
        coefs = symmetric_case(20, 14, 13)
        rng = np.random.default_rng(14)
        b = np.asfortranarray(rng.standard_normal((20, 14)).astype(np.float32))
        x0 = np.zeros((22, 16), order="F")
    lam = np.linalg.eigvals(interior_matrix(coefs, periodic).toarray()).real
    mn, mx = float(lam.min()), float(lam.max())
    # a fixed number of steps, then (symmetric interval, where P-CSI
    # converges) to the reference's tolerance
    for tol in ((0.0,) if periodic else (0.0, oracle.PCSI_TOL)):
        xo = x0.copy(order="F")
        it_o, res_o = oracle.pcsi(coefs, mn, mx, b, xo, periodic=periodic, max_iter=kmax if tol == 0 else 1000,
                                  tol=tol)
        with plan_for(coefs) as plan:
            bt, xt = dev(b, x0)
            st = plan.pcsi_dev(mn, mx, bt, xt, periodic=periodic, ref_order=True,
                               max_iter=kmax if tol == 0 else 1000, tol=tol)
        assert st.iterations == it_o and fields_equal(st.residual, res_o), (st.iterations, it_o)
        assert fields_equal(host(xt), xo)


def test_pcsi_fast_matches(oracle, device):
    coefs = symmetric_case(40, 30, 15)
    lam = np.linalg.eigvalsh(interior_matrix(coefs, False).toarray())
    b = np.asfortranarray(np.random.default_rng(16).standard_normal((40, 30)).astype(np.float32))
    xo = np.zeros((42, 32), order="F")
    it_o, _ = oracle.pcsi(coefs, lam[0], lam[-1], b, xo, periodic=False)
    with plan_for(coefs) as plan:
        bt, xt = dev(b, np.zeros((42, 32), order="F"))
        st = plan.pcsi_dev(lam[0], lam[-1], bt, xt, periodic=False)
    assert st.converged and abs(st.iterations - it_o) <= 2
    assert np.abs(host(xt) - xo).max() <= 1e-6 * np.abs(xo).max()


@pytest.mark.parametrize("kind", ["synth", "symmetric"])
def test_lanczos_ref_order_bitwise(oracle, device, kind):
    print(f"[lanczos-ref] {kind}", flush=True)
    if kind == "synth":  # nonsymmetric, its halos feed the products
        coefs, _, _ = case(40, 24, 17, True)
    else:
        coefs = symmetric_case(16, 12, 18)
    mn_o, mx_o, a_o, _, s_o = oracle.lanczos(coefs, 10)
    with plan_for(coefs) as plan:
        mn, mx, a, s = plan.lanczos_dev(10, ref_order=True)
        mn_f, mx_f, a_f, s_f = plan.lanczos_dev(10)
    assert fields_equal(a, a_o) and fields_equal(s, s_o)
    assert fields_equal(np.array([mn, mx]), np.array([mn_o, mx_o]))
    if kind == "symmetric":
        assert mn_f == pytest.approx(mn_o, rel=1e-9) and mx_f == pytest.approx(mx_o, rel=1e-9)


def test_one_to_one_calls_equal_plan_calls(oracle, device):
    """sevp_p_bicgstab_le / sevp_p_csi / sevp_lanczos (host arrays, operator
    uploaded per call) give the plan calls' bits."""
    coefs, b, x0 = hard_case(32, 20, 5, True)
    x_h = x0.copy(order="F")
    st_h = S.p_bicgstab_le(coefs, b, x_h)
    with plan_for(coefs) as plan:
        bt, xt = dev(b, x0)
        st = plan.bicgstab_dev(bt, xt)
        assert (st.iterations, st.residual) == (st_h.iterations, st_h.residual)
        assert fields_equal(host(xt), x_h)
        lz = plan.lanczos_dev(10)
        lam = np.linalg.eigvals(interior_matrix(coefs, True).toarray()).real
        bt, xt = dev(b, x0)
        sp = plan.pcsi_dev(lam.min(), lam.max(), bt, xt, max_iter=30, tol=0.0)
    lz_h = S.lanczos(coefs, 10)
    assert fields_equal(np.array(lz[:2]), np.array(lz_h[:2])) and fields_equal(lz[2], lz_h[2])
    x_h = x0.copy(order="F")
    sp_h = S.p_csi(coefs, lam.min(), lam.max(), b, x_h, max_iter=30, tol=0.0)
    assert sp.iterations == sp_h.iterations == 30 and fields_equal(host(xt), x_h)


def test_bicgstab_cross_checks_sevp_at_full_size(device):
    """The benchmark's operator (1440x720, +-80 deg, REP1 form): BiCGSTAB and
    the refined SEVP solve agree, as FS_SOLVER's comparison expects."""
    import torch
    i2, j2 = 1440, 720
    print(f"[bicg-vs-sevp] {i2}x{j2}", flush=True)
    c = synth.make_case(i2, j2, 3000, metric="latlon", lat_max=80.0, h_land=50.0)
    s = synth.make_rhs(c.mask, 3001, 1)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    with S.SevpPlan(c.coefs, ie, form=S.FORM_FULL_PERIODIC, precision=S.PREC_FP32) as plan:
        plan.build_influence()
        s_d = torch.from_numpy(np.ascontiguousarray(s.transpose(2, 1, 0))).cuda()
        x_e = torch.zeros((1, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
        rel, _, ok = plan.solve_refined_dev(s_d, x_e, tol=1e-12, max_iter=10)
        assert ok
        x_k = torch.zeros((j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
        st = plan.bicgstab_dev(s_d[0], x_k)
        torch.cuda.synchronize()
    assert st.converged, (st.iterations, st.residual)
    xe = x_e[0].cpu().numpy()
    xk = x_k.cpu().numpy()
    d = np.abs(xe - xk)[1:-1, 1:-1].max() / np.abs(xe[1:-1, 1:-1]).max()
    print(f"[bicg-vs-sevp] iterations {st.iterations} residual {st.residual:.3e} "
          f"max diff {d:.3e} ({st.seconds * 1e3:.1f} ms)", flush=True)
    assert d <= 1e-8
