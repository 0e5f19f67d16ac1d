"""The oracle's step-14 caller (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1154-1283)
pinned by what its arithmetic must achieve.

The step projects the barotropic flow: the RHS S is the depth-integrated
divergence of the layer velocities, the solve gives X with A X = S, and the
velocity correction -grad X (per face, on wet layers) changes the divergence
by -D X, where D is the divergence-of-gradient part of the operator the PREP
assembly builds from the same metrics and masks (musoc_solver_amd.ocean3d).
So after one correction the recomputed divergence is exactly what A leaves
over: -gamma X with the free surface (AC = -sum A - gamma), zero with the
rigid lid, except at the one cell where PREP pins the pressure.  A wrong sign,
index or metric row in either the RHS assembly, the correction or the
operator assembly breaks these identities far beyond the FP32 rounding the
REAL(4) velocities allow.  The cases where the reference's arithmetic is
trivially closed form (the shipped harness override, the land masking) are
checked exactly.
"""
import numpy as np
import pytest

from musoc_solver_amd import ocean3d, synth


def solve_rep(oracle, oc, s, x, passes=2):
    """X with A X = S by REP passes refined in FP64 (non-periodic operator)."""
    ie = synth.reference_ie(oc.j2, synth.nb0_for_height(oc.j2, 2))
    r, r1, _ = oracle.pre(oc.coefs, ie)
    s64 = s.astype(np.float64)
    for _ in range(passes):
        res = s64 - oracle.sqprod(oc.coefs, x)[1:-1, 1:-1]
        d = np.zeros_like(x)
        oracle.rep(oc.coefs, r, r1, np.asfortranarray(res.astype(np.float32)), d, ie)
        x += d
    return oracle.rel_residual(oc.coefs, s, x)


def divergence_after_one_correction(oracle, free_surface, seed):
    oc = ocean3d.make_ocean3d(48, 32, 4, seed, free_surface=free_surface, periodic=False)
    u, v, w, p0 = ocean3d.make_state(oc, seed + 1)
    s = np.zeros((oc.i2, oc.j2), np.float32, order="F")
    scr = np.zeros((oc.i2 + 2, oc.j2 + 2, 4), np.float32, order="F")
    oracle.step14_prepare(oc, u, v, w, s, mask=True, free_surface=free_surface)
    x = np.zeros((oc.i2 + 2, oc.j2 + 2), order="F")
    assert solve_rep(oracle, oc, s, x) < 1e-11
    oracle.step14_correct(oc, x, p0, scr, u, v, odt=1.0)
    s_new = np.zeros_like(s)
    w2 = np.zeros_like(w)
    oracle.step14_prepare(oc, u, v, w2, s_new, mask=False, free_surface=free_surface)
    keep = np.ones(s.shape, bool)
    keep[oc.i2 // 2 - 1, 0] = False  # the pinned cell AB(I2/2, 1)
    return oc, s, s_new, x, keep


@pytest.mark.parametrize("seed", [1, 2])
def test_free_surface_correction_leaves_minus_gamma_x(oracle, seed):
    oc, s, s_new, x, keep = divergence_after_one_correction(oracle, True, seed)
    target = -oc.gamma * x[1:-1, 1:-1]
    err = np.abs(s_new - target)[keep].max()
    assert err <= 1e-4 * np.abs(s).max(), err / np.abs(s).max()
    # the target is not trivially small: the identity is a real constraint
    assert np.abs(target)[keep].max() > 1e-3 * np.abs(s).max()


@pytest.mark.parametrize("seed", [3, 4])
def test_rigid_lid_correction_removes_divergence(oracle, seed):
    oc, s, s_new, x, keep = divergence_after_one_correction(oracle, False, seed)
    assert np.abs(s_new)[keep].max() <= 1e-4 * np.abs(s).max()


def test_harness_override_closed_forms(oracle):
    """The shipped source sets U = 1, V = 0, W(:,:,K0) = 0 before the RHS:
    the free-surface RHS is then exactly zero, and the rigid-lid RHS is
    -W(I,J,1) of the incoming W (every layer's divergence is zero)."""
    oc = ocean3d.make_ocean3d(40, 24, 3, 5, periodic=True)
    u, v, w, _ = ocean3d.make_state(oc, 6)
    w[:, :, 0] = np.random.default_rng(1).standard_normal(w.shape[:2]).astype(np.float32)
    w0 = w[:, :, 0].copy()
    s = np.zeros((oc.i2, oc.j2), np.float32, order="F")
    oracle.step14_prepare(oc, u, v, w, s, harness=True, free_surface=True)
    assert (u == 1.0).all() and (v == 0.0).all() and not s.any()
    u, v, w, _ = ocean3d.make_state(oc, 6)
    w[:, :, 0] = w0
    oracle.step14_prepare(oc, u, v, w, s, harness=True, free_surface=False)
    assert np.array_equal(s, -w0[1:-1, 1:-1])


def test_land_masking(oracle):
    oc = ocean3d.make_ocean3d(40, 24, 3, 7, periodic=True)
    u, v, w, _ = ocean3d.make_state(oc, 8)
    u_in, v_in = u.copy(), v.copy()
    s = np.zeros((oc.i2, oc.j2), np.float32, order="F")
    oracle.step14_prepare(oc, u, v, w, s, mask=True, lopenw=1)
    i1, j1, j2 = oc.i2 + 1, oc.j2 + 1, oc.j2
    # top-layer faces: zero where IU0 / IV0 say land, untouched elsewhere
    uu = u[:i1, 1:j1, 0]
    assert np.array_equal(uu, np.where(oc.iu0[:i1, 1:j1] == 0, 0.0, u_in[:i1, 1:j1, 0]))
    vv = v[1:i1, 1:j2, 0]
    assert np.array_equal(vv, np.where(oc.iv0[1:i1, 1:j2] == 0, 0.0, v_in[1:i1, 1:j2, 0]))
    assert np.array_equal(u[:, :, 1:], u_in[:, :, 1:]) and np.array_equal(v[:, :, 1:], v_in[:, :, 1:])


def test_itmask_iterations_reduce_land_speed(oracle):
    """Each iteration lowers the top-layer speed on land faces (the loop's
    stopping criterion, :1263-1283); the alternating masking and projection
    converge slowly, hence the reference's iteration cap MXMASK."""
    oc = ocean3d.make_ocean3d(48, 32, 4, 9, free_surface=True, periodic=False)
    u, v, w, p0 = ocean3d.make_state(oc, 10)
    s = np.zeros((oc.i2, oc.j2), np.float32, order="F")
    scr = np.zeros((oc.i2 + 2, oc.j2 + 2, 4), np.float32, order="F")
    x = np.zeros((oc.i2 + 2, oc.j2 + 2), order="F")
    tmps = []
    for _ in range(4):
        oracle.step14_prepare(oc, u, v, w, s, mask=True)
        solve_rep(oracle, oc, s, x, passes=1)
        tmps.append(oracle.step14_correct(oc, x, p0, scr, u, v, odt=1.0))
    assert all(b < a for a, b in zip(tmps, tmps[1:])), tmps


def test_planted_slip_breaks_the_identity(oracle, tmp_path):
    """A plausible slip planted in a copy of the oracle (the U correction
    reading the metric of the next row) must break the free-surface
    identity: the pin has teeth."""
    import ctypes
    import os
    import subprocess
    from oracle import oracle as O
    so = str(tmp_path / "liboracle_planted.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-std=c99",
                    "-DORC_PLANT_STEP14_ROW", os.path.join(O.HERE, "sevp_oracle.c"), "-o", so],
                   check=True)
    bad = ctypes.CDLL(so)
    bad.orc_step14_correct.restype = ctypes.c_float
    oc = ocean3d.make_ocean3d(48, 32, 4, 1, free_surface=True, periodic=False)
    u, v, w, p0 = ocean3d.make_state(oc, 2)
    s = np.zeros((oc.i2, oc.j2), np.float32, order="F")
    scr = np.zeros((oc.i2 + 2, oc.j2 + 2, 4), np.float32, order="F")
    oracle.step14_prepare(oc, u, v, w, s, mask=True)
    x = np.zeros((oc.i2 + 2, oc.j2 + 2), order="F")
    solve_rep(oracle, oc, s, x)
    P, ci = O.P, O._ci
    bad.orc_step14_correct(ci(oc.i2), ci(oc.j2), ci(oc.k1), ci(0), ctypes.c_float(1.0),
                           P(oc.iu), P(oc.iv), P(oc.iu0), P(oc.iv0), P(oc.odx), P(oc.odyv),
                           P(x), P(p0), P(scr), P(u), P(v))
    s_new = np.zeros_like(s)
    oracle.step14_prepare(oc, u, v, np.zeros_like(w), s_new, mask=False)
    keep = np.ones(s.shape, bool)
    keep[oc.i2 // 2 - 1, 0] = False
    err = np.abs(s_new + oc.gamma * x[1:-1, 1:-1])[keep].max()
    assert err > 1e-3 * np.abs(s).max(), err / np.abs(s).max()  # 10x the identity tolerance
