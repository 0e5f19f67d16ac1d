"""Synthetic layered ocean in the reference's array shapes, for the step-14
caller (barotropic correction around the solve).

Arrays follow MAIN/template/GLOBAL/src/grid_var_GLOBAL.f90 (Fortran order,
1-based in the comments): metrics ODX, ODY, OCS (J0), CSV, ODYV (J1), ODZ
(K1) as REAL(4); masks KB, IU0, IV0 (I0, J0), IU (I0, J0, K1), IV (I0, J1, K1)
as INTEGER(2); state U (I0, J0, K1), V (I0, J1, K1), W (I0, J0, K0),
P0 (I0, J0) as REAL(4).

The masks are derived as PREP derives them from KB
(PREP/template/GLOBAL/src/indata_GLOBAL.f90:480-500 for IN,
:866-925 IU_V_W_CAL for IU, IV, IU0, IV0, including the all-wet top layer of
IU/IV that regularises the solver's operator), and the operator AL..AT, AC by
the layered assembly of PREP/template/GLOBAL/src/prep_GLOBAL.f90:170-244
(DIMEN = 3), in REAL(4) arithmetic in the source's evaluation order.  The
shipped PREP then overwrites the operator with a constant test stencil
(:246-261); assemble_operator_3d stops before that, so the operator is the
discrete depth-integrated divergence of the pressure-gradient correction that
the step-14 caller applies, which is what makes its corrections consistent.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .metgen import metgen_sphere
from .synth import EARTH_RADIUS, GRAVITY, make_grid, make_mask

F32 = np.float32


@dataclasses.dataclass
class Ocean3D:
    i2: int
    j2: int
    k1: int
    periodic: bool
    odx: np.ndarray   # (J0,)
    ody: np.ndarray   # (J0,)
    ocs: np.ndarray   # (J0,)
    csv: np.ndarray   # (J1,)
    odyv: np.ndarray  # (J1,)
    odz: np.ndarray   # (K1,)
    kb: np.ndarray    # (I0, J0) int16
    iu0: np.ndarray   # (I0, J0) int16
    iv0: np.ndarray   # (I0, J0) int16
    iu: np.ndarray    # (I0, J0, K1) int16
    iv: np.ndarray    # (I0, J1, K1) int16
    coefs: tuple      # (AL, AB, AC, AR, AT) float32 (I2, J2)
    gamma: float      # free-surface term of AC (0 for the rigid lid)
    mask: np.ndarray  # (I2, J2) ocean cells


def _fa(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def assemble_operator_3d(odx, ody, ocs, csv, odyv, odz, iu, iv, i2, j2, nbir=1,
                         free_surface=True, dt=1800.0, gravity=GRAVITY):
    """AL, AB, AC, AR, AT (float32, (I2, J2)) by the layered assembly of
    PREP/template/GLOBAL/src/prep_GLOBAL.f90:170-244 (DIMEN = 3), with the
    NBIR = 1 east/west edge zeroing and pressure pin (:213-229) and the
    free-surface term under FLAG_FS (:231-244).  Returns (coefs, gamma)."""
    k1 = odz.shape[0]
    al = np.zeros((i2, j2), F32)
    ar = np.zeros((i2, j2), F32)
    ab = np.zeros((i2, j2), F32)
    at = np.zeros((i2, j2), F32)
    one = F32(1.0)
    for k in range(1, k1 + 1):
        tmp = one / odz[k - 1]
        for j in range(1, j2 + 1):
            dzx = (odx[j] * odx[j]) * tmp                                   # ODX(J+1)**2*TMP
            dzyb = (((csv[j - 1] * ocs[j]) * odyv[j - 1]) * ody[j]) * tmp   # CSV(J)*OCS(J+1)*ODYV(J)*ODY(J+1)*TMP
            dzyt = (((csv[j] * ocs[j]) * odyv[j]) * ody[j]) * tmp           # CSV(J+1)*OCS(J+1)*ODYV(J+1)*ODY(J+1)*TMP
            if k == 1:  # top layer: all water
                al[:, j - 1] = dzx
                ar[:, j - 1] = dzx
                ab[:, j - 1] = dzyb
                at[:, j - 1] = dzyt
            else:
                # IU(I,J+1,K), IU(I+1,J+1,K), IV(I+1,J,K), IV(I+1,J+1,K) for I = 1..I2
                al[:, j - 1] = al[:, j - 1] + dzx * iu[0:i2, j, k - 1].astype(F32)
                ar[:, j - 1] = ar[:, j - 1] + dzx * iu[1:i2 + 1, j, k - 1].astype(F32)
                ab[:, j - 1] = ab[:, j - 1] + dzyb * iv[1:i2 + 1, j - 1, k - 1].astype(F32)
                at[:, j - 1] = at[:, j - 1] + dzyt * iv[1:i2 + 1, j, k - 1].astype(F32)
    if nbir == 1:
        al[0, :] = 0.0
        ar[i2 - 1, :] = 0.0
    ab[:, 0] = 0.0
    at[:, j2 - 1] = 0.0
    if nbir == 1:  # pin down pressure: AB(I2/2, 1) = AB(I2/2, 2)
        ab[i2 // 2 - 1, 0] = ab[i2 // 2 - 1, 1]
    gamma = F32(0.0)
    if free_surface:
        gamma = one / ((F32(0.5) * (F32(dt) * F32(dt))) * F32(gravity))
    ac = (((-al - ar) - ab) - at) - gamma
    f = lambda a: _fa(a, F32)
    return (f(al), f(ab), f(ac), f(ar), f(at)), float(gamma)


def make_ocean3d(i2: int, j2: int, k1: int, seed: int, lat_max: float = 60.0,
                 land_frac: float = 0.3, periodic: bool = False, free_surface: bool = True,
                 dz=None, dt: float = 1800.0, nbir: int = 1, metric: str = "latlon",
                 dxmnut: float | None = None, y0deg: float | None = None) -> Ocean3D:
    """A layered ocean.  ``metric``: "latlon" (uniform latitude spacing over
    +-lat_max) or "metgen" (PREP's METGEN spherical grid with DYDX = 1,
    square cells, from y0deg north, zonal spacing dxmnut minutes; its R0 in
    metres here, so the metrics agree with GRAVITY and dz in SI units)."""
    i0, j0 = i2 + 2, j2 + 2
    i1, j1 = i0 - 1, j0 - 1
    if metric == "metgen":
        mg = metgen_sphere(i2, j2, dxmnut if dxmnut is not None else 21600.0 / i2,
                           y0deg if y0deg is not None else -lat_max, r0=EARTH_RADIUS)
        odx, ody, ocs = mg["odx"].copy(), mg["ody"].copy(), mg["ocs"].copy()
        csv, odyv = mg["csv"].copy(), mg["odyv"].copy()
    elif metric == "latlon":
        g = make_grid(i2, j2, "latlon", lat_max)
        # metrics (halo rows copy their neighbours)
        odx = np.empty(j0, F32)
        ody = np.empty(j0, F32)
        ocs = np.empty(j0, F32)
        odx[1:j1] = 1.0 / g.dx
        ody[1:j1] = 1.0 / g.dy
        ocs[1:j1] = 1.0 / np.cos(g.lat_c)
        for a in (odx, ody, ocs):
            a[0], a[j0 - 1] = a[1], a[j1 - 1]
        csv = np.cos(g.lat_v).astype(F32)           # (J1,): face J between X rows J and J+1
        odyv = (1.0 / g.dyv).astype(F32)
    else:
        raise ValueError(f"unknown metric {metric!r}")
    if dz is None:
        dz = [50.0, 150.0, 300.0, 500.0, 1000.0, 1500.0, 2000.0][:k1]
    dz = np.asarray(dz, np.float64)
    odz = (1.0 / dz).astype(F32)
    # depth and layer counts
    mask = make_mask(i2, j2, seed, land_frac)
    rng = np.random.default_rng(seed + 7)
    depth = np.exp(rng.uniform(np.log(dz[0]), np.log(dz.sum()), size=(i2, j2)))
    edges = np.cumsum(dz)
    kb_in = np.clip(np.searchsorted(edges, depth, side="right") + 1, 1, k1)
    # IN(I,J,K): 1 on wet layers of ocean cells (land: KB = 1, IN = 0)
    inn = np.zeros((i0, j0, k1), np.int16)
    kb = np.zeros((i0, j0), np.int16)
    for i in range(i2):
        for j in range(j2):
            if mask[i, j]:
                kb[i + 1, j + 1] = kb_in[i, j]
                inn[i + 1, j + 1, :kb_in[i, j]] = 1
            else:
                kb[i + 1, j + 1] = 1
    if periodic:
        inn[0, :, :] = inn[i1 - 1, :, :]
        inn[i0 - 1, :, :] = inn[1, :, :]
        kb[0, :], kb[i0 - 1, :] = kb[i1 - 1, :], kb[1, :]
    # IU_V_W_CAL
    iu = np.zeros((i0, j0, k1), np.int16)
    iv = np.zeros((i0, j1, k1), np.int16)
    iu[0:i1, :, :] = inn[0:i1, :, :] * inn[1:i0, :, :]
    iv[:, 0:j1, :] = inn[:, 0:j1, :] * inn[:, 1:j0, :]
    if periodic:
        iu[i0 - 1, :, :] = iu[1, :, :]
    iu0 = np.zeros((i0, j0), np.int16)
    iv0 = np.zeros((i0, j0), np.int16)
    irange = range(0, i0) if periodic else range(1, i2)
    for j in range(1, j1):
        for i in irange:
            iu0[i, j] = iu[i, j, 0]
            iu[i, j, 0] = 1
    for j in range(1, j2):
        for i in range(1, i1):
            iv0[i, j] = iv[i, j, 0]
            iv[i, j, 0] = 1
    coefs, gamma = assemble_operator_3d(odx, ody, ocs, csv, odyv, odz, iu, iv, i2, j2,
                                        nbir=nbir, free_surface=free_surface, dt=dt)
    f = lambda a, t: _fa(a, t)
    return Ocean3D(i2, j2, k1, periodic, f(odx, F32), f(ody, F32), f(ocs, F32), f(csv, F32),
                   f(odyv, F32), f(odz, F32), f(kb, np.int16), f(iu0, np.int16),
                   f(iv0, np.int16), f(iu, np.int16), f(iv, np.int16), coefs, gamma, mask)


def make_state(oc: Ocean3D, seed: int, scale: float = 0.1):
    """Random layer velocities, zero W, random P0 (REAL(4), reference shapes)."""
    rng = np.random.default_rng(seed)
    i0, j0, j1, k1 = oc.i2 + 2, oc.j2 + 2, oc.j2 + 1, oc.k1
    u = _fa(scale * rng.standard_normal((i0, j0, k1)), F32)
    v = _fa(scale * rng.standard_normal((i0, j1, k1)), F32)
    w = np.zeros((i0, j0, k1 + 1), F32, order="F")
    p0 = _fa(rng.standard_normal((i0, j0)), F32)
    return u, v, w, p0
