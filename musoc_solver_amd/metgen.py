"""Spherical grid metrics as PREP's METGEN builds them (SPHERE branch,
PREP/template/GLOBAL/src/metgen_GLOBAL.f90:127-201).

Each row's latitude step is found by fixed-point iteration so that the
meridional spacing is DYDX times the zonal spacing at the row's mid
latitude, dy = DYDX * dx(phi): with DYDX = 1 the cells are square, the
Mercator spacing that keeps the SEVP march's amplification per row the same
at every latitude (the anisotropy of a uniform latitude-longitude grid makes
it grow without bound toward the poles).  Every statement is evaluated in
REAL(4) in the source's order; the intrinsics COS, SIN and TAN are numpy's
float32 ones, which may differ from a Fortran library's by an ulp.

Arrays carry the reference's names and 1-based meaning, stored 0-based:
``yvdeg[j-1]`` is YVDEG(J).  Rows the SPHERE branch leaves unset (index 1 of
DX, CS, OCS, ODX, ...) are zero, as in the zero-initialised module arrays.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
R0 = F32(6.4e8)  # earth radius in cm (PREP/template/GLOBAL/src/grid_var_GLOBAL.f90:753)


def metgen_sphere(i2: int, j2: int, dxmnut: float, y0deg: float, x0deg: float = 0.0,
                  dydx: float = 1.0, dx0std: int = 1, dx0: float | None = None,
                  r0: float = float(R0), tol: float = 1e-4) -> dict:
    """The metric arrays of METGEN's SPHERE branch for an I2 x J2 grid.

    dxmnut  zonal spacing in minutes (DXMNUT; 15 for 0.25 degree)
    y0deg   latitude of the southern face of row 2 (Y0DEG)
    dydx    meridional / zonal spacing ratio (DYDX)
    dx0std  0: DX0 is the southernmost DX, 1: the equatorial DX, 2: ``dx0``
            given (DX0STD)
    """
    i0, j0 = i2 + 2, j2 + 2
    j1 = j0 - 1
    f = F32
    pi_180 = f(3.141592654) / f(180.0)
    omega2 = f(3.141592654) / f(21600.0)
    r0 = f(r0)
    dydx = f(dydx)
    dxdeg = f(dxmnut) / f(60.0)
    z = lambda n: np.zeros(n, F32)
    xdeg = z(i0)
    yvdeg, ydeg, yv, y, dy, ody, dx, cs = z(j0), z(j0), z(j0), z(j0), z(j0), z(j0), z(j0), z(j0)
    ocs, odx = z(j0), z(j0)
    dyv, odyv, csv, ocsv, dxv, odxv = z(j1), z(j1), z(j1), z(j1), z(j1), z(j1)
    tanphi, fcor = z(j2), z(j2)
    xdeg[0] = f(x0deg) - f(0.5) * dxdeg
    for i in range(1, i0):
        xdeg[i] = xdeg[i - 1] + dxdeg
    yvdeg[0] = f(y0deg)
    ydeg[0] = f(y0deg)
    if dx0std == 0:
        dx0v = r0 * dxdeg * pi_180 * np.cos(yvdeg[0] * pi_180)
    elif dx0std == 1:
        dx0v = r0 * dxdeg * pi_180
    elif dx0std == 2:
        if dx0 is None:
            raise ValueError("DX0STD = 2 needs dx0")
        dx0v = f(dx0)
    else:
        raise ValueError("DX0STD must be 0, 1 or 2")
    tol = f(tol)
    for J in range(2, j0 + 1):
        j = J - 1
        yvdeg[j] = yvdeg[j - 1] + dxdeg * dydx * np.cos(yvdeg[j - 1] * pi_180)
        dx[j] = dx0v * np.cos(f(0.5) * (yvdeg[j - 1] + yvdeg[j]) * pi_180)
        for _ in range(1000):
            tmp = yvdeg[j]
            yvdeg[j] = yvdeg[j - 1] + dydx * dx[j] / (r0 * pi_180)
            ydeg[j] = f(0.5) * (yvdeg[j - 1] + yvdeg[j])
            cs[j] = np.cos(ydeg[j] * pi_180)
            dxrad = dxdeg * pi_180 * cs[j]
            dx[j] = dxrad * r0
            if abs(tmp - yvdeg[j]) < tol:
                break
        else:
            raise RuntimeError(f"METGEN iteration did not settle at row {J}")
        dydeg = yvdeg[j] - yvdeg[j - 1]
        dy[j] = dydeg * r0 * pi_180
        dyv[j - 1] = (ydeg[j] - ydeg[j - 1]) * r0 * pi_180
        ody[j] = f(1.0) / dy[j]
        odyv[j - 1] = f(1.0) / dyv[j - 1]
        yv[j] = yv[j - 1] + dy[j]
        y[j] = f(0.5) * (yv[j] + yv[j - 1])
        ocs[j] = f(1.0) / cs[j]
        odx[j] = f(1.0) / dx[j]
        csv[j - 1] = np.cos(pi_180 * yvdeg[j - 1])
        ocsv[j - 1] = f(1.0) / csv[j - 1]
        dxv[j - 1] = dx0v * csv[j - 1]
        odxv[j - 1] = f(1.0) / dxv[j - 1]
    for J in range(2, j1 + 1):
        tanphi[J - 2] = np.tan(pi_180 * ydeg[J - 1]) / r0
        fcor[J - 2] = omega2 * np.sin(pi_180 * ydeg[J - 1])
    return dict(xdeg=xdeg, yvdeg=yvdeg, ydeg=ydeg, yv=yv, y=y, dx=dx, dy=dy, dyv=dyv, ody=ody,
                odyv=odyv, cs=cs, ocs=ocs, odx=odx, csv=csv, ocsv=ocsv, dxv=dxv, odxv=odxv,
                tanphi=tanphi, f=fcor, y1deg=yvdeg[j1 - 1], dx0=dx0v)


def rows_to_latitude(dxmnut: float, y0deg: float, lat_north: float, dydx: float = 1.0) -> int:
    """J2 of a Mercator grid (DYDX spacing) from y0deg up to lat_north:
    the number of rows whose northern face stays at or below lat_north."""
    dlam = np.deg2rad(dxmnut / 60.0) * dydx
    psi = lambda d: np.arctanh(np.sin(np.deg2rad(d)))
    return int(np.floor((psi(lat_north) - psi(y0deg)) / dlam))
