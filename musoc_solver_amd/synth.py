"""Synthetic SEVP workloads: grid metric, land mask, depth, operator, RHS.

The operator has the form the TIMCOM preprocessor assembles
(PREP/template/GLOBAL/src/prep_GLOBAL.f90:183-244):

    AL(i,j) = Hw / dx_j**2                      west face
    AR(i,j) = He / dx_j**2                      east face
    AB(i,j) = cos(phi_v[j-1]) Hs / (cos(phi_j) dy_j dyv[j-1])   south face
    AT(i,j) = cos(phi_v[j])   Hn / (cos(phi_j) dy_j dyv[j])     north face
    AC      = -(AL + AR + AB + AT) - gamma,  gamma = 1 / (0.5 g dt**2)

with AB(:,1) = AT(:,J2) = 0 (prep_GLOBAL.f90:227-230) and the face depth the
shallower of the two cells it joins (the layer count of the reference's
Σ_k dz_k·mask sum).  Land is not removed from the domain: as the reference
keeps its top layer wet everywhere (prep_GLOBAL.f90:184-196), land cells keep a
small regularising depth ``h_land`` so that every north coefficient is
positive and the row march never divides by zero.

Coefficients are stored as float32 (the reference's default REAL) in Fortran
(column-major) order with shape (I2, J2).  Everything is seeded.

Metrics:
  "latlon"   uniform latitude spacing between +-lat_max.
  "mercator" the reference's METGEN form (PREP/template/GLOBAL/src/
             metgen_GLOBAL.f90:164-194): cells of constant aspect ratio
             dy = dydx * dx, spaced uniformly in Mercator y and centred on the
             equator.
"""
from __future__ import annotations

import dataclasses

import numpy as np

EARTH_RADIUS = 6.371e6
GRAVITY = 9.81


@dataclasses.dataclass
class Grid:
    i2: int
    j2: int
    lat_c: np.ndarray   # (j2,) cell-centre latitude, radians
    lat_v: np.ndarray   # (j2+1,) face latitude, radians (south face of row j is lat_v[j-1])
    dx: np.ndarray      # (j2,) zonal cell width, m
    dy: np.ndarray      # (j2,) meridional cell height, m
    dyv: np.ndarray     # (j2+1,) centre-to-centre distance across each face, m


def make_grid(i2: int, j2: int, metric: str = "latlon", lat_max: float = 90.0,
              dydx: float = 1.0) -> Grid:
    dlam = 2.0 * np.pi / i2
    if metric == "latlon":
        phim = np.deg2rad(lat_max)
        lat_v = np.linspace(-phim, phim, j2 + 1)
    elif metric == "mercator":
        dym = dydx * dlam
        ys = (np.arange(j2 + 1) - j2 / 2.0) * dym
        lat_v = 2.0 * np.arctan(np.exp(ys)) - np.pi / 2.0
    else:
        raise ValueError(f"unknown metric {metric!r}")
    lat_c = 0.5 * (lat_v[:-1] + lat_v[1:])
    dx = EARTH_RADIUS * np.cos(lat_c) * dlam
    dy = EARTH_RADIUS * np.diff(lat_v)
    dyv = np.empty(j2 + 1)
    dyv[1:-1] = EARTH_RADIUS * np.diff(lat_c)
    dyv[0] = dy[0]
    dyv[-1] = dy[-1]
    return Grid(i2, j2, lat_c, lat_v, dx, dy, dyv)


def make_mask(i2: int, j2: int, seed: int, land_frac: float = 0.30,
              kmax: int = 6) -> np.ndarray:
    """Boolean ocean mask (True = ocean), shape (i2, j2).

    A random Fourier field with integer wavenumbers |kx|, |ky| <= kmax and
    N(0,1) coefficients, thresholded at its ``land_frac`` quantile."""
    rng = np.random.default_rng(seed)
    x = (np.arange(i2) + 0.5) / i2 * 2.0 * np.pi
    y = (np.arange(j2) + 0.5) / j2 * np.pi
    field = np.zeros((i2, j2))
    for kx in range(0, kmax + 1):
        for ky in range(0, kmax + 1):
            if kx == 0 and ky == 0:
                continue
            a, b, c, d = rng.standard_normal(4)
            cx, sx = np.cos(kx * x), np.sin(kx * x)
            cy, sy = np.cos(ky * y), np.sin(ky * y)
            field += np.outer(a * cx + b * sx, c * cy + d * sy)
    thr = np.quantile(field, land_frac)
    return field > thr


def _smooth5(h: np.ndarray) -> np.ndarray:
    """One pass of a 5-point box average, periodic in i, clamped in j."""
    up = np.concatenate([h[:, 1:], h[:, -1:]], axis=1)
    dn = np.concatenate([h[:, :1], h[:, :-1]], axis=1)
    return (h + np.roll(h, 1, axis=0) + np.roll(h, -1, axis=0) + up + dn) / 5.0


def make_depth(mask: np.ndarray, seed: int, hmin: float = 50.0,
               hmax: float = 5500.0, passes: int = 3,
               h_land: float = 10.0) -> np.ndarray:
    """Log-uniform depth on [hmin, hmax], smoothed, land set to h_land."""
    rng = np.random.default_rng(seed)
    h = np.exp(rng.uniform(np.log(hmin), np.log(hmax), size=mask.shape))
    for _ in range(passes):
        h = _smooth5(h)
    return np.where(mask, h, h_land)


def make_operator(depth: np.ndarray, grid: Grid, dt: float = 1800.0,
                  gravity: float = GRAVITY):
    """Return (AL, AB, AC, AR, AT) float32 Fortran-ordered (I2, J2)."""
    i2, j2 = depth.shape
    hw = np.minimum(depth, np.roll(depth, 1, axis=0))
    he = np.minimum(depth, np.roll(depth, -1, axis=0))
    hs = np.zeros_like(depth)
    hn = np.zeros_like(depth)
    hs[:, 1:] = np.minimum(depth[:, 1:], depth[:, :-1])
    hn[:, :-1] = np.minimum(depth[:, :-1], depth[:, 1:])
    odx2 = 1.0 / grid.dx ** 2
    cosc = np.cos(grid.lat_c)
    cosv = np.cos(grid.lat_v)
    fb = cosv[:-1] / (cosc * grid.dy * grid.dyv[:-1])
    ft = cosv[1:] / (cosc * grid.dy * grid.dyv[1:])
    al = hw * odx2[None, :]
    ar = he * odx2[None, :]
    ab = hs * fb[None, :]
    at = hn * ft[None, :]
    ab[:, 0] = 0.0
    at[:, -1] = 0.0
    gamma = 1.0 / (0.5 * dt * dt * gravity)
    ac = -(al + ar + ab + at) - gamma
    f32 = lambda a: np.asfortranarray(a.astype(np.float32))
    return f32(al), f32(ab), f32(ac), f32(ar), f32(at)


def ideal_operator(i2: int, j2: int):
    """The reference's constant test operator with Neumann north/south edges
    (PREP/template/GLOBAL/src/prep_GLOBAL.f90:246-261)."""
    al = np.full((i2, j2), -0.25, np.float32, order="F")
    ab = al.copy(order="F")
    ar = al.copy(order="F")
    at = al.copy(order="F")
    ac = np.full((i2, j2), 1.0, np.float32, order="F")
    ac[:, -1] += at[:, -1]
    at[:, -1] = 0.0
    ac[:, 0] += ab[:, 0]
    ab[:, 0] = 0.0
    return al, ab, ac, ar, at


def make_rhs(mask: np.ndarray, seed: int, nrhs: int = 1,
             lowfreq: float = 0.0) -> np.ndarray:
    """float32 RHS of shape (I2, J2, nrhs), Fortran order: N(0,1) on ocean
    cells, zero on land, zero mean over the ocean.  ``lowfreq`` adds a slowly
    drifting low-wavenumber component (successive time steps' forcing)."""
    i2, j2 = mask.shape
    rng = np.random.default_rng(seed)
    out = np.zeros((i2, j2, nrhs), np.float32, order="F")
    x = (np.arange(i2) + 0.5) / i2 * 2.0 * np.pi
    y = (np.arange(j2) + 0.5) / j2 * np.pi
    for k in range(nrhs):
        s = rng.standard_normal((i2, j2))
        if lowfreq:
            ph = 0.05 * k
            s += lowfreq * np.outer(np.sin(x + ph), np.sin(y) * np.cos(2 * y + 0.5 * ph))
        s = np.where(mask, s, 0.0)
        s[mask] -= s[mask].mean()
        out[:, :, k] = s
    return out


def reference_ie(j2: int, nb0: int) -> np.ndarray:
    """Block row ends IE(1..NB0) by the reference rule
    (PREP/template/GLOBAL/src/prep_GLOBAL.f90:147-168): NN = ANINT(J2/NB0),
    IE(N) = MIN(1+NN*N, J1), IE(NB0) = J1, with the same sanity check."""
    j1 = j2 + 1
    nn = int(np.floor(j2 / nb0 + 0.5))
    ie = np.array([min(1 + nn * n, j1) for n in range(1, nb0 + 1)], np.int32)
    if nb0 > 1 and not (ie[-1] > ie[-2] and j1 < ie[-2] + nn + 3):
        raise ValueError(f"IE(NB1)={ie[-2]}, J1={j1} not properly defined")
    ie[-1] = j1
    return ie


def nb0_for_height(j2: int, n0: int) -> int:
    """NB0 = ANINT(J2/N0) (MAIN/template/GLOBAL/src/grid_var_GLOBAL.f90:28)."""
    return int(np.floor(j2 / n0 + 0.5))


@dataclasses.dataclass
class Case:
    al: np.ndarray
    ab: np.ndarray
    ac: np.ndarray
    ar: np.ndarray
    at: np.ndarray
    mask: np.ndarray
    grid: Grid | None

    @property
    def coefs(self):
        return self.al, self.ab, self.ac, self.ar, self.at

    @property
    def shape(self):
        return self.al.shape


def make_case(i2: int, j2: int, seed: int, metric: str = "latlon",
              lat_max: float = 90.0, dydx: float = 1.0, land_frac: float = 0.30,
              h_land: float = 10.0, dt: float = 1800.0) -> Case:
    grid = make_grid(i2, j2, metric, lat_max, dydx)
    mask = make_mask(i2, j2, seed, land_frac)
    depth = make_depth(mask, seed + 1, h_land=h_land)
    al, ab, ac, ar, at = make_operator(depth, grid, dt)
    return Case(al, ab, ac, ar, at, mask, grid)


def basin_specs(n: int, seed: int, i2_range=(64, 512), j2_range=(64, 384),
                land_range=(0.20, 0.45)) -> list:
    """Regional basins for a batched workload: (i2, j2, land fraction, seed)
    per basin, sizes log-uniform in i2_range x j2_range (I2 rounded to a
    multiple of 16, J2 to a multiple of 8), land fraction uniform in
    land_range, one seed per basin for its mask, depth and right-hand sides."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        i2 = int(np.exp(rng.uniform(np.log(i2_range[0]), np.log(i2_range[1]))) / 16 + 0.5) * 16
        j2 = int(np.exp(rng.uniform(np.log(j2_range[0]), np.log(j2_range[1]))) / 8 + 0.5) * 8
        i2 = min(max(i2, i2_range[0]), i2_range[1])
        j2 = min(max(j2, j2_range[0]), j2_range[1])
        out.append((i2, j2, float(rng.uniform(*land_range)), seed + 10 * k))
    return out
