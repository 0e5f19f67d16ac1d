"""Block-row placement (IE) from measured march growth.

SEVP marches each block from its free row; rounding in the free row grows by
the march's amplification over the block (roughly the square of it reaches
the solution through the influence-matrix correction).  The reference fixes
one block height everywhere (IE(N) = 1 + NN*N,
PREP/template/GLOBAL/src/prep_GLOBAL.f90:147-168); its own guidance is "IE
spacings of about 5-8 for REAL*8" (MAIN/template/GLOBAL/src/sevp.f90:351-356),
which holds only for well-conditioned operators.

``adaptive_ie`` measures the growth of the homogeneous march of the actual
operator row by row (FP64, periodic or Dirichlet east-west, several random
free rows plus the most oscillatory mode) and closes each block at the
largest height whose growth stays under ``max_growth``.  The result is a
valid IE for every plan form: strictly increasing, last entry J2+1.
"""
from __future__ import annotations

import numpy as np


def march_growth(coefs, js: int, hmax: int, periodic: bool = True, nvec: int = 4,
                 seed: int = 0) -> np.ndarray:
    """Growth g[h-1] = max over test vectors of |H(js+h)| / |H(js+1)|, for
    block heights h = 1..hmax starting at row js (1-based, the row below the
    free row), with H(js) = 0 as in the march of an influence column."""
    al, ab, ac, ar, at = [np.asarray(c, np.float64) for c in coefs]
    i2, j2 = al.shape
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((i2, nvec))
    v = np.concatenate([v, ((-1.0) ** np.arange(i2))[:, None]], axis=1)
    h0 = np.zeros_like(v)       # row js
    h1 = v / np.abs(v).max(axis=0)  # free row js+1
    base = np.abs(h1).max(axis=0)
    g = np.ones(hmax)
    hs, hc = h0, h1
    for h in range(2, hmax + 1):
        j = js + h - 2  # equation row producing X row j+2 = js+h (1-based)
        if j > j2 - 1:
            g[h - 1:] = np.inf
            break
        jj = j - 1
        if periodic:
            hw, he = np.roll(hc, 1, axis=0), np.roll(hc, -1, axis=0)
        else:
            hw = np.vstack([np.zeros((1, hc.shape[1])), hc[:-1]])
            he = np.vstack([hc[1:], np.zeros((1, hc.shape[1]))])
        hn = -(al[:, jj, None] * hw + ab[:, jj, None] * hs + ac[:, jj, None] * hc
               + ar[:, jj, None] * he) / at[:, jj, None]
        hs, hc = hc, hn
        g[h - 1] = float((np.abs(hn).max(axis=0) / base).max())
    return g


def adaptive_ie(coefs, max_growth: float = 1e3, hmax: int = 8, hmin: int = 2,
                periodic: bool = True) -> np.ndarray:
    """IE(1..NB0) with each block as tall as the growth bound allows."""
    al = coefs[0]
    j2 = al.shape[1]
    j1 = j2 + 1
    ie = []
    prev = 1  # IE(0) := 1 (the first block starts at row 1)
    while prev < j1:
        g = march_growth(coefs, prev, hmax, periodic)
        h = hmin
        for hh in range(hmin, hmax + 1):
            if g[hh - 1] <= max_growth:
                h = hh
            else:
                break
        nxt = min(prev + h, j1)
        # never leave a last block thinner than hmin
        if j1 - nxt < hmin and nxt != j1:
            nxt = j1
        ie.append(nxt)
        prev = nxt
    return np.array(ie, np.int32)


def uniform_ie(j2: int, h: int) -> np.ndarray:
    """Blocks of h rows, the last one absorbing the remainder (the
    reference's rule for NB0 = ANINT(J2/h))."""
    from .synth import nb0_for_height, reference_ie
    return reference_ie(j2, nb0_for_height(j2, h))
