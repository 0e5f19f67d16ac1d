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


def _test_rows(i2: int, nvec: int, seed: int) -> np.ndarray:
    """The march's free rows: nvec random ones and the most oscillatory mode."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((i2, nvec))
    return np.concatenate([v, ((-1.0) ** np.arange(i2))[:, None]], axis=1)


def march_growth(coefs, js: int, hmax: int, periodic: bool = True, nvec: int = 4,
                 seed: int = 0) -> np.ndarray:
    """Growth g[h-1] = max over test vectors of |H(js+h)| / |H(js+1)|, for
    block heights h = 1..hmax starting at row js (1-based, the row below the
    free row), with H(js) = 0 as in the march of an influence column."""
    c64 = [np.asarray(c, np.float64) for c in coefs]
    return _growth(c64, _test_rows(c64[0].shape[0], nvec, seed), js, hmax, periodic)


def _growth(c64, v, js: int, hmax: int, periodic: bool, stop: float = np.inf) -> np.ndarray:
    """march_growth on FP64 coefficients and given free rows v; the heights
    past the first whose growth exceeds ``stop`` are not marched (inf)."""
    al, ab, ac, ar, at = c64
    i2, j2 = al.shape
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
        if g[h - 1] > stop:
            g[h:] = np.inf
            break
    return g


def adaptive_ie(coefs, max_growth: float = 1e3, hmax: int = 8, hmin: int = 2,
                periodic: bool = True) -> np.ndarray:
    """IE(1..NB0) with each block as tall as the growth bound allows."""
    # the coefficients in FP64 and the free rows once, not per block
    c64 = [np.asarray(c, np.float64) for c in coefs]
    v = _test_rows(c64[0].shape[0], 4, 0)
    j2 = c64[0].shape[1]
    j1 = j2 + 1
    ie = []
    prev = 1  # IE(0) := 1 (the first block starts at row 1)
    while prev < j1:
        g = _growth(c64, v, prev, hmax, periodic, stop=max_growth)
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


# ------------------------------------------------------- timing on the device
# The performance parameters of the MFMA corrections are chosen by timing
# refined solves on the GPU: the pass kernels, the wavefronts per fused
# workgroup and the RHS groups by sevp_autotune inside one plan, the block
# height (the plan's IE) across plans built for each candidate.  Choices are
# cached under the operator hash (sevp_operator_hash: grid, IE, form and the
# coefficient bits) with the RHS count and precision, so a plan of the same
# operator reuses them without timing.

def _cache_key(plan, nrhs):
    from . import sevp as S
    h = S.operator_hash(plan.coefs, plan.ie, plan.form, plan.nbir)
    return f"{h:016x}:nrhs={nrhs}:prec={plan.precision}"


def tune(plan, s, x, tol=1e-10, max_iter=20, cache_dir=None):
    """Apply the fastest (pass kernels, waves, RHS groups, row tiles) for refined
    solves of s (torch device tensors; x is overwritten while timing).  With
    ``cache_dir`` the choice is read from / stored in
    ``<cache_dir>/sevp_tuning.json``.  Returns the choice as a dict with the
    time per solve call (ms) and whether it came from the cache."""
    import json
    import os
    nrhs = s.shape[0]
    path = os.path.join(str(cache_dir), "sevp_tuning.json") if cache_dir else None
    key = _cache_key(plan, nrhs) if path else None
    table = {}
    if path and os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
        if key in table:
            c = table[key]
            from . import sevp as S
            plan.set_tuning(S.PASS_FUSED if c["pass_kernels"] == "fused" else S.PASS_SPLITK,
                            c["waves"], c["rhs_groups"], c.get("rows", 0))
            return dict(c, cached=True)
    t, ms = plan.autotune(s, x, tol=tol, max_iter=max_iter)
    out = dict(t.as_dict(), ms=ms)
    if path:
        os.makedirs(str(cache_dir), exist_ok=True)
        table[key] = out
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1)
        os.replace(tmp, path)
    return dict(out, cached=False)


def tune_block_height(coefs, heights, s, make_plan, tol=1e-10, max_iter=20):
    """Block height by timing: for each h, a plan of uniform h-row blocks
    from ``make_plan(ie)`` (its influence matrices built), tuned with
    ``tune`` on the right-hand sides s; returns (best h, {h: ms per solve
    call}).  Plans whose refinement does not converge are skipped."""
    import torch
    j2 = coefs[0].shape[1]
    times = {}
    for h in heights:
        plan = make_plan(uniform_ie(j2, h))
        try:
            x = torch.zeros((s.shape[0], j2 + 2, coefs[0].shape[0] + 2), dtype=torch.float64,
                            device=s.device)
            r = tune(plan, s, x, tol=tol, max_iter=max_iter)
            x.zero_()
            rel, st, ok = plan.solve_refined_dev(s, x, tol=tol, max_iter=max_iter)
            if ok:
                times[h] = r["ms"]
        finally:
            plan.close()
    if not times:
        raise RuntimeError("no block height converged")
    return min(times, key=times.get), times
