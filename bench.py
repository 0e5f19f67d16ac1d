#!/usr/bin/env python3
"""SEVP benchmark: solves/s on the 0.25-degree global grid, mixed-precision
batched versus plain FP64, every solve taken to <= 1e-10 relative residual.

Workload (one "step" = one batch of NRHS right-hand sides, each solved from
x = 0 until ||s - A x||_2 / ||s||_2 <= 1e-10, with the inputs resident on the
GPU):

  grid       1440 x 720 cells, 0.25 deg zonal spacing, uniform latitude
             spacing over +-80 deg (polar caps beyond 80 deg excluded)
  operator   synthetic TIMCOM form (musoc_solver_amd.synth): Fourier land
             mask (30% land), log-uniform 50-5500 m depth smoothed 3 times,
             land kept at 50 m, free-surface term 1/(0.5 g dt^2), dt = 1800 s
  solver     full-width periodic SEVP (REP1/PRE1 form) with block rows of
             height 3 (NB0 = 240), FP64 iterative refinement
  batch      64 N(0,1) right-hand sides, zero mean over the ocean

Legs
  mixed      correction products as FP32 MFMA GEMMs, marching and residuals
             FP64  -> "value"
  plain      the reference operation order in FP64 (bitwise equal to the
             reference REP1), refined the same way
  cpu        the CPU oracle (C restatement, one core) on a bounded sample of
             the same workload: one REP1 pass per right-hand side, which
             already reaches the tolerance at this block height

Prints ONE JSON line (see the driver contract).  ``--gpus N`` with N > 1
exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SEVP solves/s @0.25° grid: mixed-prec batched vs FP64 plain @1e-10 rel resid"
PEAK_FP32_MFMA = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--i2", type=int, default=1440)
    ap.add_argument("--j2", type=int, default=720)
    ap.add_argument("--n0", type=int, default=3, help="block height (rows per block)")
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--lat-max", type=float, default=80.0)
    ap.add_argument("--h-land", type=float, default=50.0)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-iter", type=int, default=12)
    ap.add_argument("--seed", type=int, default=3000)
    ap.add_argument("--cpu-rhs", type=int, default=3, help="RHS in the CPU baseline sample (0: skip)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--legs", default="mixed,plain", help="comma list of GPU legs to time")
    ap.add_argument("--log", default=None, help="progress log file (default stderr)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from musoc_solver_amd import sevp as S
    from musoc_solver_amd import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    i2, j2, nrhs = a.i2, a.j2, a.nrhs
    log(f"workload {i2}x{j2} n0={a.n0} nrhs={nrhs} lat_max={a.lat_max} h_land={a.h_land}")
    case = synth.make_case(i2, j2, a.seed, metric="latlon", lat_max=a.lat_max, h_land=a.h_land)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, a.n0))
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs,
                      precision=S.PREC_FP64)
    t0 = time.perf_counter()
    zp = plan.build_influence()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log(f"influence matrices built in {t_setup:.2f} s (NB0={len(ie)}, zero pivots {zp}, "
        f"{plan.device_bytes / 2**30:.2f} GiB on device)")
    s_np = synth.make_rhs(case.mask, a.seed + 1, nrhs)
    s = torch.from_numpy(np.ascontiguousarray(s_np.transpose(2, 1, 0))).to(dev)
    x = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device=dev)

    def run_leg(prec, name):
        plan.set_precision(prec)
        for _ in range(a.warmup):
            x.zero_()
            plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
        torch.cuda.synchronize()
        plan.profile_begin(S.KCLASS_CORR)
        iters, worst, fails = [], 0.0, 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            x.zero_()
            rel, st, ok = plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
            iters.append(st.iterations)
            worst = max(worst, float(rel.max()))
            fails += 0 if ok else 1
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        corr_ms, corr_n = plan.profile_end()
        r = {"solves_per_s": a.steps * nrhs / dt, "ms_per_step": 1e3 * dt / a.steps,
             "passes": float(np.mean(iters)), "max_rel_resid": worst, "failed_steps": fails,
             "corr_avg_us": 1e3 * corr_ms / max(corr_n, 1), "corr_launches": corr_n,
             "corr_share": corr_ms / (1e3 * dt)}
        log(f"{name}: {r}")
        return r

    legs = {}
    order = [l for l in a.legs.split(",") if l]
    if a.skip_plain and "plain" in order:
        order.remove("plain")
    for name in order:
        prec = {"mixed": S.PREC_FP32, "plain": S.PREC_FP64}[name]
        legs[name] = run_leg(prec, name)
        legs[name]["x"] = x.clone() if name in ("mixed", "plain") else None

    # error of the mixed solution against the plain FP64 solution (the plain
    # path is bitwise the reference REP1 arithmetic)
    err = None
    if "mixed" in legs and "plain" in legs:
        xm, xp = legs["mixed"]["x"], legs["plain"]["x"]
        d = (xm - xp)[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
        n = xp[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
        err = float((d / n).max())
        log(f"mixed vs plain FP64: max rel error {err:.3e}")
    for v in legs.values():
        v.pop("x", None)

    main_leg = legs.get("mixed") or legs.get("plain")
    # roofline of the dominant kernel: the correction product (FP32 MFMA GEMM)
    flops = 2.0 * nrhs * i2 * i2
    achieved = flops / (main_leg["corr_avg_us"] * 1e-6)
    mixed_main = "mixed" in legs
    roof = {"bound": "mfma", "achieved": achieved / 1e12,
            "peak": (PEAK_FP32_MFMA if mixed_main else PEAK_FP64) / 1e12, "unit": "TFLOP/s",
            "frac": achieved / (PEAK_FP32_MFMA if mixed_main else PEAK_FP64),
            "traffic": None, "kernel": "gemm_f32_mfma_kernel" if mixed_main else "gemm_seq_kernel",
            "algorithmic_flops_per_launch": flops, "avg_launch_us": main_leg["corr_avg_us"]}

    cpu = None
    if a.cpu_rhs > 0:
        cpu = cpu_baseline(plan, case, ie, s_np, a)

    out = {
        "metric": METRIC,
        "value": main_leg["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_leg["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32mfma" if mixed_main else "f64",
        "data": "synthetic",
        "config": {"workload": "SEVP 1440x720 (0.25 deg zonal, +-80 deg), REP1 form, "
                               f"block height {a.n0}, {nrhs} RHS per batch, to {a.tol:g} rel resid",
                   "i2": i2, "j2": j2, "nb0": len(ie), "nrhs": nrhs, "tol": a.tol},
        "legs": legs,
        "mixed_over_plain": (legs["mixed"]["solves_per_s"] / legs["plain"]["solves_per_s"]
                             if "mixed" in legs and "plain" in legs else None),
        "mixed_vs_fp64_max_rel_err": err,
        "setup_s": t_setup,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()


def cpu_baseline(plan, case, ie, s_np, a):
    """The oracle's REP1 pass on a bounded sample of the same workload."""
    from oracle import oracle as O
    O.build()
    log("cpu baseline: copying influence matrices to the host")
    rinv, rinv1 = plan.get_influence()
    i2, j2 = case.shape
    t_all, worst = 0.0, 0.0
    for r in range(a.cpu_rhs):
        x = np.zeros((i2 + 2, j2 + 2), order="F")
        sr = np.asfortranarray(s_np[:, :, r])
        t0 = time.perf_counter()
        O.rep1(case.coefs, rinv, rinv1, sr, x, ie)
        t_all += time.perf_counter() - t0
        worst = max(worst, O.rel_residual(case.coefs, sr, x))
        log(f"cpu rhs {r}: {t_all:.2f} s cumulative, rel resid {worst:.2e}")
    return {"value": a.cpu_rhs / t_all, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{a.cpu_rhs} of the {a.nrhs} RHS, one FP64 REP1 pass each "
                      f"(max rel resid {worst:.2e}); influence matrices from the GPU build",
            "rel_resid": worst}


if __name__ == "__main__":
    main()
