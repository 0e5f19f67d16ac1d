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
  solver     full-width periodic SEVP (REP1/PRE1 form), block rows of height
             2 (NB0 = 360; --n0 / --n0-mixed), FP64 iterative refinement
  batch      64 N(0,1) right-hand sides, zero mean over the ocean

Legs
  mixed      correction products as FP32 MFMA GEMMs, marching and residuals
             FP64  -> "value"
  plain      the reference operation order in FP64 (bitwise equal to the
             reference REP1), refined the same way
  cpu        the CPU oracle (C restatement, one core) on a bounded sample of
             the same workload, refined to the same tolerance

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
    ap.add_argument("--n0", type=int, default=2, help="block height (rows per block), plain leg")
    ap.add_argument("--n0-mixed", type=int, default=None,
                    help="block height of the mixed-precision leg (default: --n0)")
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--lat-max", type=float, default=80.0)
    ap.add_argument("--h-land", type=float, default=50.0)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-iter", type=int, default=12)
    ap.add_argument("--seed", type=int, default=3000)
    ap.add_argument("--cpu-rhs", type=int, default=3, help="RHS in the CPU baseline sample (0: skip)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--legs", default="mixed,plain", help="comma list of GPU legs to time")
    ap.add_argument("--rhs-groups", default="2,2",
                    help="RHS groups whose passes run concurrently, mixed,plain legs")
    ap.add_argument("--kernel-timing", choices=("after", "timed"), default="after",
                    help="time the correction kernel with HIP events in one extra solve after "
                         "the timed steps (after) or inside them (timed: event nodes in the "
                         "captured pass lengthen the timed region)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (reference kernel for HBM counters)")
    return ap.parse_args()


def copy_calibration():
    """One device-to-device copy of exactly 2**30 bytes: FETCH_SIZE and
    WRITE_SIZE of this kernel scale the counters of the measured kernels."""
    import torch
    a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": 2 ** 30, "ok": bool(b[-1].item() == 1.0)}))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
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
    s_np = synth.make_rhs(case.mask, a.seed + 1, nrhs)
    s = torch.from_numpy(np.ascontiguousarray(s_np.transpose(2, 1, 0))).to(dev)
    x = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device=dev)
    n0_of = {"plain": a.n0, "mixed": a.n0_mixed or a.n0}

    def make_plan(n0, prec):
        ie = synth.reference_ie(j2, synth.nb0_for_height(j2, n0))
        plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs, precision=prec)
        t0 = time.perf_counter()
        zp = plan.build_influence()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"influence matrices for block height {n0} built in {dt:.2f} s (NB0={len(ie)}, "
            f"zero pivots {zp}, {plan.device_bytes / 2**30:.2f} GiB on device)")
        return plan, ie, dt

    def plan_groups(name):
        return dict(zip(("mixed", "plain"), (int(v) for v in a.rhs_groups.split(",")))).get(name, 1)

    def run_leg(plan, name):
        for _ in range(a.warmup):
            x.zero_()
            plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
        torch.cuda.synchronize()
        if a.kernel_timing == "timed":
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
        if a.kernel_timing == "timed":
            corr_ms, corr_n = plan.profile_end()
        else:
            # the same solve once more, its correction products bracketed by
            # HIP events (event-record nodes in the captured pass)
            plan.profile_begin(S.KCLASS_CORR)
            x.zero_()
            plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
            corr_ms, corr_n = plan.profile_end()
        r = {"solves_per_s": a.steps * nrhs / dt, "ms_per_step": 1e3 * dt / a.steps,
             "block_height": n0_of[name], "nb0": plan.nb0,
             "passes": float(np.mean(iters)), "max_rel_resid": worst, "failed_steps": fails,
             "corr_avg_us": 1e3 * corr_ms / max(corr_n, 1), "corr_launches": corr_n,
             "corr_share": (corr_ms * (a.steps if a.kernel_timing == "after" else 1)
                            / (1e3 * dt)),
             "kernel_timing": a.kernel_timing, "rhs_groups": plan_groups(name)}
        log(f"{name}: {r}")
        return r

    legs, setup, xs = {}, {}, {}
    order = [l for l in a.legs.split(",") if l]
    if a.skip_plain and "plain" in order:
        order.remove("plain")
    plan_keep = None
    for name in order:
        prec = {"mixed": S.PREC_FP32, "plain": S.PREC_FP64}[name]
        plan, ie, dt = make_plan(n0_of[name], prec)
        plan.set_rhs_groups(plan_groups(name))
        setup[name] = dt
        legs[name] = run_leg(plan, name)
        xs[name] = x.clone()
        if name == "plain" or plan_keep is None:
            if plan_keep is not None:
                plan_keep[0].close()
            plan_keep = (plan, ie)
        else:
            plan.close()
        torch.cuda.empty_cache()
    t_setup = setup

    # error of the mixed solution against the plain FP64 solution (the plain
    # path is bitwise the reference REP1 arithmetic)
    err = None
    if "mixed" in legs and "plain" in legs:
        xm, xp = xs["mixed"], xs["plain"]
        d = (xm - xp)[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
        n = xp[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
        err = float((d / n).max())
        log(f"mixed vs plain FP64: max rel error {err:.3e}")

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
        cpu = cpu_baseline(plan_keep[0], case, plan_keep[1], s_np, a)

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
        "config": {"workload": f"SEVP {i2}x{j2} (0.25 deg zonal, +-{a.lat_max:g} deg), REP1 form, "
                               f"{nrhs} RHS per batch, each to {a.tol:g} rel resid",
                   "i2": i2, "j2": j2, "nrhs": nrhs, "tol": a.tol,
                   "block_height": n0_of},
        "legs": legs,
        "mixed_over_plain": (legs["mixed"]["solves_per_s"] / legs["plain"]["solves_per_s"]
                             if "mixed" in legs and "plain" in legs else None),
        "mixed_vs_fp64_max_rel_err": err,
        "setup_s": t_setup,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan_keep[0].close()


def cpu_baseline(plan, case, ie, s_np, a):
    """The oracle on a bounded sample of the same workload: each RHS solved to
    the same tolerance by FP64 REP1 passes with residual refinement (the
    residual passed to the next pass as the float32 right-hand side the
    reference interface takes)."""
    from oracle import oracle as O
    O.build()
    log("cpu baseline: copying influence matrices to the host")
    rinv, rinv1 = plan.get_influence()
    i2, j2 = case.shape
    t_all, worst, passes = 0.0, 0.0, 0
    for r in range(a.cpu_rhs):
        sr = np.asarray(s_np[:, :, r], np.float64)
        x = np.zeros((i2 + 2, j2 + 2), order="F")
        t0 = time.perf_counter()
        for _ in range(a.max_iter):
            res = sr - O.sqprod(case.coefs, x)[1:-1, 1:-1]
            rel = float(np.linalg.norm(res) / np.linalg.norm(sr))
            if rel <= a.tol:
                break
            d = np.zeros((i2 + 2, j2 + 2), order="F")
            O.rep1(case.coefs, rinv, rinv1, np.asfortranarray(res.astype(np.float32)), d, ie)
            x += d
            passes += 1
        t_all += time.perf_counter() - t0
        worst = max(worst, rel)
        log(f"cpu rhs {r}: {t_all:.2f} s cumulative, rel resid {rel:.2e}")
    return {"value": a.cpu_rhs / t_all, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{a.cpu_rhs} of the {a.nrhs} RHS, each refined to {a.tol:g} with FP64 "
                      f"REP1 passes ({passes / a.cpu_rhs:.1f} per RHS, max rel resid {worst:.2e}); "
                      "influence matrices from the GPU build",
            "rel_resid": worst}


if __name__ == "__main__":
    main()
