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
  fast64     correction products on the FP64 matrix cores (FP64 arithmetic in
             a blocked summation order), refined the same way
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
    ap.add_argument("--cpu-rhs", type=int, default=16, help="RHS in the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=16, help="host threads of the CPU baseline")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--legs", default="mixed,plain,fast64", help="comma list of GPU legs to time")
    ap.add_argument("--rhs-groups", default="1,2,1",
                    help="RHS groups whose passes run concurrently, mixed,plain,fast64 legs "
                         "(the fused pass reads each influence matrix once per group)")
    ap.add_argument("--pass-kernels", choices=("fused", "splitk"), default="fused",
                    help="pass kernels of the MFMA legs: one launch per block step (fused) or "
                         "the split-K products with separate marches (splitk)")
    ap.add_argument("--fused-waves", type=int, default=0,
                    help="wavefronts per fused workgroup without the autotuner (8 or 16; 0: default)")
    ap.add_argument("--fused-rows", type=int, default=0,
                    help="16-row tiles per fused workgroup without the autotuner (1 or 2; 0: default)")
    ap.add_argument("--tenth", type=int, default=1,
                    help="also run the 0.1 degree leg: 3600x1800, 16 RHS, FP32 MFMA corrections "
                         "refined to the tolerance, products-only plan (0: skip)")
    ap.add_argument("--tenth-only", action="store_true", help="run the 0.1 degree leg alone")
    ap.add_argument("--tenth-j2", type=int, default=1800,
                    help="rows of the 0.1 degree leg (fewer rows keep every kernel's shape: counter "
                         "runs, whose per-dispatch cost would make the 900-block build too slow)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="choose pass kernels, waves per fused workgroup and RHS groups of the MFMA "
                         "legs by timing on the device (musoc_solver_amd.autotune); 0: the flags")
    ap.add_argument("--bf16-passes", type=int, default=4,
                    help="passes of the BF16 trial on the first RHS (0: skip)")
    ap.add_argument("--kernel-timing", choices=("after", "timed"), default="after",
                    help="time the correction kernel with HIP events in one extra solve after "
                         "the timed steps, launched eagerly with events from the kernel's own "
                         "timestamps (after), or inside the timed steps through event nodes in "
                         "the captured pass, which also count the dispatch gap (timed)")
    ap.add_argument("--krylov", type=int, default=1,
                    help="also time the device BiCGSTAB cross-check on the first RHS (0: skip)")
    ap.add_argument("--krylov-only", action="store_true",
                    help="only the BiCGSTAB leg (for kernel profiles of its iteration)")
    ap.add_argument("--basins", type=int, default=256,
                    help="regional basins refined together (sevp_batch) beside one by one (0: skip)")
    ap.add_argument("--basin-rhs", type=int, default=8, help="right-hand sides per basin")
    ap.add_argument("--basin-prec", choices=("mixed", "fast64"), default="mixed",
                    help="basin corrections: FP32 MFMA (mixed) or FP64 MFMA (fast64)")
    ap.add_argument("--basins-only", action="store_true", help="run the basins leg alone")
    ap.add_argument("--sequence", type=int, default=200,
                    help="sequential single-RHS solves at 720x360, warm-started (0: skip)")
    ap.add_argument("--sequence-only", action="store_true", help="run the sequence leg alone")
    ap.add_argument("--bir", type=int, default=1,
                    help="also time the reference's default form (5 BIR sweeps) at 720x360, "
                         "NBIR 60, 2-row blocks, 1, 64 and 256 RHS (0: skip)")
    ap.add_argument("--bir-only", action="store_true", help="run the BIR leg alone")
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
    n0_of = {"plain": a.n0, "mixed": a.n0_mixed or a.n0, "fast64": a.n0}

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
        return dict(zip(("mixed", "plain", "fast64"),
                        (int(v) for v in a.rhs_groups.split(",")))).get(name, 1)

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
            # the same solve once more with its kernels launched one by one,
            # every correction product carrying HIP events recorded from its
            # own start and end timestamps (hipExtLaunchKernelGGL)
            plan.set_graphs(False)
            plan.profile_begin(S.KCLASS_CORR)
            x.zero_()
            plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
            span_ms, span_n = plan.profile_spans()
            corr_ms, corr_n = plan.profile_end()
            plan.set_graphs(True)
        # HIP-event duration per launch (the kernel's own dispatch timestamps
        # in the eager solve, event nodes in the timed graph replays); the
        # span from first wavefront start to last wavefront end, from the GPU
        # clock, is reported beside it
        avg_us = 1e3 * corr_ms / max(corr_n, 1)
        share_ms = corr_ms * (a.steps if a.kernel_timing == "after" else 1)
        span_us = 1e3 * span_ms / span_n if a.kernel_timing == "after" and span_n else None
        r = {"solves_per_s": a.steps * nrhs / dt, "ms_per_step": 1e3 * dt / a.steps,
             "block_height": n0_of[name], "nb0": plan.nb0,
             "passes": float(np.mean(iters)), "max_rel_resid": worst, "failed_steps": fails,
             "corr_avg_us": avg_us, "corr_span_us": span_us,
             "corr_launches": corr_n,
             "corr_share": share_ms / (1e3 * dt),
             "kernel_timing": a.kernel_timing, "rhs_groups": plan_groups(name)}
        log(f"{name}: {r}")
        return r

    if a.basins_only:
        print(json.dumps({"basins": basins_leg(a)}), flush=True)
        return
    if a.sequence_only:
        print(json.dumps({"sequence": sequence_leg(a)}), flush=True)
        return
    if a.bir_only:
        print(json.dumps({"bir": bir_leg(a)}), flush=True)
        return
    if a.tenth_only:
        print(json.dumps({"tenth": tenth_leg(a)}), flush=True)
        return
    legs, setup, xs = {}, {}, {}
    order = [l for l in a.legs.split(",") if l]
    if a.krylov_only:  # BiCGSTAB needs the operator only, no influence matrices
        ie = synth.reference_ie(j2, synth.nb0_for_height(j2, a.n0))
        plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC)
        print(json.dumps({"krylov": krylov_leg(plan, s, None, a)}), flush=True)
        plan.close()
        return
    if a.skip_plain and "plain" in order:
        order.remove("plain")
    # one plan per block height; the legs switch its correction arithmetic
    prec_of = {"mixed": S.PREC_FP32, "plain": S.PREC_FP64, "fast64": S.PREC_FP64_MFMA}
    plans = {}
    for name in order:
        n0 = n0_of[name]
        if n0 not in plans:
            plan, ie, dt = make_plan(n0, S.PREC_FP64)
            plans[n0] = (plan, ie)
            setup[name] = dt
        plan = plans[n0][0]
        t0 = time.perf_counter()
        plan.set_pass(S.PASS_FUSED if a.pass_kernels == "fused" else S.PASS_SPLITK)
        plan.set_precision(prec_of[name])
        torch.cuda.synchronize()
        setup[name] = setup.get(name, 0.0) + time.perf_counter() - t0
        plan.set_rhs_groups(plan_groups(name))
        if name != "plain" and (a.fused_waves or a.fused_rows):
            t = plan.tuning
            plan.set_tuning(t.pass_kind, a.fused_waves or t.waves, t.rhs_groups, a.fused_rows)
        tuning = None
        if a.autotune and name != "plain":
            from musoc_solver_amd import autotune
            t0 = time.perf_counter()
            tuning = autotune.tune(plan, s, x, tol=a.tol, max_iter=a.max_iter)
            tuning["tuning_s"] = time.perf_counter() - t0
            log(f"{name}: autotuned {tuning}")
        legs[name] = run_leg(plan, name)
        legs[name]["tuning"] = tuning
        if tuning:
            legs[name]["rhs_groups"] = tuning["rhs_groups"]
        xs[name] = x.clone()
    plan_keep = plans[n0_of[order[-1]]] if order else None
    if plan_keep is not None and "plain" in order:
        plan_keep = plans[n0_of["plain"]]
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
    err64 = None
    if "fast64" in legs and "plain" in legs:
        xf, xp = xs["fast64"], xs["plain"]
        d = (xf - xp)[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
        n = xp[:, 1:-1, 1:-1].reshape(nrhs, -1).norm(dim=1)
        err64 = float((d / n).max())
        log(f"fast64 vs plain FP64: max rel error {err64:.3e}")

    # the BF16 trial: one pass at a time from x = 0 on the first right-hand
    # side, BF16 matrices and error rows with FP32 accumulation (fused pass)
    bf16 = None
    if a.bf16_passes > 0 and a.pass_kernels == "fused":
        p_b = plans[n0_of[order[0]]][0]
        try:
            p_b.set_pass(S.PASS_FUSED)
            p_b.set_precision(S.PREC_BF16)
            xb = torch.zeros((1, j2 + 2, i2 + 2), dtype=torch.float64, device=dev)
            hist = [1.0]
            for _ in range(a.bf16_passes):
                rel_b, _, _ = p_b.solve_refined_dev(s[:1], xb, tol=0.0, max_iter=1)
                hist.append(float(rel_b[0]))
            bf16 = {"rel_resid_after_pass": hist[1:],
                    "contraction": [hist[k + 1] / hist[k] for k in range(len(hist) - 1)]}
            log(f"bf16 trial: {bf16}")
        except Exception as e:  # noqa: BLE001 - reported, not hidden
            bf16 = {"error": repr(e)}
    main_name = next(n for n in ("mixed", "plain", "fast64") if n in legs)
    main_leg = legs[main_name]
    def leg_fused(name):
        t = legs.get(name, {}).get("tuning")
        return (t["pass_kernels"] == "fused") if t else a.pass_kernels == "fused"
    roof = roofline(main_leg, main_name, nrhs, i2, a, leg_fused(main_name))
    krylov = krylov_leg(plan_keep[0], s, xs.get(main_name), a) if a.krylov else None
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
        "dtype": {"mixed": "f64+f32mfma", "fast64": "f64mfma"}.get(main_name, "f64"),
        "data": "synthetic",
        "config": {"workload": f"SEVP {i2}x{j2} (0.25 deg zonal, +-{a.lat_max:g} deg), REP1 form, "
                               f"{nrhs} RHS per batch, each to {a.tol:g} rel resid",
                   "i2": i2, "j2": j2, "nrhs": nrhs, "tol": a.tol,
                   "block_height": n0_of},
        "legs": legs,
        "mixed_over_plain": (legs["mixed"]["solves_per_s"] / legs["plain"]["solves_per_s"]
                             if "mixed" in legs and "plain" in legs else None),
        "fast64_over_plain": (legs["fast64"]["solves_per_s"] / legs["plain"]["solves_per_s"]
                              if "fast64" in legs and "plain" in legs else None),
        "mixed_vs_fp64_max_rel_err": err,
        "fast64_vs_fp64_max_rel_err": err64,
        "roofline_fast64": (roofline(legs["fast64"], "fast64", nrhs, i2, a, leg_fused("fast64"))
                            if "fast64" in legs else None),
        "roofline_plain": (roofline(legs["plain"], "plain", nrhs, i2, a, False)
                           if "plain" in legs else None),
        "setup_s": t_setup,
        "roofline": roof,
        "bf16_trial": bf16,
        "cpu_baseline": cpu,
        "krylov": krylov,
        "basins": basins_or_error(a),
        "sequence": sequence_or_error(a),
        "bir": bir_or_error(a),
    }
    for plan, _ in plans.values():
        plan.close()
    torch.cuda.empty_cache()
    out["tenth"] = tenth_or_error(a)
    print(json.dumps(out), flush=True)


def corr_work(name, fused, m, k, nb0):
    """Algorithmic flops and bytes per launch of the correction kernels,
    averaged over the launches of one pass (block height 2, K = N = I2 = k,
    m right-hand sides per launch, operand elements of esz bytes).

    Split-K and plain passes: NB0-1 forward products (N = K) and NB0
    backward ones (N = 2K with the coupling row for NB0-1 blocks, N = K for
    the bottom block; the plain pass runs the coupling as a separate FP64
    product, so it has 3 NB0 - 2 products of N = K), each
    2 m K N flops and esz (m K + K N + m N) bytes.

    Fused pass: the product's operands esz (m K + K N) (its result stays in
    the workgroup) plus what the epilogue reads and writes per right-hand
    side and column -- forward: X rows south, free and north (3 x 8 B) and
    the marched row (8 B), two FP64 right-hand-side rows (16 B), the next
    error row (esz) and 2 x 5 coefficient rows per column (40 B); backward:
    the free row read and written, the two rows of the closure residual
    (32 B), the right-hand side (8 B), the error row (esz), c and cpl (16 B),
    one coefficient row (20 B per column), and the homogeneous march of the
    block above (c and cpl read, one X row read and written: 32 B, 20 B of
    coefficients per column)."""
    esz = 4 if name == "mixed" else 8
    if name == "plain" or not fused:
        if name == "plain":
            ncols = [(k, nb0 - 1), (k, nb0), (k, nb0 - 1)]
        else:
            ncols = [(k, nb0 - 1), (2 * k, nb0 - 1), (k, 1)]
        launches = sum(c for _, c in ncols)
        flops = sum(2.0 * m * k * n * c for n, c in ncols) / launches
        nbytes = sum(esz * (m * k + k * n + m * n) * c for n, c in ncols) / launches
        return flops, nbytes, launches
    fwd_f = 2.0 * m * k * k
    fwd_b = esz * (m * k + k * k) + m * k * (32 + 16 + esz) + 40 * k
    bwd_f = 2.0 * m * k * 2 * k
    bwd_b = esz * (m * k + 2 * k * k) + m * k * (16 + 32 + 8 + esz + 16 + 32) + 40 * k
    bot_f = 2.0 * m * k * k
    bot_b = esz * (m * k + k * k) + m * k * (16 + 8 + esz + 16 + 32) + 20 * k
    launches = (nb0 - 1) + nb0
    flops = ((nb0 - 1) * (fwd_f + bwd_f) + bot_f) / launches
    nbytes = ((nb0 - 1) * (fwd_b + bwd_b) + bot_b) / launches
    return flops, nbytes, launches


def roofline(leg, name, nrhs, i2, a, fused=True):
    """Roofline of the dominant kernels, the correction launches (the fused
    kernels of the fused pass, the split-K products, or the plain FP64
    products), from their average launch duration (HIP events with the
    kernels' own start and end timestamps, one solve after the timed steps)
    and the algorithmic work of corr_work.  The bound is whichever of the
    MFMA (or FP64 vector, plain) time and the HBM time at peak is larger for
    the average launch; both fractions are reported."""
    nb0, groups = leg["nb0"], leg["rhs_groups"]
    m = -(-nrhs // groups)
    fused = fused and name != "plain"
    flops, nbytes, _ = corr_work(name, fused, m, i2, nb0)
    t = leg["corr_avg_us"] * 1e-6
    peak_f = PEAK_FP32_MFMA if name == "mixed" else PEAK_FP64
    mfma_frac = flops / t / peak_f
    hbm_frac = nbytes / t / PEAK_HBM
    if nbytes / PEAK_HBM >= flops / peak_f:
        r = {"bound": "hbm", "achieved": nbytes / t / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
             "frac": hbm_frac}
    else:
        r = {"bound": "mfma", "achieved": flops / t / 1e12, "peak": peak_f / 1e12,
             "unit": "TFLOP/s", "frac": mfma_frac}
    if fused:
        kern = "fused_fwd_kernel + fused_bwd_kernel (%s)" % ("float" if name == "mixed" else "double")
    else:
        kern = {"mixed": "gemm_part_kernel", "fast64": "gemm_part64_kernel"}.get(name, "gemm_seq_kernel")
    r.update({"traffic": measured_traffic(name, m, i2, fused),
              "kernel": kern,
              "rhs_groups": groups,
              "durations": ("kernel's own: the profiled solve runs the RHS groups one after "
                            "another; the timed steps run them on %d concurrent streams, so "
                            "corr_ms_per_step may exceed ms_per_step" % groups) if groups > 1
                           else "kernel's own (one RHS group)",
              "compute_peak": "fp32 mfma" if name == "mixed" else (
                  "fp64 vector (separate multiply and add)" if name == "plain" else "fp64 mfma"),
              "mfma_frac": mfma_frac, "hbm_frac": hbm_frac,
              "algorithmic_bytes_per_launch": nbytes, "algorithmic_flops_per_launch": flops,
              "avg_launch_us": leg["corr_avg_us"], "span_us": leg.get("corr_span_us"),
              "rhs_per_launch": m,
              "corr_ms_per_step": leg["corr_avg_us"] * 1e-3 * leg["corr_launches"],
              "ms_per_step": leg["ms_per_step"]})
    return r


def kernel_source_sha():
    """Hash of the kernel sources the correction launches are built from: a
    stored traffic record applies only to the build it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in ("sevp_fused.hip", "sevp_mixed.hip", "sevp_full.hip", "sevp_march.h", "sevp_dev.h",
              "sevp_internal.h"):
        with open(os.path.join(ROOT, "musoc_solver_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def tenth_or_error(a):
    """The 0.1 degree leg after the main legs (their plans are closed first:
    this one needs 143 GiB), isolated like the basins leg."""
    if a.tenth <= 0:
        return None
    try:
        return tenth_leg(a)
    except Exception as e:  # noqa: BLE001 - reported, not hidden
        log(f"tenth leg failed: {e!r}")
        return {"error": repr(e)}


def tenth_leg(a):
    """The 0.1 degree configuration on one GPU: 3600x1800 cells (uniform
    latitude spacing over +-80 deg, the benchmark operator's form), block
    height 2 (NB0 = 900), 16 right-hand sides per call, FP32 MFMA corrections
    refined in FP64 to the tolerance.  The FP64 influence matrices of this
    layout would take 186 GB, so the plan keeps only the FP32 product
    operands (RINV, RINV1, W: 140 GB), built one block at a time
    (sevp_set_storage(SEVP_STORE_PRODUCTS)).  Reports solves/s, passes,
    residuals, and the fused correction kernels' roofline (their HBM bytes
    include the march rows of the epilogue)."""
    import torch
    from musoc_solver_amd import synth
    from musoc_solver_amd import sevp as S
    i2, j2, nrhs = 3600, a.tenth_j2, 16
    case = synth.make_case(i2, j2, a.seed + 5000, metric="latlon", lat_max=a.lat_max,
                           h_land=a.h_land)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    s_np = synth.make_rhs(case.mask, a.seed + 5001, nrhs)
    s = torch.from_numpy(np.ascontiguousarray(s_np.transpose(2, 1, 0))).cuda()
    x = torch.zeros((nrhs, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=nrhs,
                      precision=S.PREC_FP32, storage=S.STORE_PRODUCTS)
    try:
        plan.set_pass(S.PASS_FUSED if a.pass_kernels == "fused" else S.PASS_SPLITK)
        t0 = time.perf_counter()
        zp = plan.build_influence()
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        log(f"tenth: {i2}x{j2} FP32 operands of NB0={len(ie)} blocks built in {setup:.1f} s "
            f"({plan.device_bytes / 2**30:.1f} GiB, zero pivots {zp})")
        tuning = None
        if a.autotune:
            from musoc_solver_amd import autotune
            tuning = autotune.tune(plan, s, x, tol=a.tol, max_iter=a.max_iter)
            log(f"tenth: autotuned {tuning}")
        plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)  # warm-up (graph capture)
        steps = max(1, min(a.steps, 3))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        passes, worst, ok_all = [], 0.0, True
        for _ in range(steps):
            x.zero_()
            rel, st, ok = plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
            passes.append(st.iterations)
            worst = max(worst, float(rel.max()))
            ok_all = ok_all and ok
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        rel_chk = plan.residual_dev(s, x)  # independent FP64 residual of the returned X
        plan.set_graphs(False)
        plan.profile_begin(S.KCLASS_CORR)
        x.zero_()
        plan.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
        span_ms, span_n = plan.profile_spans()
        corr_ms, corr_n = plan.profile_end()
        leg = {"solves_per_s": nrhs / dt, "ms_per_step": 1e3 * dt, "block_height": 2,
               "nb0": len(ie), "passes": float(np.mean(passes)), "max_rel_resid": worst,
               "residual_check_max": float(np.max(rel_chk)), "converged": ok_all,
               "corr_avg_us": 1e3 * corr_ms / max(corr_n, 1),
               "corr_span_us": 1e3 * span_ms / span_n if span_n else None,
               "corr_launches": corr_n, "rhs_groups": plan.tuning.rhs_groups, "setup_s": setup,
               "tuning": tuning,
               "device_gib": plan.device_bytes / 2**30, "storage": "fp32 products only",
               "workload": f"SEVP {i2}x{j2} (0.1 deg zonal, +-{a.lat_max:g} deg), REP1 form, "
                           f"{nrhs} RHS per call, each to {a.tol:g} rel resid"}
        leg["roofline"] = roofline(leg, "mixed", nrhs, i2, a, plan.pass_kind == S.PASS_FUSED)
        log(f"tenth: {leg}")
        return leg
    finally:
        plan.close()


def bir_or_error(a):
    """The BIR leg after the main legs, isolated like the basins leg."""
    if a.bir <= 0:
        return None
    try:
        return bir_leg(a)
    except Exception as e:  # noqa: BLE001 - reported, not hidden
        log(f"bir leg failed: {e!r}")
        return {"error": repr(e)}


def bir_leg(a):
    """The reference's default solver form (FL_EVP_STP = 1): 5 multiplicative
    BIR sweeps over 60 overlapping 24-column strips of the 0.5 degree grid
    (720x360), REPBIR per strip, 2-row blocks, plain FP64 and bitwise the
    reference's arithmetic.  One solve is a dependent chain of 5 x 60 x 720 row
    steps (each a 77 ns FP64 chain on its own), so it is timed for one RHS and
    for 64 and 256 independent RHS (one workgroup, i.e. one CU, each); the
    first RHS of the 64-RHS batch must equal the single solve bit for bit."""
    import torch
    from musoc_solver_amd import synth
    from musoc_solver_amd import sevp as S
    i2, j2, nbir, h, nb, nbig = 720, 360, 60, 2, 64, 256
    case = synth.make_case(i2, j2, a.seed + 720, metric="latlon", lat_max=a.lat_max, h_land=a.h_land)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, h))
    plan = S.SevpPlan(case.coefs, ie, form=S.FORM_BIR, nbir=nbir, nsweep=5, nrhs_max=nbig)
    try:
        t0 = time.perf_counter()
        plan.build_influence()
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        s_np = synth.make_rhs(case.mask, a.seed + 721, nbig)
        sd = torch.from_numpy(np.ascontiguousarray(s_np.transpose(2, 1, 0))).cuda()
        out = {"workload": f"{i2}x{j2} NBIR {nbir} (24-column strips), block height {h}, 5 sweeps",
               "setup_s": setup}
        xs = {}
        for n in (1, nb, nbig):
            x = torch.zeros((n, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
            plan.solve_dev(sd[:n], x)  # warm-up
            ts = []
            for _ in range(3):
                x.zero_()
                plan.reset_seam()  # the seam scratch XX carries over between solves
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                plan.solve_dev(sd[:n], x)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t1)
            xs[n] = x
            out[f"ms_per_call_{n}rhs"] = 1e3 * min(ts)
            out[f"solves_per_s_{n}rhs"] = n / min(ts)
        row_steps = 5 * nbir * 2 * j2
        out["row_step_ns_1rhs"] = 1e6 * out["ms_per_call_1rhs"] / row_steps
        out["chain_ns"] = 77.0
        out["batch_first_equals_single"] = bool(torch.equal(xs[1][0], xs[nb][0]))
        if a.cpu_rhs > 0:
            out["cpu_baseline"] = bir_cpu_baseline(plan, case, ie, nbir, s_np, xs[1][0])
        log(f"bir: {out}")
        return out
    finally:
        plan.close()


def bir_cpu_baseline(plan, case, ie, nbir, s_np, x_gpu):
    """The BIR leg's CPU baseline: one solve of the same workload (the first
    right-hand side, 5 sweeps from x = 0) by the oracle's C FS_SOLVER loop
    around its REPBIR restatement, on one host core, with the influence
    matrices of the GPU build (bitwise the reference's).  Its result must
    equal the GPU's single solve bit for bit."""
    from oracle import oracle as O
    O.build()
    rinv, rinv1 = plan.get_influence()
    i2, j2 = case.shape
    ts = []
    for _ in range(2):
        x = np.zeros((i2 + 2, j2 + 2), order="F")
        xx = np.zeros((2 * i2 // nbir + 2, j2 + 2), order="F")
        t0 = time.perf_counter()
        O.bir_solve(case.coefs, s_np[:, :, 0], rinv, rinv1, x, xx, ie, nbir)
        ts.append(time.perf_counter() - t0)
    t = min(ts)
    xg = x_gpu.cpu().numpy().T
    return {"value": 1.0 / t, "unit": "solves/s", "cores": 1, "kind": "port",
            "ms_per_solve": 1e3 * t,
            "equals_gpu": bool(np.array_equal(x.view(np.int64), xg.view(np.int64))),
            "sample": "one 1-RHS solve (5 sweeps from x = 0), best of two, oracle C loop"}


def sequence_or_error(a):
    """The sequence leg after the main legs, isolated like the basins leg."""
    if a.sequence <= 0:
        return None
    try:
        return sequence_leg(a)
    except Exception as e:  # noqa: BLE001 - reported, not hidden
        log(f"sequence leg failed: {e!r}")
        return {"error": repr(e)}


def sequence_leg(a):
    """Sequential single-RHS solves on the 0.5 degree grid (720x360): each
    step's forcing is N(0,1) plus a slowly drifting low-wavenumber field, and
    each solve starts from the previous step's solution and is refined to the
    tolerance with the REP1 form (block height 2), as a model's time loop
    calls the solver.  Reports milliseconds per solve for each correction
    arithmetic on one plan (the MFMA ones autotuned once on the first
    forcing)."""
    import torch
    from musoc_solver_amd import synth
    from musoc_solver_amd import sevp as S
    i2, j2, n = 720, 360, a.sequence
    case = synth.make_case(i2, j2, a.seed + 2000, metric="latlon", lat_max=a.lat_max,
                           h_land=a.h_land)
    ie = synth.reference_ie(j2, synth.nb0_for_height(j2, 2))
    forcing = synth.make_rhs(case.mask, a.seed + 2001, n, lowfreq=2.0)
    sd = [torch.from_numpy(np.ascontiguousarray(forcing[:, :, k:k + 1].transpose(2, 1, 0))).cuda()
          for k in range(n)]
    out = {"grid": f"{i2}x{j2}", "solves": n, "block_height": 2, "nb0": len(ie), "tol": a.tol}
    with S.SevpPlan(case.coefs, ie, form=S.FORM_FULL_PERIODIC, nrhs_max=1) as plan:
        plan.build_influence()
        for name, prec in (("plain", S.PREC_FP64), ("fast64", S.PREC_FP64_MFMA),
                           ("mixed", S.PREC_FP32)):
            plan.set_precision(prec)
            x = torch.zeros((1, j2 + 2, i2 + 2), dtype=torch.float64, device="cuda")
            tuning = None
            if a.autotune and name != "plain":  # once per plan, as a model's setup would
                from musoc_solver_amd import autotune
                tuning = autotune.tune(plan, sd[0], x, tol=a.tol, max_iter=a.max_iter)
                x.zero_()
            plan.solve_refined_dev(sd[0], x, tol=a.tol, max_iter=a.max_iter)  # warm-up
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            passes, worst = [], 0.0
            for k in range(n):
                rel, st, ok = plan.solve_refined_dev(sd[k], x, tol=a.tol, max_iter=a.max_iter)
                passes.append(st.iterations)
                worst = max(worst, float(rel.max()))
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / n
            out[name] = {"ms_per_solve": 1e3 * dt, "passes_mean": float(np.mean(passes)),
                         "max_rel_resid": worst}
            if tuning:
                out[name]["tuning"] = tuning
    log(f"sequence: {out}")
    return out


def basins_or_error(a):
    """The basins leg after the main legs: a failure there is reported in its
    own field and leaves the main line intact."""
    if a.basins <= 0:
        return None
    try:
        return basins_leg(a)
    except Exception as e:  # noqa: BLE001 - reported, not hidden
        log(f"basins leg failed: {e!r}")
        return {"error": repr(e)}


def basins_leg(a):
    """Regional basins (sizes log-uniform in 64x64 .. 512x384, 20-45% land,
    IE from the growth autotuner) with FP32 MFMA corrections and FP64
    refinement to the tolerance (or FP64 MFMA corrections with --basin-prec
    fast64): all basins refined together by one batch
    (one grouped launch per pass step, sevp_batch_solve_refined_dev) against
    the same basins refined one plan at a time; the batch's solutions must
    equal the plans' bit for bit."""
    import torch
    from musoc_solver_amd import autotune, synth
    from musoc_solver_amd import sevp as S
    nrhs = a.basin_rhs
    t0 = time.perf_counter()
    plans, ss, nb0 = [], [], []
    for i2, j2, land, seed in synth.basin_specs(a.basins, a.seed + 4000):
        case = synth.make_case(i2, j2, seed, metric="latlon", lat_max=60.0, land_frac=land,
                               h_land=a.h_land)
        ie = autotune.adaptive_ie(case.coefs, max_growth=1e3, hmax=8, hmin=2, periodic=False)
        prec = S.PREC_FP32 if a.basin_prec == "mixed" else S.PREC_FP64_MFMA
        p = S.SevpPlan(case.coefs, ie, form=S.FORM_FULL, nrhs_max=nrhs, precision=prec)
        p.set_pass(S.PASS_SPLITK)  # the batch's pass kernels, also for the one-by-one solves
        p.build_influence()
        s_np = synth.make_rhs(case.mask, seed + 1, nrhs)
        ss.append(torch.from_numpy(np.ascontiguousarray(s_np.transpose(2, 1, 0))).cuda())
        plans.append(p)
        nb0.append(len(ie))
        if len(plans) % 32 == 0:
            log(f"basins: {len(plans)} of {a.basins} plans built")
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    log(f"basins: {len(plans)} plans built in {setup:.1f} s (NB0 max {max(nb0)}, sum {sum(nb0)}, "
        f"{sum(p.device_bytes for p in plans) / 2**30:.2f} GiB on device)")
    zeros = lambda: [torch.zeros((nrhs, p.j2 + 2, p.i2 + 2), dtype=torch.float64, device="cuda")
                     for p in plans]
    reps = max(1, min(a.steps, 3))
    # one plan after another
    x1 = zeros()
    worst1, passes1 = 0.0, []
    for rep in range(reps + 1):  # the first repetition warms up (graph capture)
        if rep == 1:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        passes1 = []
        for p, s, x in zip(plans, ss, x1):
            x.zero_()
            rel, st, ok = p.solve_refined_dev(s, x, tol=a.tol, max_iter=a.max_iter)
            passes1.append(st.iterations)
            worst1 = max(worst1, float(rel.max()))
    torch.cuda.synchronize()
    dt1 = (time.perf_counter() - t1) / reps
    # all basins together
    xb = zeros()
    with S.SevpBatch(plans) as batch:
        for rep in range(reps + 1):
            if rep == 1:
                torch.cuda.synchronize()
                t2 = time.perf_counter()
            for x in xb:
                x.zero_()
            rel, passes, st, ok = batch.solve_refined_dev(ss, xb, tol=a.tol, max_iter=a.max_iter)
        torch.cuda.synchronize()
        dt2 = (time.perf_counter() - t2) / reps
    same = all(bool(torch.equal(u, v)) for u, v in zip(x1, xb))
    worst = max(float(r.max()) for r in rel)
    n = len(plans) * nrhs
    r = {"basins": len(plans), "rhs_per_basin": nrhs,
         "precision": "f64+f32mfma" if a.basin_prec == "mixed" else "f64mfma",
         "batched_solves_per_s": n / dt2, "one_by_one_solves_per_s": n / dt1,
         "speedup": dt1 / dt2, "ms_batched": 1e3 * dt2, "ms_one_by_one": 1e3 * dt1,
         "passes_max": int(np.max(passes)), "passes_one_by_one": int(np.max(passes1)),
         "nb0_max": int(max(nb0)), "nb0_sum": int(sum(nb0)),
         "launches_per_pass_batched": 4 * int(max(nb0)) + 2,
         "max_rel_resid": worst, "converged": bool(ok), "bitwise_equal_to_plans": same,
         "setup_s": setup}
    log(f"basins: {r}")
    for p in plans:
        p.close()
    return r


def krylov_leg(plan, s, x_sevp, a):
    """The device BiCGSTAB (p_bicgstab_le, fast-order sums) on the first
    right-hand side from x = 0, to the reference's stop ||r||_2 <= 1e-8
    (REAL(4)), timed after one untimed run.  One iteration is three fused
    kernels; its algorithmic HBM traffic is 17 passes over the
    (I2+2) x (J2+2) FP64 vectors (PV: r, p, v, rh read, p, v written; ST: r,
    v read, s, t written; XR: x, p, s, t, rh read, x, r written) plus the
    five FP32 coefficient arrays read by the two 5-point products:
    136 (I2+2)(J2+2) + 40 I2 J2 bytes.  Compared with the SEVP solution
    as FS_SOLVER compares them (max difference over max |x|)."""
    import torch
    i2, j2 = a.i2, a.j2
    b = s[0]
    res = None
    for _ in range(2):
        x = torch.zeros((j2 + 2, i2 + 2), dtype=torch.float64, device=s.device)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = plan.bicgstab_dev(b, x)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    nbytes = 136.0 * (i2 + 2) * (j2 + 2) + 40.0 * i2 * j2
    per_it = dt / max(st.iterations, 1)
    r = {"solver": "BiCGSTAB (p_bicgstab_le), fixed-order sums", "iterations": st.iterations,
         "residual": st.residual, "converged": bool(st.converged), "ms": 1e3 * dt,
         "us_per_iteration": 1e6 * per_it, "bytes_per_iteration": nbytes,
         "hbm_gbs": nbytes / per_it / 1e9, "hbm_frac": nbytes / per_it / PEAK_HBM}
    if x_sevp is not None:
        xe = x_sevp[0]
        r["max_diff_vs_sevp"] = float((x - xe)[1:-1, 1:-1].abs().max() / xe[1:-1, 1:-1].abs().max())
    log(f"krylov: {r}")
    return r


def measured_traffic(name, m, i2, fused):
    """HBM bytes per correction launch as measured by the FETCH_SIZE /
    WRITE_SIZE counter runs recorded under profiles/ (scaled by the same
    counters of a copy of known size), for this configuration; null when no
    recorded run matches or the record was measured on other kernel sources
    (its source_sha differs from the current build's)."""
    path = os.path.join(ROOT, "profiles", "correction_product_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("source_sha") != kernel_source_sha():
        return None
    for r in rec.get("runs", []):
        if (r.get("leg") == name and r.get("rhs_per_launch") == m and r.get("i2") == i2
                and bool(r.get("fused")) == bool(fused)):
            return r.get("hbm_bytes_per_launch")
    return None


def cpu_baseline(plan, case, ie, s_np, a):
    """The CPU path on a bounded sample of the same workload: right-hand
    sides solved to the same tolerance by FP64 REP1 passes with residual
    refinement (the residual passed to the next pass as the float32
    right-hand side the reference interface takes), one right-hand side per
    host thread.  REP1 is the reference's own routine compiled from
    MAIN/template/GLOBAL/src/sevp.f90 (oracle/_ref): timed in its parity
    build (-O2, no contraction) and in a fast build (-O3, AVX2/FMA), whose
    rate is the reported value; the C restatement when neither library is
    present.  The influence matrices come from the GPU build (bitwise the
    reference PRE1).  REP1 is the CPU's fastest of the reference's forms at
    this tolerance: the BIR sweeps contract the residual by about 0.7 per
    sweep on this grid, some 55 sweeps (about 5 s per right-hand side) to
    1e-10, against one REP1 pass (0.9 s)."""
    import threading
    from oracle import oracle as O
    O.build()
    use_ref = O.ref_available()
    builds = ["fast", True] if (use_ref and O.ref_fast_available()) else [use_ref]
    log("cpu baseline: copying influence matrices to the host")
    rinv, rinv1 = plan.get_influence()
    i2, j2 = case.shape
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(a.cpu_threads, ncpu, a.nrhs))
    nrhs = min(a.cpu_rhs, a.nrhs)

    def run(build):
        results = [None] * nrhs

        def solve(r):
            sr = np.asarray(s_np[:, :, r], np.float64)
            x = np.zeros((i2 + 2, j2 + 2), order="F")
            passes, rel = 0, 1.0
            for _ in range(a.max_iter + 1):
                res = sr - O.sqprod(case.coefs, x)[1:-1, 1:-1]
                rel = float(np.linalg.norm(res) / np.linalg.norm(sr))
                if rel <= a.tol or passes >= a.max_iter:
                    break
                d = np.zeros((i2 + 2, j2 + 2), order="F")
                O.rep1(case.coefs, rinv, rinv1, np.asfortranarray(res.astype(np.float32)), d, ie,
                       use_ref=build)
                x += d
                passes += 1
            results[r] = (passes, rel)

        def worker(t):
            for r in range(t, nrhs, threads):
                solve(r)

        # one right-hand side alone first: the single-core time per solve
        t0 = time.perf_counter()
        solve(0)
        t_one = time.perf_counter() - t0
        t0 = time.perf_counter()
        pool = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
        for t in pool:
            t.start()
        for t in pool:
            t.join()
        t_all = time.perf_counter() - t0
        passes = [p for p, _ in results]
        worst = max(rel for _, rel in results)
        name = {"fast": "fast build", True: "parity build", False: "C restatement"}[build]
        log(f"cpu ({name}): 1 rhs in {t_one:.2f} s on one core; {nrhs} rhs in {t_all:.2f} s on "
            f"{threads} threads, max rel resid {worst:.2e}")
        return {"solves_per_s": nrhs / t_all, "single_core_s_per_solve": t_one,
                "passes_mean": float(np.mean(passes)), "rel_resid": worst, "build": name}

    runs = [run(b) for b in builds]
    best = runs[0]
    return {"value": best["solves_per_s"], "unit": "solves/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": f"{nrhs} of the {a.nrhs} RHS, one per thread, each refined to {a.tol:g} with "
                      f"FP64 REP1 passes ({best['passes_mean']:.1f} per RHS, max rel resid "
                      f"{best['rel_resid']:.2e}); REP1 "
                      f"{'compiled from the reference source, ' + best['build'] if use_ref else 'C restatement'}"
                      ", residuals by the C restatement of p_sqprod, influence matrices from the "
                      "GPU build",
            "single_core_s_per_solve": best["single_core_s_per_solve"], "rel_resid": best["rel_resid"],
            "builds": runs, "host_cpu": host_cpu()}


def host_cpu():
    """Model name and logical CPU count of the host the baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


if __name__ == "__main__":
    main()
