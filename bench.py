#!/usr/bin/env python3
"""Stokes solve time to relative residual 1e-9 on a 256^3 periodic box.

Workload: the 3D periodic variable-viscosity staggered Stokes problem
(visc_type -1, eta log-normal with contrast 100 and two passes of the
27-point mean, theta*rho in [1, 4], RHS = divergence of seeded white-noise
stress, reference-default solver: projection preconditioner, one V-cycle of
each multigrid, 2/2/8 sweeps, 6-colour Gauss-Seidel, restart 5, tol 1e-9).

One step = one complete GMRES solve from a zero initial guess with every
input already resident in HBM.  Two forms are timed on the same device
problem: plain FP64 (the reference algorithm) and FP32 V-cycles inside an
FP64 flexible GMRES.  The JSON line's ``value`` is the FP32-MG solve time;
the plain FP64 time, iteration counts, the true relative residual of each
and the FP32-MG error against the FP64 solution sit beside it.

``roofline`` is for the dominant kernel, the finest-level fused red-black
sweep (k_rb_march2x3: the three face components in one launch) in the
reported form: its algorithmic bytes per launch (read phi, rhs and alpha of
each component, beta_cc and the three edge viscosities once, write phi of
each component: 16 arrays of 256^3 values, 13 for a V-cycle's first sweep,
which starts from the zero guess without reading phi; averaged over the
launches) over the launch durations that
the dispatch itself timestamps on the solver's stream inside the timed
region.  ``traffic`` is the HBM traffic per launch that rocprofv3's
FETCH_SIZE and WRITE_SIZE counters measured for the same kernel on the same
workload (FETCH_SIZE scaled by 2, the factor that makes it read the true byte
count of a plain conversion kernel); it is read from the committed counter
summary and is null when that file is absent.

``cpu_baseline`` is the CPU oracle (the FP64 restatement of the reference
algorithm, OpenMP) solving the same problem family at a bounded size (128^3
by default, all host threads it is given) to the same tolerance; its time is
scaled by the cell ratio to a per-256^3-solve figure.  The FP64 device path
is checked against that oracle solve at the same size.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Stokes solve time to rel-res 1e-9 at 256^3: FP32-MG/learned vs plain FP64"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="cells per direction")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-n", type=int, default=128, help="size of the CPU oracle sample (0: skip)")
    ap.add_argument("--forms", default="fp64,fp32_mg")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(msg, quiet=False):
    if not quiet:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_true_residual(S, ctx, dp, tmp_u, tmp_p):
    """|b - A x| / |b| over velocity and pressure (plain sums), FP64 on the device."""
    S.apply_matrix(ctx, tmp_u, tmp_p, dp.x_u, dp.x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma,
                   dp.pb.theta_alpha)
    num = 0.0
    den = 0.0
    for f, b in zip(tmp_u + [tmp_p], dp.b_u + [dp.b_p]):
        a = f.download()
        bb = b.download()
        num += float(np.sum((bb - a) ** 2))
        den += float(np.sum(bb ** 2))
    return (num / den) ** 0.5


def rel_diff(xa, xb):
    (ua, pa), (ub, pb) = xa, xb
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(ua, ub))
    den = sum(float(np.sum(b ** 2)) for b in ub)
    err_u = (num / den) ** 0.5
    pa0, pb0 = pa - pa.mean(), pb - pb.mean()
    err_p = float(np.linalg.norm(pa0 - pb0) / max(np.linalg.norm(pb0), 1e-300))
    return err_u, err_p


COUNTERS = os.path.join(ROOT, "profiles", "smoother_l0_256cubed_counters.txt")


def measured_traffic(form, n):
    """HBM bytes per level-0 smoother launch from the committed counter summary
    (lines ``traffic <form> <n> <bytes>``), or None."""
    try:
        with open(COUNTERS) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == "traffic" and w[1] == form and int(w[2]) == n:
                    return float(w[3])
    except OSError:
        pass
    return None


def time_form(S, ctx, dp, form, steps, warmup, quiet):
    prec = {"fp64": S.FP64, "fp32_mg": S.FP32_MG}[form]
    ctx.set_params(precision=prec)
    info = None
    for w in range(warmup):
        dp.zero_solution()
        _, info = dp.solve()
        log(f"{form} warmup {w}: iters {info['total_iter']} status {info['status']} "
            f"{info['seconds'] * 1e3:.1f} ms", quiet)
    # timed region: only the fused sweeps carry events, stamped by their own
    # dispatch (profiling level 2), so the timing costs the solve nothing
    ctx.reset_stats()
    ctx.profiling(2)
    ctx.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(steps):
        dp.zero_solution()
        _, info = dp.solve()
        iters.append(info["total_iter"])
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.profiling(False)
    sm = ctx.kernel_stats("stag_smooth_l0")
    # breakdown by kernel family from one more solve outside the timed region
    # (level 1 brackets every launch with events, which slows the host side)
    ctx.reset_stats()
    ctx.profiling(1)
    dp.zero_solution()
    dp.solve()
    ctx.synchronize()
    ctx.profiling(False)
    fams = {k: ctx.kernel_stats(k) for k in ("stag_smooth_l0", "stag_smooth", "stag_restrict_l0", "stag_restrict",
                                              "stag_prolong",
                                              "stag_coarse", "cc_smooth_l0", "cc_smooth", "cc_restrict_l0", "cc_coarse",
                                              "apply_matrix", "krylov")}
    ms = (t1 - t0) * 1e3 / steps
    log(f"{form}: {ms:.2f} ms/solve, iters {iters}, status {info['status']}", quiet)
    return ms, info, iters, sm, fams


def cpu_baseline(P, O, helpers, n, seed, quiet):
    """Oracle FP64 solve of the same problem family at n^3; returns (seconds, info, problem, solution)."""
    pb = P.config3(n, seed)
    b_u = P.rhs_from_noise_host(pb)
    g, p, cf = helpers.oracle_args(pb)
    t0 = time.perf_counter()
    xu, xp, info = O.gmres(g, p, cf, pb.theta_alpha, b_u, pb.b_p)
    sec = time.perf_counter() - t0
    log(f"cpu oracle {n}^3: {sec:.2f} s, iters {info['total_iter']}", quiet)
    return sec, info, pb, b_u, (xu, xp)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: this build runs the "
              "solver on a single MI355X (the multi-GPU decomposition is a design note only)",
              file=sys.stderr, flush=True)
        sys.exit(2)
    from lowmachmixtures_amd import problems as P
    from lowmachmixtures_amd import stokes as S

    S.lib()  # fails loudly when the gfx950 library is missing
    forms = [f for f in a.forms.split(",") if f]
    n = a.n
    N = n ** 3
    t = time.perf_counter()
    pb = P.config3(n, a.seed)
    log(f"host inputs {pb.name}: {time.perf_counter() - t:.1f} s", a.quiet)

    results = {}
    with S.Context(3, pb.n, pb.dx, **pb.solver) as ctx:
        dp = S.DeviceProblem(ctx, pb)
        tmp_u = ctx.faces()
        tmp_p = ctx.field(S.CELL)
        sols = {}
        for form in forms:
            ms, info, iters, sm, fams = time_form(S, ctx, dp, form, a.steps, a.warmup, a.quiet)
            res = rel_true_residual(S, ctx, dp, tmp_u, tmp_p)
            sols[form] = dp.solution()
            achieved = sm["bytes"] / (sm["ms"] * 1e-3) / 1e9 if sm["ms"] > 0 else 0.0
            results[form] = {
                "ms_per_solve": round(ms, 3),
                "solves_per_s": round(1e3 / ms, 4),
                "cells_per_s": round(N * 1e3 / ms, 1),
                "gmres_iters": iters[-1],
                "vcycles": info["vcycles"],
                "status": info["status"],
                "rel_true_residual": res,
                "smoother_l0": {
                    "launches": sm["launches"],
                    "avg_us": round(sm["ms"] * 1e3 / max(sm["launches"], 1), 2),
                    "algorithmic_bytes_per_launch": sm["bytes"] / max(sm["launches"], 1),
                    "achieved_GBs": round(achieved, 1),
                    "frac_hbm": round(achieved / HBM_PEAK_GBS, 4),
                },
                "time_by_family_ms": {k: round(v["ms"], 3) for k, v in fams.items()},
            }
        if "fp64" in sols:
            for form in forms:
                if form != "fp64":
                    eu, ep = rel_diff(sols[form], sols["fp64"])
                    results[form]["err_vs_fp64_u"] = eu
                    results[form]["err_vs_fp64_p"] = ep
        for f in tmp_u + [tmp_p]:
            f.free()

    head = "fp32_mg" if "fp32_mg" in results else forms[0]
    hr = results[head]
    sm = hr["smoother_l0"]
    out = {
        "metric": METRIC,
        "value": round(hr["ms_per_solve"] / 1e3, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": hr["ms_per_solve"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "fp32_mg" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D periodic {n}^3 variable-viscosity staggered Stokes solve to rel-res 1e-9 "
                        f"(visc_type -1, log-normal eta contrast 100, theta*rho in [1,4], white-noise RHS)",
            "form": head,
            "cells": N,
            "seed": a.seed,
            "solver": "GMRES restart 5, projection preconditioner, 1 V-cycle each (2/2/8 sweeps, 6-colour GS)",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_rb_march2x3 fused red-black sweep of the three face components, level 0 ({head})",
            "achieved": sm["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": sm["frac_hbm"],
            "algorithmic_bytes": sm["algorithmic_bytes_per_launch"],
            "traffic": measured_traffic(head, n),
        },
        "forms": results,
    }
    if "fp64" in results and head != "fp64":
        out["speedup_vs_fp64"] = round(results["fp64"]["ms_per_solve"] / hr["ms_per_solve"], 3)

    if a.cpu_n > 0:
        from oracle import oracle as O
        from tests import helpers

        sec, oinfo, pbc, b_c, (ou, op) = cpu_baseline(P, O, helpers, a.cpu_n, a.seed, a.quiet)
        scale = (n / a.cpu_n) ** 3
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {
            "value": round(sec * scale, 3),
            "unit": "s",
            "cores": cores,
            "kind": "port",
            "sample": f"oracle FP64 GMRES on the {a.cpu_n}^3 instance of the same problem family "
                      f"({oinfo['total_iter']} iterations, {sec:.2f} s), scaled by the cell ratio {scale:g}",
        }
        # the FP64 device path against the oracle on that instance
        with S.Context(3, pbc.n, pbc.dx, **pbc.solver) as ctx:
            dp = S.DeviceProblem(ctx, pbc, rhs=b_c)
            ctx.set_params(precision=S.FP64)
            _, ginfo = dp.solve()
            gu, gp = dp.solution()
        eu, ep = rel_diff((gu, gp), (ou, op))
        out["parity_fp64_vs_oracle"] = {
            "n": a.cpu_n, "err_u": eu, "err_p": ep,
            "iters_gpu": ginfo["total_iter"], "iters_oracle": oinfo["total_iter"],
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
