"""Reduce rocprofv3 output of bench.py runs into the summaries kept in profiles/.

Three runs of the same program give the inputs (from the repository root on an
MI355X; the bench arguments make the dispatch sequence of the three runs the
same, so their Dispatch_Ids match):

  cd /tmp && export TMPDIR=/tmp
  rocprofv3 --kernel-trace --stats --output-format csv -d <dir>/trace -o run -- python3 <repo>/bench.py <args>
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/pmc_f -o run -- python3 <repo>/bench.py <args>
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/pmc_w -o run -- python3 <repo>/bench.py <args>

    python -m lowmachmixtures_amd.profiling <dir> <n> "<command text>" <out_trace.txt> <out_counters.txt>

writes a kernel-trace summary (top kernels, level-0 fused sweeps) and a counter
summary (``python -m lowmachmixtures_amd.profiling steps <dir> <n> "<command>" <out.txt>``
instead summarises the kernel trace of a run of the bench's inertial_step line per time
step): per level-0 kernel its mean duration, HBM read (FETCH_SIZE scaled by
the factor that makes a plain conversion kernel read its true byte count) and
write bytes, algorithmic bytes and the two bandwidths, and the
``traffic <form> <n> <bytes>`` lines bench.py reads.
"""
from __future__ import annotations

import csv
import glob
import os
import sys
from collections import defaultdict

HBM_PEAK_GBS = 8000.0


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def load_trace(d):
    """Dispatch_Id -> (kernel name, duration ns, grid size)."""
    out = {}
    for r in _rows(_find(d, "kernel_trace.csv")):
        out[int(r["Dispatch_Id"])] = (r["Kernel_Name"], int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                                      int(r.get("Grid_Size", r.get("Grid_Size_X", 0)) or 0))
    return out


def load_counter(d, counter):
    """Dispatch_Id -> (kernel name, counter value in bytes, grid size); values are KiB per dispatch."""
    acc = defaultdict(float)
    meta = {}
    for r in _rows(_find(d, "counter_collection.csv")):
        if r["Counter_Name"] != counter:
            continue
        k = int(r["Dispatch_Id"])
        acc[k] += float(r["Counter_Value"]) * 1024.0
        meta[k] = (r["Kernel_Name"], int(r.get("Grid_Size", 0) or 0))
    return {k: (meta[k][0], acc[k], meta[k][1]) for k in acc}


def short(name, width=60):
    name = name.replace("lmm::(anonymous namespace)::", "").replace("lmm::", "").replace("void ", "")
    name = name.split("(")[0] if name.startswith("k_") else name
    return name[:width]


def fetch_factor(fetch):
    """True read bytes / FETCH_SIZE of k_convert<float, double> (one 8-byte value per work-item)."""
    ratios = []
    for name, val, grid in fetch.values():
        if "k_convert<float, double>" in name and val > 0 and grid >= (1 << 22):
            ratios.append(grid * 8.0 / val)
    if not ratios:
        return 2.0, []
    ratios.sort()
    return ratios[len(ratios) // 2], ratios


# Algorithmic arrays (of N level-0 values) of the level-0 kernels: one launch
# reads every array it needs once and writes each output once.  A V-cycle's
# first finest-level sweep starts from the zero guess without reading phi:
# the sweeps are averaged over a V-cycle's four (13 + 3 * 16 arrays for the
# three-component sweep, 5 + 3 * 6 for the cell-centred one).  The fused
# residual + restriction writes 1/8 of an array per component.
def level0_arrays(name):
    if "k_rb_march2x3" in name and ", -1>" in name:
        return 15.25, 3
    if "k_rb_march2<" in name and "CCRB" in name:
        return 5.75, 1
    if "k_resid_restrict3" in name and ", -1>" in name:
        # three components: phi, rhs, alpha each, beta and the three edge
        # arrays once, the three coarse right-hand sides written
        return 13.375, 0.375
    if "k_resid_restrict" in name and ", -1>" in name:
        return 6.125, 0.125
    return None, None


def summarize(d, n, command, out_trace, out_counters):
    N = n ** 3
    trace = load_trace(os.path.join(d, "trace"))
    fetch = load_counter(os.path.join(d, "pmc_f"), "FETCH_SIZE")
    write = load_counter(os.path.join(d, "pmc_w"), "WRITE_SIZE")
    factor, ratios = fetch_factor(fetch)
    # level-0 launches: the ones that write at least half the arrays of a level-0 launch
    lv0 = defaultdict(list)
    for did, (name, dur, grid) in trace.items():
        arrays, outs = level0_arrays(name)
        if arrays is None or did not in write or did not in fetch:
            continue
        es = 8 if "double" in name else 4
        if write[did][1] >= 0.5 * outs * N * es:
            lv0[name].append((dur, fetch[did][1] * factor, write[did][1], arrays * N * es))
    lines = [f"rocprofv3 counters: level-0 ({n}^3) fused kernels of the 3D periodic {n}^3 Stokes solve (bench.py)",
             "", "Commands (from the repository root on an MI355X, gfx950), three runs of the same program:",
             "  cd /tmp && export TMPDIR=/tmp"]
    for what in ("--kernel-trace --stats", "--pmc FETCH_SIZE", "--pmc WRITE_SIZE"):
        lines.append(f"  rocprofv3 {what} --output-format csv -d <dir> -o run -- {command}")
    lines += ["Reduced by python -m lowmachmixtures_amd.profiling (joined on Dispatch_Id; a launch is level 0 when",
              "its WRITE_SIZE is at least half the arrays a level-0 launch writes).", "",
              "Calibration with the plain conversion kernel k_convert<float, double> (one 8-byte value per work-item):",
              f"  true read / FETCH_SIZE = {factor:.3f} (median of {len(ratios)} dispatches"
              + (f", min {ratios[0]:.3f} max {ratios[-1]:.3f})" if ratios else ")"),
              f"  => HBM read bytes = {factor:.2f} x FETCH_SIZE; WRITE_SIZE is taken as is", "",
              f"{'kernel (level 0)':60s} {'n':>4s} {'avg_us':>8s} {'read MB':>9s} {'write MB':>9s} {'traffic MB':>10s}"
              f" {'alg MB':>8s} {'ratio':>6s} {'alg GB/s':>9s} {'HBM GB/s':>9s}"]
    traffic = {}
    for name in sorted(lv0):
        v = lv0[name]
        k = len(v)
        dur = sum(x[0] for x in v) / k * 1e-9
        rd = sum(x[1] for x in v) / k
        wr = sum(x[2] for x in v) / k
        alg = sum(x[3] for x in v) / k
        lines.append(f"{short(name):60s} {k:4d} {dur * 1e6:8.2f} {rd / 1e6:9.2f} {wr / 1e6:9.2f} {(rd + wr) / 1e6:10.2f}"
                     f" {alg / 1e6:8.2f} {(rd + wr) / alg:6.3f} {alg / dur / 1e9:9.1f} {(rd + wr) / dur / 1e9:9.1f}")
        if "k_rb_march2x3" in name:
            traffic["fp64" if "double" in name else "fp32_mg"] = (rd + wr, dur)
    lines += ["", "Machine-readable lines read by bench.py (HBM bytes per launch of the three-component sweep):"]
    for form, (t, dur) in sorted(traffic.items()):
        lines.append(f"traffic {form} {n} {int(round(t))}")
    with open(out_counters, "w") as f:
        f.write("\n".join(lines) + "\n")

    tot = defaultdict(lambda: [0, 0])
    for name, dur, _ in trace.values():
        tot[name][0] += 1
        tot[name][1] += dur
    total_ms = sum(v[1] for v in tot.values()) * 1e-6
    t = [f"rocprofv3 kernel trace: 3D periodic {n}^3 variable-viscosity Stokes solve (bench.py), MI355X (gfx950)", "",
         "Command (from the repository root on an MI355X):",
         f"  cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o run -- {command}",
         "Reduced by python -m lowmachmixtures_amd.profiling from run_kernel_trace.csv (End - Start per dispatch).", "",
         "Level-0 fused sweeps (level 0 by the WRITE_SIZE run of the same command):"]
    for name in sorted(lv0):
        v = lv0[name]
        t.append(f"  {short(name, 70):70s} {len(v):5d} launches  avg {sum(x[0] for x in v) / len(v) / 1e3:9.2f} us")
    t += ["", f"Whole run, top kernels; total kernel time {total_ms:.1f} ms:",
          f"  {'kernel':52s} {'calls':>6s} {'total_ms':>9s} {'avg_us':>8s} {'pct':>6s}"]
    for name, (c, dsum) in sorted(tot.items(), key=lambda kv: -kv[1][1])[:30]:
        t.append(f"  {short(name, 52):52s} {c:6d} {dsum * 1e-6:9.3f} {dsum / c / 1e3:8.2f} {dsum * 1e-6 / total_ms * 100:6.2f}")
    with open(out_trace, "w") as f:
        f.write("\n".join(t) + "\n")


def inertial_steps(d, n, command, out_trace):
    """Kernel-trace summary of the bench's inertial_step line: inertial steps
    are delimited by their first k_adv_s launch (two per step: predictor and
    corrector), overdamped steps by their first k_zero_minus; the bench runs 4 FP64 inertial steps, 4 FP32-MG inertial
    steps, then 4 + 4 overdamped steps, the first of each group a warm-up.
    Per-step wall time (first to last dispatch), kernel busy time and the
    per-step kernel breakdown of each group's timed steps."""
    rows = sorted(((s_, e_, name) for name, s_, e_ in _trace_rows(d)), key=lambda r: r[0])
    # the overdamped steps are the ones with k_zero_minus (0 - grad pi, once
    # per stage); an inertial step starts a few dispatches before its first
    # k_adv_s (the face averages of rho), an overdamped one a few before its
    # first k_zero_minus (face averages, noise fill, grad pi)
    zm = [i for i, r in enumerate(rows) if "k_zero_minus" in r[2]]
    first_od = zm[0] if zm else len(rows)
    adv = [i for i, r in enumerate(rows) if "k_adv_s" in r[2] and i < first_od]
    bounds = [max(0, i - 8) for i in adv[0::2]] + [max(0, i - 12) for i in zm[0::2]] + [len(rows)]
    steps = [rows[bounds[k]:bounds[k + 1]] for k in range(len(bounds) - 1)]
    t = [f"rocprofv3 kernel trace: the inertial and overdamped time steps, 3D periodic {n}^3, 4 species "
         "(bench.py inertial_step line)",
         "", "Command (from the repository root on an MI355X, gfx950):",
         f"  cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o run -- {command}",
         "Reduced by python -m lowmachmixtures_amd.profiling steps <dir> <n> <command> <out> from run_kernel_trace.csv:",
         "inertial steps delimited by their first k_adv_s dispatch, overdamped ones by their first k_zero_minus;",
         "groups of inertial FP64, inertial FP32-MG, overdamped FP64, overdamped FP32-MG steps, the first of each a",
         "warm-up; a group's last step runs to the next group's first dispatch (downloads, the next context's set-up).",
         ""]
    labels = ["inertial, FP64 solves", "inertial, FP32-MG solves", "overdamped, FP64 solves",
              "overdamped, FP32-MG solves"]
    ngroups = 4 if len(steps) >= 16 else 2
    per = max(1, len(steps) // ngroups)
    groups = [steps[g * per:(g + 1) * per] for g in range(ngroups)]
    for g, grp in enumerate(groups):
        for k, st in enumerate(grp):
            wall = (st[-1][1] - st[0][0]) * 1e-6
            busy = sum(e_ - s_ for s_, e_, _ in st) * 1e-6
            t.append(f"  step {k} ({labels[g]}): wall {wall:7.2f} ms, kernel busy {busy:7.2f} ms, "
                     f"{len(st)} dispatches")
    for label, sel in ((labels[g], grp[1:]) for g, grp in enumerate(groups)):
        if not sel:
            continue
        tot = defaultdict(lambda: [0, 0])
        for st in sel:
            for s_, e_, name in st:
                tot[name][0] += 1
                tot[name][1] += e_ - s_
        busy = sum(v[1] for v in tot.values()) / len(sel) * 1e-6
        t += ["", f"{label}: per step ({len(sel)} timed steps), kernel busy {busy:.2f} ms:",
              f"  {'kernel':60s} {'calls':>7s} {'ms':>8s} {'pct':>6s}"]
        for name, (c, dsum) in sorted(tot.items(), key=lambda kv: -kv[1][1])[:24]:
            ms = dsum / len(sel) * 1e-6
            t.append(f"  {short(name, 60):60s} {c / len(sel):7.1f} {ms:8.3f} {ms / busy * 100:6.2f}")
    with open(out_trace, "w") as f:
        f.write("\n".join(t) + "\n")


def _trace_rows(d):
    for r in _rows(_find(d, "kernel_trace.csv")):
        yield r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"])


if __name__ == "__main__":
    if len(sys.argv) == 6 and sys.argv[1] == "steps":
        inertial_steps(sys.argv[2], int(sys.argv[3]), sys.argv[4], sys.argv[5])
        sys.exit(0)
    if len(sys.argv) != 6:
        print(__doc__)
        sys.exit(2)
    summarize(sys.argv[1], int(sys.argv[2]), sys.argv[3], sys.argv[4], sys.argv[5])
