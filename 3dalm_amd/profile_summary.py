"""Summaries of rocprofv3 output for the hot-path kernels.

    python 3dalm_amd/profile_summary.py trace <run>_kernel_trace.csv [--cells N] [--match REGEX]
    python 3dalm_amd/profile_summary.py traffic <fetch>_counter_collection.csv \\
        <write>_counter_collection.csv --cells N --padded-cells L [--json out.json]

``trace`` tabulates every kernel of a ``rocprofv3 --kernel-trace --output-format csv``
run.  The Krylov solvers enqueue iterations ahead of the device's stopping test,
and the launches past it return at once; they are counted apart ("idle": shorter
than --idle-us) so that the mean duration is that of launches that did the work.
With --cells, the algorithmic bytes per interior cell of each known kernel give
its achieved bandwidth and fraction of the 8 TB/s HBM peak (--padded-cells for the
kernels that sweep the whole padded box); --match keeps the kernels whose names
match a regular expression (the shares are then of those kernels' time), and --span
keeps the dispatches from the first to the last one of the kernels matching another
(e.g. a reverse sweep, replays included).

``traffic`` turns two ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs
(kilobytes per dispatch) into bytes per interior cell per launch, scaled by what
the same counters report for a kernel of known traffic: the library's k_copy3,
which ``bench.py --calibration N`` runs N times after the forms and which reads and
writes 24 B per padded cell (--padded-cells); the library's read-only reduction
k_scalar_sum (8 B per interior cell) cross-checks the FETCH_SIZE scale.  Without
--padded-cells the largest device-to-device copy of ``--copy-mib`` MiB calibrates.  With --json the per-kernel figures are also written as JSON, which
bench.py reads to fill roofline.traffic.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re

import numpy as np

HBM_PEAK_GBS = 8000.0

# algorithmic bytes per interior cell and launch (compulsory reads + writes,
# operator matrix-free); the FP32 kernels are those of the precision form
ALGO_BYTES = {
    "k_jacobi": 24, "k_bi_spmv<0>": 24, "k_bi_spmv<1>": 24, "k_bi_p": 40, "k_bi_s": 32, "k_bi_xr": 64,
    "k_bi_init": 32, "k_sor": 16, "k_resnorm": 16, "k_predictor": 104, "k_project": 104, "k_post": 64,
    "k_flux": 48, "k_rhs": 32,
    "kf_jacobi": 12, "kg_spmv<0>": 12, "kg_spmv<1>": 12, "kg_p": 20, "kg_s": 16, "kg_xr": 40,
    "kg_resid<true>": 28, "kg_resid<false>": 24, "km_predictor": 64, "km_project": 80, "km_post": 48,
    "km_flux_rhs": 44,
    "k_calc_q<double>": 32, "k_calc_q<float>": 20,
    "kf_sor": 10, "ks_resid": 24, "ks_add": 20,
    "k_jacobi4": 24, "k_bi_spmv4<0>": 24, "k_bi_spmv4<1>": 24, "k_bi_p4": 40, "k_bi_s4": 32,
    # the x/r update: r only (the fused iteration), x and r; the fused iteration's x
    # updates of eight iterations
    "k_bi_xr4<0>": 32, "k_bi_xr4<1>": 64, "k_x_group4": 144,
    "kf_jacobi4": 12, "kg_spmv4<0>": 12, "kg_spmv4<1>": 12, "kg_p4": 20, "kg_s4": 16, "kg_xr4": 40,
    # fused half-iterations (rows of 256 cells, Jacobi x3)
    "kg_half4<0>": 36, "kg_half4<1>": 24, "k_half_pc4<0>": 40, "k_half_pc4<1>": 32,
    # the fused iteration's x updates of eight iterations (working launches)
    "kg_apply_pending4": 80,
    # the predictor's column walk with the LDS-staged plane
    "k_predictor_star": 104, "km_predictor_star": 64, "k_post_star": 64, "km_post_star": 48,
}
# kernels whose launches alternate between two algorithmic byte counts per cell:
# per-launch bytes of each kind (none at present)
ALTERNATING = {}
# kernels that sweep every padded cell: algorithmic bytes per padded cell
# (needs --padded-cells)
ALGO_BYTES_PADDED = {"k_time_average<double>": 72, "k_time_average<float>": 60,
                     # reverse predictor: read U_b, Uold, JU (3 each), nu_t; write the bars of
                     # Uold, JU, f (3 each) and nu_t
                     "ka_predictor": 160}


def short_name(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    m = re.match(r"([\w:]+(<[^>]*>)?)", name)
    return m.group(1) if m else name[:40]


def trace_table(path: str, cells: int | None, idle_us: float, padded: int | None = None,
                match: str | None = None, span: str | None = None) -> str:
    by = collections.defaultdict(list)
    keep = re.compile(match) if match else None
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    head = []
    if span:
        # the dispatches from the first to the last one whose kernel matches `span`
        sp = re.compile(span)
        hit = [n for n, r in enumerate(rows) if sp.search(short_name(r["Kernel_Name"]))]
        if hit:
            rows = rows[hit[0]:hit[-1] + 1]
            wall = (int(rows[-1]["End_Timestamp"]) - int(rows[0]["Start_Timestamp"])) / 1e6
            busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rows) / 1e6
            head = [f"span of /{span}/: {len(rows)} dispatches, {wall:.2f} ms wall, {busy:.2f} ms in kernels", ""]
    for row in rows:
        name = short_name(row["Kernel_Name"])
        if keep and not keep.search(name):
            continue
        us = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3
        by[name].append(us)
    total = sum(sum(v) for v in by.values())
    lines = head + ["| kernel | launches | idle | mean µs (working) | median µs | total ms | share | B/cell | GB/s | of 8 TB/s |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for name, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        v = np.asarray(v)
        work = v[v >= idle_us] if name in ALGO_BYTES and (v >= idle_us).any() else v
        mean = float(work.mean())
        bpc, n = ALGO_BYTES.get(name), cells
        if bpc is None and padded and name in ALGO_BYTES_PADDED:
            bpc, n = ALGO_BYTES_PADDED[name], padded
        gbs = frac = ""
        if n and bpc:
            g = bpc * n / (mean * 1e-6) / 1e9
            gbs, frac = f"{g:.0f}", f"{g / HBM_PEAK_GBS:.3f}"
        lines.append(f"| {name} | {len(v)} | {len(v) - len(work)} | {mean:.2f} | {float(np.median(work)):.2f} | "
                     f"{v.sum() / 1e3:.1f} | {100 * v.sum() / total:.1f}% | {bpc or ''} | {gbs} | {frac} |")
        if 100 * v.sum() / total < 0.05:
            break
    return "\n".join(lines)


def _counter(path: str, counter: str):
    by = collections.defaultdict(list)
    for row in csv.DictReader(open(path)):
        if row.get("Counter_Name") != counter:
            continue
        by[short_name(row["Kernel_Name"])].append(float(row["Counter_Value"]))
    return by


def traffic_table(fetch: str, write: str, cells: int, padded: int | None, copy_mib: float,
                  json_out: str | None = None) -> str:
    f, w = _counter(fetch, "FETCH_SIZE"), _counter(write, "WRITE_SIZE")
    record = {"cells": cells, "kernels": {}}
    if padded and f.get("k_copy3"):
        # the library's copy of the three velocity components over the padded
        # array: 24 B read and 24 B written per padded cell
        cal = "k_copy3"
        want_kib = 24 * padded / 1024
        what = f"k_copy3 (U to Uold, 3 x {padded} padded cells) reads and writes {want_kib:.0f} KiB"
        record["calibration"] = {"kernel": cal, "padded_cells": padded}
    else:
        cal = "__amd_rocclr_copyBuffer"
        want_kib = copy_mib * 1024
        what = f"a {copy_mib:g} MiB copy reads and writes {want_kib:.0f} KiB"
        record["calibration"] = {"kernel": cal, "copy_mib": copy_mib}
    cf = float(np.median(f.get(cal, [0])))
    cw = float(np.median(w.get(cal, [0])))
    sf = want_kib / cf if cf else 1.0
    sw = want_kib / cw if cw else 1.0
    lines = [f"FETCH_SIZE scale {sf:.3f}, WRITE_SIZE scale {sw:.3f} ({what}; the counters report "
             f"{cf:.0f} and {cw:.0f} KiB)"]
    # cross-check of the FETCH_SIZE scale with a read-only kernel of known
    # traffic: the library's reduction k_scalar_sum reads one FP64 field over
    # the interior, 8 B per interior cell
    ks = f.get("k_scalar_sum")
    if ks:
        got = float(np.median(ks)) * 1024 * sf
        lines.append(f"cross-check: k_scalar_sum reads {cells * 8 / 2 ** 20:.1f} MiB (8 B x {cells} interior "
                     f"cells); FETCH_SIZE x scale gives {got / 2 ** 20:.1f} MiB ({got / (cells * 8):.3f}x)")
        record["cross_check_k_scalar_sum"] = round(got / (cells * 8), 4)
    lines += ["",
             "| kernel | dispatches | read B/cell | write B/cell | total B/cell | algorithmic B/cell | ratio |",
             "|---|---|---|---|---|---|---|"]
    for name in sorted(f, key=lambda k: -np.median(f[k])):
        if name not in ALGO_BYTES:
            continue
        vf, vw = np.asarray(f[name]), np.asarray(w.get(name, [0.0]))
        # launches past a solver's stopping test move almost nothing: use the
        # median of the dispatches that did work
        vf = vf[vf > 0.2 * vf.max()] if vf.max() > 0 else vf
        vw = vw[vw > 0.2 * vw.max()] if vw.max() > 0 else vw
        if name in ALTERNATING:
            # one row per kind of launch (split at the middle of the range), and the
            # kernel's record: the mean of the two, as they alternate
            modes = []
            for (lo_b, label), pick in zip(zip(ALTERNATING[name], ("x deferred", "two x updates")),
                                           (np.less, np.greater_equal)):
                mf = vf[pick(vf, 0.5 * (vf.min() + vf.max()))]
                mw = vw[pick(vw, 0.5 * (vw.min() + vw.max()))]
                rd = float(np.median(mf)) * 1024 * sf / cells
                wr = float(np.median(mw)) * 1024 * sw / cells
                modes.append((rd, wr))
                lines.append(f"| {name} ({label}) | {len(mf)} | {rd:.1f} | {wr:.1f} | {rd + wr:.1f} | {lo_b} | "
                             f"{(rd + wr) / lo_b:.2f} |")
            rd = 0.5 * (modes[0][0] + modes[1][0])
            wr = 0.5 * (modes[0][1] + modes[1][1])
        else:
            rd = float(np.median(vf)) * 1024 * sf / cells
            wr = float(np.median(vw)) * 1024 * sw / cells
        algo = ALGO_BYTES[name]
        lines.append(f"| {name} | {len(f[name])} | {rd:.1f} | {wr:.1f} | {rd + wr:.1f} | {algo} | "
                     f"{(rd + wr) / algo:.2f} |")
        record["kernels"][name] = {"read_B_per_cell": round(rd, 2), "write_B_per_cell": round(wr, 2),
                                   "algorithmic_B_per_cell": algo}
    record["fetch_scale"], record["write_scale"] = round(sf, 4), round(sw, 4)
    if json_out:
        with open(json_out, "w") as fh:
            json.dump(record, fh, indent=1)
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--cells", type=int)
    t.add_argument("--idle-us", type=float, default=6.0)
    t.add_argument("--padded-cells", type=int, help="padded cell count, for the kernels that sweep it")
    t.add_argument("--match", help="keep only the kernels whose names match this regular expression")
    t.add_argument("--span", help="keep only the dispatches from the first to the last kernel matching this")
    r = sub.add_parser("traffic")
    r.add_argument("fetch_csv")
    r.add_argument("write_csv")
    r.add_argument("--cells", type=int, required=True)
    r.add_argument("--padded-cells", type=int, help="padded cell count: calibrate on the library's k_copy3")
    r.add_argument("--copy-mib", type=float, default=256, help="else calibrate on a copy of this many MiB")
    r.add_argument("--json")
    a = ap.parse_args()
    if a.cmd == "trace":
        print(trace_table(a.csv, a.cells, a.idle_us, a.padded_cells, a.match, a.span))
    else:
        print(traffic_table(a.fetch_csv, a.write_csv, a.cells, a.padded_cells, a.copy_mib, a.json))


if __name__ == "__main__":
    main()
