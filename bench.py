"""Benchmark of the 3dalm per-time-step hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--forms fp64,mixed]
                    [--grid 256,128,128] [--no-cpu] [--probe jacobi]
                    [--adjoint-steps 20] [--adjoint-every 10] [--no-adjoint] [--no-extras]
                    [--no-target] [--tol 1e-8]

Workload (one "step" = one Solver::main_loop_once): the single-turbine ALM case on
256x128x128 cells (box 16x8x8, y/z tanh beta=1.5, rotor D=1 at x=3, 3 blades x 64
actuator points, TSR 6, eps = 2 dx, 5% seeded noise, dt = 0.005), Poisson by
BiCGStab with 3 Jacobi sweeps as preconditioner to 1e-8.  Synthetic inputs,
resident in HBM before the timed region.  The run is 1000 steps long; by default
steps 11-110 are timed (the first steps after the impulsive start need two to
three times the Poisson iterations of the rest and are warm-up).

Forms: "fp64" is the FP64 drop-in; "mixed" is the precision form (FP32 stencils,
FP32 BiCGStab vectors with FP64 dot products, norms, pressure and residual
replacements, to the same true-residual test).

Prints ONE JSON line (rank 0).  value = Mcell-steps/s = N_interior * K / t / 1e6
of the precision form when it is built, else of the FP64 drop-in; every form is
reported under "forms".  roofline: the dominant kernel's algorithmic bytes per
launch over its mean launch time, timed with HIP events on the library's stream.
cpu_baseline: the reference's own code compiled from its headers (oracle/_ref) when
that library is present and loads, else the C restatement in oracle/ (one core),
on a bounded sample of this workload (see "sample").
target_size: both forms at 512x256x256 (the size the precision form's targets are
stated at; two rotors 5 D apart, an initial turbulence of 200 seeded Fourier modes at
10 % intensity), steps 11-110, with the Poisson tolerance 4e-9 that keeps avg div <= 1e-6
there; ratio, iterations, whole-step bandwidth and the turbine-power difference.
forms.<form>.step_roofline: the whole step's algorithmic bytes (fixed kernels + bytes
per BiCGStab iteration x iterations per step) over the time per step.
output_path: mean device time and achieved bandwidth of the output-path kernels
(Q-criterion, time average of U) at the bench grid, both forms.  sor_forms: the
64x48x48 SOR case (no turbine) in both forms.
adjoint_forward_ratio: on the differentiation-form workload (192x96x96, three
rotors, FP64, BiCGStab(Jacobi 3) to 1e-8), the wall time per step of the reverse
sweep (segment replays from checkpoints plus reverse steps, adjoint solves to 1e-8
relative) over the wall time per step of the recorded forward run of the same
steps (which includes keeping the pressure of every step).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s mixed vs FP64 at div<=1e-6; adjoint/forward step-cost ratio"
HBM_PEAK_GBS = 8000.0
# per-kernel HBM traffic of this workload from rocprofv3 FETCH_SIZE / WRITE_SIZE runs
# (3dalm_amd/profile_summary.py traffic --json; the command is in the .md beside it)
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hbm_traffic_alm_256x128x128.json")
PROBE_KERNEL = {"jacobi": "k_jacobi", "jacobi32": "kf_jacobi", "spmv": "k_bi_spmv<0>", "update": "k_bi_xr",
                "pupdate": "k_bi_p", "supdate": "k_bi_s", "predictor": "k_predictor", "spmv_mixed": "kg_spmv<0>",
                "update_mixed": "kg_xr", "pupdate_mixed": "kg_p", "supdate_mixed": "kg_s"}

# algorithmic bytes per interior cell for one launch of each probed kernel (FP64):
# compulsory reads + writes, with the operator applied matrix-free from 1-D coefficients
PROBE_BYTES = {
    "jacobi": 24,     # read z_in, v; write z_out
    "spmv": 24,       # read z, a1 (r0 or s); write y
    "update": 64,     # read x, pp, ss, s, t, r0; write x, r
    "pupdate": 40,    # read r, p, q; write p, z
    "supdate": 32,    # read r, q; write s, z
    "predictor": 104,  # read Uold(3), JU(3), nut, f(3); write U(3)
    "sor": 16,        # half-sweep: read x, b over the colour + neighbours ~ 2w, write half
    "resnorm": 16,    # read x, b
    # precision form's BiCGStab (FP32 Krylov vectors, FP64 pressure)
    "jacobi32": 12,   # read z_in, v; write z_out (FP32)
    "spmv_mixed": 12,  # read z, r0 or s; write y (FP32)
    "update_mixed": 40,  # read x (FP64), z, z', s, t, r0 (FP32); write x (FP64), r (FP32)
    "pupdate_mixed": 20,  # read r, p, q; write p, z (FP32)
    "supdate_mixed": 16,  # read r, q; write s, z (FP32)
}
# algorithmic bytes per interior cell of one whole step: the fixed kernels (predictor,
# flux + right-hand side, projection, post with div stored) and one BiCGStab(Jacobi 3)
# iteration, plus the solve's start (residual and shadow residual)
STEP_BYTES = {"fp64": {"fixed": 104 + 48 + 32 + 104 + 64, "per_it": 280, "per_solve": 32},
              "mixed": {"fixed": 64 + 44 + 80 + 48, "per_it": 148, "per_solve": 28}}
PROBE_IDS = {"jacobi": 1, "spmv": 2, "update": 3, "predictor": 4, "sor": 5, "resnorm": 6, "pupdate": 7,
             "supdate": 8, "jacobi32": 9, "spmv_mixed": 10, "update_mixed": 11, "pupdate_mixed": 12,
             "supdate_mixed": 13}
FORMS = ("fp64", "mixed")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_case(grid, steps, tol=1e-8):
    cases = importlib.import_module("3dalm_amd.cases")
    nx, ny, nz = grid
    c = cases.case_c2(nx, ny, nz, steps=steps)
    c.tol = tol
    return c


def run_form(case, precision, steps, warmup, probe, form):
    solver = importlib.import_module("3dalm_amd.solver")
    import ctypes as C
    log(f"[bench] form={form} grid={case.mesh.interior} warmup={warmup} steps={steps}")
    s = solver.Solver(case, precision)
    try:
        for w in range(warmup):
            st = s.step()
            log(f"[bench] warmup step {st['step']} it={st['it']} err={st['err']:.3e} div={st['avg_div']:.3e}")
        sync()
        t0 = time.perf_counter()
        its, divs, power = [], [], []
        for k in range(steps):
            st = s.step()
            its.append(st["it"])
            divs.append(st["avg_div"])
            if "power" in st:
                power.append(float(st["power"].sum()))
        sync()
        t1 = time.perf_counter()
        log(f"[bench] timed {steps} steps in {t1 - t0:.3f} s; it/step mean {np.mean(its):.1f}; "
            f"last div {divs[-1]:.3e}")
        res = {"seconds": t1 - t0, "it_mean": float(np.mean(its)), "it": its,
               "avg_div_max": float(np.max(divs)), "power_last": power[-1] if power else None}
        if probe:
            # kernel durations with HIP events on the library's stream, over a few
            # more steps (launched one by one while the probe is on)
            s.ctx.check(s.ctx.L.dalm_probe_kernel(s.ctx.h, PROBE_IDS[probe]), "probe")
            for k in range(3):
                s.step()
            n, ms = C.c_int64(), C.c_double()
            s.ctx.check(s.ctx.L.dalm_probe_read(s.ctx.h, C.byref(n), C.byref(ms)), "probe_read")
            s.ctx.check(s.ctx.L.dalm_probe_kernel(s.ctx.h, 0), "probe off")
            res["probe"] = {"kernel": probe, "launches": n.value, "total_ms": ms.value}
        res["U"] = None
        return res, s
    except Exception:
        s.close()
        raise


_torch = None


def sync():
    global _torch
    if _torch is None:
        try:
            import torch
            _torch = torch if torch.cuda.is_available() else False
        except Exception:
            _torch = False
    if _torch:
        _torch.cuda.synchronize()


class _StdoutToStderr:
    """The reference prints its per-step log with printf; keep bench's stdout to
    the one JSON line by pointing file descriptor 1 at stderr meanwhile."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        import ctypes
        ctypes.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)


def cpu_baseline(case, gpu_it_mean):
    with _StdoutToStderr():
        return _cpu_baseline(case, gpu_it_mean)


def _cpu_baseline(case, gpu_it_mean):
    """Reference (or restatement) on the host: one initialisation + one step of this
    workload with the Poisson solve capped, then a separately timed capped solve;
    a full step at the GPU run's mean iteration count is extrapolated from the fixed
    and per-iteration costs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    po = importlib.import_module("pyoracle")
    kind = "port"
    try:
        if po.ref_available():
            po.lib("ref")
            kind = "reference"
    except OSError as e:
        log(f"[bench] reference build does not load here ({e}); timing the restatement")
        kind = "port"
    cap = 4
    case.max_it = cap
    case.tol = 0.0
    t0 = time.perf_counter()
    S = po.RefSolver(case) if kind == "reference" else po.OracleSolver(case)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    S.step()
    t_step_cap = time.perf_counter() - t0
    # per-iteration cost: a capped solve on the same data
    K = po.Kernels("ref" if kind == "reference" else "oracle")
    m = case.mesh
    g = po.make_grid(m)
    A = np.zeros((m.L, 7))
    K.build_A(g, A)
    x = np.zeros(m.L)
    b = np.ascontiguousarray(S.field("b"))
    work = np.zeros(9 * m.L)
    it, err = po.C.c_int64(), po.C.c_double()
    fn = po.lib(kind if kind == "oracle" else "ref")[0]
    pref = "ref_" if kind == "reference" else "or_"
    t0 = time.perf_counter()
    getattr(fn, pref + "run_bicgstab")(po.C.byref(g), po.ptr(A), po.ptr(x), po.ptr(b), po.ptr(work), 0, 1.0,
                                       case.pc_it, po.C.byref(it), cap, po.C.byref(err), 0.0)
    t_solve = time.perf_counter() - t0
    t_iter = t_solve / max(it.value, 1)
    t_fixed = max(t_step_cap - cap * t_iter, 0.0)
    t_step = t_fixed + gpu_it_mean * t_iter
    if hasattr(S, "close"):
        S.close()
    return {"value": m.N / t_step / 1e6, "unit": "Mcell-steps/s", "cores": 1, "kind": kind,
            "sample": (f"{m.interior[0]}x{m.interior[1]}x{m.interior[2]}: init {t_init:.1f}s + 1 step with "
                       f"the solve capped at {cap} it ({t_step_cap:.1f}s) + a {it.value}-it solve "
                       f"({t_solve:.1f}s); step cost extrapolated to {gpu_it_mean:.1f} it/step "
                       f"(fixed {t_fixed:.2f}s + {t_iter:.3f}s/it)"),
            "seconds_per_step": t_step}


def run_adjoint(steps, every, warmup=2):
    """Forward vs reverse-sweep cost on the differentiation-form workload."""
    cases = importlib.import_module("3dalm_amd.cases")
    solver = importlib.import_module("3dalm_amd.solver")
    case = cases.case_c4(steps=steps)
    omega = np.array([st.rot_speed for st in case.sites])
    log(f"[bench] adjoint: grid={case.mesh.interior} steps={steps} every={every}")
    s = solver.Solver(case)
    try:
        for _ in range(warmup):
            s.step()
        its = []
        s.ctx.adjoint_begin(steps, every, omega / steps, True)
        sync()
        t1 = time.perf_counter()
        for _ in range(steps):
            its.append(s.step()["it"])
        sync()
        t_rec = (time.perf_counter() - t1) / steps
        t2 = time.perf_counter()
        J, ap_g, wt_g, st = s.ctx.adjoint_run(1e-8)
        sync()
        t_rev = (time.perf_counter() - t2) / steps
        s.ctx.adjoint_end()
    finally:
        s.close()
    log(f"[bench] adjoint: recorded forward {t_rec * 1e3:.2f} ms/step, reverse "
        f"{t_rev * 1e3:.2f} ms/step (replay {st['replay_ms'] / steps:.2f} + reverse steps "
        f"{st['reverse_ms'] / steps:.2f}); adjoint solve {st['adj_it'] / steps:.1f} it/step, forward "
        f"{np.mean(its):.1f} it/step")
    return {"ratio": t_rev / t_rec, "forward_ms": t_rec * 1e3,
            "reverse_ms": t_rev * 1e3, "replay_ms": st["replay_ms"] / steps,
            "reverse_steps_ms": st["reverse_ms"] / steps, "adjoint_it_per_step": st["adj_it"] / steps,
            "forward_it_per_step": float(np.mean(its)), "steps": steps, "checkpoint_every": every,
            "workload": (f"{case.mesh.interior[0]}x{case.mesh.interior[1]}x{case.mesh.interior[2]}, "
                         f"{len(case.sites)} rotors x 3 blades x {case.ap_per_blade} points, FP64, "
                         "pressure kept per step"),
            "J": J, "dJ_dyaw_per_turbine": [float(v) for v in wt_g[:, 0]]}


# output path kernels: id, algorithmic bytes per cell (FP64 form, precision form) and
# whether the count is over interior cells (calc_q) or every padded cell (the average)
OUTPUT_KERNELS = {"calc_q": (14, 32, 20, "interior"),        # read U(3); write q
                  "time_average": (15, 72, 60, "padded")}    # read Utavg(3), U(3); write Utavg(3)


def run_output_path(grid, reps=20):
    """Mean device time of the output-path kernels (Q-criterion, time average of U)
    at the bench grid, both forms, HIP events on the library's stream."""
    import ctypes as C
    hip = importlib.import_module("3dalm_amd.hip")
    case = make_case(grid, 1)
    m = case.mesh
    out = {}
    for form, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        c = hip.Context(m, prec)
        try:
            c.upload("U", case.U0())
            c.calc_q()
            c.time_average(1.0)
            res = {}
            for name, (kid, b64, b32, over) in OUTPUT_KERNELS.items():
                c.check(c.L.dalm_probe_kernel(c.h, kid), "probe")
                for n in range(reps):
                    c.calc_q() if name == "calc_q" else c.time_average(1.0 / (n + 2))
                launches, ms = C.c_int64(), C.c_double()
                c.check(c.L.dalm_probe_read(c.h, C.byref(launches), C.byref(ms)), "probe_read")
                c.check(c.L.dalm_probe_kernel(c.h, 0), "probe off")
                cells = m.N if over == "interior" else m.L
                nbytes = (b64 if form == "fp64" else b32) * cells
                mean_s = ms.value / launches.value / 1e3
                res[name] = {"mean_us": round(mean_s * 1e6, 2), "bytes_per_launch": nbytes,
                             "GB/s": round(nbytes / mean_s / 1e9, 1),
                             "frac": round(nbytes / mean_s / 1e9 / HBM_PEAK_GBS, 4), "launches": launches.value}
            out[form] = res
        finally:
            c.close()
    log(f"[bench] output path: {out}")
    return out


# the size the precision form's targets are stated at, with the Poisson tolerance
# that keeps avg div <= 1e-6 there: avg div = dt * max_diag * err and dt * max_diag =
# 230 at 512x256x256 (err <= 4.35e-9), where 1e-8 would leave 2.3e-6 in both forms
TARGET_GRID, TARGET_TOL = (512, 256, 256), 4e-9


def run_target_size(steps=100, warmup=10):
    """Both forms at the target size (two rotors 5 D apart, an initial turbulence of
    200 seeded Fourier modes at 10 % intensity):
    Mcell-steps/s, iterations, avg div, the whole-step algorithmic bandwidth and
    the turbine power after the last step."""
    hip = importlib.import_module("3dalm_amd.hip")
    cases = importlib.import_module("3dalm_amd.cases")
    res = {}
    for form, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        case = cases.case_c3(*TARGET_GRID, steps=warmup + steps, tol=TARGET_TOL)
        r, s = run_form(case, prec, steps, warmup, None, f"target-{form}")
        s.close()
        N = case.mesh.N
        b = STEP_BYTES[form]
        ms = 1e3 * r["seconds"] / steps
        bw = (b["fixed"] + b["per_solve"] + b["per_it"] * r["it_mean"]) * N / (ms * 1e-3) / 1e9
        res[form] = {"mcell_steps_s": round(N * steps / r["seconds"] / 1e6, 3), "ms_per_step": round(ms, 2),
                     "it_per_step": round(r["it_mean"], 1), "avg_div_max": r["avg_div_max"],
                     "power_last": r["power_last"], "step_bandwidth_GBs": round(bw, 1),
                     "step_frac": round(bw / HBM_PEAK_GBS, 4)}
    res["mixed_vs_fp64"] = round(res["mixed"]["mcell_steps_s"] / res["fp64"]["mcell_steps_s"], 3)
    p64, pm = res["fp64"]["power_last"], res["mixed"]["power_last"]
    res["power_rel_diff_last"] = abs(pm - p64) / abs(p64) if p64 else None
    res["workload"] = (f"{TARGET_GRID[0]}x{TARGET_GRID[1]}x{TARGET_GRID[2]}, two rotors 5 D apart, 200-mode "
                       f"Fourier turbulence at 10 %, BiCGStab(Jacobi x3) to {TARGET_TOL:g}, steps {warmup + 1}-{warmup + steps}")
    log(f"[bench] target size: {res}")
    return res


def run_sor_forms(steps=5, warmup=2):
    """The reference's own CPU-runnable case (64x48x48, no turbine, SOR 1.2 to 1e-6,
    at most 1000 iterations) in both forms: Mcell-steps/s and iterations per step."""
    cases = importlib.import_module("3dalm_amd.cases")
    hip = importlib.import_module("3dalm_amd.hip")
    res = {}
    for form, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        case = cases.case_c1(steps=warmup + steps)
        r, s = run_form(case, prec, steps, warmup, None, f"sor-{form}")
        s.close()
        res[form] = {"mcell_steps_s": round(case.mesh.N * steps / r["seconds"] / 1e6, 3),
                     "it_per_step": round(r["it_mean"], 1), "avg_div_max": r["avg_div_max"]}
    res["mixed_vs_fp64"] = round(res["mixed"]["mcell_steps_s"] / res["fp64"]["mcell_steps_s"], 3)
    res["workload"] = "64x48x48, no turbine, SOR 1.2 to 1e-6 (max 1000 it)"
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--grid", default="256,128,128")
    ap.add_argument("--tol", type=float, default=1e-8, help="Poisson tolerance (the workload's is 1e-8)")
    ap.add_argument("--probe", default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=20)
    ap.add_argument("--adjoint-every", type=int, default=10)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the output-path and SOR legs")
    ap.add_argument("--no-target", action="store_true", help="skip the 512x256x256 leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a kernel of known traffic for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the whole domain lives on one MI355X and there is "
              "no multi-GPU path (it would shard as x-slabs with one- and two-plane halo exchanges). Run with "
              "--gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    hip = importlib.import_module("3dalm_amd.hip")
    if a.copy_calibration:
        import torch
        n = a.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {a.copy_calibration} MiB done")
        del src, dst
    grid = tuple(int(v) for v in a.grid.split(","))
    forms = [f for f in a.forms.split(",") if f]
    unknown = [f for f in forms if f not in FORMS]
    if unknown:
        print(f"bench.py: unknown form(s) {unknown}; choose from {list(FORMS)}", file=sys.stderr)
        sys.exit(2)
    have_mixed = getattr(hip, "MIXED_BUILT", False)
    if not have_mixed and "mixed" in forms:
        log("[bench] precision form not built in this library; measuring the FP64 drop-in only")
        forms = [f for f in forms if f != "mixed"]
    results = {}
    N = None
    for form in forms:
        case = make_case(grid, a.steps, a.tol)
        N = case.mesh.N
        if a.probe == "auto":
            probe = "jacobi" if form == "fp64" else "jacobi32"
        else:
            probe = a.probe if a.probe != "none" else None
        prec = hip.FP64 if form == "fp64" else hip.MIXED
        res, s = run_form(case, prec, a.steps, a.warmup, probe, form)
        s.close()
        res["mcell_steps_s"] = N * a.steps / res["seconds"] / 1e6
        res["ms_per_step"] = 1e3 * res["seconds"] / a.steps
        results[form] = res
    main_form = "mixed" if "mixed" in results else next(iter(results))
    R = results[main_form]

    def roofline(res):
        if "probe" not in res or not res["probe"]["launches"]:
            return None
        pr = res["probe"]
        mean_s = pr["total_ms"] / pr["launches"] / 1e3
        bytes_per_launch = PROBE_BYTES[pr["kernel"]] * N
        ach = bytes_per_launch / mean_s / 1e9
        traffic = None
        try:
            rec = json.load(open(TRAFFIC_JSON))
            k = rec["kernels"].get(PROBE_KERNEL.get(pr["kernel"], ""))
            if k and rec["cells"] == N:
                traffic = round((k["read_B_per_cell"] + k["write_B_per_cell"]) * N)
        except (OSError, ValueError, KeyError):
            pass
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": pr["kernel"],
                "launches": pr["launches"], "mean_us": round(mean_s * 1e6, 2),
                "bytes_per_launch": bytes_per_launch,
                "traffic_source": os.path.relpath(TRAFFIC_JSON, ROOT) if traffic else None}
    roof = roofline(R)
    roof_fp64 = roofline(results["fp64"]) if "fp64" in results and main_form != "fp64" else None
    adj = None
    if not a.no_adjoint:
        try:
            adj = run_adjoint(a.adjoint_steps, a.adjoint_every)
        except Exception as e:
            log(f"[bench] adjoint leg failed: {e}")
    output_path = sor = None
    if not a.no_extras:
        try:
            output_path = run_output_path(grid)
            sor = run_sor_forms()
        except Exception as e:
            log(f"[bench] output-path / SOR legs failed: {e}")
    target = None
    if not a.no_target:
        try:
            target = run_target_size()
        except Exception as e:
            log(f"[bench] target-size leg failed: {e}")
    cpu = None
    if not a.no_cpu:
        try:
            # the reference solves in FP64: extrapolate with the FP64 form's iterations
            it_ref = results["fp64"]["it_mean"] if "fp64" in results else R["it_mean"]
            cpu = cpu_baseline(make_case(grid, 1), it_ref)
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")
    def step_roofline(form, r):
        """whole-step algorithmic bytes (fixed + per iteration x iterations) over the
        time per step (the ALM projection's box and the precision form's residual
        replacements are left out: a lower bound)"""
        b = STEP_BYTES[form]
        nbytes = (b["fixed"] + b["per_solve"] + b["per_it"] * r["it_mean"]) * N
        ach = nbytes / (r["ms_per_step"] * 1e-3) / 1e9
        return {"bytes_per_step": round(nbytes), "achieved": round(ach, 1), "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4)}

    forms_out = {f: {"mcell_steps_s": round(r["mcell_steps_s"], 3), "ms_per_step": round(r["ms_per_step"], 3),
                     "it_per_step": round(r["it_mean"], 2), "avg_div_max": r["avg_div_max"],
                     "div_within_1e-6": bool(r["avg_div_max"] <= 1e-6),
                     "power_last": r["power_last"], "step_roofline": step_roofline(f, r)}
                 for f, r in results.items()}

    line = {
        "metric": METRIC,
        "value": round(R["mcell_steps_s"], 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(R["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"single 3-blade turbine ALM, {grid[0]}x{grid[1]}x{grid[2]} cells, "
                                f"BiCGStab(Jacobi x3) to {a.tol:g}, form={main_form}"),
                   "cells": N, "turbines": 1, "actuator_points": 192, "poisson": "BiCGStab+Jacobi3",
                   "tol": a.tol},
        "forms": forms_out,
        "mixed_vs_fp64": (round(results["mixed"]["mcell_steps_s"] / results["fp64"]["mcell_steps_s"], 3)
                          if "mixed" in results and "fp64" in results else None),
        "adjoint_forward_ratio": round(adj["ratio"], 3) if adj else None,
        "adjoint": adj,
        "roofline": roof,
        "roofline_fp64": roof_fp64,
        "cpu_baseline": cpu,
        "target_size": target,
        "output_path": output_path,
        "sor_forms": sor,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
