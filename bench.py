"""Benchmark of the 3dalm per-time-step hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--forms fp64,mixed] [--tol 4e-9]
                    [--no-cpu] [--cpu-steps 1] [--probe auto]
                    [--adjoint-steps 200] [--adjoint-every 20] [--adjoint-tol 1e-6] [--no-adjoint]
                    [--mixed-adjoint-steps 20] [--c5-adjoint-steps 0] [--adjoint-tol-study TOLS]
                    [--no-extras] [--calibration N]

Headline workload (one "step" = one Solver::main_loop_once): 512x256x256 cells on a
16x8x8 box (y/z tanh beta=1.5), two 3-blade rotors (D = 1) 5 D apart at x = 3 and
8 with 3 x 64 actuator points each, TSR 6, eps = 2 dx, a seeded isotropic turbulence
of 200 random Fourier modes at 10 % intensity as the initial field and, frozen and
convected at the inflow speed, as the x- inflow, dt = 0.005, Poisson by BiCGStab with
3 Jacobi sweeps as preconditioner to 4e-9 (the tolerance that keeps avg div <= 1e-6
at this size).  Synthetic inputs, resident in HBM before the timed region.  W
untimed warm-up steps from the impulsive start, then steps W+1 .. W+K timed, in
each form:

  fp64   the FP64 drop-in;
  mixed  the precision form (FP32 stencils, FP32 BiCGStab vectors with FP64 dot
         products, norms, pressure and residual replacements, to the same
         true-residual test).

Prints ONE JSON line.  value = Mcell-steps/s = N_interior * K / t / 1e6 of the
precision form; forms holds both (ms/step, iterations per step and how many of them
ran in FP64, avg div, turbine power, the whole step's algorithmic bandwidth);
mixed_vs_fp64 is their ratio and mixed_errors_vs_fp64 the relative L2 distance of
the precision form's U, p and div from the FP64 drop-in's after the last timed step,
with the turbine-power difference.

roofline: the dominant kernel of the precision form on rows of 256 cells, its fused
p-half-iteration kg_half4<0> (the previous iteration's r update, the p update, three
FP32 Jacobi sweeps, FP64 SpMV and dot product in one pass; roofline_fp64: the FP64
drop-in's fused update and two sweeps, k_half_pc4<0>): algorithmic bytes per launch
(36 / 40 B per interior cell x N) over
the mean launch duration, timed with HIP events on the library's stream over the
timed steps.  traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE
runs of this workload (profiles/hbm_traffic_alm_512x256x256.json), each counter
scaled by what it reads for the library's own copy kernel (--calibration).

cpu_baseline: the oracle's OpenMP build on every host core, full steps of the
headline workload continuing from the FP64 GPU run's final state.

adjoint_forward_ratio: on the differentiation-form workload (192x96x96, three
rotors, FP64, BiCGStab(Jacobi 3) to 1e-8), the wall time per step of the reverse
sweep (segment replays from checkpoints plus reverse steps, adjoint Poisson solves
to --adjoint-tol relative, 1e-6 by default) over the wall time per step of the
recorded forward run of the same steps; 200 steps with checkpoints every 20 by
default.  --adjoint-tol-study 1e-10,1e-8,1e-6 runs a short window (30 steps, then 20
recorded with checkpoints every 10) at each tolerance and reports the cost ratio
and the largest relative difference of the yaw and pitch gradients from those of
the first (tightest) tolerance: at 1e-6 they agree with 1e-10's to about 1e-7,
a thousandth of the 1e-4 finite-difference bar.

c2_256x128x128: the single-rotor 256x128x128 workload (tolerance 1e-8) in both
forms.  output_path: mean device time and achieved bandwidth of the output-path
kernels (Q-criterion, time average of U) at 256x128x128, both forms.  sor_forms:
the 64x48x48 SOR case (no turbine) in both forms.
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
# the headline workload: the size the precision form's targets are stated at, with the
# Poisson tolerance that keeps avg div <= 1e-6 there (avg div = dt * max_diag * err and
# dt * max_diag = 230 at 512x256x256, so err <= 4.35e-9)
HEADLINE_GRID, HEADLINE_TOL = (512, 256, 256), 4e-9
HBM_PEAK_GBS = 8000.0
# per-kernel HBM traffic of this workload from rocprofv3 FETCH_SIZE / WRITE_SIZE runs
# (3dalm_amd/profile_summary.py traffic --json; the command is in the .md beside it)
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hbm_traffic_alm_512x256x256.json")
# the kernels behind each probe on rows of a multiple of 256 interior cells (the
# headline grid), where the iteration kernels run four cells per lane
PROBE_KERNEL = {"jacobi": "k_jacobi4", "jacobi32": "kf_jacobi4", "spmv": "k_bi_spmv4<0>", "update": "k_bi_xr4<0>",
                "pupdate": "k_bi_p4", "supdate": "k_bi_s4", "predictor": "k_predictor",
                "spmv_mixed": "kg_spmv4<0>", "update_mixed": "kg_xr4", "pupdate_mixed": "kg_p4",
                "supdate_mixed": "kg_s4",
                # rows of exactly 256 cells: the fused half-iterations
                "half_p_mixed": "kg_half4<0>", "half_s_mixed": "kg_half4<1>",
                "half_p": "k_half_pc4<0>", "half_s": "k_half_pc4<1>"}

# algorithmic bytes per interior cell for one launch of each probed kernel (FP64):
# compulsory reads + writes, with the operator applied matrix-free from 1-D coefficients
PROBE_BYTES = {
    "jacobi": 24,     # read z_in, v; write z_out
    "spmv": 24,       # read z, a1 (r0 or s); write y
    # read s, t, r0, write r (k_bi_xr4<0>: the fused iteration's x updates are applied
    # eight at a time by k_x_group4, 144 B; 64 B with x on rows that do not take the
    # fused iteration)
    "update": 32,
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
    # fused half-iterations (rows of 256 cells, Jacobi x3): every intermediate level
    # stays on chip
    # p-half: read s, t, p, q, r0; write r, p, z3, q (FP32; the pressure updates are
    # applied by kg_apply_pending4, eight iterations' at a time)
    "half_p_mixed": 36,
    "half_s_mixed": 24,  # read r, q, r0; write s, z3, t (FP32)
    "half_p": 40,        # read r, p, q; write p, z3 (FP64)
    "half_s": 32,        # read r, q; write s, z3 (FP64)
}
# algorithmic bytes per interior cell of one whole step: the fixed kernels (predictor,
# flux + right-hand side, projection, post with div stored) and one BiCGStab(Jacobi 3)
# iteration, plus the solve's start (residual and shadow residual)
STEP_BYTES = {"fp64": {"fixed": 104 + 48 + 32 + 104 + 64, "per_it": 280, "per_solve": 32},
              "mixed": {"fixed": 64 + 44 + 80 + 48, "per_it": 148, "per_solve": 28}}
# with the fused half-iterations (rows of exactly 256 cells, Jacobi x3): FP64 40 + 24
# (SpMV) + 32 + 24 + 32 (r update) + 18 (the x updates of eight iterations every eighth
# time, 144 B); precision form 36 (p-half with the previous iteration's r update) + 24
# + 10 (the x updates of eight iterations every eighth time, 80 B)
STEP_BYTES_FUSED = {"fp64": {"fixed": 104 + 48 + 32 + 104 + 64, "per_it": 170, "per_solve": 32},
                    "mixed": {"fixed": 64 + 44 + 80 + 48, "per_it": 70, "per_solve": 28}}
PROBE_IDS = {"jacobi": 1, "spmv": 2, "update": 3, "predictor": 4, "sor": 5, "resnorm": 6, "pupdate": 7,
             "supdate": 8, "jacobi32": 9, "spmv_mixed": 10, "update_mixed": 11, "pupdate_mixed": 12,
             "supdate_mixed": 13, "half_p_mixed": 16, "half_s_mixed": 17, "half_p": 18, "half_s": 19}


def fused_rows(mesh, case):
    """The half-iterations run fused on rows of exactly 256 cells with BiCGStab(Jacobi x3)."""
    return mesh.interior[2] == 256 and case.method == "BiCG" and case.pc == "Jacobi" and case.pc_it == 3
FORMS = ("fp64", "mixed")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_case(grid, steps, tol=1e-8):
    """The single-rotor workload (256x128x128 by default)."""
    cases = importlib.import_module("3dalm_amd.cases")
    nx, ny, nz = grid
    c = cases.case_c2(nx, ny, nz, steps=steps)
    c.tol = tol
    return c


def make_headline_case(steps, tol):
    cases = importlib.import_module("3dalm_amd.cases")
    return cases.case_c3(*HEADLINE_GRID, steps=steps, tol=tol)


def run_form(case, precision, steps, warmup, probe, form, keep_fields=False):
    """warmup untimed steps, then `steps` timed ones.  With `probe`, every launch of
    that kernel inside the timed region is timed with HIP events on the library's
    stream (dalm_probe_kernel).  With keep_fields, U, p and div after the last step
    are returned (for the comparison of the forms)."""
    solver = importlib.import_module("3dalm_amd.solver")
    import ctypes as C
    log(f"[bench] form={form} grid={case.mesh.interior} warmup={warmup} steps={steps}")
    s = solver.Solver(case, precision)
    try:
        for w in range(warmup):
            st = s.step()
            log(f"[bench] warmup step {st['step']} it={st['it']} err={st['err']:.3e} div={st['avg_div']:.3e}")
        if probe:
            s.ctx.check(s.ctx.L.dalm_probe_kernel(s.ctx.h, PROBE_IDS[probe]), "probe")
        sync()
        t0 = time.perf_counter()
        its, it64, divs, power = [], [], [], []
        for k in range(steps):
            st = s.step()
            its.append(st["it"])
            it64.append(s.ctx.last_solve_info()[1])
            divs.append(st["avg_div"])
            if "power" in st:
                power.append(float(st["power"].sum()))
        sync()
        t1 = time.perf_counter()
        log(f"[bench] timed {steps} steps in {t1 - t0:.3f} s; it/step mean {np.mean(its):.1f} "
            f"(FP64 {np.mean(it64):.1f}); last div {divs[-1]:.3e}")
        res = {"seconds": t1 - t0, "it_mean": float(np.mean(its)), "it64_mean": float(np.mean(it64)), "it": its,
               "avg_div_max": float(np.max(divs)), "avg_div_last": float(divs[-1]),
               "power_last": power[-1] if power else None,
               "window": [warmup + 1, warmup + steps]}
        if probe:
            n, ms = C.c_int64(), C.c_double()
            s.ctx.check(s.ctx.L.dalm_probe_read(s.ctx.h, C.byref(n), C.byref(ms)), "probe_read")
            s.ctx.check(s.ctx.L.dalm_probe_kernel(s.ctx.h, 0), "probe off")
            res["probe"] = {"kernel": probe, "launches": n.value, "total_ms": ms.value}
        if keep_fields:
            res["fields"] = {name: s.field(name) for name in ("U", "p", "div")}
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


def cpu_baseline(case, state, steps, fp64_res):
    with _StdoutToStderr():
        return _cpu_baseline(case, state, steps, fp64_res)


def _omp_threads():
    import ctypes
    try:
        return int(ctypes.CDLL("libgomp.so.1").omp_get_max_threads())
    except OSError:
        return os.cpu_count() or 1


def _cpu_baseline(case, state, steps, fp64_res):
    """The oracle's OpenMP build (the C restatement of the reference's step with its
    cell loops shared among threads) on every host core: `steps` full steps of the
    headline workload, continuing from the FP64 GPU run's state after its timed
    steps, so the CPU integrates the next step(s) of the same run, at the same point
    of the flow, with the same BiCGStab.  The reference's Gaussian projection is
    dense over all cells and actuator points (src/alm.h:416-448), as here."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    po = importlib.import_module("pyoracle")
    cores = _omp_threads()
    t0 = time.perf_counter()
    S = po.OracleSolver(case, kind="omp", U0=state["U"], initialize=False)
    S.load_state({k: state[k] for k in ("Uold", "JU", "p", "nut", "f")}, state["step"])
    t_setup = time.perf_counter() - t0
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = S.step()
        its.append(st["it"])
        log(f"[bench] cpu step {st['step']}: it {st['it']} err {st['err']:.3e} ({time.perf_counter() - t0:.1f} s)")
    t = time.perf_counter() - t0
    m = case.mesh
    return {"value": m.N * steps / t / 1e6, "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
            "sample": (f"{steps} full step(s) of the headline workload ({m.interior[0]}x{m.interior[1]}x"
                       f"{m.interior[2]}, step {state['step'] + 1}, continuing from the FP64 GPU run's state) "
                       f"on {cores} OpenMP threads: {t:.1f} s, {np.mean(its):.0f} BiCGStab iterations per step "
                       f"(the GPU's FP64 form: {fp64_res['it_mean']:.0f} over its timed steps); set-up "
                       f"{t_setup:.1f} s untimed"),
            "seconds_per_step": t / steps, "it_per_step": float(np.mean(its))}


def run_adjoint(steps, every, warmup=2, tol=1e-6, grads=False):
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
        J, ap_g, wt_g, st = s.ctx.adjoint_run(tol)
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
            "warmup": warmup, "tol_rel": tol,
            "workload": (f"{case.mesh.interior[0]}x{case.mesh.interior[1]}x{case.mesh.interior[2]}, "
                         f"{len(case.sites)} rotors x 3 blades x {case.ap_per_blade} points, FP64, "
                         "pressure kept per step"),
            "J": J, "dJ_dyaw_per_turbine": [float(v) for v in wt_g[:, 0]],
            "dJ_dpitch_per_turbine": [float(v) for v in wt_g[:, 1]],
            **({"grads": np.concatenate([wt_g[:, 0], wt_g[:, 1]])} if grads else {})}


def adjoint_tol_study(tols, warmup=30, steps=20, every=10):
    """The adjoint leg on a short window at each tolerance: cost ratio, adjoint
    iterations per step and the largest relative difference of dJ/d(yaw, pitch)
    from the first tolerance's."""
    out, ref = [], None
    for tol in tols:
        r = run_adjoint(steps, every, warmup, tol, grads=True)
        g = r.pop("grads")
        ref = g if ref is None else ref
        out.append({"tol_rel": tol, "ratio": round(r["ratio"], 3), "forward_ms": round(r["forward_ms"], 2),
                    "reverse_ms": round(r["reverse_ms"], 2), "adjoint_it_per_step": r["adjoint_it_per_step"],
                    "forward_it_per_step": r["forward_it_per_step"],
                    "grad_max_rel_vs_first": float(np.max(np.abs(g - ref) / np.abs(ref)))})
        log(f"[bench] adjoint tolerance {tol:g}: {out[-1]}")
    return {"warmup": warmup, "steps": steps, "checkpoint_every": every, "runs": out}


def run_precision_adjoint(case, steps, every, warmup=2, tol=1e-6):
    """The precision form's adjoint: `warmup` steps of the precision form, then
    `steps` recorded ones (dalm_adjoint_begin_from into an FP64 context, checkpoints
    every `every` steps, no pressure tape: the replays re-solve in FP64) and the
    reverse sweep.  Wall time per reverse step (FP64 segment replays plus reverse
    steps) over wall time per recorded precision-form step."""
    solver = importlib.import_module("3dalm_amd.solver")
    hip = importlib.import_module("3dalm_amd.hip")
    omega = np.array([st.rot_speed for st in case.sites])
    log(f"[bench] precision-form adjoint: grid={case.mesh.interior} warmup={warmup} steps={steps} every={every}")
    sm = solver.Solver(case, hip.MIXED)
    s64 = None
    try:
        for _ in range(warmup):
            sm.step()
        s64 = solver.Solver(case)
        s64.ctx.adjoint_begin_from(sm.ctx, steps, every, omega / steps, False)
        its = []
        t1 = time.perf_counter()
        for n in range(steps):
            its.append(sm.step()["it"])
            if (n + 1) % 10 == 0:
                log(f"[bench] precision-form adjoint: recorded {n + 1} of {steps} steps")
        t_rec = (time.perf_counter() - t1) / steps
        t2 = time.perf_counter()
        J, ap_g, wt_g, st = s64.ctx.adjoint_run(tol)
        t_rev = (time.perf_counter() - t2) / steps
        s64.ctx.adjoint_end()
    finally:
        sm.close()
        if s64 is not None:
            s64.close()
    log(f"[bench] precision-form adjoint: recorded forward {t_rec * 1e3:.2f} ms/step, reverse {t_rev * 1e3:.2f} "
        f"ms/step (replay {st['replay_ms'] / steps:.2f} + reverse steps {st['reverse_ms'] / steps:.2f})")
    m = case.mesh
    return {"ratio": t_rev / t_rec, "forward_ms": t_rec * 1e3, "reverse_ms": t_rev * 1e3,
            "replay_ms": st["replay_ms"] / steps, "reverse_steps_ms": st["reverse_ms"] / steps,
            "adjoint_it_per_step": st["adj_it"] / steps, "forward_it_per_step": float(np.mean(its)),
            "steps": steps, "checkpoint_every": every, "warmup": warmup, "tol_rel": tol,
            "workload": (f"{m.interior[0]}x{m.interior[1]}x{m.interior[2]}, {len(case.sites)} rotors x 3 blades x "
                         f"{case.ap_per_blade} points, precision form recorded into an FP64 context, FP64 "
                         "replays re-solving the pressure"),
            "J": J, "dJ_dyaw_per_turbine": [float(v) for v in wt_g[:, 0]],
            "dJ_dpitch_per_turbine": [float(v) for v in wt_g[:, 1]]}


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


def run_c2_forms(steps=100, warmup=10):
    """The single-rotor workload at 256x128x128 (BiCGStab(Jacobi 3) to 1e-8) in both
    forms: Mcell-steps/s, iterations per step, avg div."""
    hip = importlib.import_module("3dalm_amd.hip")
    res = {}
    for form, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        case = make_case((256, 128, 128), warmup + steps)
        r, s = run_form(case, prec, steps, warmup, None, f"c2-{form}")
        s.close()
        res[form] = {"mcell_steps_s": round(case.mesh.N * steps / r["seconds"] / 1e6, 3),
                     "ms_per_step": round(1e3 * r["seconds"] / steps, 3), "it_per_step": round(r["it_mean"], 1),
                     "fp64_it_per_step": round(r["it64_mean"], 1), "avg_div_max": r["avg_div_max"]}
    res["mixed_vs_fp64"] = round(res["mixed"]["mcell_steps_s"] / res["fp64"]["mcell_steps_s"], 3)
    res["workload"] = (f"256x128x128, one rotor (3 x 64 actuator points), 5 % uniform noise, BiCGStab(Jacobi x3) "
                       f"to 1e-8, steps {warmup + 1}-{warmup + steps}")
    log(f"[bench] c2: {res}")
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


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def interior(mesh, a):
    nx, ny, nz = mesh.size
    g = mesh.gc
    return a.reshape((nx, ny, nz) + a.shape[1:])[g:nx - g, g:ny - g, g:nz - g]


def calibration_kernels(reps, mesh):
    """Kernels of known HBM traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE, all of
    them the library's own (no second HIP runtime in the process): `reps` calls each of
    dalm_copy_u_to_uold (k_copy3 reads 24 B and writes 24 B per padded cell) and of
    dalm_l2_norm (k_scalar_sum reads 8 B per interior cell), on a context of `mesh`."""
    hip = importlib.import_module("3dalm_amd.hip")
    c = hip.Context(mesh)
    try:
        rng = np.random.default_rng(5)
        c.upload("U", rng.uniform(-1.0, 1.0, 3 * mesh.L))
        c.upload("p", rng.uniform(-1.0, 1.0, mesh.L))
        for _ in range(reps):
            c.copy_u_to_uold()
        for _ in range(reps):
            c.l2_norm("p")
    finally:
        c.close()
    log(f"[bench] calibration: {reps} copies of U ({mesh.L} padded cells x 3 fields) and {reps} l2 norms")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--tol", type=float, default=HEADLINE_TOL,
                    help="Poisson tolerance (4e-9 keeps avg div <= 1e-6 at 512x256x256)")
    ap.add_argument("--probe", default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1, help="full steps of the CPU baseline")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-tol", type=float, default=1e-6,
                    help="relative tolerance of the adjoint Poisson solves")
    ap.add_argument("--adjoint-tol-study", default="", metavar="TOLS",
                    help="comma-separated tolerances: the adjoint leg on a short window at each")
    ap.add_argument("--mixed-adjoint-steps", type=int, default=20,
                    help="recorded steps of the precision form's adjoint on the 192x96x96 geometry (0: skip)")
    ap.add_argument("--c5-adjoint-steps", type=int, default=0,
                    help="recorded steps of the precision form's adjoint at 768x384x384 (0: skip; 100 is the "
                         "configuration's window)")
    ap.add_argument("--c5-warmup", type=int, default=2)
    ap.add_argument("--no-extras", action="store_true", help="skip the 256x128x128, output-path and SOR legs")
    ap.add_argument("--calibration", type=int, default=0, metavar="N",
                    help="after the forms, run N copies of the velocity field and N l2 norms of one field "
                         "(kernels of known traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--c2-only", action="store_true",
                    help="run only the single-rotor 256x128x128 workload in both forms (--steps, --warmup; "
                         "then --calibration on its grid) and print its record: for profiling that size")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the whole domain lives on one MI355X and there is "
              "no multi-GPU path (it would shard as x-slabs with one- and two-plane halo exchanges). Run with "
              "--gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if a.c2_only:
        rec = run_c2_forms(a.steps, a.warmup)
        if a.calibration:
            calibration_kernels(a.calibration, make_case((256, 128, 128), 1).mesh)
        print(json.dumps(rec), flush=True)
        return
    hip = importlib.import_module("3dalm_amd.hip")
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
    fp64_state = None
    N = mesh = None
    fused = False
    for form in forms:
        case = make_headline_case(a.warmup + a.steps, a.tol)
        N, mesh = case.mesh.N, case.mesh
        fused = fused_rows(mesh, case)
        if a.probe == "auto":
            if fused_rows(case.mesh, case):
                probe = "half_p" if form == "fp64" else "half_p_mixed"
            else:
                probe = "jacobi" if form == "fp64" else "jacobi32"
        else:
            probe = a.probe if a.probe != "none" else None
        prec = hip.FP64 if form == "fp64" else hip.MIXED
        res, s = run_form(case, prec, a.steps, a.warmup, probe, form, keep_fields=len(forms) > 1)
        if form == "fp64" and not a.no_cpu:
            # the state after the timed steps: the CPU baseline continues from it
            fp64_state = {name: s.field(name) for name in ("U", "Uold", "p", "JU", "nut", "f")}
            fp64_state["step"] = s.stats["step"]
        s.close()
        res["mcell_steps_s"] = N * a.steps / res["seconds"] / 1e6
        res["ms_per_step"] = 1e3 * res["seconds"] / a.steps
        results[form] = res
    if a.calibration and mesh is not None:
        calibration_kernels(a.calibration, mesh)
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
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": PROBE_KERNEL[pr["kernel"]],
                "launches": pr["launches"], "mean_us": round(mean_s * 1e6, 2),
                "bytes_per_launch": bytes_per_launch, "timing": "HIP events on the library's stream, timed steps",
                "traffic_source": os.path.relpath(TRAFFIC_JSON, ROOT) if traffic else None}
    roof = roofline(R)
    roof_fp64 = roofline(results["fp64"]) if "fp64" in results and main_form != "fp64" else None
    errors = None
    if "fp64" in results and "mixed" in results:
        f64, fm = results["fp64"].pop("fields"), results["mixed"].pop("fields")
        pm_i, p64_i = interior(mesh, fm["p"]), interior(mesh, f64["p"])
        dm_i, d64_i = interior(mesh, fm["div"]), interior(mesh, f64["div"])
        errors = {"U_rel_l2": rel_l2(interior(mesh, fm["U"]), interior(mesh, f64["U"])),
                  "p_rel_l2": rel_l2(pm_i, p64_i),
                  # the pressure enters the step only through its differences: their
                  # distance (forward differences along i, j, k over the interior)
                  "grad_p_rel_l2": rel_l2(np.concatenate([np.diff(pm_i, axis=ax).reshape(-1) for ax in range(3)]),
                                          np.concatenate([np.diff(p64_i, axis=ax).reshape(-1) for ax in range(3)])),
                  # div is dt * max_diag * (b - A p): both fields are at the solve's
                  # tolerance, so their RMS values are stated rather than their distance
                  "div_rms_fp64": float(np.sqrt(np.mean(d64_i ** 2))),
                  "div_rms_mixed": float(np.sqrt(np.mean(dm_i ** 2))),
                  "div_rel_l2": rel_l2(dm_i, d64_i),
                  "at_step": a.warmup + a.steps}
        del pm_i, p64_i, dm_i, d64_i
        p64, pm = results["fp64"]["power_last"], results["mixed"]["power_last"]
        errors["power_rel"] = abs(pm - p64) / abs(p64) if p64 else None
        del f64, fm
        log(f"[bench] precision form against FP64: {errors}")
    for r in results.values():
        r.pop("fields", None)
    adj = None
    if not a.no_adjoint:
        try:
            adj = run_adjoint(a.adjoint_steps, a.adjoint_every, tol=a.adjoint_tol)
        except Exception as e:
            log(f"[bench] adjoint leg failed: {e}")
    adj_mixed = adj_c5 = None
    if not a.no_adjoint and a.mixed_adjoint_steps > 0 and have_mixed:
        try:
            cases = importlib.import_module("3dalm_amd.cases")
            n = a.mixed_adjoint_steps
            adj_mixed = run_precision_adjoint(cases.case_c4(steps=2 + n), n, max(1, n // 2), tol=a.adjoint_tol)
        except Exception as e:
            log(f"[bench] precision-form adjoint leg failed: {e}")
    if a.c5_adjoint_steps > 0 and have_mixed:
        try:
            cases = importlib.import_module("3dalm_amd.cases")
            n = a.c5_adjoint_steps
            adj_c5 = run_precision_adjoint(cases.case_c5(steps=a.c5_warmup + n), n, max(1, min(12, n // 2)),
                                           warmup=a.c5_warmup, tol=a.adjoint_tol)
        except Exception as e:
            log(f"[bench] 768x384x384 adjoint leg failed: {e}")
    tol_study = None
    if a.adjoint_tol_study:
        try:
            tol_study = adjoint_tol_study([float(v) for v in a.adjoint_tol_study.split(",") if v])
        except Exception as e:
            log(f"[bench] adjoint tolerance study failed: {e}")
    c2 = output_path = sor = None
    if not a.no_extras:
        try:
            c2 = run_c2_forms(min(a.steps, 100), a.warmup)
            output_path = run_output_path((256, 128, 128))
            sor = run_sor_forms()
        except Exception as e:
            log(f"[bench] 256x128x128 / output-path / SOR legs failed: {e}")
    cpu = None
    if not a.no_cpu and fp64_state is not None:
        try:
            cpu = cpu_baseline(make_headline_case(a.warmup + a.steps + a.cpu_steps, a.tol), fp64_state,
                               a.cpu_steps, results["fp64"])
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")

    def step_roofline(form, r):
        """whole-step algorithmic bytes (fixed + per iteration x iterations) over the
        time per step (the ALM projection's box and the precision form's residual
        replacements are left out; its FP64 iterations are counted at FP64 bytes)"""
        table = STEP_BYTES_FUSED if fused else STEP_BYTES
        b = table[form]
        if form == "mixed":
            per_it = (b["per_it"] * (r["it_mean"] - r["it64_mean"]) + table["fp64"]["per_it"] * r["it64_mean"])
        else:
            per_it = b["per_it"] * r["it_mean"]
        nbytes = (b["fixed"] + b["per_solve"] + per_it) * N
        ach = nbytes / (r["ms_per_step"] * 1e-3) / 1e9
        return {"bytes_per_step": round(nbytes), "achieved": round(ach, 1), "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4)}

    forms_out = {f: {"mcell_steps_s": round(r["mcell_steps_s"], 3), "ms_per_step": round(r["ms_per_step"], 3),
                     "it_per_step": round(r["it_mean"], 2), "fp64_it_per_step": round(r["it64_mean"], 2),
                     "avg_div_max": r["avg_div_max"], "div_within_1e-6": bool(r["avg_div_max"] <= 1e-6),
                     "power_last": r["power_last"], "window": r["window"], "step_roofline": step_roofline(f, r)}
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
        "config": {"workload": (f"two 3-blade rotors 5 D apart (ALM, 2 x 3 x 64 actuator points), "
                                f"{HEADLINE_GRID[0]}x{HEADLINE_GRID[1]}x{HEADLINE_GRID[2]} cells, 200-mode Fourier "
                                f"turbulence at 10 %, BiCGStab(Jacobi x3) to {a.tol:g}, steps "
                                f"{a.warmup + 1}-{a.warmup + a.steps}, form={main_form}"),
                   "cells": N, "turbines": 2, "actuator_points": 384, "poisson": "BiCGStab+Jacobi3",
                   "tol": a.tol},
        "forms": forms_out,
        "mixed_vs_fp64": (round(results["mixed"]["mcell_steps_s"] / results["fp64"]["mcell_steps_s"], 3)
                          if "mixed" in results and "fp64" in results else None),
        "mixed_errors_vs_fp64": errors,
        "adjoint_forward_ratio": round(adj["ratio"], 3) if adj else None,
        "adjoint": adj,
        "adjoint_mixed": adj_mixed,
        "adjoint_mixed_768x384x384": adj_c5,
        "adjoint_tol_study": tol_study,
        "roofline": roof,
        "roofline_fp64": roof_fp64,
        "cpu_baseline": cpu,
        "c2_256x128x128": c2,
        "output_path": output_path,
        "sor_forms": sor,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
