"""Whole-step parity of the FP64 drop-in (Solver::main_loop_once order) against the
CPU oracle on identical seeded inputs.

* SOR paths: the sweeps are bit-exact, so iteration counts are equal and every
  field agrees to <= 1e-10 relative L2.  div is a sum of nearly cancelling face
  fluxes, so its error is measured on the flux scale max|U|/min(dx) instead of
  relative to its own (tiny) norm.
* BiCGStab paths: alpha, omega and beta come from dot products whose summation
  order (a block tree on the GPU, a serial loop in the reference, src/eq.h:62-73)
  changes their last bits; BiCGStab amplifies that, so at the production tolerance
  1e-8 iteration counts differ by a few and the solution by up to the tolerance.
  The 1e-10 bar is therefore checked with the solve converged to 1e-15, where the
  solution no longer depends on the Krylov path (p itself needs that: its smooth
  component is barely seen by the residual test, so two exact-arithmetic-equivalent
  solvers still differ by 1e-9 in p at 1e-13 and by 1e-11 at 1e-15, while U agrees
  to 1e-13 at either); at 1e-8 the fields must agree within what that tolerance
  allows.
"""
import importlib
import os

import numpy as np
import pytest

from helpers import interior_mask, rel_l2

hip = importlib.import_module("3dalm_amd.hip")
cases = importlib.import_module("3dalm_amd.cases")
solver = importlib.import_module("3dalm_amd.solver")
po = importlib.import_module("pyoracle")

pytestmark = pytest.mark.gpu
TOL = 1e-10


def field_errors(case, fields):
    m = interior_mask(case.mesh)
    U = fields["U"][1]
    dmin = min(np.min(np.diff(case.mesh.x)), np.min(np.diff(case.mesh.y)), np.min(np.diff(case.mesh.z)))
    flux_scale = np.max(np.abs(U)) / dmin
    errs = {}
    for name, (a, b) in fields.items():
        if name == "div":
            errs[name] = np.sqrt(np.mean((a[m] - b[m]) ** 2)) / flux_scale
        elif name in ("p", "nut"):
            errs[name] = rel_l2(a[m], b[m])
        else:
            errs[name] = rel_l2(a, b)
    print("field errors:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    return errs


def run_pair(case, steps, exact_it, power_rtol=TOL):
    print(f"[case {case.name} {case.mesh.interior} {case.method}/{case.pc} tol={case.tol} steps={steps}]",
          flush=True)
    o = po.OracleSolver(case)
    s = solver.Solver(case)
    try:
        for n in range(steps):
            so, sg = o.step(), s.step()
            print(f"  step {n + 1}: it {sg['it']} vs {so['it']}, err {sg['err']:.6e} vs {so['err']:.6e}",
                  flush=True)
            assert sg["step"] == so["step"] == n + 1
            if exact_it:
                assert sg["it"] == so["it"], (n, sg["it"], so["it"])
                assert abs(sg["err"] - so["err"]) <= 1e-9 * abs(so["err"]) + 1e-300
            else:
                assert abs(sg["it"] - so["it"]) <= max(4, 0.2 * so["it"]), (n, sg["it"], so["it"])
                assert sg["err"] <= case.tol
            assert abs(sg["max_cfl"] - so["max_cfl"]) <= (1e-12 if exact_it else 1e-6) * so["max_cfl"]
            if "torque" in so:
                assert np.allclose(sg["torque"], so["torque"], rtol=power_rtol, atol=0)
                assert np.allclose(sg["thrust"], so["thrust"], rtol=power_rtol, atol=0)
        fields = {name: (s.field(name), o.field(name)) for name in ("U", "JU", "p", "nut", "div", "f")}
    finally:
        s.close()
    return field_errors(case, fields), fields


def test_c1_small_sor():
    case = cases.case_c1(16, 12, 12, steps=6)
    errs, _ = run_pair(case, 6, exact_it=True)
    assert max(errs.values()) <= TOL, errs


def test_c1_small_sor_converging():
    c = cases.case_c1(12, 10, 10, steps=8)
    c.tol = 1e-3
    errs, _ = run_pair(c, 8, exact_it=True)
    assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize("pc", ["Jacobi", "SOR"])
def test_c2_small_bicgstab_tight(pc):
    """BiCGStab converged to 1e-15, Jacobi or SOR preconditioner, one turbine."""
    c = cases.case_c2(32, 16, 16, steps=4, ap_per_blade=8)
    c.pc, c.pc_omega, c.pc_it = pc, 1.1, (3 if pc == "Jacobi" else 2)
    c.tol, c.max_it = 1e-15, 5000
    errs, _ = run_pair(c, 4, exact_it=False)
    assert max(errs.values()) <= TOL, errs


def test_rows256_bicgstab_tight():
    """Rows of 256 interior cells (32x16x256 on a 4x2x8 box, one rotor): the FP64
    BiCGStab runs all its iteration kernels with four cells per lane there, with the
    same per-cell arithmetic, so the bar is the one-cell kernels' (solve to 1e-15) for
    every field but p.  Along 256 stretched cells the operator is worse conditioned
    than in the 32x16x16 case: two solves to 1e-15 whose dot products are summed in
    different orders leave p 2.6e-10 apart (measured on MI355X), so p is held to
    1e-9.  The kernels themselves are pinned to rounding by the first-iterations
    test in test_gpu_kernels.py."""
    m = cases.make_mesh(32, 16, 256, ((0.0, 4.0), (-1.0, 1.0), (0.0, 8.0)), 1.2, 1.2)
    c = cases.Case("rows256", m, method="BiCG", pc="Jacobi", pc_it=3, tol=1e-15, max_it=5000, steps=3,
                   turbine=cases.default_turbine(R=0.5, tower=4.0), sites=[cases.TurbineSite((1.0, 0.0, 0.0), 12.0)],
                   ap_per_blade=4, width=0.25)
    errs, _ = run_pair(c, 3, exact_it=False)
    assert errs["p"] <= 1e-9 and max(v for k, v in errs.items() if k != "p") <= TOL, errs


def test_c2_small_bicgstab_production_tolerance():
    """BiCGStab at the production tolerance 1e-8 (Jacobi-3 preconditioner).

    Both solutions stop somewhere inside the tolerance, so each is compared with
    the oracle converged to 1e-13: the GPU's distance from it may not exceed five
    times the oracle's own distance at 1e-8 (plus 1e-8).  U and the
    power are fixed much more tightly than p, whose smooth error the residual test
    hardly sees.
    """
    c = cases.case_c2(32, 16, 16, steps=4, ap_per_blade=8)
    errs, fields = run_pair(c, 4, exact_it=False, power_rtol=1e-7)
    t = cases.case_c2(32, 16, 16, steps=4, ap_per_blade=8)
    t.tol, t.max_it = 1e-13, 5000
    o = po.OracleSolver(t)
    for _ in range(4):
        o.step()
    m = interior_mask(c.mesh)
    for name in ("U", "p"):
        truth = o.field(name)[m]
        e_gpu = rel_l2(fields[name][0][m], truth)
        e_ora = rel_l2(fields[name][1][m], truth)
        print(f"{name}: gpu {e_gpu:.2e}, oracle {e_ora:.2e} from the 1e-13 solution", flush=True)
        assert e_gpu <= 5 * e_ora + 1e-8, (name, e_gpu, e_ora)
    for name in ("U", "JU", "f"):
        assert errs[name] <= 1e-6, (name, errs[name])


def test_c1_full_against_reference_trace():
    """The 64x48x48 C1 case, 100 steps, against the reference's own per-step log
    and final fields (tests/golden/trace_c1.npz)."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trace_c1.npz"))
    case = cases.case_c1(64, 48, 48, steps=100)
    print("[case c1 full 64x48x48 SOR 100 steps]", flush=True)
    s = solver.Solver(case)
    try:
        for n in range(100):
            st = s.step()
            row = d["log"][n]
            assert st["step"] == row[0] and st["it"] == row[1], (n, st["it"], row[1])
            assert abs(st["err"] - row[2]) <= 1e-10 * row[2]
            assert abs(st["avg_div"] - row[3]) <= 1e-8 * row[3]
            assert abs(st["max_cfl"] - row[4]) <= 1e-12 * row[4]
        m = interior_mask(case.mesh)
        nx = case.mesh.size[0]
        for name in ("U", "p", "JU", "nut", "div"):
            a = s.field(name)
            assert abs(np.linalg.norm(a[m]) - d[f"{name}_norm"][0]) <= TOL * d[f"{name}_norm"][0], name
            plane = a.reshape(case.mesh.size + (-1,))[nx // 2]
            assert rel_l2(plane, d[f"{name}_plane_i"]) <= TOL, name
    finally:
        s.close()


@pytest.mark.parametrize("which", ["yaw", "pitch"])
def test_c2_small_time_varying_angle(which):
    """Whole steps with a time-varying turbine angle, {amp, T, phase, offset}
    (src/alm.h:195-200, 236-244): yaw or pitch oscillating with a period of 20
    steps, so the angular rate omega_angle is comparable to the rotor speed and the
    frame transform's rate terms (quirks Q2, Q3; src/euler_angle.h:64-127) drive the
    blade forces.  BiCGStab converged to 1e-15; every field within 1e-10."""
    c = cases.case_c2(32, 16, 16, steps=4, ap_per_blade=8)
    osc = {"amp": 10.0, "T": 0.1, "phase": 30.0, "offset": 5.0}
    c.sites[0].angle = (0.0, 0.0, osc) if which == "yaw" else (0.0, osc, 0.0)
    c.tol, c.max_it = 1e-15, 5000
    errs, _ = run_pair(c, 4, exact_it=False)
    assert max(errs.values()) <= TOL, errs
