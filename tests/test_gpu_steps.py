"""Whole-step parity of the FP64 drop-in (Solver::main_loop_once order) against the
CPU oracle on identical seeded inputs.

Bar: per-step iteration counts equal, per-field relative L2 <= 1e-10 (BASELINE
north star for the FP64 drop-in).  SOR paths are expected to be bit-exact apart
from cbrt in nu_t; BiCGStab paths differ through reduction order only.
"""
import importlib

import numpy as np
import pytest

from helpers import interior_mask, rel_l2

hip = importlib.import_module("3dalm_amd.hip")
cases = importlib.import_module("3dalm_amd.cases")
solver = importlib.import_module("3dalm_amd.solver")
po = importlib.import_module("pyoracle")

pytestmark = pytest.mark.gpu
TOL = 1e-10


def compare(case, steps, fields=("U", "JU", "p", "nut", "div", "f")):
    print(f"[case {case.name} {case.mesh.interior} {case.method} steps={steps}]", flush=True)
    o = po.OracleSolver(case)
    s = solver.Solver(case)
    try:
        for n in range(steps):
            so = o.step()
            sg = s.step()
            assert sg["step"] == so["step"] == n + 1
            assert sg["it"] == so["it"], (n, sg["it"], so["it"])
            assert abs(sg["err"] - so["err"]) <= 1e-9 * abs(so["err"]) + 1e-300
            assert abs(sg["avg_div"] - so["avg_div"]) <= 1e-8 * abs(so["avg_div"]) + 1e-300
            assert abs(sg["max_cfl"] - so["max_cfl"]) <= 1e-12 * so["max_cfl"]
            if "torque" in so:
                assert np.allclose(sg["torque"], so["torque"], rtol=TOL, atol=0)
                assert np.allclose(sg["thrust"], so["thrust"], rtol=TOL, atol=0)
        m = interior_mask(case.mesh)
        for name in fields:
            a, b = s.field(name), o.field(name)
            if name in ("p", "nut", "div"):
                a, b = a[m], b[m]
            assert rel_l2(a, b) <= TOL, (name, rel_l2(a, b))
    finally:
        s.close()


def test_c1_small_sor():
    compare(cases.case_c1(16, 12, 12, steps=6), 6)


def test_c1_small_sor_converging():
    c = cases.case_c1(12, 10, 10, steps=8)
    c.tol = 1e-3
    compare(c, 8)


def test_c2_small_bicgstab_alm():
    compare(cases.case_c2(32, 16, 16, steps=4, ap_per_blade=8), 4)


def test_c2_small_bicgstab_sor_pc():
    c = cases.case_c2(24, 12, 12, steps=3, ap_per_blade=6)
    c.pc, c.pc_omega, c.pc_it = "SOR", 1.1, 2
    compare(c, 3)


def test_c1_full_against_reference_trace():
    """The 64x48x48 C1 case, 100 steps, against the reference's own per-step log
    and final fields (tests/golden/trace_c1.npz)."""
    import os
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
