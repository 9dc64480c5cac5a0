"""The turbulent inflow on the HIP path (dalm_set_inflow, k_inflow and the x- boundary
of k_ubc / k_flux / k_jubc / k_project and their precision-form counterparts)
against the oracle's or_inflow_eval, which tests/test_inflow.py pins by closed forms.

  * the boundary values themselves: dalm_eval_inflow + dalm_apply_Ubc /
    dalm_apply_JUbc against or_inflow_eval at three times, to 1e-14 (the device's
    cos against glibc's);
  * whole FP64 steps of a two-rotor case with the inflow, BiCGStab converged to
    1e-15 so that the Krylov path does not matter: every field within 1e-10
    relative L2, turbine torque and thrust within 1e-10;
  * the precision form with the inflow against the FP64 oracle at the stated
    tolerances (err <= tol on the true residual, avg div <= 1e-6, power within
    0.1 %, U within 1e-4)."""
import ctypes as C
import importlib

import numpy as np
import pytest

from helpers import interior_mask, rel_l2

hip = importlib.import_module("3dalm_amd.hip")
cases = importlib.import_module("3dalm_amd.cases")
solver = importlib.import_module("3dalm_amd.solver")
po = importlib.import_module("pyoracle")

pytestmark = pytest.mark.gpu


def _case(tol=1e-15, steps=3):
    """48x24x24 on a 12x3x3 box, two rotors 5 D apart, 24 Fourier modes at 10 % as
    the initial field and the inflow."""
    m = cases.make_mesh(48, 24, 24, ((0.0, 12.0), (-1.5, 1.5), (0.0, 3.0)), 1.3, 1.3)
    t = cases.default_turbine(R=0.5, tower=1.5)
    sites = [cases.TurbineSite((2.0, 0.0, 0.0), 12.0), cases.TurbineSite((7.0, 0.0, 0.0), 12.0)]
    return cases.Case("inflow2", m, method="BiCG", pc="Jacobi", pc_it=3, tol=tol, max_it=20000, steps=steps,
                      noise=0.10, modes=24, inflow=True, turbine=t, sites=sites, ap_per_blade=8,
                      width=2 * 12.0 / 48)


def _oracle_bc(case, shift):
    L, _ = po.lib("oracle")
    m = case.mesh
    g = m.gc
    ny, nz = m.size[1] - 2 * g, m.size[2] - 2 * g
    cells, face = np.zeros(g * ny * nz * 3), np.zeros(ny * nz)
    modes = np.ascontiguousarray(case.inflow_modes())
    uin = np.array(case.uin, dtype=np.float64)
    grid = po.make_grid(m)
    L.or_inflow_eval(C.byref(grid), po.ptr(modes), len(modes), po.ptr(uin), shift, po.ptr(cells), po.ptr(face))
    return cells.reshape(g, ny, nz, 3), face.reshape(ny, nz)


def test_inflow_boundary_values():
    case = _case()
    m = case.mesh
    g = m.gc
    nx, ny, nz = m.size
    print(f"[inflow boundary values {m.interior}, {len(case.inflow_modes())} modes]", flush=True)
    ctx = hip.Context(m)
    try:
        ctx.upload("U", case.U0())
        ctx.set_params(solver.params_for(case))
        ctx.set_inflow(case.inflow_modes(), case.inflow_speed())
        for t in (0.0, 0.37, 2.5):
            ctx.eval_inflow(t)
            ctx.apply_Ubc(case.uin, case.dt)
            ctx.interpolate_JU()
            ctx.apply_JUbc(case.uin)
            U = ctx.download("U").reshape(nx, ny, nz, 3)
            JU = ctx.download("JU").reshape(nx, ny, nz, 3)
            cells, face = _oracle_bc(case, case.inflow_speed() * t)
            ayz = m.dy[g:ny - g, None] * m.dz[None, g:nz - g]
            e1 = np.max(np.abs(U[:g, g:ny - g, g:nz - g] - cells))
            e2 = np.max(np.abs(JU[g - 1, g:ny - g, g:nz - g, 0] - ayz * face) / ayz)
            print(f"  t={t}: ghost cells {e1:.2e}, face {e2:.2e}", flush=True)
            assert e1 <= 1e-14 and e2 <= 1e-14
        # n_modes = 0 restores the constant inflow
        ctx.set_inflow(np.zeros((0, 7)), 0.0)
        ctx.apply_Ubc(case.uin, case.dt)
        U = ctx.download("U").reshape(nx, ny, nz, 3)
        assert np.all(U[:g, g:ny - g, g:nz - g] == np.asarray(case.uin))
    finally:
        ctx.close()


def test_inflow_whole_steps_fp64():
    case = _case(tol=1e-15)
    print(f"[inflow whole steps FP64 {case.mesh.interior}]", flush=True)
    o = po.OracleSolver(case)
    s = solver.Solver(case)
    try:
        for n in range(case.steps):
            so, sg = o.step(), s.step()
            print(f"  step {n + 1}: it {sg['it']} vs {so['it']}, err {sg['err']:.3e}", flush=True)
            assert sg["err"] <= case.tol
            assert np.allclose(sg["torque"], so["torque"], rtol=1e-10, atol=0)
            assert np.allclose(sg["thrust"], so["thrust"], rtol=1e-10, atol=0)
        m = interior_mask(case.mesh)
        for name in ("U", "JU", "f"):
            e = rel_l2(s.field(name), o.field(name))
            print(f"  {name}: {e:.2e}", flush=True)
            assert e <= 1e-10, (name, e)
        e = rel_l2(s.field("p")[m], o.field("p")[m])
        print(f"  p: {e:.2e}", flush=True)
        assert e <= 1e-10
    finally:
        s.close()


def test_inflow_precision_form():
    case = _case(tol=1e-8, steps=4)
    print(f"[inflow precision form {case.mesh.interior}]", flush=True)
    o = po.OracleSolver(case)
    s = solver.Solver(case, hip.MIXED)
    try:
        for n in range(case.steps):
            so, sm = o.step(), s.step()
            print(f"  step {n + 1}: it {sm['it']} vs {so['it']}, err {sm['err']:.3e}, div {sm['avg_div']:.2e}",
                  flush=True)
            assert sm["status"] == hip.OK, sm
            assert sm["err"] <= case.tol and sm["avg_div"] <= 1e-6
            p64 = np.array([st.rot_speed for st in case.sites]) * so["torque"]
            assert np.all(np.abs(sm["power"] - p64) <= 1e-3 * np.abs(p64))
        m = interior_mask(case.mesh)
        eu = rel_l2(s.field("U")[m], o.field("U")[m])
        print(f"  U vs FP64 oracle: {eu:.2e}", flush=True)
        assert eu <= 1e-4
    finally:
        s.close()
