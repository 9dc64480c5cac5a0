"""Host-side mirror of the reference's Solver (src/solver.h:519-1310) on the HIP path.

``Solver(case)`` does what ``Solver::initialize`` does after parsing the case
(allocation, operator, initial fields, ALM set-up, start-up pass) and
``main_loop_once`` advances one time step on the GPU; ``can_continue`` follows
``step <= max_step`` (src/solver.h:1308-1310).  Every numerical call goes through
the C ABI (include/dalm_hip.h); there is no CPU fallback.
"""
from __future__ import annotations

import numpy as np

from . import hip
from .cases import Case, turbine_records


def params_for(case: Case) -> hip.Params:
    p = hip.Params()
    p.Uin[:] = case.uin
    p.Re, p.Cs, p.dt = case.Re, case.Cs, case.dt
    p.method = hip.BICGSTAB if case.method == "BiCG" else hip.SOR
    p.pc_kind = hip.PC_SOR if case.pc == "SOR" else hip.PC_JACOBI
    p.omega, p.pc_omega, p.pc_it = case.omega, case.pc_omega, case.pc_it
    p.max_it, p.tol = case.max_it, case.tol
    p.retracted_floor = 0 if case.floor_start is None else 1
    p.floor_start_x = 0.0 if case.floor_start is None else case.floor_start
    return p


class Solver:
    """``ap_hook(ap, wt)``, when given, may change the actuator-point and turbine
    records (chord, twist, pitch, angle formula) before the start-up pass.  With
    ``output_dir`` the reference's output path runs after the start-up pass and
    after every step (snapshots, time average, probes, on the case's schedule;
    3dalm_amd/output.py), as main_loop_once does (src/solver.h:1193-1250)."""

    def __init__(self, case: Case, precision: int = hip.FP64, U0: np.ndarray | None = None, ap_hook=None,
                 output_dir: str | None = None):
        self.case = case
        self.ctx = hip.Context(case.mesh, precision)
        L = case.mesh.L
        U = case.U0() if U0 is None else U0
        Uold = np.empty((L, 3))
        Uold[:] = case.uin
        self.ctx.upload("U", U)
        self.ctx.upload("Uold", Uold)
        self.ctx.set_params(params_for(case))
        if case.tgg is not None:  # src/solver.h:740-750
            t = case.tgg
            self.ctx.set_turbulence_grid(t["thick"], t["mesh"], t["bar"], t["placement"])
        modes = case.inflow_modes()
        if modes is not None and len(modes):  # turbulent inflow (not in the reference)
            self.ctx.set_inflow(modes, case.inflow_speed())
        self.ap = None
        if case.turbine is not None and case.sites:
            nwt = len(case.sites)
            ap, cd, cl, atk = hip.build_ap_props(case.turbine, nwt, case.ap_per_blade)
            per_wt = case.turbine.blades * case.ap_per_blade
            for a in range(len(ap)):
                ap[a, 13] = case.sites[a // per_wt].pitch
            wt = turbine_records(case.turbine, case.sites)
            if ap_hook is not None:
                ap_hook(ap, wt)
            self.ctx.alm_setup(wt, ap, case.turbine.blades, case.ap_per_blade, cd, cl, atk, case.width)
            self.ap = ap
            alm_tables = (ap, cd, cl, atk)
        else:
            alm_tables = None
        self.output = None
        if output_dir is not None:
            from .output import OutputWriter
            self.output = OutputWriter(self.ctx, case, output_dir, alm_tables)
        self.stats = self.ctx.initialize()
        self.max_step = case.steps
        if self.output:
            self.output.after_step(0)

    def main_loop_once(self):
        self.stats = self.ctx.step()
        if self.output:
            self.output.after_step(int(self.stats["step"]))
        return self.stats

    step = main_loop_once

    def can_continue(self):
        return self.stats["step"] < self.max_step

    def field(self, name):
        return self.ctx.download(name)

    def close(self):
        if self.output:
            self.output.close()
            self.output = None
        self.ctx.close()
