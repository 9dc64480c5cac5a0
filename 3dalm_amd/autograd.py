"""torch.autograd.Function over the HIP forward run and its discrete adjoint.

``TurbinePower.apply(yaw, pitch, chord, twist, case, steps, every)`` runs ``steps``
FP64 steps of ``case`` on the GPU with the turbines' yaw angles (degrees, angle
type forced to yaw), collective blade pitch (degrees) and blade chord / twist at
the airfoil stations of each turbine (absolute chord, twist in degrees: the
windturbine.json stations, src/alm.h:39-103) set from the tensors, and returns
the time-averaged power

    J = (1/steps) sum_{n=1..steps} sum_w Omega_w torque_w(n)

(torque = sum over the turbine's actuator points of |F_t| r, src/alm.h:362,
384-393).  ``backward`` replays the run from checkpoints every ``every`` steps and
runs the reverse sweep (dalm_adjoint_run), giving dJ/d(every input).  The station
values enter through the same linear interpolation onto the actuator points as
build_ap_props (src/alm.h:66-95), whose weights are taken from the library's own
dalm_build_ap_props applied to unit vectors.

A thin validation harness: tensors are float64 on the CPU, one evaluation runs a
whole simulation, and the work itself happens in the HIP library.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import hip
from .cases import Case, turbine_records
from .solver import params_for

DEG = math.pi / 180


def station_weights(turbine, ap_per_blade):
    """W [blades * ap_per_blade, n_stations] with chord_ap = W @ chord_stations (and
    the same for twist), from dalm_build_ap_props on unit station vectors."""
    ns = len(turbine.stations_r)
    W = np.zeros((turbine.blades * ap_per_blade, ns))
    for s in range(ns):
        e = np.zeros(ns)
        e[s] = 1.0
        t = dataclasses.replace(turbine, stations_chord=e)
        ap, _, _, _ = hip.build_ap_props(t, 1, ap_per_blade)
        W[:, s] = ap[:, 4]
    return W


class PowerRun:
    """One recorded forward run of a case with given turbine parameters."""

    def __init__(self, case: Case, yaw_deg, pitch_deg, chord, twist, U0=None):
        nwt = len(case.sites)
        t = case.turbine
        self.case, self.nwt = case, nwt
        self.per_wt = t.blades * case.ap_per_blade
        self.W = station_weights(t, case.ap_per_blade)
        ap, cd, cl, atk = hip.build_ap_props(t, nwt, case.ap_per_blade)
        for w in range(nwt):
            sl = slice(w * self.per_wt, (w + 1) * self.per_wt)
            ap[sl, 4] = self.W @ np.asarray(chord[w], dtype=np.float64)
            ap[sl, 5] = self.W @ np.asarray(twist[w], dtype=np.float64)
            ap[sl, 13] = float(pitch_deg[w])
        wt = turbine_records(t, case.sites)
        for w in range(nwt):
            wt[w, 6] = 2                      # yaw
            wt[w, 9:13] = (0.0, 0.0, 0.0, float(yaw_deg[w]) * DEG)
        self.ctx = hip.Context(case.mesh, hip.FP64)
        L = case.mesh.L
        Uold = np.empty((L, 3))
        Uold[:] = case.uin
        self.ctx.upload("U", case.U0() if U0 is None else U0)
        self.ctx.upload("Uold", Uold)
        self.ctx.set_params(params_for(case))
        self.ctx.alm_setup(wt, ap, t.blades, case.ap_per_blade, cd, cl, atk, case.width)
        self.ctx.initialize()
        self.omega = np.array([s.rot_speed for s in case.sites])

    def forward(self, steps, every, keep_pressure=True):
        self.steps = steps
        self.ctx.adjoint_begin(steps, every, self.omega / steps, keep_pressure)
        J = 0.0
        for _ in range(steps):
            st = self.ctx.step()
            J += float(np.sum(self.omega / steps * st["torque"]))
        return J

    def backward(self, tol_rel=1e-10):
        J, ap_g, wt_g, stats = self.ctx.adjoint_run(tol_rel)
        self.stats = stats
        d_yaw = wt_g[:, 0] * DEG
        d_pitch = wt_g[:, 1]
        d_chord = np.stack([self.W.T @ ap_g[w * self.per_wt:(w + 1) * self.per_wt, 2] for w in range(self.nwt)])
        d_twist = np.stack([self.W.T @ ap_g[w * self.per_wt:(w + 1) * self.per_wt, 3] for w in range(self.nwt)])
        return J, d_yaw, d_pitch, d_chord, d_twist

    def close(self):
        self.ctx.close()


class TurbinePower(torch.autograd.Function):
    @staticmethod
    def forward(ctx, yaw, pitch, chord, twist, case, steps, every):
        run = PowerRun(case, yaw.detach().cpu().numpy(), pitch.detach().cpu().numpy(),
                       chord.detach().cpu().numpy(), twist.detach().cpu().numpy())
        J = run.forward(steps, every)
        ctx.run = run
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        run = ctx.run
        try:
            _, d_yaw, d_pitch, d_chord, d_twist = run.backward()
        finally:
            run.close()
        g = float(gJ)
        as_t = lambda a: torch.tensor(g * a, dtype=torch.float64)
        return as_t(d_yaw), as_t(d_pitch), as_t(d_chord), as_t(d_twist), None, None, None


def power(yaw, pitch, chord, twist, case: Case, steps: int, every: int):
    """J(yaw, pitch, chord, twist) as a differentiable torch scalar."""
    return TurbinePower.apply(yaw, pitch, chord, twist, case, steps, every)


def power_value(case: Case, yaw, pitch, chord, twist, steps: int) -> float:
    """J without recording (for finite differences)."""
    run = PowerRun(case, yaw, pitch, chord, twist)
    try:
        J = 0.0
        for _ in range(steps):
            st = run.ctx.step()
            J += float(np.sum(run.omega / steps * st["torque"]))
        return J
    finally:
        run.close()
