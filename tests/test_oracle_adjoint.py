"""The oracle's discrete adjoint (oracle/adjoint.c) pinned by central differences of
the oracle's forward routines.

Per kernel: the dot-product identity <y_b, K'(x) x_d> = <K'(x)^T y_b, x_d>, with
K'(x) x_d taken by central differences of the forward routine (exact up to rounding
for the linear kernels).  Whole run: d(time-averaged power)/d(yaw, pitch, chord,
twist) and the sensitivities to the initial state against central differences with
h = 1e-6 in FP64.

Directions avoid the one place where the forward is not differentiable by
construction: the x- inflow ghosts hold the same value in both layers, so the MUSCL
limiter of the first interior cell sits on its kink there (a plain perturbation of
those ghosts is a one-sided question; the ghosts are reset to the inflow every step,
so no parameter moves them).
"""
import importlib

import numpy as np
import pytest

from helpers import face_mask, interior_mask, rand_field, small_mesh

cases = importlib.import_module("3dalm_amd.cases")
po = importlib.import_module("pyoracle")


@pytest.fixture(scope="module")
def K():
    return po.Kernels("oracle")


@pytest.fixture(scope="module")
def mesh():
    return small_mesh(9, 8, 7, 1.3)


def vdot(a, b):
    return float(np.sum(np.asarray(a) * np.asarray(b)))


def check(lhs, rhs, rtol):
    assert abs(lhs - rhs) <= rtol * max(abs(lhs), abs(rhs)), (lhs, rhs)


def not_inflow_ghost(mesh):
    nx, ny, nz = mesh.size
    m = np.ones((nx, ny, nz), dtype=bool)
    m[:mesh.gc] = False
    return m.reshape(-1)


def test_predictor_adjoint(K, mesh):
    g = po.make_grid(mesh)
    L = mesh.L
    Uold = rand_field(mesh, 1, 3, 0.3, 1.0)
    JU = rand_field(mesh, 2, 3, 0.05, 0.02)
    nut = rand_field(mesh, 3, 1, 1e-3, 2e-3)
    f = rand_field(mesh, 4, 3, 0.1)
    keep = not_inflow_ghost(mesh)
    dU = rand_field(mesh, 5, 3, 1.0) * keep[:, None]
    dJ = rand_field(mesh, 6, 3, 1.0)
    dn = rand_field(mesh, 7, 1, 1e-3)
    df = rand_field(mesh, 8, 3, 1.0)
    Ub = rand_field(mesh, 9, 3, 1.0) * interior_mask(mesh)[:, None]
    Re, dt, h = 1e3, 0.01, 1e-7

    def fwd(s):
        U = np.zeros((L, 3))
        K.predictor(g, U, Uold + s * dU, JU + s * dJ, nut + s * dn, f + s * df, Re, dt)
        return U
    lhs = vdot(Ub, (fwd(h) - fwd(-h)) / (2 * h))
    bU, bJ, bn, bf = np.zeros((L, 3)), np.zeros((L, 3)), np.zeros(L), np.zeros((L, 3))
    K.predictor_adj(g, Ub, Uold, JU, nut, Re, dt, bU, bJ, bn, bf)
    rhs = vdot(bU, dU) + vdot(bJ, dJ) + vdot(bn, dn) + vdot(bf, df)
    check(lhs, rhs, 1e-6)
    # f and nut separately, so that no term can hide behind another (U* is affine
    # in each, so a unit step is exact)
    h = 1.0
    for name, d, b in (("f", df, bf), ("nut", dn, bn)):
        def one(s):
            U = np.zeros((L, 3))
            args = dict(f=f, nut=nut)
            args[name] = args[name] + s * d
            K.predictor(g, U, Uold, JU, args["nut"], args["f"], Re, dt)
            return U
        check(vdot(Ub, (one(h) - one(-h)) / (2 * h)), vdot(b, d), 1e-11)


def test_flux_rhs_adjoint(K, mesh):
    """interpolate_JU + apply_JUbc + poisson_rhs are linear in U."""
    g = po.make_grid(mesh)
    L = mesh.L
    dU = rand_field(mesh, 11, 3, 1.0)
    bb = rand_field(mesh, 12, 1, 1.0) * interior_mask(mesh)
    dt, scale = 0.01, 37.0

    def fwd(U):
        JU = np.zeros((L, 3))
        K.interpolate_JU(g, U, JU)
        K.apply_JUbc(g, JU, np.zeros(3))
        b = np.zeros(L)
        K.poisson_rhs(g, JU, b, dt, scale)
        return b
    JUb = np.zeros((L, 3))
    K.poisson_rhs_adj(g, bb, JUb, dt, scale)
    K.apply_JUbc_adj(g, JUb)
    Ub = np.zeros((L, 3))
    K.interpolate_JU_adj(g, JUb, Ub)
    check(vdot(bb, fwd(dU)), vdot(Ub, dU), 1e-12)


def test_pbc_project_adjoint(K, mesh):
    """p -> pbc -> (U, JU) projection is linear in p."""
    g = po.make_grid(mesh)
    L = mesh.L
    dp = rand_field(mesh, 21, 1, 1.0) * interior_mask(mesh)
    Ub = rand_field(mesh, 22, 3, 1.0) * interior_mask(mesh)[:, None]
    JUb = rand_field(mesh, 23, 3, 1.0) * face_mask(mesh)[:, None]
    dt = 0.02
    p = dp.copy()
    K.apply_pbc(g, p)
    U, JU = np.zeros((L, 3)), np.zeros((L, 3))
    K.project(g, U, JU, p, dt)
    lhs = vdot(Ub, U) + vdot(JUb, JU)
    pb = np.zeros(L)
    K.project_adj(g, Ub, JUb, pb, dt)
    K.apply_pbc_adj(g, pb)
    check(lhs, vdot(pb, dp), 1e-12)
    # the ghosts apply_pbc writes (one layer, interior-facing) are folded away;
    # edge ghosts are never written forward, so their bar is simply unused
    cross = interior_mask(mesh, grow=1) & ~interior_mask(mesh)
    nx, ny, nz = mesh.size
    I, J, Kk = [a.reshape(-1) for a in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")]
    g0, n = mesh.gc, mesh.size
    inside = [(A >= g0) & (A < nn - g0) for A, nn in zip((I, J, Kk), n)]
    one_face = cross & ((inside[0] & inside[1]) | (inside[0] & inside[2]) | (inside[1] & inside[2]))
    assert np.all(pb[one_face] == 0)


@pytest.mark.parametrize("retracted", [0, 1])
def test_ubc_adjoint(K, mesh, retracted):
    g = po.make_grid(mesh)
    L = mesh.L
    Upre = rand_field(mesh, 31, 3, 0.2, 1.0)
    Uold = rand_field(mesh, 32, 3, 0.2, 1.0)
    dUp = rand_field(mesh, 33, 3, 1.0)
    dUo = rand_field(mesh, 34, 3, 1.0)
    Ub = rand_field(mesh, 35, 3, 1.0)
    dt, h, fsx = 0.03, 1e-7, float(mesh.x[4])

    def fwd(s):
        U = Upre + s * dUp
        K.apply_Ubc(g, U, Uold + s * dUo, np.array([1.0, 0.1, -0.2]), dt, retracted, fsx)
        return U
    lhs = vdot(Ub, (fwd(h) - fwd(-h)) / (2 * h))
    bo = np.zeros((L, 3))
    bu = Ub.copy()
    K.apply_Ubc_adj(g, bu, Upre, Uold, bo, dt, retracted, fsx)
    check(lhs, vdot(bu, dUp) + vdot(bo, dUo), 1e-7)


def test_solid_adjoint(K, mesh):
    g = po.make_grid(mesh)
    solid = (rand_field(mesh, 41, 1, 1.0) > 0.3).astype(np.float64)
    dU = rand_field(mesh, 42, 3, 1.0)
    Ub = rand_field(mesh, 43, 3, 1.0) * interior_mask(mesh)[:, None]
    U = dU.copy()
    K.set_solid(g, U, solid)
    b = Ub.copy()
    K.set_solid_adj(g, b, solid)
    check(vdot(Ub, U * interior_mask(mesh)[:, None]), vdot(b, dU), 1e-13)


def test_nut_adjoint(K, mesh):
    g = po.make_grid(mesh)
    L = mesh.L
    U = rand_field(mesh, 51, 3, 0.3, 1.0)
    dU = rand_field(mesh, 52, 3, 1.0)
    nb = rand_field(mesh, 53, 1, 1.0) * interior_mask(mesh)
    h, Cs = 1e-7, 0.13

    def fwd(s):
        n = np.zeros(L)
        K.nut(g, U + s * dU, n, Cs)
        return n
    Ub = np.zeros((L, 3))
    K.nut_adj(g, U, nb, Ub, Cs)
    check(vdot(nb, (fwd(h) - fwd(-h)) / (2 * h)), vdot(Ub, dU), 1e-6)


def test_poisson_adjoint_transpose(K, mesh):
    """b_b = A^-T p_b: <p_b, A^-1 b> = <b_b, b> with both solves converged."""
    g = po.make_grid(mesh)
    L = mesh.L
    A = np.zeros((L, 7))
    K.build_A(g, A)
    m = interior_mask(mesh)
    b = rand_field(mesh, 61, 1, 1.0) * m
    pb = rand_field(mesh, 62, 1, 1.0) * m
    x = np.zeros(L)
    work = np.zeros(9 * L)
    it, err = po.C.c_int64(), po.C.c_double()
    po.lib("oracle")[0].or_run_bicgstab(po.C.byref(g), po.ptr(A), po.ptr(x), po.ptr(b), po.ptr(work), 1, 1.2,
                                        2, po.C.byref(it), 20000, po.C.byref(err), 1e-14)
    y, bb, work10 = np.zeros(L), np.zeros(L), np.zeros(10 * L)
    K.poisson_adj(g, A, pb, y, bb, work10, 1e-14, 20000, po.C.byref(it), po.C.byref(err))
    check(vdot(pb, x), vdot(bb, b), 1e-10)


def _alm_setup(U_seed=71):
    case = cases.case_adjoint_small(14, 12, 12, ap_per_blade=5)
    m = case.mesh
    o = po.OracleSolver(case)
    U = rand_field(m, U_seed, 3, 0.1, 0.0)
    U[:, 0] += 1.0
    return case, m, o, U


def _alm_forward(case, m, o, U, ap, wt, t):
    g = po.make_grid(m)
    K = po.Kernels("oracle")
    nb, napb = case.turbine.blades, case.ap_per_blade
    K.update_ap_position(wt, len(wt), ap, len(ap), nb, napb, t)
    K.calc_ap_force(g, U, wt, len(wt), ap, len(ap), nb, napb, o.cd, o.cl, o.atk, len(o.atk))
    f = np.zeros((m.L, 3))
    K.distribute_ap_force(g, f, ap, len(ap), case.width)
    return f, wt[:, 14].copy()


def test_alm_adjoint_velocity_and_parameters(K):
    case, m, o, U = _alm_setup()
    g = po.make_grid(m)
    t, h = 0.037, 1e-7
    fb = rand_field(m, 72, 3, 1.0) * interior_mask(m)[:, None]
    tq_w = np.array([0.7])
    ap0, wt0 = o.ap.copy(), o.wt.copy()

    def objective(Uv, ap, wt):
        f, tq = _alm_forward(case, m, o, Uv, ap.copy(), wt.copy(), t)
        return vdot(fb, f) + vdot(tq_w, tq)
    Ub = np.zeros((m.L, 3))
    grad = np.zeros((len(ap0), 4))
    ap, wt = ap0.copy(), wt0.copy()
    K.alm_adj(g, U, wt, len(wt), ap, len(ap), case.turbine.blades, case.ap_per_blade, o.cd, o.cl, o.atk,
              len(o.atk), case.width, t, fb, tq_w, Ub, grad)
    dU = rand_field(m, 73, 3, 1.0)
    lhs = (objective(U + h * dU, ap0, wt0) - objective(U - h * dU, ap0, wt0)) / (2 * h)
    check(lhs, vdot(Ub, dU), 1e-6)
    # turbine angle, collective pitch, one chord and one twist
    for col, setter in ((0, lambda ap, wt, s: wt.__setitem__((0, 12), wt[0, 12] + s)),
                        (1, lambda ap, wt, s: ap.__setitem__((slice(None), 13), ap[:, 13] + s)),
                        (2, lambda ap, wt, s: ap.__setitem__((3, 4), ap[3, 4] + s)),
                        (3, lambda ap, wt, s: ap.__setitem__((7, 5), ap[7, 5] + s))):
        ap_p, wt_p, ap_m, wt_m = ap0.copy(), wt0.copy(), ap0.copy(), wt0.copy()
        setter(ap_p, wt_p, h)
        setter(ap_m, wt_m, -h)
        fd = (objective(U, ap_p, wt_p) - objective(U, ap_m, wt_m)) / (2 * h)
        adj = grad[:, col].sum() if col < 2 else grad[3 if col == 2 else 7, col]
        check(fd, adj, 1e-6)


def _J(case, hook, steps, tq_w):
    o = po.OracleSolver(case, ap_hook=hook)
    J = 0.0
    for _ in range(steps):
        J += vdot(tq_w, o.step()["torque"])
    return J


def test_whole_run_parameter_gradients():
    """d(time-averaged power)/d(yaw, pitch, chord, twist) against central FD."""
    steps = 4
    case = cases.case_adjoint_small(steps=steps)
    tq_w = np.array([case.sites[0].rot_speed / steps])
    J, grad, _ = po.OracleSolver(case).adjoint(steps, tq_w)
    assert abs(J - _J(case, None, steps, tq_w)) <= 1e-14 * abs(J)
    h = 1e-6

    def fd(setter):
        def hk(s):
            return lambda ap, wt: setter(ap, wt, s)
        return (_J(case, hk(h), steps, tq_w) - _J(case, hk(-h), steps, tq_w)) / (2 * h)
    check(fd(lambda ap, wt, s: wt.__setitem__((0, 12), wt[0, 12] + s)), grad[:, 0].sum(), 1e-6)
    check(fd(lambda ap, wt, s: ap.__setitem__((slice(None), 13), ap[:, 13] + s)), grad[:, 1].sum(), 1e-6)
    for a in (2, 9, 16):
        check(fd(lambda ap, wt, s, a=a: ap.__setitem__((a, 4), ap[a, 4] + s)), grad[a, 2], 1e-6)
    # twist sensitivities are small next to their FD noise floor (~1e-11)
    for a in (9, 16):
        assert abs(fd(lambda ap, wt, s, a=a: ap.__setitem__((a, 5), ap[a, 5] + s)) - grad[a, 3]) <= \
            1e-6 * abs(grad[a, 3]) + 1e-10


def test_whole_run_initial_state_sensitivities():
    """The fluid part of the reverse sweep: bars of the initial U, JU, nut, f (taken
    as independent inputs after the initialisation pass) against central FD."""
    steps = 3
    case = cases.case_adjoint_small(steps=steps)
    tq_w = np.array([case.sites[0].rot_speed / steps])
    _, _, bars = po.OracleSolver(case).adjoint(steps, tq_w)
    m = case.mesh
    rng = np.random.default_rng(81)
    keep = not_inflow_ghost(m)
    for name, h, rtol in (("U", 1e-6, 1e-6), ("JU", 1e-8, 1e-5), ("nut", 1e-6, 1e-6), ("f", 1e-6, 1e-6)):
        d = rng.uniform(-1, 1, size=bars[name].shape)
        if name == "U":
            d *= keep[:, None]

        def J_of(s):
            o = po.OracleSolver(case)
            getattr(o, name)[...] += s * d
            return sum(vdot(tq_w, o.step()["torque"]) for _ in range(steps))
        check((J_of(h) - J_of(-h)) / (2 * h), vdot(bars[name], d), rtol)
