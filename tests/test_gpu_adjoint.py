"""The differentiation form on the GPU: the HIP discrete-adjoint kernels against the
oracle's adjoint (oracle/adjoint.c, itself pinned by central differences in
test_oracle_adjoint.py), and whole recorded runs against the oracle's reverse sweep
and against central differences (h = 1e-6) of the HIP forward.

Kernel bars are gathers on the GPU and scatters in the oracle, so they agree to
rounding (1e-12 of the largest value); the actuator-point corner bars use FP64
atomics.  Whole runs converge every Poisson solve to 1e-13, forward and adjoint.
"""
import importlib

import numpy as np
import pytest

from helpers import face_mask, interior_mask, rand_field, small_mesh

hip = importlib.import_module("3dalm_amd.hip")
cases = importlib.import_module("3dalm_amd.cases")
solver = importlib.import_module("3dalm_amd.solver")
po = importlib.import_module("pyoracle")

pytestmark = pytest.mark.gpu

MESHES = {
    "tiny": lambda: small_mesh(10, 9, 7),
    "ragged": lambda: small_mesh(6, 7, 70, 1.1),
}


@pytest.fixture(params=list(MESHES))
def mesh(request):
    print(f"[adjoint mesh {request.param}]", flush=True)
    return MESHES[request.param]()


@pytest.fixture(scope="module")
def ok():
    return po.Kernels("oracle")


def ctx_with(mesh, **fields):
    c = hip.Context(mesh)
    for k, v in fields.items():
        c.upload(k, v)
    return c


def close(a, b, rtol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b)) / scale
    assert err <= rtol, err
    return err


def test_predictor_adj(mesh, ok):
    g = po.make_grid(mesh)
    L = mesh.L
    Uold = rand_field(mesh, 1, 3, 0.3, 1.0)
    JU = rand_field(mesh, 2, 3, 0.05, 0.02)
    nut = rand_field(mesh, 3, 1, 1e-3, 2e-3)
    Ub = rand_field(mesh, 9, 3, 1.0) * interior_mask(mesh)[:, None]
    Re, dt = 1e3, 0.01
    bU, bJ, bn, bf = np.zeros((L, 3)), np.zeros((L, 3)), np.zeros(L), np.zeros((L, 3))
    ok.predictor_adj(g, Ub, Uold, JU, nut, Re, dt, bU, bJ, bn, bf)
    c = ctx_with(mesh, Uold=Uold, JU=JU, nut=nut, U_b=Ub)
    c.predictor_adj(Re, dt)
    close(c.download("Uold_b"), bU)
    close(c.download("JUold_b"), bJ)
    close(c.download("nutold_b"), bn)
    close(c.download("fold_b"), bf)


def test_rhs_jubc_interp_adj(mesh, ok):
    g = po.make_grid(mesh)
    L = mesh.L
    bb = rand_field(mesh, 12, 1, 1.0) * interior_mask(mesh)
    JUb0 = rand_field(mesh, 13, 3, 1.0) * face_mask(mesh)[:, None]
    dt, scale = 0.01, 37.0
    JUb = JUb0.copy()
    ok.poisson_rhs_adj(g, bb, JUb, dt, scale)
    ok.apply_JUbc_adj(g, JUb)
    Ub = np.zeros((L, 3))
    ok.interpolate_JU_adj(g, JUb, Ub)
    c = ctx_with(mesh, b_b=bb, JU_b=JUb0, U_b=np.zeros((L, 3)))
    c.poisson_rhs_adj(dt, scale)
    c.apply_JUbc_adj()
    close(c.download("JU_b"), JUb)
    c.interpolate_JU_adj()
    close(c.download("U_b"), Ub)


def test_project_pbc_adj(mesh, ok):
    g = po.make_grid(mesh)
    L = mesh.L
    Ub = rand_field(mesh, 22, 3, 1.0) * interior_mask(mesh)[:, None]
    JUb = rand_field(mesh, 23, 3, 1.0) * face_mask(mesh)[:, None]
    dt = 0.02
    pb = np.zeros(L)
    ok.project_adj(g, Ub, JUb, pb, dt)
    ok.apply_pbc_adj(g, pb)
    c = ctx_with(mesh, U_b=Ub, JU_b=JUb, p_b=np.zeros(L))
    c.project_p_adj(dt)
    c.apply_pbc_adj()
    close(c.download("p_b"), pb)


@pytest.mark.parametrize("retracted", [0, 1])
def test_ubc_solid_adj(mesh, ok, retracted):
    g = po.make_grid(mesh)
    L = mesh.L
    Upre = rand_field(mesh, 31, 3, 0.2, 1.0)
    Uold = rand_field(mesh, 32, 3, 0.2, 1.0)
    Ub = rand_field(mesh, 35, 3, 1.0)
    solid = (rand_field(mesh, 36, 1, 1.0) > 0.5).astype(np.float64)
    dt, fsx = 0.03, float(mesh.x[4])
    bu, bo = Ub.copy(), np.zeros((L, 3))
    ok.apply_Ubc_adj(g, bu, Upre, Uold, bo, dt, retracted, fsx)
    c = ctx_with(mesh, U=Upre, Uold=Uold, U_b=Ub, Uold_b=np.zeros((L, 3)), solid=solid)
    c.apply_Ubc_adj(dt, retracted, fsx)
    close(c.download("U_b"), bu)
    close(c.download("Uold_b"), bo)
    ok.set_solid_adj(g, bu, solid)
    c.set_solid_U_adj()
    close(c.download("U_b"), bu)


def test_nut_adj(mesh, ok):
    g = po.make_grid(mesh)
    L = mesh.L
    U = rand_field(mesh, 51, 3, 0.3, 1.0)
    nb = rand_field(mesh, 53, 1, 1.0) * interior_mask(mesh)
    Ub0 = rand_field(mesh, 54, 3, 1.0)
    Ub = Ub0.copy()
    ok.nut_adj(g, U, nb, Ub, 0.13)
    c = ctx_with(mesh, U=U, nut_b=nb, U_b=Ub0)
    c.calc_nut_adj(0.13)
    close(c.download("U_b"), Ub)


@pytest.mark.parametrize("method,pc", [("BiCG", "Jacobi"), ("BiCG", "SOR"), ("SOR", "Jacobi")])
def test_poisson_adj(mesh, ok, method, pc):
    """b_b = A^-T p_b against the oracle's, and the transpose identity
    <p_b, A^-1 b> = <b_b, b> with the GPU's own forward solve.  Both solves stop
    on a residual test (1e-13 relative), which leaves an error of up to cond(A)
    times that in the solution: far below 1e-9 for BiCGStab here, but plain SOR
    reduces the smooth error slowly and ends ~1e-6 from the converged answer."""
    g = po.make_grid(mesh)
    L = mesh.L
    A = np.zeros((L, 7))
    ok.build_A(g, A)
    m = interior_mask(mesh)
    pb = rand_field(mesh, 62, 1, 1.0) * m
    y, bb, work = np.zeros(L), np.zeros(L), np.zeros(10 * L)
    it, err = po.C.c_int64(), po.C.c_double()
    ok.poisson_adj(g, A, pb, y, bb, work, 1e-13, 20000, po.C.byref(it), po.C.byref(err))
    case = cases.case_c1(6, 5, 5)
    case.mesh = mesh
    case.method, case.pc, case.pc_it, case.pc_omega, case.omega = method, pc, 2, 1.1, 1.4
    c = ctx_with(mesh, p_b=pb)
    c.set_params(solver.params_for(case))
    its, _ = c.poisson_adj(1e-13, 20000)
    print(f"adjoint solve {method}/{pc}: {its} iterations (oracle {it.value})", flush=True)
    tol = 1e-9 if method == "BiCG" else 2e-5
    close(c.download("b_b")[m], bb[m], tol)
    b = rand_field(mesh, 61, 1, 1.0) * m
    c.upload("b", b)
    c.upload("p", np.zeros(L))
    if method == "BiCG":
        c.run_bicgstab(hip.PC_SOR, 1.2, 2, 20000, 1e-13)
    else:
        c.run_sor(1.4, 20000, 1e-13)
    x = c.download("p")
    lhs, rhs = float(np.sum(pb[m] * x[m])), float(np.sum(c.download("b_b")[m] * b[m]))
    assert abs(lhs - rhs) <= tol * abs(lhs), (lhs, rhs)


def test_alm_adj(ok):
    case = cases.case_adjoint_small(14, 12, 12, ap_per_blade=5)
    m = case.mesh
    g = po.make_grid(m)
    o = po.OracleSolver(case)
    U = rand_field(m, 71, 3, 0.1, 0.0)
    U[:, 0] += 1.0
    fb = rand_field(m, 72, 3, 1.0) * interior_mask(m)[:, None]
    tq_w = np.array([0.7])
    t = 0.037
    Ub = np.zeros((m.L, 3))
    grad = np.zeros((len(o.ap), 4))
    ap, wt = o.ap.copy(), o.wt.copy()
    ok.alm_adj(g, U, wt, len(wt), ap, len(ap), case.turbine.blades, case.ap_per_blade, o.cd, o.cl, o.atk,
               len(o.atk), case.width, t, fb, tq_w, Ub, grad)
    s = solver.Solver(case)
    try:
        s.ctx.upload("U", U)
        s.ctx.upload("f_b", fb)
        s.ctx.upload("U_b", np.zeros((m.L, 3)))
        g_gpu = s.ctx.run_alm_adj(t, tq_w)
        # the GPU projection skips pairs beyond 8 eps (weight < 1.6e-28)
        close(s.ctx.download("U_b"), Ub, 1e-10)
        close(g_gpu, grad, 1e-10)
    finally:
        s.close()


def _oracle_adjoint(case, steps, tq_w):
    J, g, _ = po.OracleSolver(case).adjoint(steps, tq_w, 1e-13)
    return J, g


def _gpu_adjoint(case, steps, every, tq_w, keep_pressure=True, hook=None):
    s = solver.Solver(case, ap_hook=hook)
    try:
        s.ctx.adjoint_begin(steps, every, tq_w, keep_pressure)
        J = 0.0
        for _ in range(steps):
            J += float(np.sum(tq_w * s.step()["torque"]))
        Ja, ap_g, wt_g, st = s.ctx.adjoint_run(1e-13)
        assert abs(Ja - J) <= 1e-13 * abs(J)
        U_end = s.field("U")
        return J, ap_g, wt_g, st, U_end
    finally:
        s.close()


def test_adjoint_run_matches_oracle():
    steps = 4
    case = cases.case_adjoint_small(steps=steps)
    tq_w = np.array([case.sites[0].rot_speed / steps])
    print("[adjoint run: case_adjoint_small 20x12x12, 4 steps]", flush=True)
    J_o, g_o = _oracle_adjoint(case, steps, tq_w)
    J, ap_g, wt_g, st, _ = _gpu_adjoint(case, steps, 2, tq_w)
    print("adjoint stats", st, flush=True)
    assert abs(J - J_o) <= 1e-9 * abs(J_o), (J, J_o)
    for col in range(4):
        close(ap_g[:, col], g_o[:, col], 1e-6)
    assert np.allclose(wt_g[0], [ap_g[:, 0].sum(), ap_g[:, 1].sum()], rtol=1e-14, atol=0)


def test_adjoint_checkpointing_is_exact():
    """The reverse sweep does not depend on the checkpoint interval or on whether
    the pressure is kept; the forward state is restored afterwards."""
    steps = 5
    case = cases.case_adjoint_small(steps=steps)
    tq_w = np.array([case.sites[0].rot_speed / steps])
    ref = _gpu_adjoint(case, steps, steps, tq_w, True)
    for every, keep in ((1, True), (2, False), (3, True)):
        print(f"[checkpoint every {every}, keep pressure {keep}]", flush=True)
        r = _gpu_adjoint(case, steps, every, tq_w, keep)
        assert r[0] == ref[0]
        close(r[1], ref[1], 1e-10)
        assert np.array_equal(r[4], ref[4])


def _J_gpu(case, steps, tq_w, hook):
    s = solver.Solver(case, ap_hook=hook)
    try:
        return sum(float(np.sum(tq_w * s.step()["torque"])) for _ in range(steps))
    finally:
        s.close()


def test_adjoint_against_finite_differences():
    """dJ/d(yaw, pitch, chord, twist) against central differences of the HIP
    forward with h = 1e-6: the 1e-4 relative bar."""
    steps = 4
    case = cases.case_adjoint_small(steps=steps)
    tq_w = np.array([case.sites[0].rot_speed / steps])
    _, ap_g, wt_g, _, _ = _gpu_adjoint(case, steps, 2, tq_w)
    h = 1e-6

    def fd(setter):
        def hk(s):
            return lambda ap, wt: setter(ap, wt, s)
        return (_J_gpu(case, steps, tq_w, hk(h)) - _J_gpu(case, steps, tq_w, hk(-h))) / (2 * h)
    checks = [("yaw", fd(lambda ap, wt, s: wt.__setitem__((0, 12), wt[0, 12] + s)), wt_g[0, 0]),
              ("pitch", fd(lambda ap, wt, s: ap.__setitem__((slice(None), 13), ap[:, 13] + s)), wt_g[0, 1])]
    for a in (3, 10):
        checks.append((f"chord[{a}]", fd(lambda ap, wt, s, a=a: ap.__setitem__((a, 4), ap[a, 4] + s)), ap_g[a, 2]))
        checks.append((f"twist[{a}]", fd(lambda ap, wt, s, a=a: ap.__setitem__((a, 5), ap[a, 5] + s)), ap_g[a, 3]))
    for name, f, g in checks:
        print(f"{name}: fd {f:.10e} adjoint {g:.10e}", flush=True)
        assert abs(f - g) <= 1e-4 * abs(g) + 1e-10, (name, f, g)


def test_autograd_harness():
    """The torch.autograd.Function: gradients of the time-averaged power with
    respect to per-turbine yaw, pitch and station chord / twist against central
    differences of the same function."""
    torch = pytest.importorskip("torch")
    ag = importlib.import_module("3dalm_amd.autograd")
    steps = 3
    case = cases.case_adjoint_small(steps=steps)
    t = case.turbine
    yaw = torch.tensor([10.0], dtype=torch.float64, requires_grad=True)
    pitch = torch.tensor([2.0], dtype=torch.float64, requires_grad=True)
    chord = torch.tensor(np.array([t.stations_chord]), dtype=torch.float64, requires_grad=True)
    twist = torch.tensor(np.array([t.stations_twist]), dtype=torch.float64, requires_grad=True)
    J = ag.power(yaw, pitch, chord, twist, case, steps, 2)
    J.backward()
    base = dict(yaw=yaw.detach().numpy().copy(), pitch=pitch.detach().numpy().copy(),
                chord=chord.detach().numpy().copy(), twist=twist.detach().numpy().copy())
    h = 1e-6

    def Jv(**kw):
        p = {k: v.copy() for k, v in base.items()}
        for k, (idx, dv) in kw.items():
            p[k][idx] += dv
        return ag.power_value(case, p["yaw"], p["pitch"], p["chord"], p["twist"], steps)
    assert abs(Jv() - J.item()) <= 1e-13 * abs(J.item())
    for name, idx, grad in (("yaw", (0,), yaw.grad[0]), ("pitch", (0,), pitch.grad[0]),
                            ("chord", (0, 9), chord.grad[0, 9]), ("twist", (0, 12), twist.grad[0, 12])):
        f = (Jv(**{name: (idx, h)}) - Jv(**{name: (idx, -h)})) / (2 * h)
        print(f"autograd {name}{idx}: fd {f:.10e} adjoint {float(grad):.10e}", flush=True)
        assert abs(f - float(grad)) <= 1e-4 * abs(float(grad)) + 1e-10, (name, f, float(grad))


def test_recording_through_breakdowns():
    """A BiCGStab breakdown (|rho| < FLT_MIN at the first inner product) ends the
    step's solve and the reference carries on (src/eq.h:349-353); such a step must be
    recorded like any other.  Blades of chord 1e-30 turning in fluid at rest (zero
    inflow, no noise: the walls, the inflow and the outflow all hold U = 0 exactly)
    give a force field of order 1e-30, so every step's right-hand side has
    <r0, r> ~ 1e-60 and breaks down at once, while dJ/d(chord) stays of order 0.1.
    The reverse sweep over three such steps must match the oracle's.  A recording
    that holds a non-finite step is refused."""
    steps = 3
    case = cases.case_adjoint_small(steps=steps)
    case.noise = 0.0
    case.uin = (0.0, 0.0, 0.0)
    case.turbine.stations_chord = np.full_like(case.turbine.stations_chord, 1e-30)
    tq_w = np.array([case.sites[0].rot_speed / steps])
    print("[adjoint run through breakdowns: 20x12x12, chord 1e-30, fluid at rest]", flush=True)
    J_o, g_o = _oracle_adjoint(case, steps, tq_w)
    s = solver.Solver(case)
    try:
        s.ctx.adjoint_begin(steps, 2, tq_w, False)
        for _ in range(steps):
            st = s.step()
            print(f"  step {st['step']}: status {st['status']} it {st['it']}", flush=True)
            assert st["status"] == hip.BREAKDOWN and st["it"] == 0, st
        J, ap_g, wt_g, _ = s.ctx.adjoint_run(1e-13)
    finally:
        s.close()
    assert abs(J - J_o) <= 1e-12 * abs(J_o), (J, J_o)
    assert np.max(np.abs(g_o[:, 2])) > 1e-3  # the chord sensitivity is of order one
    close(ap_g[:, 2], g_o[:, 2], 1e-12)
    # a recording with a non-finite step cannot be differentiated
    s = solver.Solver(case)
    try:
        s.ctx.adjoint_begin(2, 1, tq_w, False)
        s.step()
        U = s.field("U")
        n0, n1, n2 = case.mesh.size
        U[((n0 // 2) * n1 + n1 // 2) * n2 + n2 // 2, 0] = np.nan  # an interior cell at the box centre
        s.ctx.upload("U", U)
        st = s.step()
        assert st["status"] == hip.NONFINITE, st
        with pytest.raises(Exception):
            s.ctx.adjoint_run(1e-13)
    finally:
        s.close()


def test_c4_adjoint_against_finite_differences():
    """The differentiation-form geometry at its full size: 192x96x96 cells, three
    rotors in tandem 5 D apart, 3 x 3 x 64 actuator points, yaw 5 / 0 / -5 degrees.
    A recorded window of three steps (checkpoints every two) with every forward
    solve converged to 1e-12 and the adjoint solves to 1e-13 relative; dJ/d(yaw)
    and dJ/d(collective pitch) of each turbine and dJ/d(chord), dJ/d(twist) at two
    actuator points against central differences of the HIP forward with h = 1e-6,
    to 1e-4 relative.

    A solve stopped by its tolerance makes J(theta) a step function: the two
    evaluations of a central difference take different Krylov paths.  At this size
    a residual of 1e-12 still leaves J uncertain by ~1e-10 (measured: J at pitch
    offsets of 1e-9 scatters by 3e-10), which is 1e-3 of a difference quotient with
    h = 1e-6.  The same holds 60 iterations past that residual (651 in all).  From
    about 800 iterations the BiCGStab iterate sits at its rounding floor and J is a
    smooth function of theta (the offsets of 1e-9 give J linear to 1e-16, and the
    quotients for h = 1e-5 and 1e-6 agree with the adjoint to 1e-7).  The
    difference quotients therefore run every solve for a fixed 800 iterations (at
    least 250 past the recorded run's largest count), tolerance 0."""
    steps = 3
    case = cases.case_c4(steps=steps)
    case.tol, case.max_it = 1e-12, 20000
    omega = np.array([st.rot_speed for st in case.sites])
    tq_w = omega / steps
    per_wt = case.turbine.blades * case.ap_per_blade
    print(f"[adjoint at full size: {case.mesh.interior}, {len(case.sites)} rotors, {steps} steps]", flush=True)
    s = solver.Solver(case)
    its = []
    try:
        s.ctx.adjoint_begin(steps, 2, tq_w, True)
        J = 0.0
        for _ in range(steps):
            st = s.step()
            print(f"  step {st['step']}: it {st['it']} err {st['err']:.3e}", flush=True)
            assert st["status"] == hip.OK, st
            its.append(st["it"])
            J += float(np.sum(tq_w * st["torque"]))
        Ja, ap_g, wt_g, stats = s.ctx.adjoint_run(1e-13)
        print(f"  J {Ja:.12e}; adjoint stats {stats}", flush=True)
        assert abs(Ja - J) <= 1e-13 * abs(J)
    finally:
        s.close()
    h = 1e-6
    fixed = cases.case_c4(steps=steps)
    fixed.tol, fixed.max_it = 0.0, max(800, max(its) + 250)

    def fd(setter):
        def hk(sg):
            return lambda ap, wt: setter(ap, wt, sg)
        return (_J_gpu(fixed, steps, tq_w, hk(h)) - _J_gpu(fixed, steps, tq_w, hk(-h))) / (2 * h)

    checks = []
    for w in range(3):
        checks.append((f"yaw[{w}]", fd(lambda ap, wt, sg, w=w: wt.__setitem__((w, 12), wt[w, 12] + sg)), wt_g[w, 0]))
        sl = slice(w * per_wt, (w + 1) * per_wt)
        checks.append((f"pitch[{w}]", fd(lambda ap, wt, sg, sl=sl: ap.__setitem__((sl, 13), ap[sl, 13] + sg)),
                       wt_g[w, 1]))
    for a in (40, 2 * per_wt + 100):
        checks.append((f"chord[{a}]", fd(lambda ap, wt, sg, a=a: ap.__setitem__((a, 4), ap[a, 4] + sg)), ap_g[a, 2]))
        checks.append((f"twist[{a}]", fd(lambda ap, wt, sg, a=a: ap.__setitem__((a, 5), ap[a, 5] + sg)), ap_g[a, 3]))
    bad = []
    for name, f, g in checks:
        print(f"  {name}: fd {f:.10e} adjoint {g:.10e} rel {abs(f - g) / abs(g):.2e}", flush=True)
        if not abs(f - g) <= 1e-4 * abs(g) + 1e-10:
            bad.append((name, f, g))
    assert not bad, bad


def test_precision_form_adjoint_against_fp64():
    """The precision form's recording (dalm_adjoint_begin_from): a DALM_MIXED run of
    four steps records widened checkpoints (every two steps) and its pressures into
    an FP64 context, whose reverse sweep replays the segments in FP64 and runs the
    FP64 reverse steps.  Against the FP64 form's own recording of the same case:
    J within 1e-6 relative (the precision form's torques), dJ/d(yaw, pitch) per
    turbine and dJ/d(chord, twist) per actuator point within 1e-4 relative of the
    largest of each kind.  The recording context cannot step meanwhile."""
    steps = 4
    case = cases.case_adjoint_small(steps=steps)
    case.tol = 1e-12
    tq_w = np.array([case.sites[0].rot_speed / steps])
    print("[precision-form adjoint: case_adjoint_small 20x12x12, 4 steps]", flush=True)
    J64, ap64, wt64, _, _ = _gpu_adjoint(case, steps, 2, tq_w)
    sm = solver.Solver(case, hip.MIXED)
    s64 = solver.Solver(case)
    try:
        s64.ctx.adjoint_begin_from(sm.ctx, steps, 2, tq_w, True)
        with pytest.raises(Exception):
            s64.step()  # the FP64 context holds the recording
        Jm = 0.0
        for _ in range(steps):
            st = sm.step()
            assert st["status"] == hip.OK, st
            Jm += float(np.sum(tq_w * st["torque"]))
        J, ap_g, wt_g, stats = s64.ctx.adjoint_run(1e-13)
        print(f"  J mixed {J:.12e} (steps {Jm:.12e}) fp64 {J64:.12e}; stats {stats}", flush=True)
        assert abs(J - Jm) <= 1e-13 * abs(Jm)
        assert abs(J - J64) <= 1e-6 * abs(J64)
        for name, a, b in (("yaw/pitch", wt_g, wt64), ("chord", ap_g[:, 2], ap64[:, 2]),
                           ("twist", ap_g[:, 3], ap64[:, 3])):
            err = np.max(np.abs(a - b)) / np.max(np.abs(b))
            print(f"  {name}: max deviation {err:.2e} of the largest", flush=True)
            assert err <= 1e-4, (name, err)
    finally:
        sm.close()
        s64.close()
