"""Per-kernel parity of the HIP library (FP64 drop-in) against the CPU oracle.

The oracle (oracle/oracle.c) is pinned against the reference build and its golden
vectors by the CPU tests.  Here the HIP kernels, called through the C ABI, must
reproduce it: bit-exact for every kernel without libm (the FP64 kernels keep the
reference's operation order and are compiled without FMA contraction), and within
the stated tolerance where the device libm (cbrt, exp, sin, cos, atan) or a
reduction order enters.
"""
import importlib
import os

import numpy as np
import pytest

from helpers import face_mask, interior_mask, max_rel, rand_field, rel_l2, small_mesh

hip = importlib.import_module("3dalm_amd.hip")
cases = importlib.import_module("3dalm_amd.cases")
po = importlib.import_module("pyoracle")

pytestmark = pytest.mark.gpu

MESHES = {
    "tiny": lambda: small_mesh(10, 9, 7),
    "ragged": lambda: small_mesh(6, 7, 70, 1.1),      # two k-tiles, ragged j
}


@pytest.fixture(params=list(MESHES))
def mesh(request):
    print(f"[mesh {request.param}]", flush=True)
    return MESHES[request.param]()


@pytest.fixture(scope="module")
def ok():
    return po.Kernels("oracle")


def ctx_with(mesh, **fields):
    c = hip.Context(mesh)
    for k, v in fields.items():
        c.upload(k, v)
    return c


def test_predictor(mesh, ok):
    Uold = rand_field(mesh, 1, 3, 0.3, 1.0)
    JU = rand_field(mesh, 2, 3, 0.05, 0.0)
    nut = np.abs(rand_field(mesh, 3, 1, 1e-3))
    f = rand_field(mesh, 4, 3, 0.2)
    U0 = rand_field(mesh, 5, 3)
    c = ctx_with(mesh, U=U0, Uold=Uold, JU=JU, nut=nut, f=f)
    c.predictor(1e4, 0.005)
    got = c.download("U")
    exp = U0.copy()
    ok.predictor(po.make_grid(mesh), exp, Uold, JU, nut, f, 1e4, 0.005)
    assert np.array_equal(got, exp), max_rel(got, exp)


def test_predictor_column_walk(ok):
    """The predictor's column walk with the LDS-staged plane (k_predictor_star) on a
    mesh large enough to select it (256 units of 16 planes): ragged along i (250
    planes), j (60 rows: a last tile of 4) and k (100 cells: a second tile of 36
    lanes), bit-exact against the oracle."""
    mesh = small_mesh(250, 60, 100, 1.2)
    print(f"[predictor column walk {mesh.size}]", flush=True)
    Uold = rand_field(mesh, 11, 3, 0.3, 1.0)
    JU = rand_field(mesh, 12, 3, 0.05, 0.0)
    nut = np.abs(rand_field(mesh, 13, 1, 1e-3))
    f = rand_field(mesh, 14, 3, 0.2)
    U0 = rand_field(mesh, 15, 3)
    c = ctx_with(mesh, U=U0, Uold=Uold, JU=JU, nut=nut, f=f)
    try:
        c.predictor(1e4, 0.005)
        got = c.download("U")
    finally:
        c.close()
    exp = U0.copy()
    ok.predictor(po.make_grid(mesh), exp, Uold, JU, nut, f, 1e4, 0.005)
    assert np.array_equal(got, exp), max_rel(got, exp)


def test_flux_rhs_bc(mesh, ok):
    U = rand_field(mesh, 7, 3, 0.4, 1.0)
    g = po.make_grid(mesh)
    c = ctx_with(mesh, U=U, JU=np.zeros((mesh.L, 3)))
    c.interpolate_JU()
    JU = np.zeros((mesh.L, 3))
    ok.interpolate_JU(g, U, JU)
    assert np.array_equal(c.download("JU"), JU)
    uin = np.array([1.0, 0.0, 0.0])
    c.apply_JUbc(uin)
    ok.apply_JUbc(g, JU, uin)
    assert np.array_equal(c.download("JU"), JU)
    c.poisson_rhs(0.005, 12.5)
    b = np.zeros(mesh.L)
    ok.poisson_rhs(g, JU, b, 0.005, 12.5)
    assert np.array_equal(c.download("b"), b)


def test_operator_and_sor(mesh, ok):
    g = po.make_grid(mesh)
    A = np.zeros((mesh.L, 7))
    md = ok.build_A(g, A)
    p0 = rand_field(mesh, 11, 1, 0.1)
    b = rand_field(mesh, 12, 1, 1.0) * interior_mask(mesh)
    c = ctx_with(mesh, p=p0, b=b)
    assert c.build_A() == md
    r_ref = np.zeros(mesh.L)
    ok.residual(g, A, p0, b, r_ref)
    r = c.residual()
    m = interior_mask(mesh)
    assert np.array_equal(r[m], r_ref[m])
    # fixed 7 iterations (tol 0): identical iterates
    x = p0.copy()
    it, err = po.C.c_int64(), po.C.c_double()
    rbuf = np.zeros(mesh.L)
    po.lib("oracle")[0].or_run_sor(po.C.byref(g), po.ptr(A), po.ptr(x), po.ptr(b), po.ptr(rbuf), 1.2,
                                   po.C.byref(it), 7, po.C.byref(err), 0.0)
    git, gerr, st = c.run_sor(1.2, 7, 0.0)
    assert git == it.value == 7 and st == hip.NONCONVERGED
    assert np.array_equal(c.download("p"), x)
    assert abs(gerr - err.value) <= 1e-13 * err.value


@pytest.mark.parametrize("pc", [hip.PC_JACOBI, hip.PC_SOR])
def test_bicgstab(mesh, ok, pc):
    g = po.make_grid(mesh)
    A = np.zeros((mesh.L, 7))
    ok.build_A(g, A)
    p0 = rand_field(mesh, 21, 1, 0.1)
    b = rand_field(mesh, 22, 1, 1.0) * interior_mask(mesh)
    c = ctx_with(mesh, p=p0, b=b)
    x = p0.copy()
    work = np.zeros(9 * mesh.L)
    it, err = po.C.c_int64(), po.C.c_double()
    # converged to 1e-13 so that the solution, not the Krylov path, is compared:
    # the dot products differ from the reference's serial loop in summation order,
    # which BiCGStab amplifies into a different iteration count
    po.lib("oracle")[0].or_run_bicgstab(po.C.byref(g), po.ptr(A), po.ptr(x), po.ptr(b), po.ptr(work), pc,
                                        1.1, 3, po.C.byref(it), 5000, po.C.byref(err), 1e-13)
    git, gerr, st = c.run_bicgstab(pc, 1.1, 3, 5000, 1e-13)
    print(f"bicgstab pc={pc}: it {git} vs {it.value}, err {gerr:.3e} vs {err.value:.3e}", flush=True)
    assert st == hip.OK and err.value <= 1e-13 and gerr <= 1e-13
    assert abs(git - it.value) <= max(4, 0.2 * it.value), (git, it.value)
    m = interior_mask(mesh)
    assert rel_l2(c.download("p")[m], x[m]) <= 1e-10


@pytest.mark.parametrize("iters", [2, 3, 8, 9])
@pytest.mark.parametrize("shape", [(5, 6, 256), (40, 20, 256), (6, 5, 384)])
def test_bicgstab_rows256_first_iterations(ok, shape, iters):
    """Rows of 256 interior cells run the FP64 BiCGStab kernels (p and s updates
    with the first Jacobi sweep, Jacobi sweeps, SpMV with its dot products, x/r
    update with its norms) with four cells per lane; with at least 8 rows the
    update and the three sweeps of each half-iteration are one kernel
    (k_half_pc4: 40x20x256 has two plane chunks, the second overlapping the first,
    and a last j tile of 4 of its 8 rows).  Rows of 384 cells run the four-cell
    kernels unfused, with a second, half-empty tile of lanes per row.  The fused
    iteration applies its x updates eight at a time: after 8 iterations in the solve,
    after 2, 3 and 9 (partly) at its end.  Over the first iterations the Krylov
    path has not yet amplified the dot products' summation order (the second
    iteration reads p and q from the partner buffers), so the iterate
    must agree with the oracle's to rounding (every stored vector is bit-identical
    to the one-cell kernels' for the same scalars)."""
    print(f"[bicgstab rows256 first iterations {shape} {iters}]", flush=True)
    mesh = small_mesh(*shape, 1.2)
    g = po.make_grid(mesh)
    A = np.zeros((mesh.L, 7))
    ok.build_A(g, A)
    p0 = rand_field(mesh, 31, 1, 0.1)
    b = rand_field(mesh, 32, 1, 1.0) * interior_mask(mesh)
    c = ctx_with(mesh, p=p0, b=b)
    x = p0.copy()
    work = np.zeros(9 * mesh.L)
    it, err = po.C.c_int64(), po.C.c_double()
    po.lib("oracle")[0].or_run_bicgstab(po.C.byref(g), po.ptr(A), po.ptr(x), po.ptr(b), po.ptr(work),
                                        hip.PC_JACOBI, 1.0, 3, po.C.byref(it), iters, po.C.byref(err), 1e-30)
    git, gerr, st = c.run_bicgstab(hip.PC_JACOBI, 1.0, 3, iters, 1e-30)
    print(f"bicgstab rows256, {iters} iterations: err {gerr:.6e} vs {err.value:.6e}", flush=True)
    assert git == it.value == iters and st == hip.NONCONVERGED
    # the summation order's rounding grows with the iterations (allowed 10x per
    # iteration past the second); a pressure update left out would be off by far more
    grow = 10.0 ** (iters - 2)
    assert abs(gerr - err.value) <= 1e-12 * grow * err.value
    m = interior_mask(mesh)
    dp = rel_l2(c.download("p")[m], x[m])
    print(f"  pressure against the oracle's: {dp:.3e}", flush=True)
    assert dp <= 1e-13 * grow
    c.close()


def test_pbc_project_ubc(mesh, ok):
    g = po.make_grid(mesh)
    p = rand_field(mesh, 31, 1, 0.2)
    U = rand_field(mesh, 32, 3, 0.3, 1.0)
    Uold = rand_field(mesh, 33, 3, 0.3, 1.0)
    JU = rand_field(mesh, 34, 3, 0.1)
    c = ctx_with(mesh, p=p, U=U, Uold=Uold, JU=JU)
    c.apply_pbc()
    ok.apply_pbc(g, p)
    assert np.array_equal(c.download("p"), p)
    c.project_p(0.005)
    ok.project(g, U, JU, p, 0.005)
    assert np.array_equal(c.download("U"), U)
    assert np.array_equal(c.download("JU"), JU)
    uin = np.array([1.0, 0.0, 0.0])
    for retracted in (0, 1):
        c.apply_Ubc(uin, 0.005, retracted, 0.7)
        ok.apply_Ubc(g, U, Uold, uin, 0.005, retracted, 0.7)
        assert np.array_equal(c.download("U"), U)


def test_post(mesh, ok):
    g = po.make_grid(mesh)
    U = rand_field(mesh, 41, 3, 0.3, 1.0)
    JU = rand_field(mesh, 42, 3, 0.1)
    c = ctx_with(mesh, U=U, JU=JU)
    c.calc_nut(0.1)
    nut = np.zeros(mesh.L)
    ok.nut(g, U, nut, 0.1)
    m = interior_mask(mesh)
    # cbrt is the only libm call: device and glibc differ by at most an ulp or two
    assert max_rel(c.download("nut")[m], nut[m]) <= 1e-14
    c.calc_divergence()
    div = np.zeros(mesh.L)
    ok.divergence(g, JU, div)
    assert np.array_equal(c.download("div")[m], div[m])
    assert abs(c.l2_norm("div") - ok.l2_norm(g, div)) <= 1e-14 * ok.l2_norm(g, div)
    assert c.max_cfl(0.005) == ok.max_cfl(g, U, 0.005)


def test_post_column_walk(ok):
    """nu_t through the post pass's column walk (k_post_star: 4-row blocks of the LDS
    walk, selected on this mesh) against the oracle, ragged along i, j and k."""
    mesh = small_mesh(250, 60, 100, 1.2)
    print(f"[post column walk {mesh.size}]", flush=True)
    g = po.make_grid(mesh)
    U = rand_field(mesh, 43, 3, 0.3, 1.0)
    JU = rand_field(mesh, 44, 3, 0.1)
    c = ctx_with(mesh, U=U, JU=JU)
    try:
        c.calc_nut(0.1)
        got = c.download("nut")
    finally:
        c.close()
    nut = np.zeros(mesh.L)
    ok.nut(g, U, nut, 0.1)
    m = interior_mask(mesh)
    assert max_rel(got[m], nut[m]) <= 1e-14


def test_alm(ok):
    case = cases.case_c2(32, 16, 16, steps=1, ap_per_blade=8)
    case.sites[0].angle = (0.0, 0.0, 7.0)   # yawed rotor
    mesh = case.mesh
    g = po.make_grid(mesh)
    U = rand_field(mesh, 51, 3, 0.1, 1.0)
    U[:, 1:] *= 0.2
    ap, cd, cl, atk = po.ap_props_oracle(case.turbine, 1, case.ap_per_blade)
    hap, hcd, hcl, _ = hip.build_ap_props(case.turbine, 1, case.ap_per_blade)
    assert np.array_equal(ap, hap) and np.array_equal(cd, hcd) and np.array_equal(cl, hcl)
    wt = cases.turbine_records(case.turbine, case.sites)
    c = ctx_with(mesh, U=U, f=np.zeros((mesh.L, 3)))
    c.alm_setup(wt, ap, 3, case.ap_per_blade, cd, cl, atk, case.width)
    t = 0.37
    tq, th = c.run_alm(t)
    owt, oap = wt.copy(), ap.copy()
    ok.update_ap_position(owt, 1, oap, len(oap), 3, case.ap_per_blade, t)
    ok.calc_ap_force(g, U, owt, 1, oap, len(oap), 3, case.ap_per_blade, cd, cl, atk, len(atk))
    f = np.zeros((mesh.L, 3))
    ok.distribute_ap_force(g, f, oap, len(oap), case.width)
    gap, gwt = c.alm_download()
    assert max_rel(gap[:, :3], oap[:, :3]) <= 1e-14
    assert max_rel(gap[:, 8:13], oap[:, 8:13]) <= 1e-12
    assert abs(tq[0] - owt[0, 14]) <= 1e-12 * abs(owt[0, 14])
    assert abs(th[0] - owt[0, 15]) <= 1e-12 * abs(owt[0, 15])
    assert rel_l2(c.download("f"), f) <= 1e-12



GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALM_BRANCHES = ("none", "roll", "pitch", "yaw", "yaw_osc", "pitch_osc", "roll_osc")


@pytest.mark.parametrize("name", ALM_BRANCHES)
def test_alm_angle_branches(name, ok):
    """Every angle branch of update_ap_position and calc_ap_force
    (src/alm.h:233-414, src/euler_angle.h:34-127) on the GPU against the reference's
    own outputs in tests/golden/alm_tiny.npz (written by the reference build): no
    angle, static roll (frame_transform_U's roll algebra, quirk Q1), pitch (the
    pitch sign, Q3), yaw, and the time-varying {amp, T, phase, offset} formula on
    yaw (omega_angle != 0: the yaw-rate term, Q2), pitch and roll, at t = 0.37.
    Positions, angle and angular rate, forces, torque, thrust and the projected
    force field within 1e-12 (device sin / cos / atan / exp against glibc)."""
    d = np.load(os.path.join(GOLDEN, "alm_tiny.npz"))
    mesh = cases.Mesh(d["x"], d["y"], d["z"])
    n = d["ap0"].shape[0]
    per_blade = n // 3
    wt0 = d[f"{name}_wt0"]
    print(f"[alm branch {name}: angle type {int(wt0[0, 6])}, formula {wt0[0, 9:13]}]", flush=True)
    c = ctx_with(mesh, U=d["U"], f=np.zeros((mesh.L, 3)))
    try:
        c.alm_setup(wt0, d["ap0"], 3, per_blade, d["cd"], d["cl"], d["atk"], float(d["width"][0]))
        tq, th = c.run_alm(0.37)
        gap, gwt = c.alm_download()
        ref_ap, ref_wt = d[f"{name}_ap"], d[f"{name}_wt"]
        assert max_rel(gap[:, :8], d[f"{name}_pos"]) <= 1e-13
        assert max_rel(gap[:, 8:13], ref_ap[:, 8:13]) <= 1e-12
        assert np.allclose(gwt[0, 7:9], ref_wt[0, 7:9], rtol=1e-13, atol=1e-15), (gwt[0, 7:9], ref_wt[0, 7:9])
        assert abs(tq[0] - ref_wt[0, 14]) <= 1e-12 * abs(ref_wt[0, 14])
        assert abs(th[0] - ref_wt[0, 15]) <= 1e-12 * abs(ref_wt[0, 15])
        if f"{name}_f" in d.files:
            fexp = d[f"{name}_f"]
        else:  # the oracle's projection (pinned to the reference) of the reference's forces
            fexp = np.zeros((mesh.L, 3))
            ok.distribute_ap_force(po.make_grid(mesh), fexp, ref_ap, n, float(d["width"][0]))
        assert rel_l2(c.download("f"), fexp) <= 1e-12
    finally:
        c.close()
