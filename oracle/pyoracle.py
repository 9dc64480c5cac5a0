"""ctypes bindings for the test oracle (oracle.c) and the reference build (_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.

Both libraries expose the same per-kernel C ABI (oracle.h); ``Kernels('oracle')``
and ``Kernels('ref')`` therefore share one Python surface.  ``OracleSolver`` and
``RefSolver`` run whole steps (initialisation pass + Solver::main_loop_once order).
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
ORACLE_OMP_SO = os.path.join(HERE, "_build", "liboracle_omp.so")
REF_SO = os.path.join(HERE, "_ref", "libref3dalm.so")

i64 = C.c_int64
dbl = C.c_double
P = C.POINTER(C.c_double)
WT_STRIDE = 16
AP_STRIDE = 16


class OrGrid(C.Structure):
    _fields_ = [("n", i64 * 3), ("gc", i64), ("c", P * 3), ("d", P * 3)]


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(P)


def make_grid(mesh):
    g = OrGrid()
    for a in range(3):
        g.n[a] = mesh.size[a]
    g.gc = mesh.gc
    keep = [mesh.x, mesh.y, mesh.z, mesh.dx, mesh.dy, mesh.dz]
    for a in range(3):
        g.c[a] = ptr(keep[a])
        g.d[a] = ptr(keep[3 + a])
    g._keep = keep
    return g


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run: make -C oracle [ref])")
    return C.CDLL(path)


_libs = {}


def lib(kind="oracle"):
    if kind not in _libs:
        L = _load({"oracle": ORACLE_SO, "omp": ORACLE_OMP_SO, "ref": REF_SO}[kind])
        pre = "ref_" if kind == "ref" else "or_"
        G = C.POINTER(OrGrid)
        sig = {
            "predictor": (None, [G, P, P, P, P, P, dbl, dbl]),
            "interpolate_JU": (None, [G, P, P]),
            "apply_JUbc": (None, [G, P, P]),
            "poisson_rhs": (None, [G, P, P, dbl, dbl]),
            "build_A": (dbl, [G, P]),
            "dot": (dbl, [G, P, P]),
            "l2_norm": (dbl, [G, P]),
            "residual": (None, [G, P, P, P, P]),
            "Ax": (None, [G, P, P, P]),
            "sor_sweep": (None, [G, P, P, P, dbl, i64]),
            "jacobi_sweep": (None, [G, P, P, P, P]),
            "run_sor": (None, [G, P, P, P, P, dbl, C.POINTER(i64), i64, P, dbl]),
            "run_bicgstab": (None, [G, P, P, P, P, i64, dbl, i64, C.POINTER(i64), i64, P, dbl]),
            "apply_pbc": (None, [G, P]),
            "project": (None, [G, P, P, P, dbl]),
            "set_solid": (None, [G, P, P]),
            "apply_Ubc": (None, [G, P, P, P, dbl, i64, dbl]),
            "nut": (None, [G, P, P, dbl]),
            "divergence": (None, [G, P, P]),
            "max_cfl": (dbl, [G, P, dbl]),
            "calc_q": (None, [G, P, P]),
            "tgg_solid": (None, [G, P, dbl, dbl, dbl, P]),
            "time_average": (None, [G, P, P, dbl]),
            "update_ap_position": (None, [P, i64, P, i64, i64, i64, dbl]),
            "calc_ap_force": (None, [G, P, P, i64, P, i64, i64, i64, P, P, P, i64]),
            "distribute_ap_force": (None, [G, P, P, i64, dbl]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, pre + name)
            f.restype = res
            f.argtypes = args
        if kind in ("oracle", "omp"):
            L.or_mesh_widths.argtypes = [P, i64, P]
            L.or_build_ap_props.argtypes = [P, P, P, P, P, i64, i64, dbl, dbl, i64, i64, i64, P, P, P]
            L.or_state_initialize.argtypes = [G, C.c_void_p]
            L.or_state_step.argtypes = [G, C.c_void_p]
            L.or_inflow_eval.argtypes = [G, P, i64, P, dbl, P, P]
            L.or_apply_inflow_U.argtypes = [G, P, P]
            L.or_apply_inflow_JU.argtypes = [G, P, P]
        if kind == "oracle":
            adj = {
                "predictor_adj": [G, P, P, P, P, dbl, dbl, P, P, P, P],
                "interpolate_JU_adj": [G, P, P],
                "apply_JUbc_adj": [G, P],
                "poisson_rhs_adj": [G, P, P, dbl, dbl],
                "apply_pbc_adj": [G, P],
                "project_adj": [G, P, P, P, dbl],
                "set_solid_adj": [G, P, P],
                "apply_Ubc_adj": [G, P, P, P, P, dbl, i64, dbl],
                "nut_adj": [G, P, P, P, dbl],
                "alm_adj": [G, P, P, i64, P, i64, i64, i64, P, P, P, i64, dbl, dbl, P, P, P, P],
                "poisson_adj": [G, P, P, P, P, P, dbl, i64, C.POINTER(i64), P],
            }
            for name, args in adj.items():
                f = getattr(L, "or_" + name)
                f.restype = None
                f.argtypes = args
            L.or_adjoint_run.argtypes = [G, C.c_void_p, i64, P, dbl, P, P, P]
        elif kind == "ref":
            L.ref_build_mesh.argtypes = [C.c_char_p, P, P, P, P, P, P, C.POINTER(i64), C.POINTER(i64)]
            L.ref_build_ap_props.restype = i64
            L.ref_build_ap_props.argtypes = [C.c_char_p, i64, i64, P, P, P, P, C.POINTER(i64)]
            L.ref_build_wt_props.restype = i64
            L.ref_build_wt_props.argtypes = [C.c_char_p, C.c_char_p, P]
            L.ref_solver_create.restype = C.c_void_p
            L.ref_solver_create.argtypes = [C.c_char_p, P, dbl, dbl, dbl, i64, i64, dbl, dbl, i64, i64,
                                            dbl, C.c_char_p, C.c_char_p, i64, dbl]
            L.ref_solver_perturb.argtypes = [C.c_void_p, P]
            L.ref_solver_output.argtypes = [C.c_void_p, C.c_char_p, dbl, dbl, dbl, P, i64]
            L.ref_solver_tgg.argtypes = [C.c_void_p, dbl, dbl, dbl, P]
            L.ref_solver_floor.argtypes = [C.c_void_p, dbl]
            L.ref_solver_initialize.argtypes = [C.c_void_p]
            L.ref_solver_step.argtypes = [C.c_void_p]
            L.ref_solver_get.restype = C.c_int
            L.ref_solver_get.argtypes = [C.c_void_p, C.c_char_p, P]
            L.ref_solver_stats.argtypes = [C.c_void_p, C.POINTER(i64), C.POINTER(i64), P, P, P]
            L.ref_solver_destroy.argtypes = [C.c_void_p]
        _libs[kind] = (L, pre)
    return _libs[kind]


def ref_available():
    return os.path.exists(REF_SO)


class Kernels:
    """Per-kernel calls into the oracle ('oracle') or the reference build ('ref')."""

    def __init__(self, kind="oracle"):
        self.kind = kind
        self.L, self.pre = lib(kind)

    def f(self, name):
        return getattr(self.L, self.pre + name)

    def __getattr__(self, name):
        fn = self.f(name)

        def call(*args):
            conv = []
            for a in args:
                if isinstance(a, np.ndarray):
                    conv.append(ptr(a))
                elif isinstance(a, OrGrid):
                    conv.append(C.byref(a))
                else:
                    conv.append(a)
            return fn(*conv)
        return call


# ----------------------------------------------------------------------------- ALM props
def ap_props_oracle(turbine, wt_count, ap_per_blade):
    """AP records and Cd/Cl rows from a TurbineModel (oracle restatement)."""
    L, _ = lib("oracle")
    ns, na = len(turbine.stations_r), len(turbine.atk)
    n = wt_count * turbine.blades * ap_per_blade
    ap = np.zeros((n, AP_STRIDE))
    cd = np.zeros((n, na))
    cl = np.zeros((n, na))
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in
            (turbine.stations_r, turbine.stations_chord, turbine.stations_twist,
             np.asarray(turbine.cd).reshape(ns, na), np.asarray(turbine.cl).reshape(ns, na))]
    L.or_build_ap_props(*[ptr(a) for a in arrs], ns, na, turbine.R, turbine.hub_R, turbine.blades,
                        wt_count, ap_per_blade, ptr(ap), ptr(cd), ptr(cl))
    return ap, cd, cl, np.ascontiguousarray(turbine.atk, dtype=np.float64)


def wt_records(turbine, sites):
    from importlib import import_module
    return import_module("3dalm_amd.cases").turbine_records(turbine, sites)


# ----------------------------------------------------------------------------- solvers
class OrState(C.Structure):
    _fields_ = [(n, P) for n in ("U", "Uold", "JU", "p", "nut", "div", "f", "solid", "A", "b", "work")] + [
        ("Uin", dbl * 3), ("Re", dbl), ("Cs", dbl), ("dt", dbl), ("method", i64), ("pc_kind", i64),
        ("omega", dbl), ("pc_omega", dbl), ("pc_it", i64), ("max_it", i64), ("tol", dbl),
        ("retracted_floor", i64), ("floor_start_x", dbl),
        ("wt", P), ("ap", P), ("cd_tbl", P), ("cl_tbl", P), ("atk", P),
        ("wt_count", i64), ("ap_count", i64), ("blade_per_wt", i64), ("ap_per_blade", i64),
        ("atk_count", i64), ("width", dbl),
        ("max_diag", dbl), ("step", i64), ("it", i64), ("err", dbl), ("avg_div", dbl), ("max_cfl", dbl),
        ("inflow_modes", P), ("inflow_n", i64), ("inflow_speed", dbl), ("inflow_cells", P), ("inflow_face", P)]


class OracleSolver:
    """Whole-step CPU oracle on a Case (cases.Case).  kind "omp" runs the OpenMP
    build (the CPU baseline); with initialize=False the start-up pass is left to
    load_state (only the operator is built)."""

    def __init__(self, case, ap_hook=None, kind="oracle", U0=None, initialize=True):
        self.case = case
        m = case.mesh
        self.g = make_grid(m)
        L = m.L
        self.U = np.ascontiguousarray(case.U0() if U0 is None else U0, dtype=np.float64)
        self.Uold = np.empty((L, 3))
        self.Uold[:] = case.uin
        self.JU = np.zeros((L, 3))
        self.p = np.zeros(L)
        self.nut = np.zeros(L)
        self.div = np.zeros(L)
        self.f = np.zeros((L, 3))
        self.solid = np.zeros(L)
        self.A = np.zeros((L, 7))
        self.b = np.zeros(L)
        self.work = np.zeros(9 * L)
        s = OrState()
        for n in ("U", "Uold", "JU", "p", "nut", "div", "f", "solid", "A", "b", "work"):
            setattr(s, n, ptr(getattr(self, n)))
        s.Uin[:] = case.uin
        s.Re, s.Cs, s.dt = case.Re, case.Cs, case.dt
        s.method = 1 if case.method == "BiCG" else 0
        s.pc_kind = 1 if case.pc == "SOR" else 0
        s.omega, s.pc_omega, s.pc_it, s.max_it, s.tol = case.omega, case.pc_omega, case.pc_it, case.max_it, case.tol
        if case.floor_start is not None:
            s.retracted_floor, s.floor_start_x = 1, case.floor_start
        if case.turbine is not None and case.sites:
            self.ap, self.cd, self.cl, self.atk = ap_props_oracle(case.turbine, len(case.sites), case.ap_per_blade)
            for a in range(len(self.ap)):
                self.ap[a, 13] = case.sites[a // (case.turbine.blades * case.ap_per_blade)].pitch
            self.wt = wt_records(case.turbine, case.sites)
            s.wt, s.ap, s.cd_tbl, s.cl_tbl, s.atk = (ptr(self.wt), ptr(self.ap), ptr(self.cd),
                                                     ptr(self.cl), ptr(self.atk))
            s.wt_count, s.ap_count = len(self.wt), len(self.ap)
            s.blade_per_wt, s.ap_per_blade = case.turbine.blades, case.ap_per_blade
            s.atk_count, s.width = len(self.atk), case.width
            if ap_hook is not None:
                ap_hook(self.ap, self.wt)
        modes = case.inflow_modes() if hasattr(case, "inflow_modes") else None
        if modes is not None and len(modes):
            g0 = m.gc
            ny, nz = m.size[1] - 2 * g0, m.size[2] - 2 * g0
            self.inflow_modes = np.ascontiguousarray(modes, dtype=np.float64)
            self.inflow_cells = np.zeros(g0 * ny * nz * 3)
            self.inflow_face = np.zeros(ny * nz)
            s.inflow_modes, s.inflow_n = ptr(self.inflow_modes), len(self.inflow_modes)
            s.inflow_speed = case.inflow_speed()
            s.inflow_cells, s.inflow_face = ptr(self.inflow_cells), ptr(self.inflow_face)
        self.s = s
        self.L, _ = lib(kind)
        if case.tgg is not None:
            plc = np.array(case.tgg["placement"], dtype=np.float64)
            self.L.or_tgg_solid(C.byref(self.g), ptr(self.solid), case.tgg["thick"], case.tgg["mesh"],
                                case.tgg["bar"], ptr(plc))
        if initialize:
            self.L.or_state_initialize(C.byref(self.g), C.byref(self.s))
        else:
            self.s.max_diag = self.L.or_build_A(C.byref(self.g), ptr(self.A))

    def load_state(self, fields, step):
        """Continue from a given state (U, Uold, JU, p, nut, f as [L][3] / [L] arrays)
        after step `step`."""
        for name, v in fields.items():
            getattr(self, name)[...] = np.asarray(v, dtype=np.float64).reshape(getattr(self, name).shape)
        self.s.step = int(step)

    def step(self):
        self.L.or_state_step(C.byref(self.g), C.byref(self.s))
        return self.stats()

    def adjoint(self, steps, tq_w, adj_tol=1e-13):
        """Run `steps` steps and the reverse sweep (adjoint.c): returns J, the
        per-AP sensitivities [ap][4] (turbine angle, pitch, chord, twist) and the
        bars of the initial state {U, JU, nut, f}."""
        tq = np.ascontiguousarray(tq_w, dtype=np.float64)
        J = dbl()
        g = np.zeros((max(len(getattr(self, "ap", [])), 1), 4))
        L = self.case.mesh.L
        sb = np.zeros(10 * L)
        self.L.or_adjoint_run(C.byref(self.g), C.byref(self.s), steps, ptr(tq), adj_tol, C.byref(J), ptr(g),
                              ptr(sb))
        bars = {"U": sb[:3 * L].reshape(L, 3), "JU": sb[3 * L:6 * L].reshape(L, 3),
                "nut": sb[6 * L:7 * L], "f": sb[7 * L:].reshape(L, 3)}
        return J.value, g, bars

    def stats(self):
        s = self.s
        d = dict(step=s.step, it=s.it, err=s.err, avg_div=s.avg_div, max_cfl=s.max_cfl, max_diag=s.max_diag)
        if s.wt_count:
            d["torque"] = self.wt[:, 14].copy()
            d["thrust"] = self.wt[:, 15].copy()
        return d

    def field(self, name):
        return getattr(self, name).copy()


class RefSolver:
    """The reference's own Solver driven through ref_harness.cpp."""

    def __init__(self, case, workdir=None, output_dir=None):
        """With output_dir, the reference's own output path is switched on with the
        case's schedule and probes (files appear there; probe files are completed
        by close())."""
        L, _ = lib("ref")
        self.L = L
        self.case = case
        self.tmp = tempfile.TemporaryDirectory(dir=workdir)
        mesh_path = os.path.join(self.tmp.name, "mesh.json")
        with open(mesh_path, "w") as fh:
            fh.write(case.mesh.to_json())
        uin = np.array(case.uin, dtype=np.float64)
        wt_prop = case.turbine.to_json().encode() if case.turbine is not None and case.sites else None
        from importlib import import_module
        sites = import_module("3dalm_amd.cases").sites_json(case.sites).encode() if wt_prop else None
        self.h = L.ref_solver_create(mesh_path.encode(), ptr(uin), case.Re, case.Cs, case.dt,
                                     1 if case.method == "BiCG" else 0, 1 if case.pc == "SOR" else 0,
                                     case.omega, case.pc_omega, case.pc_it, case.max_it, case.tol,
                                     wt_prop, sites, case.ap_per_blade, case.width)
        self.nwt = len(case.sites) if wt_prop else 0
        U0 = np.ascontiguousarray(case.U0())
        L.ref_solver_perturb(self.h, ptr(U0))
        if case.floor_start is not None:
            L.ref_solver_floor(self.h, case.floor_start)
        if case.tgg is not None:
            plc = np.array(case.tgg["placement"], dtype=np.float64)
            L.ref_solver_tgg(self.h, case.tgg["thick"], case.tgg["mesh"], case.tgg["bar"], ptr(plc))
        if output_dir is not None:
            pr = np.ascontiguousarray(np.array(case.probes, dtype=np.float64).reshape(-1, 3))
            L.ref_solver_output(self.h, str(output_dir).encode(), case.output_start, case.output_interval,
                                case.tavg_start, ptr(pr), len(pr))
        L.ref_solver_initialize(self.h)

    def step(self):
        self.L.ref_solver_step(self.h)
        return self.stats()

    def stats(self):
        st, it = i64(), i64()
        vals = np.zeros(4)
        tq = np.zeros(max(self.nwt, 1))
        th = np.zeros(max(self.nwt, 1))
        self.L.ref_solver_stats(self.h, C.byref(st), C.byref(it), ptr(vals), ptr(tq), ptr(th))
        d = dict(step=st.value, it=it.value, err=vals[0], avg_div=vals[1], max_cfl=vals[2], max_diag=vals[3])
        if self.nwt:
            d["torque"], d["thrust"] = tq[:self.nwt], th[:self.nwt]
        return d

    def field(self, name):
        n = self.case.mesh.L
        w = 3 if name in ("U", "Uold", "JU", "f", "Utavg") else 1  # p nut div b q solid: scalars
        out = np.zeros(n * w)
        assert self.L.ref_solver_get(self.h, name.encode(), ptr(out)) == 0
        return out.reshape(n, 3) if w == 3 else out

    def close(self):
        if self.h:
            self.L.ref_solver_destroy(self.h)
            self.h = None
            self.tmp.cleanup()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
