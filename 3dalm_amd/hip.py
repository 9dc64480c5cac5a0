"""ctypes binding of the C ABI in include/dalm_hip.h (the product path).

Loading fails loudly when the HIP library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(PKG, "lib", "lib3dalm_hip.so")

i64 = C.c_int64
dbl = C.c_double
P = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)

OK, NONCONVERGED, BREAKDOWN, NONFINITE, HIP_ERROR, BAD_ARG = range(6)
STATUS_NAMES = {0: "OK", 1: "NONCONVERGED", 2: "BREAKDOWN", 3: "NONFINITE", 4: "HIP_ERROR", 5: "BAD_ARG"}
FP64, MIXED = 0, 1
MIXED_BUILT = True  # the precision form is part of this library
F_U, F_UOLD, F_JU, F_P, F_NUT, F_DIV, F_F, F_SOLID, F_B = range(9)
# sensitivities of the discrete adjoint (FP64 contexts)
(F_U_B, F_UOLD_B, F_JU_B, F_JUOLD_B, F_NUT_B, F_NUTOLD_B, F_F_B, F_FOLD_B, F_P_B, F_B_B) = range(9, 19)
# output path (either form)
F_Q, F_UTAVG = 19, 20
FIELDS = {"U": F_U, "Uold": F_UOLD, "JU": F_JU, "p": F_P, "nut": F_NUT, "div": F_DIV, "f": F_F,
          "solid": F_SOLID, "b": F_B,
          "U_b": F_U_B, "Uold_b": F_UOLD_B, "JU_b": F_JU_B, "JUold_b": F_JUOLD_B, "nut_b": F_NUT_B,
          "nutold_b": F_NUTOLD_B, "f_b": F_F_B, "fold_b": F_FOLD_B, "p_b": F_P_B, "b_b": F_B_B,
          "q": F_Q, "Utavg": F_UTAVG}
VECTOR_FIELDS = {F_U, F_UOLD, F_JU, F_F, F_U_B, F_UOLD_B, F_JU_B, F_JUOLD_B, F_F_B, F_FOLD_B, F_UTAVG}
ADJ_KEEP_PRESSURE = 1
SOR, BICGSTAB = 0, 1
PC_JACOBI, PC_SOR = 0, 1
WT_STRIDE = 16
AP_STRIDE = 16
MAX_TURBINES = 64


class Params(C.Structure):
    _fields_ = [("Uin", dbl * 3), ("Re", dbl), ("Cs", dbl), ("dt", dbl), ("method", i64),
                ("pc_kind", i64), ("omega", dbl), ("pc_omega", dbl), ("pc_it", i64), ("max_it", i64),
                ("tol", dbl), ("retracted_floor", i64), ("floor_start_x", dbl)]


class StepStats(C.Structure):
    _fields_ = [("step", i64), ("it", i64), ("time", dbl), ("err", dbl), ("avg_div", dbl),
                ("max_cfl", dbl), ("status", i64), ("wt_count", i64),
                ("torque", dbl * MAX_TURBINES), ("thrust", dbl * MAX_TURBINES),
                ("power", dbl * MAX_TURBINES)]

    def as_dict(self):
        n = self.wt_count
        d = dict(step=self.step, it=self.it, time=self.time, err=self.err, avg_div=self.avg_div,
                 max_cfl=self.max_cfl, status=self.status)
        if n:
            d["torque"] = np.array(self.torque[:n])
            d["thrust"] = np.array(self.thrust[:n])
            d["power"] = np.array(self.power[:n])
        return d


class AdjointStats(C.Structure):
    _fields_ = [("steps", i64), ("every", i64), ("adj_it", i64), ("replay_ms", dbl), ("reverse_ms", dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HipError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(SO):
        raise HipError(f"{SO} is not built: run __graft_entry__.build() (no CPU fallback exists)")
    L = C.CDLL(SO)
    V = C.c_void_p
    sig = {
        "dalm_create": [C.POINTER(V), PI64, i64, P, P, P, i64],
        "dalm_destroy": [V],
        "dalm_set_device": [i64],
        "dalm_device_count": [PI64],
        "dalm_mesh_widths": [P, i64, P],
        "dalm_upload": [V, i64, P],
        "dalm_download": [V, i64, P],
        "dalm_set_params": [V, C.POINTER(Params)],
        "dalm_copy_u_to_uold": [V],
        "dalm_predictor": [V, dbl, dbl],
        "dalm_interpolate_JU": [V],
        "dalm_apply_JUbc": [V, P],
        "dalm_poisson_rhs": [V, dbl, dbl],
        "dalm_build_A": [V, P],
        "dalm_run_sor": [V, dbl, i64, dbl, PI64, P],
        "dalm_run_bicgstab": [V, i64, dbl, i64, i64, dbl, PI64, P],
        "dalm_residual": [V, P],
        "dalm_apply_pbc": [V],
        "dalm_project_p": [V, dbl],
        "dalm_set_solid_U": [V],
        "dalm_apply_Ubc": [V, P, dbl, i64, dbl],
        "dalm_calc_nut": [V, dbl],
        "dalm_calc_divergence": [V],
        "dalm_l2_norm": [V, i64, P],
        "dalm_set_inflow": [V, P, i64, dbl],
        "dalm_eval_inflow": [V, dbl],
        "dalm_max_cfl": [V, dbl, P],
        "dalm_calc_q": [V],
        "dalm_set_turbulence_grid": [V, dbl, dbl, dbl, P],
        "dalm_time_average": [V, dbl],
        "dalm_sample_U": [V, PI64, i64, P],
        "dalm_build_ap_props": [P, P, P, P, P, i64, i64, dbl, dbl, i64, i64, i64, P, P, P],
        "dalm_alm_setup": [V, P, i64, P, i64, i64, i64, P, P, P, i64, dbl],
        "dalm_run_alm": [V, dbl, P, P],
        "dalm_alm_download": [V, P, P],
        "dalm_initialize": [V, C.POINTER(StepStats)],
        "dalm_step": [V, C.POINTER(StepStats)],
        "dalm_run": [V, i64, C.POINTER(StepStats)],
        "dalm_probe_kernel": [V, i64],
        "dalm_probe_read": [V, PI64, P],
        "dalm_set_fp32_pass_cap": [V, i64],
        "dalm_last_solve_info": [V, PI64],
        "dalm_predictor_adj": [V, dbl, dbl],
        "dalm_interpolate_JU_adj": [V],
        "dalm_apply_JUbc_adj": [V],
        "dalm_poisson_rhs_adj": [V, dbl, dbl],
        "dalm_apply_pbc_adj": [V],
        "dalm_project_p_adj": [V, dbl],
        "dalm_set_solid_U_adj": [V],
        "dalm_apply_Ubc_adj": [V, dbl, i64, dbl],
        "dalm_calc_nut_adj": [V, dbl],
        "dalm_poisson_adj": [V, dbl, i64, PI64, P],
        "dalm_run_alm_adj": [V, dbl, P, P],
        "dalm_adjoint_begin": [V, i64, i64, P, i64],
        "dalm_adjoint_begin_from": [V, V, i64, i64, P, i64],
        "dalm_adjoint_run": [V, dbl, P, P, P, C.POINTER(AdjointStats)],
        "dalm_adjoint_end": [V],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.dalm_last_error.argtypes = [V]
    L.dalm_last_error.restype = C.c_char_p
    _lib = L
    return L


def exported_symbols():
    """Names the header declares (used by the ABI test)."""
    hdr = os.path.join(os.path.dirname(PKG), "include", "dalm_hip.h")
    import re
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(dalm_[a-z0-9_A-Z]+)\s*\(", text)))


def ptr(a):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data_as(P)


class Context:
    """One device context (include/dalm_hip.h)."""

    def __init__(self, mesh, precision=FP64):
        L = lib()
        self.L = L
        self.mesh = mesh
        self.h = C.c_void_p()
        size = (i64 * 3)(*mesh.size)
        rc = L.dalm_create(C.byref(self.h), size, mesh.gc, ptr(mesh.x), ptr(mesh.y), ptr(mesh.z), precision)
        if rc != OK:
            raise HipError(f"dalm_create failed with status {STATUS_NAMES.get(rc, rc)}")
        self.precision = precision
        self.n_wt, self.n_ap = 0, 0

    def check(self, rc, what, allow=(OK,)):
        if rc not in allow:
            msg = self.L.dalm_last_error(self.h)
            raise HipError(f"{what}: {STATUS_NAMES.get(rc, rc)} {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if self.h:
            self.L.dalm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields
    def upload(self, name, arr):
        fid = FIELDS[name]
        n = self.mesh.L * (3 if fid in VECTOR_FIELDS else 1)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        assert a.size == n, (name, a.size, n)
        self.check(self.L.dalm_upload(self.h, fid, ptr(a)), f"upload {name}")

    def download(self, name):
        fid = FIELDS[name]
        vec = fid in VECTOR_FIELDS
        out = np.zeros(self.mesh.L * (3 if vec else 1))
        self.check(self.L.dalm_download(self.h, fid, ptr(out)), f"download {name}")
        return out.reshape(-1, 3) if vec else out

    # ---- per-kernel calls
    def copy_u_to_uold(self):
        self.check(self.L.dalm_copy_u_to_uold(self.h), "copy")

    def predictor(self, Re, dt):
        self.check(self.L.dalm_predictor(self.h, Re, dt), "predictor")

    def interpolate_JU(self):
        self.check(self.L.dalm_interpolate_JU(self.h), "interpolate_JU")

    def apply_JUbc(self, uin):
        u = np.array(uin, dtype=np.float64)
        self.check(self.L.dalm_apply_JUbc(self.h, ptr(u)), "apply_JUbc")

    def poisson_rhs(self, dt, scale):
        self.check(self.L.dalm_poisson_rhs(self.h, dt, scale), "poisson_rhs")

    def build_A(self):
        m = dbl()
        self.check(self.L.dalm_build_A(self.h, C.byref(m)), "build_A")
        return m.value

    def run_sor(self, omega, max_it, tol):
        it, err = i64(), dbl()
        rc = self.check(self.L.dalm_run_sor(self.h, omega, max_it, tol, C.byref(it), C.byref(err)),
                        "run_sor", allow=(OK, NONCONVERGED, NONFINITE))
        return it.value, err.value, rc

    def run_bicgstab(self, pc_kind, pc_omega, pc_it, max_it, tol):
        it, err = i64(), dbl()
        rc = self.check(self.L.dalm_run_bicgstab(self.h, pc_kind, pc_omega, pc_it, max_it, tol, C.byref(it),
                                                 C.byref(err)),
                        "run_bicgstab", allow=(OK, NONCONVERGED, BREAKDOWN, NONFINITE))
        return it.value, err.value, rc

    def residual(self):
        out = np.zeros(self.mesh.L)
        self.check(self.L.dalm_residual(self.h, ptr(out)), "residual")
        return out

    def apply_pbc(self):
        self.check(self.L.dalm_apply_pbc(self.h), "apply_pbc")

    def project_p(self, dt):
        self.check(self.L.dalm_project_p(self.h, dt), "project_p")

    def set_solid_U(self):
        self.check(self.L.dalm_set_solid_U(self.h), "set_solid_U")

    def apply_Ubc(self, uin, dt, retracted=0, floor_x=0.0):
        u = np.array(uin, dtype=np.float64)
        self.check(self.L.dalm_apply_Ubc(self.h, ptr(u), dt, retracted, floor_x), "apply_Ubc")

    def calc_nut(self, Cs):
        self.check(self.L.dalm_calc_nut(self.h, Cs), "calc_nut")

    def calc_divergence(self):
        self.check(self.L.dalm_calc_divergence(self.h), "calc_divergence")

    def l2_norm(self, name):
        v = dbl()
        self.check(self.L.dalm_l2_norm(self.h, FIELDS[name], C.byref(v)), "l2_norm")
        return v.value

    def max_cfl(self, dt):
        v = dbl()
        self.check(self.L.dalm_max_cfl(self.h, dt, C.byref(v)), "max_cfl")
        return v.value

    @staticmethod
    def device_count():
        n = C.c_int64()
        if lib().dalm_device_count(C.byref(n)) != OK:
            raise HipError("dalm_device_count failed")
        return n.value

    def set_turbulence_grid(self, thick, mesh, bar, placement):
        plc = np.array(placement, dtype=np.float64)
        self.check(self.L.dalm_set_turbulence_grid(self.h, thick, mesh, bar, ptr(plc)), "set_turbulence_grid")

    # ---- output path
    def calc_q(self):
        self.check(self.L.dalm_calc_q(self.h), "calc_q")

    def time_average(self, weight):
        self.check(self.L.dalm_time_average(self.h, weight), "time_average")

    def sample_U(self, cells):
        """U at padded linear cell indices, [n][3]."""
        c = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
        out = np.zeros((len(c), 3))
        self.check(self.L.dalm_sample_U(self.h, c.ctypes.data_as(PI64), len(c), ptr(out)), "sample_U")
        return out

    # ---- ALM
    def alm_setup(self, wt, ap, blade_per_wt, ap_per_blade, cd, cl, atk, width):
        wt = np.ascontiguousarray(wt, dtype=np.float64)
        ap = np.ascontiguousarray(ap, dtype=np.float64)
        cd = np.ascontiguousarray(cd, dtype=np.float64)
        cl = np.ascontiguousarray(cl, dtype=np.float64)
        atk = np.ascontiguousarray(atk, dtype=np.float64)
        self.n_wt, self.n_ap = len(wt), len(ap)
        self.check(self.L.dalm_alm_setup(self.h, ptr(wt), len(wt), ptr(ap), len(ap), blade_per_wt,
                                         ap_per_blade, ptr(cd), ptr(cl), ptr(atk), len(atk), width),
                   "alm_setup")

    def run_alm(self, t):
        tq = np.zeros(max(self.n_wt, 1))
        th = np.zeros(max(self.n_wt, 1))
        self.check(self.L.dalm_run_alm(self.h, t, ptr(tq), ptr(th)), "run_alm")
        return tq[:self.n_wt], th[:self.n_wt]

    def alm_download(self):
        ap = np.zeros((self.n_ap, AP_STRIDE))
        wt = np.zeros((self.n_wt, WT_STRIDE))
        self.check(self.L.dalm_alm_download(self.h, ptr(ap), ptr(wt)), "alm_download")
        return ap, wt

    # ---- whole steps
    def set_inflow(self, modes, speed):
        """Turbulent inflow: modes [n][7] = (kx, ky, kz, psi, ax, ay, az), frozen and
        convected at `speed` (an empty table restores the constant inflow)."""
        m = np.ascontiguousarray(modes, dtype=np.float64).reshape(-1, 7)
        self.check(self.L.dalm_set_inflow(self.h, ptr(m) if len(m) else None, len(m), float(speed)), "set_inflow")

    def eval_inflow(self, t):
        self.check(self.L.dalm_eval_inflow(self.h, float(t)), "eval_inflow")

    def set_params(self, prm: Params):
        self.check(self.L.dalm_set_params(self.h, C.byref(prm)), "set_params")

    def initialize(self):
        st = StepStats()
        self.check(self.L.dalm_initialize(self.h, C.byref(st)), "initialize")
        return st.as_dict()

    def step(self):
        st = StepStats()
        self.check(self.L.dalm_step(self.h, C.byref(st)), "step", allow=(OK, BREAKDOWN, NONFINITE))
        return st.as_dict()

    def run(self, n):
        st = StepStats()
        self.check(self.L.dalm_run(self.h, n, C.byref(st)), "run")
        return st.as_dict()

    def last_solve_info(self):
        """(FP32 iterations, FP64 iterations, residual replacements) of the last step's solve."""
        out = np.zeros(3, dtype=np.int64)
        self.check(self.L.dalm_last_solve_info(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "last_solve_info")
        return int(out[0]), int(out[1]), int(out[2])

    def set_fp32_pass_cap(self, n):
        """Testing aid of the precision form: FP32 passes end after n iterations."""
        self.check(self.L.dalm_set_fp32_pass_cap(self.h, int(n)), "set_fp32_pass_cap")

    # ---- discrete adjoint (bars are the *_b fields)
    def predictor_adj(self, Re, dt):
        self.check(self.L.dalm_predictor_adj(self.h, Re, dt), "predictor_adj")

    def interpolate_JU_adj(self):
        self.check(self.L.dalm_interpolate_JU_adj(self.h), "interpolate_JU_adj")

    def apply_JUbc_adj(self):
        self.check(self.L.dalm_apply_JUbc_adj(self.h), "apply_JUbc_adj")

    def poisson_rhs_adj(self, dt, scale):
        self.check(self.L.dalm_poisson_rhs_adj(self.h, dt, scale), "poisson_rhs_adj")

    def apply_pbc_adj(self):
        self.check(self.L.dalm_apply_pbc_adj(self.h), "apply_pbc_adj")

    def project_p_adj(self, dt):
        self.check(self.L.dalm_project_p_adj(self.h, dt), "project_p_adj")

    def set_solid_U_adj(self):
        self.check(self.L.dalm_set_solid_U_adj(self.h), "set_solid_U_adj")

    def apply_Ubc_adj(self, dt, retracted=0, floor_x=0.0):
        self.check(self.L.dalm_apply_Ubc_adj(self.h, dt, retracted, floor_x), "apply_Ubc_adj")

    def calc_nut_adj(self, Cs):
        self.check(self.L.dalm_calc_nut_adj(self.h, Cs), "calc_nut_adj")

    def poisson_adj(self, tol_rel, max_it):
        it, err = i64(), dbl()
        self.check(self.L.dalm_poisson_adj(self.h, tol_rel, max_it, C.byref(it), C.byref(err)), "poisson_adj")
        return it.value, err.value

    def run_alm_adj(self, t, tq_w):
        tq = np.zeros(MAX_TURBINES)
        tq[:len(tq_w)] = tq_w
        g = np.zeros((self.n_ap, 4))
        self.check(self.L.dalm_run_alm_adj(self.h, t, ptr(tq), ptr(g)), "run_alm_adj")
        return g

    def adjoint_begin(self, steps, every, tq_w, keep_pressure=True):
        tq = np.zeros(MAX_TURBINES)
        tq[:len(tq_w)] = tq_w
        self.check(self.L.dalm_adjoint_begin(self.h, steps, every, ptr(tq),
                                             ADJ_KEEP_PRESSURE if keep_pressure else 0), "adjoint_begin")

    def adjoint_begin_from(self, src, steps, every, tq_w, keep_pressure=True):
        """Record the next `steps` steps of the precision-form context `src` into this
        FP64 context's tape (dalm_adjoint_begin_from)."""
        tq = np.zeros(MAX_TURBINES)
        tq[:len(tq_w)] = tq_w
        self.check(self.L.dalm_adjoint_begin_from(self.h, src.h, steps, every, ptr(tq),
                                                  ADJ_KEEP_PRESSURE if keep_pressure else 0), "adjoint_begin_from")

    def adjoint_run(self, tol_rel=1e-10):
        J = dbl()
        ap_g = np.zeros((max(self.n_ap, 1), 4))
        wt_g = np.zeros((max(self.n_wt, 1), 2))
        st = AdjointStats()
        self.check(self.L.dalm_adjoint_run(self.h, tol_rel, C.byref(J), ptr(ap_g), ptr(wt_g), C.byref(st)),
                   "adjoint_run")
        return J.value, ap_g[:self.n_ap], wt_g[:self.n_wt], st.as_dict()

    def adjoint_end(self):
        self.check(self.L.dalm_adjoint_end(self.h), "adjoint_end")


def build_ap_props(turbine, wt_count, ap_per_blade):
    """dalm_build_ap_props on a cases.TurbineModel."""
    L = lib()
    ns, na = len(turbine.stations_r), len(turbine.atk)
    n = wt_count * turbine.blades * ap_per_blade
    ap = np.zeros((n, AP_STRIDE))
    cd = np.zeros((n, na))
    cl = np.zeros((n, na))
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in
            (turbine.stations_r, turbine.stations_chord, turbine.stations_twist,
             np.asarray(turbine.cd).reshape(ns, na), np.asarray(turbine.cl).reshape(ns, na))]
    rc = L.dalm_build_ap_props(*[ptr(a) for a in arrs], ns, na, turbine.R, turbine.hub_R, turbine.blades,
                               wt_count, ap_per_blade, ptr(ap), ptr(cd), ptr(cl))
    if rc != OK:
        raise HipError(f"dalm_build_ap_props: {STATUS_NAMES.get(rc, rc)}")
    return ap, cd, cl, np.ascontiguousarray(turbine.atk, dtype=np.float64)
