// ref_harness.cpp -- builds the reference's own hot-path code (its headers under
// /root/reference/src, included by path, never copied) into oracle/_ref/ as a
// shared library with a C ABI shaped like oracle.h, so tests can compare the
// restatement in oracle.c against the real thing on identical inputs.
//
// TEST INFRASTRUCTURE ONLY (built only where /root/reference exists).
//
// The per-kernel wrappers call the reference's static functions with
// MpiInfo{size=1, rank=0}.  The whole-solver wrapper drives the reference's own
// Solver::main_loop_once (src/solver.h:1034-1306); the start-up pass replays the
// call order of Solver::initialize (src/solver.h:727-833) without its MPI/OpenACC
// device setup, printing and turbulence grid, and with an optional seeded
// interior perturbation of U applied after the fills (src/solver.h:737-740).
#include "solver.h"

#include <cstring>
#include <string>

typedef int64_t i64;

extern "C" {

typedef struct {
    i64 n[3];
    i64 gc;
    const double *c[3];
    const double *d[3];
} or_grid;
}

namespace {
MpiInfo one_rank() {
    MpiInfo m;
    m.size = 1;
    m.rank = 0;
    return m;
}
struct G {
    Int size[3];
    Int off[3] = {0, 0, 0};
    Int gc;
    Real *x, *y, *z, *dx, *dy, *dz;
    explicit G(const or_grid *g) {
        for (int a = 0; a < 3; ++a) size[a] = g->n[a];
        gc = g->gc;
        x = const_cast<Real *>(g->c[0]);
        y = const_cast<Real *>(g->c[1]);
        z = const_cast<Real *>(g->c[2]);
        dx = const_cast<Real *>(g->d[0]);
        dy = const_cast<Real *>(g->d[1]);
        dz = const_cast<Real *>(g->d[2]);
    }
};
typedef Real (*V3)[3];
typedef Real (*V7)[7];
inline V3 v3(const double *p) { return reinterpret_cast<V3>(const_cast<double *>(p)); }
inline Real *v1(const double *p) { return const_cast<Real *>(p); }

// flat record layouts shared with oracle.h
enum { WT = 16, AP = 16 };

void wt_to_ref(const double *w, WindTurbine &t) {
    for (int m = 0; m < 3; ++m) { t.base[m] = w[m]; t.rot_center[m] = w[3 + m]; }
    int ty = (int)w[6];
    t.angle_type = ty == 0 ? EulerAngle::Roll : ty == 1 ? EulerAngle::Pitch : ty == 2 ? EulerAngle::Yaw : EulerAngle::None;
    t.angle = w[7];
    t.ow = w[8];
    for (int m = 0; m < 4; ++m) t.formula[m] = w[9 + m];
    t.rot_speed = w[13];
    t.torque = w[14];
    t.thrust = w[15];
}
void wt_from_ref(const WindTurbine &t, double *w) {
    for (int m = 0; m < 3; ++m) { w[m] = t.base[m]; w[3 + m] = t.rot_center[m]; }
    w[6] = t.angle_type == EulerAngle::Roll ? 0 : t.angle_type == EulerAngle::Pitch ? 1 : t.angle_type == EulerAngle::Yaw ? 2 : 3;
    w[7] = t.angle;
    w[8] = t.ow;
    for (int m = 0; m < 4; ++m) w[9 + m] = t.formula[m];
    w[13] = t.rot_speed;
    w[14] = t.torque;
    w[15] = t.thrust;
}
void ap_to_ref(const double *a, ActuatorPoint &p) {
    for (int m = 0; m < 3; ++m) { p.U[m] = 0; p.xyz[m] = a[m]; p.f[m] = a[8 + m]; }
    p.r = a[3]; p.chord = a[4]; p.twist = a[5]; p.theta = a[6]; p.dr = a[7];
    p.torque = a[11]; p.thrust = a[12];
}
void ap_from_ref(const ActuatorPoint &p, double *a) {
    for (int m = 0; m < 3; ++m) { a[m] = p.xyz[m]; a[8 + m] = p.f[m]; }
    a[3] = p.r; a[4] = p.chord; a[5] = p.twist; a[6] = p.theta; a[7] = p.dr;
    a[11] = p.torque; a[12] = p.thrust;
}

// Q4: the reference leaves angle_type/formula unset when all three angles are 0.
void define_zero_angles(const nlohmann::json &arr, WindTurbine *wt, Int n) {
    for (Int i = 0; i < n; ++i) {
        const auto &ang = arr[i]["angle"];
        bool any = false;
        for (int j = 0; j < 3; ++j) {
            const auto &a = ang[j];
            if ((a.is_number() && a.get<Real>() != 0) || a.is_object()) any = true;
        }
        if (!any) {
            wt[i].angle_type = EulerAngle::None;
            for (int m = 0; m < 4; ++m) wt[i].formula[m] = 0;
        }
    }
}
}  // namespace

extern "C" {

// ---- mesh: src/io.h:98-164 --------------------------------------------------
int ref_build_mesh(const char *path, double *c0, double *c1, double *c2, double *d0, double *d1,
                   double *d2, i64 size[3], i64 *gc) {
    Real *x, *y, *z, *dx, *dy, *dz;
    Int sz[3], g;
    MpiInfo m = one_rank();
    build_mesh_from_json(path, x, y, z, dx, dy, dz, sz, g, &m);
    double *cc[3] = {c0, c1, c2}, *dd[3] = {d0, d1, d2};
    Real *xs[3] = {x, y, z}, *ds[3] = {dx, dy, dz};
    for (int a = 0; a < 3; ++a) {
        if (cc[a]) std::memcpy(cc[a], xs[a], sizeof(double) * sz[a]);
        if (dd[a]) std::memcpy(dd[a], ds[a], sizeof(double) * sz[a]);
        size[a] = sz[a];
        delete[] xs[a];
        delete[] ds[a];
    }
    *gc = g;
    return 0;
}

// ---- per-kernel wrappers ------------------------------------------------------
void ref_predictor(const or_grid *g, double *U, const double *Uold, const double *JU,
                   const double *nut, const double *f, double Re, double dt) {
    G a(g);
    MpiInfo m = one_rank();
    calc_intermediate_U(v3(U), v3(Uold), v3(JU), v1(nut), v3(f), a.x, a.y, a.z, a.dx, a.dy, a.dz,
                        Re, dt, a.size, a.gc, &m);
}
void ref_interpolate_JU(const or_grid *g, const double *U, double *JU) {
    G a(g);
    MpiInfo m = one_rank();
    interpolate_JU(v3(U), v3(JU), a.dx, a.dy, a.dz, a.size, a.gc, &m);
}
void ref_apply_JUbc(const or_grid *g, double *JU, const double Uin[3]) {
    G a(g);
    MpiInfo m = one_rank();
    apply_JUbc(v3(JU), v1(Uin), a.dy, a.dz, a.size, a.gc, &m);
}
void ref_poisson_rhs(const or_grid *g, const double *JU, double *b, double dt, double scale) {
    G a(g);
    MpiInfo m = one_rank();
    calc_poisson_rhs(v3(JU), b, a.dx, a.dy, a.dz, dt, scale, a.size, a.gc, &m);
}
double ref_build_A(const or_grid *g, double *A) {
    G a(g);
    MpiInfo m = one_rank();
    return build_A(reinterpret_cast<V7>(A), a.x, a.y, a.z, a.dx, a.dy, a.dz, a.size, a.size, a.off,
                   a.gc, &m);
}
double ref_dot(const or_grid *g, const double *x, const double *y) {
    G a(g);
    MpiInfo m = one_rank();
    return calc_inner_product(v1(x), v1(y), a.size, a.gc, &m);
}
double ref_l2_norm(const or_grid *g, const double *v) {
    G a(g);
    MpiInfo m = one_rank();
    return calc_l2_norm(v1(v), a.size, a.gc, &m);
}
void ref_residual(const or_grid *g, const double *A, const double *x, const double *b, double *r) {
    G a(g);
    MpiInfo m = one_rank();
    calc_residual(reinterpret_cast<V7>(const_cast<double *>(A)), v1(x), v1(b), r, a.size, a.gc, &m);
}
void ref_Ax(const or_grid *g, const double *A, const double *x, double *y) {
    G a(g);
    MpiInfo m = one_rank();
    calc_Ax(reinterpret_cast<V7>(const_cast<double *>(A)), v1(x), y, a.size, a.gc, &m);
}
void ref_sor_sweep(const or_grid *g, const double *A, double *x, const double *b, double omega,
                   i64 color) {
    G a(g);
    MpiInfo m = one_rank();
    sweep_sor(reinterpret_cast<V7>(const_cast<double *>(A)), x, v1(b), omega, color, a.size, a.off,
              a.gc, &m);
}
void ref_jacobi_sweep(const or_grid *g, const double *A, double *x, const double *xold,
                      const double *b) {
    G a(g);
    MpiInfo m = one_rank();
    sweep_jacobi(reinterpret_cast<V7>(const_cast<double *>(A)), x, v1(xold), v1(b), a.size, a.gc,
                 &m);
}
void ref_run_sor(const or_grid *g, const double *A, double *x, const double *b, double *r,
                 double omega, i64 *it, i64 max_it, double *err, double tol) {
    G a(g);
    MpiInfo m = one_rank();
    Int itr = 0;
    Real e = 0;
    run_sor(reinterpret_cast<V7>(const_cast<double *>(A)), x, v1(b), r, omega, itr, max_it, e, tol,
            a.size, a.size, a.off, a.gc, &m);
    *it = itr;
    *err = e;
}
void ref_run_bicgstab(const or_grid *g, const double *A, double *x, const double *b, double *work,
                      i64 pc_kind, double pc_omega, i64 pc_it, i64 *it, i64 max_it, double *err,
                      double tol) {
    G a(g);
    MpiInfo m = one_rank();
    const i64 L = g->n[0] * g->n[1] * g->n[2];
    double *r = work, *r0 = work + L, *p = work + 2 * L, *pp = work + 3 * L, *q = work + 4 * L;
    double *s = work + 5 * L, *ss = work + 6 * L, *t = work + 7 * L, *tmp = work + 8 * L;
    Int itr = 0;
    Real e = 0;
    V7 AA = reinterpret_cast<V7>(const_cast<double *>(A));
    if (pc_kind == 0)
        run_pbicgstab(AA, x, v1(b), r, r0, p, pp, q, s, ss, t, tmp, itr, max_it, e, tol, pc_it, a.size,
                      a.size, a.gc, &m);
    else
        run_pbicgstab(AA, x, v1(b), r, r0, p, pp, q, s, ss, t, itr, max_it, e, tol, pc_omega, pc_it,
                      a.size, a.size, a.off, a.gc, &m);
    *it = itr;
    *err = e;
}
void ref_apply_pbc(const or_grid *g, double *p) {
    G a(g);
    MpiInfo m = one_rank();
    // U and nut are unused by live code (src/bc.h:304-411)
    apply_pbc(nullptr, p, nullptr, a.z, a.dx, a.dy, a.dz, 1.0, a.size, a.gc, &m);
}
void ref_project(const or_grid *g, double *U, double *JU, const double *p, double dt) {
    G a(g);
    MpiInfo m = one_rank();
    project_p(v3(U), v3(JU), v1(p), a.x, a.y, a.z, a.dx, a.dy, a.dz, dt, a.size, a.gc, &m);
}
void ref_set_solid(const or_grid *g, double *U, const double *solid) {
    G a(g);
    set_solid_U(v3(U), v1(solid), a.size, a.gc);
}
void ref_apply_Ubc(const or_grid *g, double *U, const double *Uold, const double Uin[3], double dt,
                   i64 retracted_floor, double floor_start_x) {
    G a(g);
    MpiInfo m = one_rank();
    apply_Ubc(v3(U), v3(Uold), v1(Uin), a.x, a.y, a.z, a.dx, a.dy, a.dz, dt, a.size, a.gc, &m,
              retracted_floor != 0, floor_start_x);
}
void ref_nut(const or_grid *g, const double *U, double *nut, double Cs) {
    G a(g);
    MpiInfo m = one_rank();
    calc_nut(v3(U), nut, a.x, a.y, a.z, a.dx, a.dy, a.dz, Cs, a.size, a.gc, &m);
}
void ref_divergence(const or_grid *g, const double *JU, double *div) {
    G a(g);
    MpiInfo m = one_rank();
    calc_divergence(v3(JU), div, a.dx, a.dy, a.dz, a.size, a.gc, &m);
}
void ref_tgg_solid(const or_grid *g, double *solid, double thick, double mesh, double bar, const double plc[3]) {
    G a(g);
    set_turbulence_generating_grid(solid, a.x, a.y, a.z, a.dx, a.dy, a.dz, thick, mesh, bar, plc[0], plc[1],
                                   plc[2], a.size, a.gc);
}
void ref_calc_q(const or_grid *g, const double *U, double *q) {
    G a(g);
    MpiInfo m = one_rank();
    calc_q(v3(U), q, a.x, a.y, a.z, a.size, a.gc, &m);
}
// the time average as Solver::main_loop_once computes it (src/solver.h:1197-1203)
void ref_time_average(const or_grid *g, double *Utavg, const double *U, double weight) {
    G a(g);
    const Int len = a.size[0] * a.size[1] * a.size[2];
    calc_ax_plus_by(1. - weight, v3(Utavg), weight, v3(U), v3(Utavg), len);
}
double ref_max_cfl(const or_grid *g, const double *U, double dt) {
    G a(g);
    MpiInfo m = one_rank();
    return calc_max_cfl(v3(U), a.dx, a.dy, a.dz, dt, a.size, a.gc, &m);
}

// ---- ALM --------------------------------------------------------------------
// AP properties from a windturbine.json text: src/alm.h:35-142.  Returns the AP
// count; writes records (AP stride) and the flattened Cd/Cl rows and attack list.
i64 ref_build_ap_props(const char *wt_prop_text, i64 wt_count, i64 ap_per_blade, double *ap,
                       double *cd_tbl, double *cl_tbl, double *atk, i64 *atk_count_out) {
    auto js = nlohmann::json::parse(std::string(wt_prop_text));
    ActuatorPoint *lst;
    Real *atk_lst, **cd, **cl;
    Int n, na;
    build_ap_props(build_ap_props(js, wt_count, ap_per_blade), lst, atk_lst, cd, cl, n, na);
    for (Int a = 0; a < n; ++a) {
        if (ap) {
            std::memset(ap + AP * a, 0, sizeof(double) * AP);
            lst[a].theta = 0;
            for (int m = 0; m < 3; ++m) { lst[a].xyz[m] = 0; lst[a].f[m] = 0; }
            lst[a].torque = lst[a].thrust = 0;
            ap_from_ref(lst[a], ap + AP * a);
        }
        if (cd_tbl) std::memcpy(cd_tbl + na * a, cd[a], sizeof(double) * na);
        if (cl_tbl) std::memcpy(cl_tbl + na * a, cl[a], sizeof(double) * na);
        delete[] cd[a];
        delete[] cl[a];
    }
    if (atk) std::memcpy(atk, atk_lst, sizeof(double) * na);
    *atk_count_out = na;
    delete[] cd;
    delete[] cl;
    delete[] atk_lst;
    delete[] lst;
    return n;
}
// Turbine records: src/alm.h:144-206 (+ explicit None for all-zero angles, Q4).
i64 ref_build_wt_props(const char *wt_prop_text, const char *wt_array_text, double *wt) {
    auto prop = nlohmann::json::parse(std::string(wt_prop_text));
    auto arr = nlohmann::json::parse(std::string(wt_array_text));
    WindTurbine *lst;
    Int n;
    build_wt_props(prop, arr, lst, n);
    define_zero_angles(arr, lst, n);
    for (Int i = 0; i < n; ++i) {
        lst[i].angle = lst[i].ow = lst[i].torque = lst[i].thrust = 0;
        wt_from_ref(lst[i], wt + WT * i);
    }
    delete[] lst;
    return n;
}
void ref_update_ap_position(double *wt, i64 wt_count, double *ap, i64 ap_count, i64 blade_per_wt,
                            i64 ap_per_blade, double t) {
    std::vector<WindTurbine> W(wt_count);
    std::vector<ActuatorPoint> P(ap_count);
    for (i64 i = 0; i < wt_count; ++i) wt_to_ref(wt + WT * i, W[i]);
    for (i64 i = 0; i < ap_count; ++i) ap_to_ref(ap + AP * i, P[i]);
    update_ap_position(W.data(), wt_count, P.data(), ap_count, blade_per_wt, ap_per_blade, t);
    for (i64 i = 0; i < wt_count; ++i) wt_from_ref(W[i], wt + WT * i);
    for (i64 i = 0; i < ap_count; ++i) ap_from_ref(P[i], ap + AP * i);
}
void ref_calc_ap_force(const or_grid *g, const double *U, double *wt, i64 wt_count, double *ap,
                       i64 ap_count, i64 blade_per_wt, i64 ap_per_blade, const double *cd_tbl,
                       const double *cl_tbl, const double *atk, i64 atk_count) {
    G a(g);
    MpiInfo m = one_rank();
    std::vector<WindTurbine> W(wt_count);
    std::vector<ActuatorPoint> P(ap_count);
    std::vector<Real *> cd(ap_count), cl(ap_count);
    for (i64 i = 0; i < wt_count; ++i) wt_to_ref(wt + WT * i, W[i]);
    for (i64 i = 0; i < ap_count; ++i) {
        ap_to_ref(ap + AP * i, P[i]);
        cd[i] = const_cast<Real *>(cd_tbl + atk_count * i);
        cl[i] = const_cast<Real *>(cl_tbl + atk_count * i);
    }
    calc_ap_force(v3(U), a.x, a.y, a.z, W.data(), wt_count, P.data(), ap_count, blade_per_wt,
                  ap_per_blade, cd.data(), cl.data(), const_cast<Real *>(atk), atk_count, 0.0, a.size,
                  a.gc, &m);
    for (i64 i = 0; i < wt_count; ++i) wt_from_ref(W[i], wt + WT * i);
    for (i64 i = 0; i < ap_count; ++i) ap_from_ref(P[i], ap + AP * i);
}
void ref_distribute_ap_force(const or_grid *g, double *f, const double *ap, i64 ap_count,
                             double width) {
    G a(g);
    std::vector<ActuatorPoint> P(ap_count);
    for (i64 i = 0; i < ap_count; ++i) ap_to_ref(ap + AP * i, P[i]);
    distribute_ap_force(v3(f), P.data(), ap_count, width, a.x, a.y, a.z, a.size, a.gc);
}

// ---- whole solver -------------------------------------------------------------
struct RefSolver {
    Solver s;
};

void *ref_solver_create(const char *mesh_path, const double Uin[3], double Re, double Cs, double dt,
                        i64 method, i64 pc_kind, double omega, double pc_omega, i64 pc_it,
                        i64 max_it, double tol, const char *wt_prop_text,
                        const char *wt_array_text, i64 ap_per_blade, double width) {
    RefSolver *R = new RefSolver();
    Solver &S = R->s;
    S.mpi = one_rank();
    S.rt.initialize(1e300, dt, 1e300, 1e300, 1e300);
    S.gmesh.initialize_from_file(mesh_path, S.gsize, S.gc, &S.mpi);
    S.mesh.initialize_from_global_mesh(&S.gmesh, S.gsize, S.size, S.offset, S.gc, &S.mpi);
    Real uin[3] = {Uin[0], Uin[1], Uin[2]};
    S.cfd.initialize(uin, Re, Cs, S.size);
    S.eq.initialize(max_it, tol, S.size, method == 1 ? "BiCG" : "SOR");
    if (method == 1) {
        S.eq.pc_method = pc_kind == 1 ? "SOR" : "Jacobi";
        S.eq.pc_max_it = pc_it;
        S.eq.pc_relax_rate = pc_omega;
    } else {
        S.eq.relax_rate = omega;
    }
    if (wt_prop_text && wt_array_text) {
        auto prop = nlohmann::json::parse(std::string(wt_prop_text));
        auto arr = nlohmann::json::parse(std::string(wt_array_text));
        nlohmann::json alm_json = {{"blade points", ap_per_blade}, {"projection width", width}};
        S.alm.intialize(prop, arr, alm_json);
        define_zero_angles(arr, S.alm.wt_lst, S.alm.wt_count);
    }
    S.eq.max_diag = build_A(S.eq.A, S.mesh.x, S.mesh.y, S.mesh.z, S.mesh.dx, S.mesh.dy, S.mesh.dz,
                            S.gsize, S.size, S.offset, S.gc, &S.mpi);
    Int len = S.size[0] * S.size[1] * S.size[2];
    fill_array(S.cfd.U, S.cfd.Uin, len);
    fill_array(S.cfd.Uold, S.cfd.Uin, len);
    fill_array(S.cfd.p, 0., len);
    return R;
}

// The output part of Solver::initialize (src/solver.h:569-598, 656-700, 835-846):
// output folder, output / time-average schedule, the inst and tavg variable lists
// and the probes at the nearest cell centres.  Call before ref_solver_initialize
// (whose step-0 pass already records and writes).
void ref_solver_output(void *h, const char *dir, double start, double interval, double tavg_start,
                       const double *probe_xyz, i64 nprobe) {
    Solver &S = static_cast<RefSolver *>(h)->s;
    S.output_dir = dir;
    S.rt.output_start_time = start;
    S.rt.output_interval_time = interval;
    S.rt.next_output_time = start;
    S.rt.tavg_start_time = tavg_start;
    S.out_handler.set_size(S.size, S.gc);
    S.out_handler.set_var({S.cfd.U[0], S.cfd.p, S.cfd.div, S.cfd.q, S.cfd.f[0]}, {3, 1, 1, 1, 3},
                          {"U", "p", "div", "q", "f"});
    S.tavg_out_handler.set_size(S.size, S.gc);
    S.tavg_out_handler.set_var({S.cfd.Utavg[0]}, {3}, {"U"});
    {   // partition.json (src/solver.h:982-999), one rank
        nlohmann::json part;
        part["size"] = {S.gsize[0], S.gsize[1], S.gsize[2]};
        part["gc"] = S.gc;
        part["partition"] = {};
        Int rs[3], ro[3];
        calc_partition(S.gsize, rs, ro, S.gc, &S.mpi);
        nlohmann::json r;
        r["size"] = {rs[0], rs[1], rs[2]};
        r["offset"] = {ro[0], ro[1], ro[2]};
        r["rank"] = 0;
        part["partition"].push_back(r);
        std::ofstream po(S.output_dir / "partition.json");
        po << std::setw(2) << part;
    }
    if (S.alm.ap_count > 0) {  // ap_info.log (src/solver.h:950-977)
        auto &alm = S.alm;
        std::string path = S.output_dir / "ap_info.log";
        FILE *f = fopen(path.c_str(), "w");
        fprintf(f, "ap count = %ld, atk acount = %ld\n", alm.ap_count, alm.atk_count);
        for (Int i = 0; i < alm.ap_count; i++) {
            auto &ap = alm.ap_lst[i];
            fprintf(f, "ap %ld\n\tr = %lf\n\tdr = %lf\n\tchord = %lf\n\ttwist = %lf\n\tcd = \n", i, ap.r, ap.dr,
                    ap.chord, ap.twist);
            for (Int j = 0; j < alm.atk_count; j++) fprintf(f, "\t\t%lf\n", alm.cd_tbl[i][j]);
            fprintf(f, "\tcl = \n");
            for (Int j = 0; j < alm.atk_count; j++) fprintf(f, "\t\t%lf\n", alm.cl_tbl[i][j]);
        }
        fprintf(f, "atk\n");
        for (Int j = 0; j < alm.atk_count; j++) fprintf(f, "\t%lf\n", alm.atk_lst[j]);
        fclose(f);
    }
    S.probes = std::vector<Probe>(nprobe);
    const Int gc = S.gc;
    Real *x = S.mesh.x, *y = S.mesh.y, *z = S.mesh.z;
    for (i64 m = 0; m < nprobe; ++m) {
        Int ni = find_nearest_index(x + gc - 1, probe_xyz[3 * m], S.size[0] - 2 * gc + 2) + gc - 1;
        Int nj = find_nearest_index(y + gc - 1, probe_xyz[3 * m + 1], S.size[1] - 2 * gc + 2) + gc - 1;
        Int nk = find_nearest_index(z + gc - 1, probe_xyz[3 * m + 2], S.size[2] - 2 * gc + 2) + gc - 1;
        if (ni >= gc && ni < S.size[0] - gc && nj >= gc && nj < S.size[1] - gc && nk >= gc &&
            nk < S.size[2] - gc)
            S.probes[m].initialize(S.output_dir / ("probe" + std::to_string(m) + ".csv"), ni, nj, nk, x[ni],
                                   y[nj], z[nk]);
    }
}

// The retracted floor of Solver::initialize (src/solver.h:620-635).
void ref_solver_floor(void *h, double floor_start_x) {
    Solver &S = static_cast<RefSolver *>(h)->s;
    S.cfd.retracted_floor = true;
    S.cfd.floor_start_x = floor_start_x;
}

// The turbulence-generating grid of Solver::initialize (src/solver.h:740-750).
void ref_solver_tgg(void *h, double thick, double mesh, double bar, const double *plc) {
    Solver &S = static_cast<RefSolver *>(h)->s;
    set_turbulence_generating_grid(S.cfd.solid, S.mesh.x, S.mesh.y, S.mesh.z, S.mesh.dx, S.mesh.dy, S.mesh.dz,
                                   thick, mesh, bar, plc[0], plc[1], plc[2], S.size, S.gc);
}

// Overwrite the interior of U with U0 (an [L][3] array); ghosts keep Uin.
void ref_solver_perturb(void *h, const double *U0) {
    Solver &S = static_cast<RefSolver *>(h)->s;
    Int G = S.gc;
    for (Int i = G; i < S.size[0] - G; ++i)
        for (Int j = G; j < S.size[1] - G; ++j)
            for (Int k = G; k < S.size[2] - G; ++k) {
                Int id = index(i, j, k, S.size);
                for (int m = 0; m < 3; ++m) S.cfd.U[id][m] = U0[3 * id + m];
            }
}

// Start-up pass, src/solver.h:767-833, then the step-0 pass of main_loop_once.
void ref_solver_initialize(void *h) {
    Solver &S = static_cast<RefSolver *>(h)->s;
    auto &cfd = S.cfd;
    auto &mesh = S.mesh;
    set_solid_U(cfd.U, cfd.solid, S.size, S.gc);
    apply_Ubc(cfd.U, cfd.Uold, cfd.Uin, mesh.x, mesh.y, mesh.z, mesh.dx, mesh.dy, mesh.dz, S.rt.dt,
              S.size, S.gc, &S.mpi, cfd.retracted_floor, cfd.floor_start_x);
    interpolate_JU(cfd.U, cfd.JU, mesh.dx, mesh.dy, mesh.dz, S.size, S.gc, &S.mpi);
    apply_JUbc(cfd.JU, cfd.Uin, mesh.dy, mesh.dz, S.size, S.gc, &S.mpi);
    calc_nut(cfd.U, cfd.nut, mesh.x, mesh.y, mesh.z, mesh.dx, mesh.dy, mesh.dz, cfd.Cs, S.size, S.gc,
             &S.mpi);
    calc_divergence(cfd.JU, cfd.div, mesh.dx, mesh.dy, mesh.dz, S.size, S.gc, &S.mpi);
    run_alm(cfd.U, cfd.f, S.alm.projection_width, mesh.x, mesh.y, mesh.z, S.alm.wt_lst,
            S.alm.wt_count, S.alm.ap_lst, S.alm.ap_count, S.alm.blade_per_wt, S.alm.ap_per_blade,
            S.alm.cd_tbl, S.alm.cl_tbl, S.alm.atk_lst, S.alm.atk_count, S.rt.get_time(), S.size,
            S.gc, &S.mpi);
    Int nint = (S.gsize[0] - 2 * S.gc) * (S.gsize[1] - 2 * S.gc) * (S.gsize[2] - 2 * S.gc);
    cfd.avg_div = calc_l2_norm(cfd.div, S.size, S.gc, &S.mpi) / sqrt(nint);
    cfd.max_cfl = calc_max_cfl(cfd.U, mesh.dx, mesh.dy, mesh.dz, S.rt.dt, S.size, S.gc, &S.mpi);
    S.main_loop_once();  // step 0: no physics, advances the counter
}

void ref_solver_step(void *h) { static_cast<RefSolver *>(h)->s.main_loop_once(); }

// name: U Uold JU f Utavg (L*3), p nut div b q (L)
int ref_solver_get(void *h, const char *name, double *out) {
    Solver &S = static_cast<RefSolver *>(h)->s;
    Int len = S.size[0] * S.size[1] * S.size[2];
    std::string n(name);
    const double *src = nullptr;
    Int w = 1;
    if (n == "U") { src = S.cfd.U[0]; w = 3; }
    else if (n == "Uold") { src = S.cfd.Uold[0]; w = 3; }
    else if (n == "JU") { src = S.cfd.JU[0]; w = 3; }
    else if (n == "f") { src = S.cfd.f[0]; w = 3; }
    else if (n == "p") src = S.cfd.p;
    else if (n == "nut") src = S.cfd.nut;
    else if (n == "div") src = S.cfd.div;
    else if (n == "b") src = S.eq.b;
    else if (n == "q") src = S.cfd.q;
    else if (n == "solid") src = S.cfd.solid;
    else if (n == "Utavg") { src = S.cfd.Utavg[0]; w = 3; }
    else return -1;
    std::memcpy(out, src, sizeof(double) * len * w);
    return 0;
}

// stats: step just integrated, it, err, avg_div, max_cfl, max_diag; torque/thrust per turbine
void ref_solver_stats(void *h, i64 *step, i64 *it, double *vals, double *torque, double *thrust) {
    Solver &S = static_cast<RefSolver *>(h)->s;
    *step = S.rt.step - 1;
    *it = S.eq.it;
    vals[0] = S.eq.err;
    vals[1] = S.cfd.avg_div;
    vals[2] = S.cfd.max_cfl;
    vals[3] = S.eq.max_diag;
    for (Int w = 0; w < S.alm.wt_count; ++w) {
        if (torque) torque[w] = S.alm.wt_lst[w].torque;
        if (thrust) thrust[w] = S.alm.wt_lst[w].thrust;
    }
}

void ref_solver_destroy(void *h) {
    RefSolver *R = static_cast<RefSolver *>(h);
    R->s.cfd.finalize(R->s.size);
    R->s.eq.finalize(R->s.size);
    delete[] R->s.mesh.x; delete[] R->s.mesh.y; delete[] R->s.mesh.z;
    delete[] R->s.mesh.dx; delete[] R->s.mesh.dy; delete[] R->s.mesh.dz;
    delete[] R->s.gmesh.x; delete[] R->s.gmesh.y; delete[] R->s.gmesh.z;
    delete[] R->s.gmesh.dx; delete[] R->s.gmesh.dy; delete[] R->s.gmesh.dz;
    delete R;
}

}  // extern "C"
