/*
 * oracle.h -- CPU restatement (FP64, serial, plain C) of the 3dalm per-time-step
 * hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path may link, load or call
 * this library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it, and only as the checker (or as the timed CPU baseline).
 *
 * Conventions follow the reference's boundary format (src/type.h:5-6,
 * src/util.h:10-13): FP64 values, int64 indices, padded extents n[3] that include
 * gc ghost layers per side, linear index i*Ny*Nz + j*Nz + k (z fastest), vector
 * fields stored as AoS triplets double[L][3].
 *
 * Every function cites the reference routine it restates.  Arithmetic is written
 * operation-for-operation in the reference's evaluation order and the library is
 * compiled with -ffp-contract=off, so results are bit-identical to a g++ -O2 x86-64
 * build of the reference wherever libm is not involved.
 *
 * Parity status of each function (how it is pinned) is recorded in the header
 * comment of oracle.c.
 */
#ifndef DALM_ORACLE_H
#define DALM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t i64;

/* Grid: padded extents, ghost width, 1-D cell centres and widths (with ghosts). */
typedef struct {
    i64 n[3];
    i64 gc;
    const double *c[3];  /* x, y, z cell centres */
    const double *d[3];  /* dx, dy, dz cell widths */
} or_grid;

/* ---- mesh ---------------------------------------------------------------- */
void or_mesh_widths(const double *c, i64 n, double *d);

/* ---- predictor / flux / rhs ---------------------------------------------- */
void or_predictor(const or_grid *g, double *U, const double *Uold, const double *JU,
                  const double *nut, const double *f, double Re, double dt);
void or_interpolate_JU(const or_grid *g, const double *U, double *JU);
void or_apply_JUbc(const or_grid *g, double *JU, const double Uin[3]);
void or_poisson_rhs(const or_grid *g, const double *JU, double *b, double dt, double scale);

/* ---- Poisson operator and solvers ---------------------------------------- */
double or_build_A(const or_grid *g, double *A /* [L][7] */);
double or_dot(const or_grid *g, const double *a, const double *b);
double or_l2_norm(const or_grid *g, const double *v);
void or_residual(const or_grid *g, const double *A, const double *x, const double *b, double *r);
void or_Ax(const or_grid *g, const double *A, const double *x, double *y);
void or_sor_sweep(const or_grid *g, const double *A, double *x, const double *b,
                  double omega, i64 color);
void or_jacobi_sweep(const or_grid *g, const double *A, double *x, const double *xold,
                     const double *b);
void or_run_sor(const or_grid *g, const double *A, double *x, const double *b, double *r,
                double omega, i64 *it, i64 max_it, double *err, double tol);
/* work: 9 vectors of length L (r, r0, p, pp, q, s, ss, t, tmp), caller zero-fills. */
void or_run_bicgstab(const or_grid *g, const double *A, double *x, const double *b,
                     double *work, i64 pc_kind /* 0 Jacobi, 1 SOR */, double pc_omega,
                     i64 pc_it, i64 *it, i64 max_it, double *err, double tol);

/* ---- projection, boundary conditions, post ------------------------------- */
void or_apply_pbc(const or_grid *g, double *p);
void or_project(const or_grid *g, double *U, double *JU, const double *p, double dt);
void or_set_solid(const or_grid *g, double *U, const double *solid);
void or_apply_Ubc(const or_grid *g, double *U, const double *Uold, const double Uin[3],
                  double dt, i64 retracted_floor, double floor_start_x);
void or_nut(const or_grid *g, const double *U, double *nut, double Cs);
void or_divergence(const or_grid *g, const double *JU, double *div);
double or_max_cfl(const or_grid *g, const double *U, double dt);

/* ---- turbulent inflow (not in the reference; see oracle.c) --------------- */
void or_inflow_eval(const or_grid *g, const double *modes /* [n][7] */, i64 n_modes, const double Uin[3],
                    double shift, double *cells, double *face);
void or_apply_inflow_U(const or_grid *g, double *U, const double *cells);
void or_apply_inflow_JU(const or_grid *g, double *JU, const double *face);

/* ---- output path --------------------------------------------------------- */
void or_calc_q(const or_grid *g, const double *U, double *q);
/* turbulence-generating grid: solid fraction over the interior (plc = placement) */
void or_tgg_solid(const or_grid *g, double *solid, double thick, double mesh, double bar, const double plc[3]);
void or_time_average(const or_grid *g, double *Utavg, const double *U, double weight);

/* ---- actuator-line model ------------------------------------------------- */
/* Turbine record: 16 doubles
 *   [0..2] base, [3..5] rot_center, [6] angle_type (0 roll,1 pitch,2 yaw,3 none),
 *   [7] angle, [8] ow, [9..12] formula A,B,C,D, [13] rot_speed, [14] torque, [15] thrust
 * Actuator point record: 16 doubles
 *   [0..2] xyz, [3] r, [4] chord, [5] twist, [6] theta, [7] dr, [8..10] f,
 *   [11] torque, [12] thrust, [13] pitch (collective, degrees; 0 in the reference), [14..15] unused
 */
enum { OR_WT_STRIDE = 16, OR_AP_STRIDE = 16 };

void or_build_ap_props(const double *af_r, const double *af_chord, const double *af_twist,
                       const double *af_cd, const double *af_cl, i64 af_count, i64 atk_count,
                       double R, double hub_R, i64 blades, i64 wt_count, i64 ap_per_blade,
                       double *ap, double *cd_tbl, double *cl_tbl);
void or_update_ap_position(double *wt, i64 wt_count, double *ap, i64 ap_count,
                           i64 blade_per_wt, i64 ap_per_blade, double t);
void or_calc_ap_force(const or_grid *g, const double *U, double *wt, i64 wt_count,
                      double *ap, i64 ap_count, i64 blade_per_wt, i64 ap_per_blade,
                      const double *cd_tbl, const double *cl_tbl, const double *atk,
                      i64 atk_count);
void or_distribute_ap_force(const or_grid *g, double *f, const double *ap, i64 ap_count,
                            double width);

/* ---- discrete adjoint (adjoint.c) ---------------------------------------- *
 * Reverse mode of the routines above; "_b" arrays are bars (sensitivities) and
 * accumulate (+=) unless stated.  Not in the reference: pinned by central
 * differences of the forward routines (tests/test_oracle_adjoint.py). */
void or_predictor_adj(const or_grid *g, const double *Ub, const double *Uold, const double *JU,
                      const double *nut, double Re, double dt, double *Uold_b, double *JU_b,
                      double *nut_b, double *f_b);
void or_interpolate_JU_adj(const or_grid *g, const double *JU_b, double *U_b);
void or_apply_JUbc_adj(const or_grid *g, double *JU_b);
void or_poisson_rhs_adj(const or_grid *g, const double *b_b, double *JU_b, double dt, double scale);
void or_apply_pbc_adj(const or_grid *g, double *p_b);
void or_project_adj(const or_grid *g, const double *U_b, const double *JU_b, double *p_b, double dt);
void or_set_solid_adj(const or_grid *g, double *U_b, const double *solid);
void or_apply_Ubc_adj(const or_grid *g, double *U_b, const double *U_pre, const double *Uold,
                      double *Uold_b, double dt, i64 retracted_floor, double floor_start_x);
void or_nut_adj(const or_grid *g, const double *U, const double *nut_b, double *U_b, double Cs);
void or_alm_adj(const or_grid *g, const double *U, double *wt, i64 wt_count, double *ap,
                i64 ap_count, i64 blade_per_wt, i64 ap_per_blade, const double *cd_tbl,
                const double *cl_tbl, const double *atk, i64 atk_count, double width, double t,
                const double *f_b, const double *tq_w, double *U_b, double *ap_grad);
void or_poisson_adj(const or_grid *g, const double *A, const double *p_b, double *y, double *b_b,
                    double *work, double tol_rel, i64 max_it, i64 *it, double *err);

/* ---- whole step ---------------------------------------------------------- */
typedef struct {
    /* fields (caller-owned, length L or L*3) */
    double *U, *Uold, *JU, *p, *nut, *div, *f, *solid;
    /* Poisson */
    double *A, *b, *work;  /* work: 9*L */
    /* parameters */
    double Uin[3];
    double Re, Cs, dt;
    i64 method;      /* 0 SOR, 1 BiCGStab */
    i64 pc_kind;     /* 0 Jacobi, 1 SOR */
    double omega, pc_omega;
    i64 pc_it, max_it;
    double tol;
    i64 retracted_floor;
    double floor_start_x;
    /* ALM */
    double *wt, *ap, *cd_tbl, *cl_tbl, *atk;
    i64 wt_count, ap_count, blade_per_wt, ap_per_blade, atk_count;
    double width;
    /* state / outputs */
    double max_diag;
    i64 step, it;
    double err, avg_div, max_cfl;
    /* turbulent inflow (inflow_n = 0: the reference's constant Uin); the cells and
       face buffers are caller-owned, gc*ny*nz*3 and ny*nz doubles */
    const double *inflow_modes;
    i64 inflow_n;
    double inflow_speed;
    double *inflow_cells, *inflow_face;
} or_state;

void or_state_initialize(const or_grid *g, or_state *s);
void or_state_step(const or_grid *g, or_state *s);
/* forward `steps` steps keeping every state, J = sum_n sum_w tq_w[w] torque_w(n),
 * then the reverse sweep; ap_grad is [ap_count][4], state_b (optional) 10*L bars of
 * the initial U, JU, nut, f (adjoint.c) */
void or_adjoint_run(const or_grid *g, or_state *s, i64 steps, const double *tq_w, double adj_tol,
                    double *J, double *ap_grad, double *state_b);

#ifdef __cplusplus
}
#endif
#endif
