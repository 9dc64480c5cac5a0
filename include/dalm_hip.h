/*
 * dalm_hip.h -- C ABI of the MI355X-native 3dalm per-time-step hot path.
 *
 * One opaque context holds one padded grid (reference layout: n[0] x n[1] x n[2]
 * cells including gc ghost layers per side, z fastest) resident in HBM.  Host
 * arrays crossing this boundary use the reference's host formats: FP64 values,
 * int64 sizes, vector fields as AoS double[L][3] (src/type.h:5-6,
 * src/solver.h:286-317).  Every entry point returns a dalm_status; no exception
 * crosses the boundary.  A context is not thread-safe; calls are stream-ordered on
 * the context's own HIP stream and return after the result is complete unless
 * stated otherwise.
 *
 * Each entry point names the reference routine it replaces (path:line in the
 * reference tree).  The reference's functions are static and void; their solver
 * status came back through it/err out-parameters (src/eq.h:291-296), which is
 * kept here, plus explicit status codes for non-convergence, breakdown and
 * non-finite residuals.
 */
#ifndef DALM_HIP_H
#define DALM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dalm_ctx dalm_ctx;

enum dalm_status {
    DALM_OK = 0,
    DALM_NONCONVERGED = 1, /* Poisson solve stopped at max_it with err > tol */
    DALM_BREAKDOWN = 2,    /* BiCGStab |rho| < FLT_MIN (src/eq.h:350-353) */
    DALM_NONFINITE = 3,    /* residual norm is NaN or Inf */
    DALM_HIP_ERROR = 4,    /* a HIP runtime call failed; see dalm_last_error */
    DALM_BAD_ARG = 5
};

enum dalm_precision {
    DALM_FP64 = 0,  /* drop-in: FP64 everywhere, reference arithmetic order */
    DALM_MIXED = 1  /* FP32 stencils and inner Poisson sweeps, FP64 reductions,
                       FP64 face fluxes and pressure, FP64 iterative refinement */
};

enum dalm_field {
    DALM_F_U = 0,     /* [L][3] cell velocity          (Cfd::U)     */
    DALM_F_UOLD = 1,  /* [L][3] previous velocity      (Cfd::Uold)  */
    DALM_F_JU = 2,    /* [L][3] face fluxes e/n/t      (Cfd::JU)    */
    DALM_F_P = 3,     /* [L]    pressure               (Cfd::p)     */
    DALM_F_NUT = 4,   /* [L]    eddy viscosity         (Cfd::nut)   */
    DALM_F_DIV = 5,   /* [L]    divergence             (Cfd::div)   */
    DALM_F_F = 6,     /* [L][3] ALM body force         (Cfd::f)     */
    DALM_F_SOLID = 7, /* [L]    solid fraction         (Cfd::solid) */
    DALM_F_B = 8,     /* [L]    Poisson right-hand side (Eq::b)     */
    /* sensitivities (bars) of the discrete adjoint, FP64 contexts only; the
       "OLD" bars are what the predictor's adjoint produces for the previous step */
    DALM_F_U_B = 9,       /* [L][3] bar of U                        */
    DALM_F_UOLD_B = 10,   /* [L][3] bar of Uold                     */
    DALM_F_JU_B = 11,     /* [L][3] bar of JU                       */
    DALM_F_JUOLD_B = 12,  /* [L][3] bar of the previous step's JU   */
    DALM_F_NUT_B = 13,    /* [L]    bar of nu_t                     */
    DALM_F_NUTOLD_B = 14, /* [L]    bar of the previous step's nu_t */
    DALM_F_F_B = 15,      /* [L][3] bar of f                        */
    DALM_F_FOLD_B = 16,   /* [L][3] bar of the previous step's f    */
    DALM_F_P_B = 17,      /* [L]    bar of p                        */
    DALM_F_B_B = 18,      /* [L]    bar of b (Poisson right-hand side) */
    /* output-path fields, FP64 in both forms, zero until first computed */
    DALM_F_Q = 19,        /* [L]    Q-criterion            (Cfd::q)     */
    DALM_F_UTAVG = 20     /* [L][3] time-averaged velocity (Cfd::Utavg) */
};

enum dalm_method { DALM_SOR = 0, DALM_BICGSTAB = 1 };
enum dalm_pc { DALM_PC_JACOBI = 0, DALM_PC_SOR = 1 };

/* Turbine / actuator-point records (flat FP64 rows).
 * turbine: [0..2] base, [3..5] rotation centre, [6] angle type (0 roll, 1 pitch,
 *          2 yaw, 3 none), [7] angle, [8] angular rate, [9..12] angle formula
 *          A,B,C,D (angle = A sin(B t + C) + D), [13] rotor speed, [14] torque,
 *          [15] thrust                      (WindTurbine, src/alm.h:23-33)
 * point:   [0..2] position, [3] r, [4] chord, [5] twist (deg), [6] theta, [7] dr,
 *          [8..10] force, [11] torque, [12] thrust, [13] collective pitch (deg,
 *          0 in the reference), [14..15] unused (ActuatorPoint, src/alm.h:10-21) */
#define DALM_WT_STRIDE 16
#define DALM_AP_STRIDE 16
#define DALM_MAX_TURBINES 64

/* Per-run parameters (setup.json keys, src/solver.h:589-652). */
typedef struct {
    double Uin[3];         /* inflow.value */
    double Re, Cs;         /* cfd.Re, cfd.Cs */
    double dt;             /* runtime.dt */
    int64_t method;        /* eq.method: DALM_SOR | DALM_BICGSTAB */
    int64_t pc_kind;       /* eq.preconditioner.method */
    double omega;          /* eq.relaxation rate (SOR) */
    double pc_omega;       /* eq.preconditioner.relaxation rate */
    int64_t pc_it;         /* eq.preconditioner.max iteration */
    int64_t max_it;        /* eq.max iteration */
    double tol;            /* eq.tolerance */
    int64_t retracted_floor;
    double floor_start_x;  /* "floor start" */
} dalm_params;

/* One line of the reference's per-step log (src/solver.h:1188-1190) plus status
 * and per-turbine torque / thrust / power (power = rotor speed * torque). */
typedef struct {
    int64_t step;
    int64_t it;
    double time;
    double err;
    double avg_div;
    double max_cfl;
    int64_t status;
    int64_t wt_count;
    double torque[DALM_MAX_TURBINES];
    double thrust[DALM_MAX_TURBINES];
    double power[DALM_MAX_TURBINES];
} dalm_step_stats;

/* ---- context ---------------------------------------------------------------- */

/* Allocate a context for a padded grid with cell centres x[size[0]], y[size[1]],
 * z[size[2]] (ghosts included).  Widths follow the reference's mesh rule.
 * Replaces Mesh/Cfd/Eq allocation and `acc enter data` (src/solver.h:207-404,
 * src/io.h:98-164). */
int dalm_create(dalm_ctx **out, const int64_t size[3], int64_t gc, const double *x,
                const double *y, const double *z, int64_t precision);
/* Free the context and its device memory: the `finalize` methods' `acc exit data`
 * (src/solver.h:272-280, 328-341, 386). */
int dalm_destroy(dalm_ctx *ctx);
/* Human-readable text of the last failure on this context ("" if none). */
const char *dalm_last_error(const dalm_ctx *ctx);
/* Device selection for subsequent dalm_create calls (src/solver.h:560-562). */
int dalm_set_device(int64_t device);
/* Number of visible GPUs: acc_get_num_devices(acc_device_nvidia)
 * (src/solver.h:560-562), printed in the start-up report. */
int dalm_device_count(int64_t *n);

/* Mesh width rule, dx[i] = (x[i+1]-x[i-1])/2 with one-sided 2nd-order ends
 * (build_mesh_from_json, src/io.h:115-125).  Host only. */
int dalm_mesh_widths(const double *c, int64_t n, double *d);

/* Host <-> device copies of a whole field in the reference's AoS layout
 * (`acc update host/device`, src/io.h:89-95).  In DALM_MIXED, FP32 fields are
 * converted on the way. */
int dalm_upload(dalm_ctx *ctx, int64_t field, const double *host);
int dalm_download(dalm_ctx *ctx, int64_t field, double *host);

int dalm_set_params(dalm_ctx *ctx, const dalm_params *prm);

/* ---- per-kernel entry points (FP64 context) --------------------------------- */

/* cpy_array(Uold, U) over the padded length (src/util.h:36-57) */
int dalm_copy_u_to_uold(dalm_ctx *ctx);
/* calc_intermediate_U (src/cfd.h:244-392): MUSCL convection + diffusion + force */
int dalm_predictor(dalm_ctx *ctx, double Re, double dt);
/* interpolate_JU (src/cfd.h:422-567) */
int dalm_interpolate_JU(dalm_ctx *ctx);
/* apply_JUbc (src/bc.h:235-302) */
int dalm_apply_JUbc(dalm_ctx *ctx, const double Uin[3]);
/* calc_poisson_rhs (src/cfd.h:569-598) */
int dalm_poisson_rhs(dalm_ctx *ctx, double dt, double scale);
/* build_A (src/eq.h:701-798): returns max|a_c|; the operator is kept matrix-free */
int dalm_build_A(dalm_ctx *ctx, double *max_diag);
/* run_sor (src/eq.h:291-306) on x = p, b = rhs */
int dalm_run_sor(dalm_ctx *ctx, double omega, int64_t max_it, double tol, int64_t *it,
                 double *err);
/* run_pbicgstab, Jacobi PC (src/eq.h:333-402) or SOR PC (src/eq.h:404-473) */
int dalm_run_bicgstab(dalm_ctx *ctx, int64_t pc_kind, double pc_omega, int64_t pc_it,
                      int64_t max_it, double tol, int64_t *it, double *err);
/* calc_residual r = b - A p (src/mv.h:90-182), result in field B's twin: r is
 * written to `out` (host, [L]) when out != NULL */
int dalm_residual(dalm_ctx *ctx, double *out);
/* apply_pbc (src/bc.h:304-411) */
int dalm_apply_pbc(dalm_ctx *ctx);
/* project_p (src/cfd.h:620-783) */
int dalm_project_p(dalm_ctx *ctx, double dt);
/* set_solid_U (src/cfd.h:1143-1160) */
int dalm_set_solid_U(dalm_ctx *ctx);
/* set_turbulence_generating_grid (src/solver.h:52-159): the solid fraction of a
 * grid of bars (setup.json "turbulence grid": thick, mesh, bar, placement) into the
 * interior of DALM_F_SOLID, either form; set_solid_U then runs every step. */
int dalm_set_turbulence_grid(dalm_ctx *ctx, double thick, double mesh, double bar,
                             const double placement[3]);
/* apply_Ubc (src/bc.h:5-233) */
int dalm_apply_Ubc(dalm_ctx *ctx, const double Uin[3], double dt, int64_t retracted_floor,
                   double floor_start_x);
/* calc_nut (src/cfd.h:827-946) */
int dalm_calc_nut(dalm_ctx *ctx, double Cs);
/* calc_divergence (src/cfd.h:1079-1107): stores the field DALM_F_DIV (either form) */
int dalm_calc_divergence(dalm_ctx *ctx);
/* calc_l2_norm (src/mv.h:33-55) of a scalar field */
int dalm_l2_norm(dalm_ctx *ctx, int64_t field, double *out);
/* calc_max_cfl (src/cfd.h:1109-1141) */
int dalm_max_cfl(dalm_ctx *ctx, double dt, double *out);


/* ---- turbulent inflow ------------------------------------------------------------ *
 * No counterpart in the reference, whose x- inflow is the constant Uin
 * (apply_Ubc src/bc.h:16-30, apply_JUbc src/bc.h:243-257).  With n_modes > 0 the
 * x- ghost cells and the x- face flux take frozen turbulence convected at `speed`:
 * Uin + sum_n a_n cos(kx (x - speed t) + ky y + kz z + psi_n), modes [n][7] =
 * (kx, ky, kz, psi, ax, ay, az), at the ghost cell centres and at the face
 * x_f = (x[gc-1] + x[gc]) / 2, with Uin from dalm_set_params.  dalm_step and
 * dalm_initialize evaluate it at each step's time level; the per-kernel entry
 * points (dalm_apply_Ubc, dalm_apply_JUbc, dalm_interpolate_JU's callers) use the
 * values of the last evaluation, made explicit by dalm_eval_inflow.  n_modes = 0
 * restores the constant inflow (the default).  At most 4096 modes. */
int dalm_set_inflow(dalm_ctx *ctx, const double *modes, int64_t n_modes, double speed);
int dalm_eval_inflow(dalm_ctx *ctx, double t);

/* ---- output path (diagnostics) ---------------------------------------------------- */

/* calc_q (src/cfd.h:948-1077): Q-criterion of U over the interior into the field
 * DALM_F_Q (either form). */
int dalm_calc_q(dalm_ctx *ctx);
/* The time average of Solver::main_loop_once (src/solver.h:1194-1204):
 * Utavg = (1 - weight) Utavg + weight U over every padded cell, the reference's
 * calc_ax_plus_by(1. - weight, Utavg, weight, U, Utavg, len) (src/mv.h:321-335);
 * weight = 1 / (number of averaged steps), in (0, 1]. */
int dalm_time_average(dalm_ctx *ctx, double weight);
/* U at n padded linear cell indices (src/util.h:10-13) into out[n][3]: the probe
 * records' `acc update host(cfd.U[id:1])` (src/solver.h:1205-1216), one gather and
 * one copy for all probes. */
int dalm_sample_U(dalm_ctx *ctx, const int64_t *cell, int64_t n, double *out);

/* ---- actuator-line model ------------------------------------------------------ */

/* build_ap_props (src/alm.h:35-142), host only: AP records (r, dr, chord, twist)
 * and per-AP Cd/Cl rows interpolated between airfoil stations by absolute r.
 * af_cd/af_cl: [af_count][atk_count]; ap: [n][DALM_AP_STRIDE]; cd/cl: [n][atk_count]
 * with n = wt_count*blades*ap_per_blade. */
int dalm_build_ap_props(const double *af_r, const double *af_chord, const double *af_twist,
                        const double *af_cd, const double *af_cl, int64_t af_count,
                        int64_t atk_count, double R, double hub_R, int64_t blades,
                        int64_t wt_count, int64_t ap_per_blade, double *ap, double *cd,
                        double *cl);
/* Upload turbines, points and tables (Alm::intialize + acc copyin,
 * src/solver.h:475-493). */
int dalm_alm_setup(dalm_ctx *ctx, const double *wt, int64_t wt_count, const double *ap,
                   int64_t ap_count, int64_t blade_per_wt, int64_t ap_per_blade,
                   const double *cd, const double *cl, const double *atk, int64_t atk_count,
                   double width);
/* run_alm (src/alm.h:450-479): AP positions at time t, AP forces from U, Gaussian
 * projection into f.  torque/thrust per turbine (may be NULL). */
int dalm_run_alm(dalm_ctx *ctx, double t, double *torque, double *thrust);
/* Current AP / turbine records (`acc update host(ap_lst)`, src/alm.h:380). */
int dalm_alm_download(dalm_ctx *ctx, double *ap, double *wt);

/* ---- instrumentation (no counterpart in the reference) ---------------------- */

/* Kernels that can be timed individually. */
enum dalm_kernel {
    DALM_K_NONE = 0,
    DALM_K_JACOBI = 1,    /* one Jacobi preconditioner sweep          */
    DALM_K_SPMV = 2,      /* BiCGStab SpMV + dot products (q and t)   */
    DALM_K_UPDATE = 3,    /* BiCGStab x/r update + norms              */
    DALM_K_PREDICTOR = 4, /* MUSCL/diffusion predictor                */
    DALM_K_SOR = 5,       /* one red-black SOR colour sweep           */
    DALM_K_RESNORM = 6,   /* SOR residual + norm                      */
    DALM_K_PUPDATE = 7,   /* BiCGStab p update + first PC sweep       */
    DALM_K_SUPDATE = 8,   /* BiCGStab s update + first PC sweep       */
    /* the precision form's kernels of its FP32-preconditioned BiCGStab */
    DALM_K_JACOBI32 = 9,  /* FP32 Jacobi preconditioner sweep         */
    DALM_K_SPMV_MIXED = 10,    /* A z (FP64 arithmetic, FP32 result) + dots */
    DALM_K_UPDATE_MIXED = 11,  /* x (FP64) / r update + norms           */
    DALM_K_PUPDATE_MIXED = 12, /* p update, first PC sweep (FP32)       */
    DALM_K_SUPDATE_MIXED = 13, /* s update, first PC sweep (FP32)      */
    /* output path */
    DALM_K_CALC_Q = 14,        /* Q-criterion                          */
    DALM_K_TIME_AVERAGE = 15,  /* time average of U                    */
    /* the precision form's fused half-iterations (rows of 256 cells, Jacobi x3) */
    DALM_K_HALF_P_MIXED = 16,  /* p update + 3 Jacobi sweeps + A z + <r0, q> */
    DALM_K_HALF_S_MIXED = 17,  /* s update + 3 Jacobi sweeps + A z + dots  */
    /* the FP64 drop-in's fused update + preconditioner (the same rows)       */
    DALM_K_HALF_P = 18,        /* p update + 3 Jacobi sweeps (FP64)        */
    DALM_K_HALF_S = 19         /* s update + 3 Jacobi sweeps (FP64)        */
};
/* Time every launch of `kernel` with HIP events on the context's stream from now
 * on (DALM_K_NONE stops); resets the totals. */
int dalm_probe_kernel(dalm_ctx *ctx, int64_t kernel);
/* Launch count and summed device time (ms) of the probed kernel so far. */
int dalm_probe_read(dalm_ctx *ctx, int64_t *launches, double *total_ms);
/* Testing aid of the precision form: every FP32 pass of the Poisson solve ends
 * after n iterations as if it had stalled (0, the default, switches it off), so
 * that small cases exercise the restarted passes and the FP64 hand-over.  The
 * result still meets the reference's test on the true FP64 residual. */
int dalm_set_fp32_pass_cap(dalm_ctx *ctx, int64_t n);
/* The last step's Poisson solve: out[0] iterations run on FP32 vectors, out[1]
 * iterations run in FP64 (the precision form's FP64 continuation or finish; every
 * iteration in an FP64 context), out[2] residual replacements (precision form). */
int dalm_last_solve_info(dalm_ctx *ctx, int64_t out[3]);

/* ---- whole step ----------------------------------------------------------------- */

/* Start-up pass of Solver::initialize after the fields are filled
 * (src/solver.h:727-833): build_A, set_solid_U, apply_Ubc, interpolate_JU,
 * apply_JUbc, calc_nut, calc_divergence, run_alm(t = 0), avg_div, max_cfl. */
int dalm_initialize(dalm_ctx *ctx, dalm_step_stats *stats);
/* One Solver::main_loop_once (src/solver.h:1034-1190): advances the step counter
 * and integrates it (first call = step 1, t = dt).  DALM_BREAKDOWN and
 * DALM_NONFINITE report a finished step whose solve ended that way; the
 * reference continues after a breakdown (src/eq.h:349-353). */
int dalm_step(dalm_ctx *ctx, dalm_step_stats *stats);
/* n steps back to back, continuing after a breakdown; stats of the last one. */
int dalm_run(dalm_ctx *ctx, int64_t n, dalm_step_stats *stats);

/* ---- differentiation form: discrete adjoint of the FP64 step --------------------
 * No counterpart in the reference (it has no adjoint).  Each *_adj entry point is
 * the transpose of the Jacobian of the forward entry point of the same name,
 * acting on the bar fields above: outputs named "(overwrites)" are written, the
 * others accumulate.  The Poisson solve is differentiated as the exact solve. */
int dalm_predictor_adj(dalm_ctx *ctx, double Re, double dt);
    /* U_B (bar of U*) with values UOLD, JU, NUT -> UOLD_B, JUOLD_B, NUTOLD_B,
       FOLD_B (overwrites) */
int dalm_interpolate_JU_adj(dalm_ctx *ctx);                 /* U_B += I^T JU_B */
int dalm_apply_JUbc_adj(dalm_ctx *ctx);                     /* JU_B = 0 on the boundary faces */
int dalm_poisson_rhs_adj(dalm_ctx *ctx, double dt, double scale);  /* JU_B += R^T B_B */
int dalm_apply_pbc_adj(dalm_ctx *ctx);        /* P_B ghosts folded onto the interior */
int dalm_project_p_adj(dalm_ctx *ctx, double dt);
    /* P_B = -dt (Gc^T U_B + Gf^T JU_B) on the interior and one ghost layer
       (overwrites) */
int dalm_set_solid_U_adj(dalm_ctx *ctx);                    /* U_B *= 1 - solid */
int dalm_apply_Ubc_adj(dalm_ctx *ctx, double dt, int64_t retracted_floor, double floor_start_x);
    /* in place on U_B with values U (before the pass) and UOLD; UOLD_B += */
int dalm_calc_nut_adj(dalm_ctx *ctx, double Cs);            /* U_B += (dnut/dU)^T NUT_B */
int dalm_poisson_adj(dalm_ctx *ctx, double tol_rel, int64_t max_it, int64_t *it, double *err);
    /* B_B = A^-T P_B (overwrites) with the configured solver (dalm_set_params),
       to tol_rel times |V^-1 P_B| */
int dalm_run_alm_adj(dalm_ctx *ctx, double t, const double *tq_w, double *ap_grad);
    /* ALM at time t from U: U_B += bars of F_B and of sum_w tq_w[w] torque_w;
       ap_grad[ap][4] += dJ/d(turbine angle, collective pitch, chord, twist) */

/* Recorded run and reverse sweep.  dalm_adjoint_begin records the current state
 * as checkpoint 0 and sets the objective J = sum over the next `steps` steps of
 * sum_w tq_w[w] torque_w; the next `steps` dalm_step calls record checkpoints
 * every `every` steps (and, with DALM_ADJ_KEEP_PRESSURE, the pressure of every
 * step, so the replay needs no Poisson solve).  dalm_adjoint_run then replays
 * each segment from its checkpoint and runs the reverse steps, and restores the
 * final state.  ap_grad [ap][4] as above; wt_grad [wt][2]: dJ/d(turbine angle
 * offset D), dJ/d(collective pitch).  1 <= steps <= 2^24, every >= 1. */
#define DALM_ADJ_KEEP_PRESSURE 1
typedef struct {
    int64_t steps, every;  /* recorded steps and checkpoint interval */
    int64_t adj_it;        /* adjoint Poisson iterations over the sweep */
    double replay_ms;      /* wall time of the segment replays */
    double reverse_ms;     /* wall time of the reverse steps */
} dalm_adjoint_stats;
int dalm_adjoint_begin(dalm_ctx *ctx, int64_t steps, int64_t every, const double *tq_w, int64_t flags);
/* The precision form's recording.  ctx: an FP64 context of the same grid, turbine
 * set-up and parameters as src, a DALM_MIXED context; the next `steps` dalm_step
 * calls on src record into ctx's tape as dalm_adjoint_begin would: checkpoints of
 * src's state widened to FP64 (its FP32 U, nu_t, f exactly; JU and p as they are),
 * the pressure of every step with DALM_ADJ_KEEP_PRESSURE, and J from src's
 * torques.  dalm_adjoint_run(ctx, ...) then replays each segment in FP64 from
 * those checkpoints (re-solving the pressure unless it was kept) and runs the FP64
 * reverse steps: the discrete adjoint of the FP64 step, linearised along the
 * precision form's trajectory.  ctx cannot step while it records. */
int dalm_adjoint_begin_from(dalm_ctx *ctx, dalm_ctx *src, int64_t steps, int64_t every,
                            const double *tq_w, int64_t flags);
int dalm_adjoint_run(dalm_ctx *ctx, double tol_rel, double *J, double *ap_grad, double *wt_grad,
                     dalm_adjoint_stats *stats);
int dalm_adjoint_end(dalm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
