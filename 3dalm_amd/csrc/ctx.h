// ctx.h -- the device context behind the C ABI (include/dalm_hip.h).
//
// One context owns one padded box of fields in HBM, the 1-D mesh metrics, the
// matrix-free Poisson coefficients, the actuator-line tables, reduction scratch
// and one HIP stream.  The reference keeps the same state in its Cfd / Eq / Mesh /
// Alm structs (src/solver.h:207-517); here it lives on the GPU for the whole run and
// only scalars cross to the host each step.
#pragma once

#include <string>
#include <vector>

#include "common.h"
#include "../../include/dalm_hip.h"

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return dalm_fail(ctx, e_, #expr, __FILE__, __LINE__); \
    } while (0)

// The fused iterations (rows of 256 cells) update the FP64 pressure once per this many
// iterations (FP64 drop-in: kXGroup; precision form: kXGroupF), with the updates of all
// of them in turn; their z vectors rotate through as many buffers.
constexpr int kXGroup = 8;
constexpr int kXGroupF = 8;

// Device-resident scalar state of one Poisson solve (read by every kernel of the
// solve; written only by the last block of a reduction).
struct SolveState {
    double rho, rho_prev, alpha, omega, beta;
    double err;
    double tol;
    double root_n;
    double aux0, aux1;
    long long it, max_it;
    int done;       // 1 once the stopping test (or a breakdown) has ended the solve
    int status;     // DALM_OK, DALM_NONCONVERGED, DALM_BREAKDOWN, DALM_NONFINITE
    int ru;         // precision form: a residual replacement is due
    int pending;    // precision form, fused iteration: the last iteration's update of
                    // the residual (r = s - omega t) is still to be formed
    int xdef;       // precision form, fused iteration: the pressure updates of the last
                    // xdef iterations are outstanding (their alpha_d, omega_d, oldest
                    // first); kg_apply_pending4 applies them once there are kXGroupF
    double maxr;    // precision form: largest |r| since the last replacement
    long long n_ru; // precision form: replacements so far
    double alpha_d[kXGroupF], omega_d[kXGroupF];
    // FP64 BiCGStab: alpha and omega of the last kXGroup iterations, oldest first (the
    // fused iteration's x updates are applied kXGroup at a time)
    double alpha_h[kXGroup], omega_h[kXGroup];
};

// Precision form: the recursive FP32 residual is replaced by the true FP64 one
// once its norm falls below this fraction of the largest norm since the last
// replacement.
constexpr double kMixedReliableDelta = 0.1;

// Per-step quantities returned to the host in one copy.
struct StepScalars {
    double avg_div_sq;  // sum of div^2 over the interior
    double max_cfl;
    double torque[DALM_MAX_TURBINES];
    double thrust[DALM_MAX_TURBINES];
};

// 1-D Poisson coefficients reproducing build_A (src/eq.h:701-798) bit for bit:
// the normalised off-diagonals are 1-D in one index each; the diagonal is
// accumulated per cell from 1-D pieces in the reference's order and divided by
// max|a_c|, so no 3-D operator array is stored or read.
struct PoissonCoef {
    double* ae_n;   // [n0] a_e / max_diag (0 at the last interior i)
    double* aw_n;   // [n0]
    double* an_n;   // [n1]
    double* as_n;   // [n1]
    double* at_n;   // [n2]
    double* ab_n;   // [n2]
    double* acx;    // [n0] (0 - a_e|2/dx^2) - a_w   (unnormalised)
    double* an_u;   // [n1] unnormalised a_n, a_s, a_t, a_b
    double* as_u;
    double* at_u;
    double* ab_u;
    double max_diag;
};

// Device view of the Poisson coefficients passed to kernels by value.
struct PCoef {
    const double *ae, *aw, *an, *as, *at, *ab, *acx, *anu, *asu, *atu, *abu;
    double max_diag;
};

// FP32 metric reciprocals of the precision form (1-D device arrays): 1/width,
// 1/(c[i+1]-c[i]), 1/(c[i]-c[i-1]), 1/(c[i+1]-c[i-1]) per direction, and widths.
struct MetricF {
    const float *rw[3];
    const float *rp[3];
    const float *rm[3];
    const float *r2[3];
    const float *w[3];
};

// FP32 pieces of the normalised operator for the inner Poisson sweeps.
struct PCoefF {
    const float *ae, *aw, *an, *as, *at, *ab;  // normalised off-diagonals
    const float *acx, *anu, *asu, *atu, *abu;  // normalised diagonal pieces
};

struct AlmTables {
    int wt_count = 0, ap_count = 0, blade_per_wt = 0, ap_per_blade = 0, atk_count = 0;
    double width = 0;
    double* wt = nullptr;     // [wt][DALM_WT_STRIDE]
    double* ap = nullptr;     // [ap][DALM_AP_STRIDE]
    double* cd = nullptr;     // [ap][atk]
    double* cl = nullptr;     // [ap][atk]
    double* atk = nullptr;    // [atk]
    std::vector<double> h_wt; // host copy of turbine records (angles are set on the host)
    std::vector<double> h_rmax; // largest AP radius per turbine
    int* box_dev = nullptr;   // unused (kept for layout stability)
    std::vector<int> prev_box; // [wt][6] i0,i1,j0,j1,k0,k1 (half-open) boxes written last step
    bool f_cleared = false;    // interior force zeroed once before the first projection
};

// Event timing of every launch of one selected kernel on the context's stream
// (dalm_probe_kernel); pairs are resolved and recycled at host sync points.
struct KernelProbe {
    int kernel = 0;
    std::vector<hipEvent_t> ev;
    size_t used = 0;               // events recorded (two per launch)
    // Solver iterations are enqueued in chunks ahead of the device's stopping
    // test, so launches past it return at once; they are tagged with their
    // iteration and left out of the totals once the solve's count is known.
    std::vector<long long> tag;    // per recorded launch: solver iteration, -1 outside a solve
    std::vector<char> idle;        // per recorded launch: enqueued past the stopping point
    std::vector<int> count;        // per event pair: launches it brackets
    long long cur_tag = -1;        // iteration being enqueued by the running solve
    size_t solve_first = 0;        // first launch recorded by the running solve
    double total_ms = 0;
    long long launches = 0;
};

// Discrete adjoint of the FP64 step (adjoint.hip): bars of the state, the
// checkpoint / pressure tape of a recorded forward run, and its objective.
struct AdjointState {
    bool have_buffers = false;
    // bars of the step's state; "n" arrays receive the bars of the previous step
    double *Ub[3] = {}, *Ubn[3] = {}, *JUb[3] = {}, *JUbn[3] = {}, *fb[3] = {}, *fbn[3] = {};
    double *nutb = nullptr, *nutbn = nullptr;
    double *pb = nullptr;   // bar of p (interior and first ghost layer)
    double *ab = nullptr;   // adjoint Poisson right-hand side V^-1 p_b, then bar of b
    double *ya = nullptr;   // adjoint Poisson solution (warm start for the next step)
    double *xp = nullptr;   // [4][3][ny*nz] x+ outflow bars of the previous velocity
    double *apred = nullptr;  // [ap][6] per-AP projection sums
    double *apg = nullptr;    // [ap][4] dJ/d(turbine angle, pitch, chord, twist) per AP
    int ap_cap = 0;           // APs the two arrays above hold
    double *tqw = nullptr;    // [wt] objective weights on the torques
    // recorded forward run
    bool taping = false;
    // a step of the recording failed (non-finite state or a HIP error): the tape
    // cannot be replayed and dalm_adjoint_run refuses it
    bool tape_bad = false;
    long long steps = 0, every = 0, recorded = 0, start_step = 0;
    // precision-form recording (dalm_adjoint_begin_from): this FP64 context holds
    // the tape of `source`, a DALM_MIXED context whose dalm_step calls record into
    // it; `sink` is the reverse link, set on the source
    dalm_ctx* source = nullptr;
    dalm_ctx* sink = nullptr;
    // the tape came from a precision-form context (set by dalm_adjoint_begin_from
    // until the next recording): its adjoint solves and, while `replaying`, the
    // segment replays' pressure solves take the precision form's solver
    bool from_mixed = false;
    bool replaying = false;
    bool keep_p = false;
    double* ck = nullptr;     // checkpoints [steps/every + 1][11][L]: U, JU, p, nut, f
    double* ptape = nullptr;  // pressure after each step [steps][L] (keep_p)
    double* seg = nullptr;    // states of one segment [every + 1][7][L]: U, JU, nut
    double* fin = nullptr;    // final state [11][L], restored after the reverse sweep
    std::vector<double> h_tqw;
    double J = 0;
    // statistics of the last reverse sweep
    long long adj_it = 0;
    double replay_ms = 0, reverse_ms = 0;
};

struct dalm_ctx {
    Dims d;
    long long L = 0, N = 0;
    int precision = DALM_FP64;
    hipStream_t stream = nullptr;
    int grid = 0;   // blocks for the cell kernels (tile loop)
    int ncu = 256;  // compute units of an MI355X (sizes the one-block-per-CU fused kernels)
    int sgrid = 0;  // blocks of the column-streaming stencil kernels (= d.nunits)

    // mesh (device)
    double *x = nullptr, *y = nullptr, *z = nullptr, *dx = nullptr, *dy = nullptr, *dz = nullptr;
    std::vector<double> hx, hy, hz, hdx, hdy, hdz;

    // FP64 fields (structure of arrays)
    double* U[3] = {};
    double* Uold[3] = {};
    double* JU[3] = {};
    double* f[3] = {};
    double *p = nullptr, *nut = nullptr, *div = nullptr, *b = nullptr, *solid = nullptr;
    bool has_solid = false;
    // Krylov work vectors
    double *r = nullptr, *r0 = nullptr, *kp = nullptr, *pp = nullptr, *q = nullptr, *s = nullptr,
           *ss = nullptr, *t = nullptr, *tmp = nullptr, *tmp2 = nullptr;
    // the FP64 fused iteration's further z_p / z_s buffers (iteration k uses pair k mod
    // kXGroup, so that a deferred pressure update keeps its z), allocated on first use
    double* ppx[kXGroup - 1] = {};
    double* ssx[kXGroup - 1] = {};
    // iterations of the previous solve per use: 0 SOR, 1 BiCGStab (either form),
    // 2 the FP64 finish of the precision form, 3 the adjoint solve (sizes the
    // chunks enqueued per host check)
    long long prev_it[4] = {0, 0, 0, 0};

    PoissonCoef pc{};
    bool have_A = false;

    // precision form (DALM_MIXED): FP32 cell fields and the FP32 BiCGStab vectors
    // (r, r0, p, q, s, t, preconditioner iterates z, z' and their Jacobi partners);
    // JU, p, b and div stay FP64, and the FP64 Krylov set above serves the FP64
    // finish of a stalled solve
    float* U32[3] = {};
    float* Uold32[3] = {};
    float* f32[3] = {};
    float* nut32 = nullptr;
    float *fr = nullptr, *fr0 = nullptr, *fp = nullptr, *fq = nullptr, *fs = nullptr, *ft = nullptr,
          *fz = nullptr, *fz2 = nullptr, *ftmp = nullptr, *ftmp2 = nullptr;
    // the fused iteration's further z_p / z_s buffers (iteration k uses pair k mod
    // kXGroupF, so that an outstanding pressure update keeps its z), allocated on first use
    float* fzx[kXGroupF - 1] = {};
    float* fz2x[kXGroupF - 1] = {};
    MetricF mf{};
    PCoefF pcf{};
    std::vector<void*> owned_f;   // float 1-D arrays behind mf / pcf
    double* pbest = nullptr;      // best pressure seen by the running solve (its own buffer:
                                  // the FP64 continuation uses tmp / tmp2 as Jacobi spares)
    long long fp32_pass_cap = 0;  // testing aid (dalm_set_fp32_pass_cap): FP32 passes end
                                  // after this many iterations as if stalled; 0 = off
    long long outer_it = 0;       // FP64 true-residual passes of the last mixed solve
    long long fp64_finish_it = 0; // iterations of its FP64 finish (0: FP32 preconditioning sufficed)

    // reduction scratch
    double* partials = nullptr;  // [8][grid]
    unsigned* counter = nullptr;
    SolveState* solve = nullptr;   // device
    SolveState* h_solve = nullptr; // pinned host mirror
    StepScalars* scal = nullptr;   // device
    StepScalars* h_scal = nullptr; // pinned host

    // run parameters (dalm_set_params)
    dalm_params prm{};
    bool have_prm = false;
    long long step = 0;
    dalm_step_stats last{};

    AlmTables alm;
    KernelProbe probe;
    AdjointState adj;

    // turbulent inflow (dalm_set_inflow; not in the reference): the mode table
    // [n][7] and, per step, the x- boundary values it gives -- ghost cells
    // [3][gc][ny][nz] (component-major) then the face velocity [ny][nz]
    double* inflow_modes = nullptr;
    double* inflow = nullptr;
    int inflow_n = 0;
    double inflow_speed = 0;
    const double* inflow_cells() const { return inflow_n ? inflow : nullptr; }
    const double* inflow_face() const {
        return inflow_n ? inflow + 3LL * d.gc * d.ny * d.nz : nullptr;
    }

    // output-path fields (diag.hip), allocated zero-filled on first use: the
    // Q-criterion and the time-averaged velocity (FP64 in both forms), and the
    // probe-cell gather buffers
    double* qcrit = nullptr;
    double* Utavg[3] = {};
    long long* gather_idx = nullptr;
    double* gather_out = nullptr;
    long long gather_cap = 0;


    std::string err;
    int hip_err = 0;

    PCoef coef() const {
        PCoef c;
        c.ae = pc.ae_n; c.aw = pc.aw_n; c.an = pc.an_n; c.as = pc.as_n; c.at = pc.at_n; c.ab = pc.ab_n;
        c.acx = pc.acx; c.anu = pc.an_u; c.asu = pc.as_u; c.atu = pc.at_u; c.abu = pc.ab_u;
        c.max_diag = pc.max_diag;
        return c;
    }
};

// finalize modes of reduce.hip
enum FinMode { FIN_POST = 0, FIN_SUM, FIN_RESID, FIN_SOR, FIN_BI_INIT, FIN_ALPHA, FIN_OMEGA, FIN_BI_XR, FIN_MX_INIT, FIN_MX_XR, FIN_MX_RU, FIN_MX_UP, FIN_MX_SOR, FIN_MX_OMEGA };
int launch_finalize(dalm_ctx* ctx, int mode, int nb, SolveState* st, double* out);
// blocks of the cell kernels (the tile loop) for a grid of ntiles tiles
int cell_grid(int ntiles);
// interior extents, tiles and strides of a padded box (no column units)
Dims make_dims(const int64_t size[3], int64_t gc);

int dalm_fail(dalm_ctx* ctx, hipError_t e, const char* what, const char* file, int line);
int dalm_fail_msg(dalm_ctx* ctx, int status, const std::string& msg);

// kernel launchers (cfd.hip, poisson.hip, alm.hip)
int launch_predictor(dalm_ctx* ctx, double Re, double dt);
int launch_copy_uold(dalm_ctx* ctx);
int launch_ghost_copy(dalm_ctx* ctx);
int launch_inflow(dalm_ctx* ctx, double t);
int launch_flux(dalm_ctx* ctx, const double Uin[3], bool apply_bc);
int launch_jubc(dalm_ctx* ctx, const double Uin[3]);
int launch_rhs(dalm_ctx* ctx, double dt, double scale);
int launch_pbc(dalm_ctx* ctx);
int launch_project(dalm_ctx* ctx, double dt, const double* Uin_bc);
int launch_solid(dalm_ctx* ctx);
int launch_ubc(dalm_ctx* ctx, const double Uin[3], double dt, int retracted, double floor_x);
int launch_nut(dalm_ctx* ctx, double Cs);
int launch_post(dalm_ctx* ctx, double Cs, double dt, bool store_div);
int launch_divergence(dalm_ctx* ctx);
int launch_l2(dalm_ctx* ctx, const double* v, double* out_host);
int launch_max_cfl(dalm_ctx* ctx, double dt, double* out_host);
int run_sor(dalm_ctx* ctx, double omega, long long max_it, double tol, long long* it, double* err);
int run_bicgstab(dalm_ctx* ctx, int pc_kind, double pc_omega, int pc_it, long long max_it, double tol,
                 long long* it, double* err);
// the same solvers on any (x, b) pair; `slot` indexes prev_it
int run_sor_on(dalm_ctx* ctx, double* x, const double* b, int slot, double omega, long long max_it, double tol,
               long long* it, double* err);
// resume: continue the iteration held in the device state and in r, r0, kp, q
// (the precision form's hand-over to FP64) instead of starting one
int run_bicgstab_on(dalm_ctx* ctx, double* x, const double* b, int slot, int pc_kind, double pc_omega, int pc_it,
                    long long max_it, double tol, long long* it, double* err, bool resume = false);
int launch_residual(dalm_ctx* ctx, const double* x, const double* b, double* r);
int launch_Ax(dalm_ctx* ctx, const double* x, double* y);
int launch_sor_sweep(dalm_ctx* ctx, double* x, const double* b, double omega, int color);
int alm_run(dalm_ctx* ctx, double t, bool sync_scalars);
// AP positions and forces at time t from the FP64 velocity U (no projection)
int alm_forces(dalm_ctx* ctx, double t, double* const U[3]);
// one FP64 step; with p_given the Poisson solve is replaced by copying p_given
int step_fp64(dalm_ctx* ctx, dalm_step_stats* stats, const double* p_given);
// adjoint.hip: record a forward step on the tape, free the adjoint memory
// status: the step's dalm_status (OK, BREAKDOWN or NONFINITE)
int adjoint_record(dalm_ctx* ctx, int status);
int adjoint_record_from(dalm_ctx* src, int status);
void adjoint_free(dalm_ctx* ctx);
int ensure_adjoint_buffers(dalm_ctx* ctx);
int sync_stream(dalm_ctx* ctx);
int zero_device(dalm_ctx* ctx, double* a, long long n);
// fused half-iterations of the precision form's BiCGStab (fused_mixed.hip)
bool fused_half_applies(const dalm_ctx* ctx, int pc_kind, int pc_it);
int fused_half_grid(const dalm_ctx* ctx, int* ni);
// planes per unit of a fused kernel that runs one block per compute unit over
// `tiles` j tiles: as many plane chunks as fill the compute units once, chunks of at
// least 16 planes
int fused_planes(const dalm_ctx* ctx, int tiles);
// the fused half-iterations: z_p into zp_out, z_s into zs_out
int launch_half_p(dalm_ctx* ctx, const float* p_in, const float* q_in, float* p_out, float* q_out, float* zp_out);
int launch_half_s(dalm_ctx* ctx, const float* q, float* zs_out);

// the pressure update of a precision-form iteration, x + alpha z_p + omega z_s, in one
// expression wherever it is evaluated (the fused p-half, the replacement and the flush)
__device__ __forceinline__ double x_update(double x, float zp, float zs, double alpha, double omega) {
    return x + alpha * (double)zp + omega * (double)zs;
}
void probe_begin(dalm_ctx* ctx, int kernel);
// launches: back-to-back launches of the kernel between the two events
void probe_end(dalm_ctx* ctx, int kernel, int launches = 1);
void probe_solve_begin(dalm_ctx* ctx);
void probe_iteration(dalm_ctx* ctx, long long i);
void probe_solve_end(dalm_ctx* ctx, long long iterations);
int probe_collect(dalm_ctx* ctx);

// output-path fields (diag.hip)
int ensure_diag_buffers(dalm_ctx* ctx);
void free_diag_buffers(dalm_ctx* ctx);
int launch_calc_q(dalm_ctx* ctx);
int launch_div_store(dalm_ctx* ctx);
int launch_time_average(dalm_ctx* ctx, double weight);

// precision form (cfd_mixed.hip, poisson_mixed.hip, alm.hip)
int setup_mixed(dalm_ctx* ctx);
// an FP64 context that records a precision-form run solves its adjoint Poisson
// equations with the precision form's solver: its FP32 tables and Krylov vectors
int ensure_mixed_solver(dalm_ctx* ctx);
// the fused iterations' further z_p / z_s buffers (fzx, fz2x; FP64: ppx, ssx),
// allocated on first use
int ensure_fused_z(dalm_ctx* ctx);
int ensure_fused_z64(dalm_ctx* ctx);
int launch_ghost_copy_m(dalm_ctx* ctx);
int launch_predictor_m(dalm_ctx* ctx, double Re, double dt);
int launch_flux_rhs_m(dalm_ctx* ctx, const double Uin[3], double dt, double scale, bool rhs);
int launch_project_m(dalm_ctx* ctx, double dt, const double Uin[3]);
int launch_ubc_m(dalm_ctx* ctx, const double Uin[3], double dt, int retracted, double floor_x);
int launch_solid_m(dalm_ctx* ctx);
int launch_post_m(dalm_ctx* ctx, double Cs, double dt, bool store_div);
int run_poisson_mixed(dalm_ctx* ctx, int pc_kind, double pc_omega, int pc_it, long long max_it, double tol,
                      long long* it, double* err);
int run_sor_mixed(dalm_ctx* ctx, double omega, long long max_it, double tol, long long* it, double* err);
