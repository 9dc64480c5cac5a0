// ctx.hip -- context lifetime, host<->device field copies, matrix-free operator
// set-up, and the C ABI entry points of include/dalm_hip.h.
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>

#include "ctx.h"

int dalm_fail(dalm_ctx* ctx, hipError_t e, const char* what, const char* file, int line) {
    if (ctx) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
        ctx->err = buf;
        ctx->hip_err = (int)e;
    }
    return DALM_HIP_ERROR;
}

int dalm_fail_msg(dalm_ctx* ctx, int status, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return status;
}

int sync_stream(dalm_ctx* ctx) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

void probe_begin(dalm_ctx* ctx, int kernel) {
    KernelProbe& P = ctx->probe;
    if (P.kernel == 0 || P.kernel != kernel) return;
    if (P.used + 2 > P.ev.size()) {
        for (int q = 0; q < 64; ++q) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            P.ev.push_back(e);
        }
    }
    (void)hipEventRecord(P.ev[P.used], ctx->stream);
}

void probe_end(dalm_ctx* ctx, int kernel, int launches) {
    KernelProbe& P = ctx->probe;
    if (P.kernel == 0 || P.kernel != kernel || P.used + 2 > P.ev.size()) return;
    (void)hipEventRecord(P.ev[P.used + 1], ctx->stream);
    P.used += 2;
    P.tag.push_back(P.cur_tag);
    P.idle.push_back(0);
    P.count.push_back(launches);
}

void probe_solve_begin(dalm_ctx* ctx) {
    ctx->probe.solve_first = ctx->probe.used / 2;
    ctx->probe.cur_tag = -1;
}

void probe_iteration(dalm_ctx* ctx, long long i) { ctx->probe.cur_tag = i; }

void probe_solve_end(dalm_ctx* ctx, long long iterations) {
    KernelProbe& P = ctx->probe;
    for (size_t q = P.solve_first; q < P.used / 2; ++q)
        if (P.tag[q] >= iterations) P.idle[q] = 1;
    P.cur_tag = -1;
    P.solve_first = P.used / 2;
}

int probe_collect(dalm_ctx* ctx) {
    KernelProbe& P = ctx->probe;
    if (P.used == 0) return DALM_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (size_t q = 0; q + 1 < P.used; q += 2) {
        if (P.idle[q / 2]) continue;
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, P.ev[q], P.ev[q + 1]));
        P.total_ms += ms;
        P.launches += P.count[q / 2];
    }
    P.used = 0;
    P.solve_first = 0;
    P.tag.clear();
    P.idle.clear();
    P.count.clear();
    return DALM_OK;
}

namespace {

template <typename T>
int dev_alloc(dalm_ctx* ctx, T** p, size_t n) {
    HIP_TRY(hipMalloc((void**)p, n * sizeof(T) + 256));
    HIP_TRY(hipMemset(*p, 0, n * sizeof(T) + 256));
    return DALM_OK;
}

int upload_vec(dalm_ctx* ctx, double* dst, const std::vector<double>& src) {
    HIP_TRY(hipMemcpy(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    return DALM_OK;
}

// Host computation of the 1-D pieces of build_A (src/eq.h:716-795), single domain.
int build_operator(dalm_ctx* ctx) {
    const Dims& d = ctx->d;
    const int g = d.gc;
    std::vector<double> ae(d.n0, 0), aw(d.n0, 0), acx(d.n0, 0), an(d.n1, 0), as(d.n1, 0),
        at(d.n2, 0), ab(d.n2, 0);
    for (int i = g; i < d.n0 - g; ++i) {
        double c = 0;
        if (i < d.n0 - g - 1) {
            ae[i] = 1 / (ctx->hdx[i] * (ctx->hx[i + 1] - ctx->hx[i]));
            c -= ae[i];
        } else {
            c -= 2 / (ctx->hdx[i] * ctx->hdx[i]);
        }
        if (i > g) {
            aw[i] = 1 / (ctx->hdx[i] * (ctx->hx[i] - ctx->hx[i - 1]));
            c -= aw[i];
        }
        acx[i] = c;
    }
    for (int j = g; j < d.n1 - g; ++j) {
        if (j < d.n1 - g - 1) an[j] = 1 / (ctx->hdy[j] * (ctx->hy[j + 1] - ctx->hy[j]));
        if (j > g) as[j] = 1 / (ctx->hdy[j] * (ctx->hy[j] - ctx->hy[j - 1]));
    }
    for (int k = g; k < d.n2 - g; ++k) {
        if (k < d.n2 - g - 1) at[k] = 1 / (ctx->hdz[k] * (ctx->hz[k + 1] - ctx->hz[k]));
        if (k > g) ab[k] = 1 / (ctx->hdz[k] * (ctx->hz[k] - ctx->hz[k - 1]));
    }
    // max |a_c| over the interior, a_c accumulated in the reference's order
    double mx = 0;
    for (int i = g; i < d.n0 - g; ++i)
        for (int j = g; j < d.n1 - g; ++j) {
            double cij = (acx[i] - an[j]) - as[j];
            for (int k = g; k < d.n2 - g; ++k) {
                double c = (cij - at[k]) - ab[k];
                if (fabs(c) > mx) mx = fabs(c);
            }
        }
    auto norm = [&](const std::vector<double>& v) {
        std::vector<double> o(v.size());
        for (size_t q = 0; q < v.size(); ++q) o[q] = v[q] / mx;
        return o;
    };
    PoissonCoef& P = ctx->pc;
    int rc;
    if ((rc = upload_vec(ctx, P.ae_n, norm(ae))) || (rc = upload_vec(ctx, P.aw_n, norm(aw))) ||
        (rc = upload_vec(ctx, P.an_n, norm(an))) || (rc = upload_vec(ctx, P.as_n, norm(as))) ||
        (rc = upload_vec(ctx, P.at_n, norm(at))) || (rc = upload_vec(ctx, P.ab_n, norm(ab))) ||
        (rc = upload_vec(ctx, P.acx, acx)) || (rc = upload_vec(ctx, P.an_u, an)) ||
        (rc = upload_vec(ctx, P.as_u, as)) || (rc = upload_vec(ctx, P.at_u, at)) ||
        (rc = upload_vec(ctx, P.ab_u, ab)))
        return rc;
    P.max_diag = mx;
    ctx->have_A = true;
    return DALM_OK;
}

// field id -> device component pointers; *fp32 tells whether they hold floats
int field_ptrs(dalm_ctx* ctx, int64_t field, void** c, int* ncomp, bool* fp32) {
    *ncomp = 1;
    const bool m = ctx->precision == DALM_MIXED;
    *fp32 = false;
    switch (field) {
    case DALM_F_U:
    case DALM_F_UOLD:
    case DALM_F_F: {
        *ncomp = 3;
        *fp32 = m;
        for (int q = 0; q < 3; ++q) {
            if (field == DALM_F_U) c[q] = m ? (void*)ctx->U32[q] : (void*)ctx->U[q];
            else if (field == DALM_F_UOLD) c[q] = m ? (void*)ctx->Uold32[q] : (void*)ctx->Uold[q];
            else c[q] = m ? (void*)ctx->f32[q] : (void*)ctx->f[q];
        }
        break;
    }
    case DALM_F_JU: c[0] = ctx->JU[0]; c[1] = ctx->JU[1]; c[2] = ctx->JU[2]; *ncomp = 3; break;
    case DALM_F_P: c[0] = ctx->p; break;
    case DALM_F_NUT: c[0] = m ? (void*)ctx->nut32 : (void*)ctx->nut; *fp32 = m; break;
    case DALM_F_DIV: c[0] = ctx->div; break;
    case DALM_F_SOLID: c[0] = ctx->solid; break;
    case DALM_F_B: c[0] = ctx->b; break;
    case DALM_F_Q:
    case DALM_F_UTAVG: {
        if (int rc = ensure_diag_buffers(ctx)) return rc;
        if (field == DALM_F_Q) {
            c[0] = ctx->qcrit;
        } else {
            *ncomp = 3;
            for (int q = 0; q < 3; ++q) c[q] = ctx->Utavg[q];
        }
        break;
    }
    default: {
        if (field < DALM_F_U_B || field > DALM_F_B_B) return DALM_BAD_ARG;
        if (int rc = ensure_adjoint_buffers(ctx)) return rc;
        AdjointState& S = ctx->adj;
        double* const* v3 = nullptr;
        double* one = nullptr;
        switch (field) {
        case DALM_F_U_B: v3 = S.Ub; break;
        case DALM_F_UOLD_B: v3 = S.Ubn; break;
        case DALM_F_JU_B: v3 = S.JUb; break;
        case DALM_F_JUOLD_B: v3 = S.JUbn; break;
        case DALM_F_F_B: v3 = S.fb; break;
        case DALM_F_FOLD_B: v3 = S.fbn; break;
        case DALM_F_NUT_B: one = S.nutb; break;
        case DALM_F_NUTOLD_B: one = S.nutbn; break;
        case DALM_F_P_B: one = S.pb; break;
        default: one = S.ab; break;
        }
        if (v3) {
            *ncomp = 3;
            for (int q = 0; q < 3; ++q) c[q] = v3[q];
        } else {
            c[0] = one;
        }
        break;
    }
    }
    return DALM_OK;
}

void fill_stats(dalm_ctx* ctx, dalm_step_stats* st, long long it, double err, int status) {
    std::memset(st, 0, sizeof *st);
    st->step = ctx->step;
    st->time = ctx->prm.dt * (double)ctx->step;
    st->it = it;
    st->err = err;
    st->avg_div = sqrt(ctx->h_scal->avg_div_sq) / sqrt((double)ctx->N);
    st->max_cfl = ctx->h_scal->max_cfl;
    st->status = status;
    st->wt_count = ctx->alm.wt_count;
    for (int w = 0; w < ctx->alm.wt_count && w < DALM_MAX_TURBINES; ++w) {
        st->torque[w] = ctx->h_scal->torque[w];
        st->thrust[w] = ctx->h_scal->thrust[w];
        st->power[w] = ctx->alm.h_wt[DALM_WT_STRIDE * w + 13] * st->torque[w];
    }
    ctx->last = *st;
}

int read_scalars(dalm_ctx* ctx) {
    HIP_TRY(hipMemcpyAsync(ctx->h_scal, ctx->scal, sizeof(StepScalars), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

}  // namespace

#define CHECK_CTX()                             \
    do {                                        \
        if (!ctx) return DALM_BAD_ARG;          \
    } while (0)
#define CHECK_FP64()                                                                        \
    do {                                                                                    \
        if (!ctx) return DALM_BAD_ARG;                                                      \
        if (ctx->precision != DALM_FP64)                                                    \
            return dalm_fail_msg(ctx, DALM_BAD_ARG, "per-kernel entry points need a DALM_FP64 context"); \
    } while (0)
#define TRY(x)              \
    do {                    \
        int rc_ = (x);      \
        if (rc_) return rc_; \
    } while (0)

// One FP64 step in Solver::main_loop_once order (src/solver.h:1042-1186).  With
// p_given the Poisson solve is replaced by a copy of p_given (replay of a recorded
// step during the reverse sweep).
int step_fp64(dalm_ctx* ctx, dalm_step_stats* stats, const double* p_given) {
    const dalm_params& P = ctx->prm;
    ctx->step += 1;
    const double t = P.dt * (double)ctx->step;
    TRY(launch_inflow(ctx, t));
    // Uold <- U: swap the buffers, then restore U's ghost shell from Uold
    for (int q = 0; q < 3; ++q) std::swap(ctx->U[q], ctx->Uold[q]);
    TRY(launch_ghost_copy(ctx));
    TRY(launch_predictor(ctx, P.Re, P.dt));
    TRY(launch_flux(ctx, P.Uin, true));
    long long it = 0;
    double err = 0;
    int status = DALM_OK;
    if (p_given) {
        HIP_TRY(hipMemcpyAsync(ctx->p, p_given, sizeof(double) * ctx->L, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        TRY(launch_rhs(ctx, P.dt, ctx->pc.max_diag));
        // the segment replays of a precision-form recording: their pressure takes the
        // precision form's solver too
        if (P.method == DALM_BICGSTAB && ctx->adj.replaying && ctx->adj.from_mixed)
            status = run_poisson_mixed(ctx, (int)P.pc_kind, P.pc_omega, (int)P.pc_it, P.max_it, P.tol, &it, &err);
        else if (P.method == DALM_BICGSTAB)
            status = run_bicgstab(ctx, (int)P.pc_kind, P.pc_omega, (int)P.pc_it, P.max_it, P.tol, &it, &err);
        else
            status = run_sor(ctx, P.omega, P.max_it, P.tol, &it, &err);
        if (status == DALM_HIP_ERROR) return status;
    }
    TRY(launch_pbc(ctx));
    TRY(launch_project(ctx, P.dt, P.Uin));
    if (ctx->has_solid) TRY(launch_solid(ctx));
    TRY(launch_ubc(ctx, P.Uin, P.dt, (int)P.retracted_floor, P.floor_start_x));
    TRY(launch_post(ctx, P.Cs, P.dt, true));
    TRY(alm_run(ctx, t, false));
    TRY(read_scalars(ctx));
    TRY(probe_collect(ctx));
    dalm_step_stats st;
    fill_stats(ctx, &st, it, err, status);
    if (stats) *stats = st;
    return status == DALM_NONCONVERGED ? DALM_OK : status;
}

Dims make_dims(const int64_t size[3], int64_t gc) {
    Dims d{};
    d.n0 = (int)size[0];
    d.n1 = (int)size[1];
    d.n2 = (int)size[2];
    d.gc = (int)gc;
    d.nx = d.n0 - 2 * d.gc;
    d.ny = d.n1 - 2 * d.gc;
    d.nz = d.n2 - 2 * d.gc;
    d.si = d.n1 * d.n2;
    d.sj = d.n2;
    d.tk = (d.nz + TK - 1) / TK;
    d.tj = (d.ny + TJ - 1) / TJ;
    d.ntiles = d.nx * d.tj * d.tk;
    return d;
}

int cell_grid(int ntiles) {
    const int g = ntiles < 2048 ? ntiles : 2048;
    return g >= 64 ? (g / 8) * 8 : g;
}

int ensure_mixed_solver(dalm_ctx* ctx) {
    if (ctx->precision == DALM_MIXED || ctx->fr) return DALM_OK;
    int rc;
    if ((rc = setup_mixed(ctx))) return rc;
    const size_t L = (size_t)ctx->L;
    float** scf[10] = {&ctx->fr, &ctx->fr0, &ctx->fp, &ctx->fq, &ctx->fs,
                       &ctx->ft, &ctx->fz, &ctx->fz2, &ctx->ftmp, &ctx->ftmp2};
    for (float** p : scf)
        if ((rc = dev_alloc(ctx, p, L))) return rc;
    if (!ctx->pbest && (rc = dev_alloc(ctx, &ctx->pbest, L))) return rc;
    return DALM_OK;
}

int ensure_fused_z(dalm_ctx* ctx) {
    int rc;
    for (int q = 0; q < kXGroupF - 1; ++q) {
        if (!ctx->fzx[q] && (rc = dev_alloc(ctx, &ctx->fzx[q], (size_t)ctx->L))) return rc;
        if (!ctx->fz2x[q] && (rc = dev_alloc(ctx, &ctx->fz2x[q], (size_t)ctx->L))) return rc;
    }
    return DALM_OK;
}

int ensure_fused_z64(dalm_ctx* ctx) {
    int rc;
    for (int q = 0; q < kXGroup - 1; ++q) {
        if (!ctx->ppx[q] && (rc = dev_alloc(ctx, &ctx->ppx[q], (size_t)ctx->L))) return rc;
        if (!ctx->ssx[q] && (rc = dev_alloc(ctx, &ctx->ssx[q], (size_t)ctx->L))) return rc;
    }
    return DALM_OK;
}

extern "C" {

int dalm_set_device(int64_t device) {
    return hipSetDevice((int)device) == hipSuccess ? DALM_OK : DALM_HIP_ERROR;
}

int dalm_device_count(int64_t* n) {
    if (!n) return DALM_BAD_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        *n = 0;
        return DALM_HIP_ERROR;
    }
    *n = c;
    return DALM_OK;
}

int dalm_mesh_widths(const double* c, int64_t n, double* d) {
    if (!c || !d || n < 3) return DALM_BAD_ARG;
    for (int64_t i = 1; i < n - 1; ++i) d[i] = 0.5 * (c[i + 1] - c[i - 1]);
    d[0] = -1.5 * c[0] + 2.0 * c[1] - 0.5 * c[2];
    d[n - 1] = 1.5 * c[n - 1] - 2.0 * c[n - 2] + 0.5 * c[n - 3];
    return DALM_OK;
}

const char* dalm_last_error(const dalm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dalm_create(dalm_ctx** out, const int64_t size[3], int64_t gc, const double* x, const double* y,
                const double* z, int64_t precision) {
    if (!out || !size || !x || !y || !z) return DALM_BAD_ARG;
    *out = nullptr;
    if (gc < 2 || gc > 16) return DALM_BAD_ARG;  // the reference asserts gc >= 2 (src/solver.h:615)
    // 32-bit cell indices: the padded box stays below 2^31 - 2^20 cells (checked
    // factor by factor, so that no product overflows)
    constexpr long long kMaxCells = (1LL << 31) - (1LL << 20);
    for (int a = 0; a < 3; ++a)
        if (size[a] < 2 * gc + 3 || size[a] >= kMaxCells) return DALM_BAD_ARG;
    long long L = size[0] * size[1];
    if (L >= kMaxCells) return DALM_BAD_ARG;
    L *= size[2];
    if (L >= kMaxCells) return DALM_BAD_ARG;
    if (precision != DALM_FP64 && precision != DALM_MIXED) return DALM_BAD_ARG;

    dalm_ctx* ctx = new dalm_ctx();
    ctx->d = make_dims(size, gc);
    Dims& d = ctx->d;
    ctx->L = L;
    ctx->N = (long long)d.nx * d.ny * d.nz;
    ctx->precision = (int)precision;
    ctx->grid = cell_grid(d.ntiles);
    // column units: about 2048 of them, 1 to kMaxColumnPlanes planes each -- small
    // grids get short columns rather than too few workgroups, large ones long columns
    // (each unit re-reads a halo plane at either end: 2 per 8 planes cost the
    // four-cell SpMV and x/r update 4-5 % at 512x256x256 against 2 per 32)
    {
        long long per_plane = (long long)d.tj * d.tk;
        long long ni = (d.nx * per_plane + 2047) / 2048;
        if (ni < 1) ni = 1;
        if (ni > kMaxColumnPlanes) ni = kMaxColumnPlanes;
        if (ni > d.nx) ni = d.nx;
        d.ni = (int)ni;
        d.nch = (d.nx + d.ni - 1) / d.ni;
        d.nunits = (int)(per_plane * d.nch);
        ctx->sgrid = d.nunits;
    }

    ctx->hx.assign(x, x + d.n0);
    ctx->hy.assign(y, y + d.n1);
    ctx->hz.assign(z, z + d.n2);
    ctx->hdx.resize(d.n0);
    ctx->hdy.resize(d.n1);
    ctx->hdz.resize(d.n2);
    dalm_mesh_widths(x, d.n0, ctx->hdx.data());
    dalm_mesh_widths(y, d.n1, ctx->hdy.data());
    dalm_mesh_widths(z, d.n2, ctx->hdz.data());

    int rc = DALM_OK;
    auto fail = [&](int r) {
        dalm_destroy(ctx);
        return r;
    };
    if (hipStreamCreate(&ctx->stream) != hipSuccess) return fail(DALM_HIP_ERROR);
    double** one_d[6] = {&ctx->x, &ctx->y, &ctx->z, &ctx->dx, &ctx->dy, &ctx->dz};
    const std::vector<double>* src[6] = {&ctx->hx, &ctx->hy, &ctx->hz, &ctx->hdx, &ctx->hdy, &ctx->hdz};
    for (int q = 0; q < 6; ++q) {
        if ((rc = dev_alloc(ctx, one_d[q], src[q]->size()))) return fail(rc);
        if ((rc = upload_vec(ctx, *one_d[q], *src[q]))) return fail(rc);
    }
    PoissonCoef& P = ctx->pc;
    double** pc1[11] = {&P.ae_n, &P.aw_n, &P.an_n, &P.as_n, &P.at_n, &P.ab_n, &P.acx, &P.an_u, &P.as_u,
                        &P.at_u, &P.ab_u};
    int pcn[11] = {d.n0, d.n0, d.n1, d.n1, d.n2, d.n2, d.n0, d.n1, d.n1, d.n2, d.n2};
    for (int q = 0; q < 11; ++q)
        if ((rc = dev_alloc(ctx, pc1[q], pcn[q]))) return fail(rc);

    const bool mixed = precision == DALM_MIXED;
    for (int q = 0; q < 3; ++q) {
        if ((rc = dev_alloc(ctx, &ctx->JU[q], L))) return fail(rc);
        if (mixed) {
            if ((rc = dev_alloc(ctx, &ctx->U32[q], L)) || (rc = dev_alloc(ctx, &ctx->Uold32[q], L)) ||
                (rc = dev_alloc(ctx, &ctx->f32[q], L)))
                return fail(rc);
        } else {
            if ((rc = dev_alloc(ctx, &ctx->U[q], L)) || (rc = dev_alloc(ctx, &ctx->Uold[q], L)) ||
                (rc = dev_alloc(ctx, &ctx->f[q], L)))
                return fail(rc);
        }
    }
    double** sc64[5] = {&ctx->p, &ctx->div, &ctx->b, &ctx->solid, &ctx->r};
    for (int q = 0; q < 5; ++q)
        if ((rc = dev_alloc(ctx, sc64[q], L))) return fail(rc);
    if (mixed) {
        // FP32 BiCGStab vectors (ghost shells stay zero: kernels write interiors only)
        float** scf[11] = {&ctx->nut32, &ctx->fr, &ctx->fr0, &ctx->fp, &ctx->fq,   &ctx->fs,
                           &ctx->ft,    &ctx->fz, &ctx->fz2, &ctx->ftmp, &ctx->ftmp2};
        for (int q = 0; q < 11; ++q)
            if ((rc = dev_alloc(ctx, scf[q], L))) return fail(rc);
        // FP64 vectors for the FP64 finish of a stalled solve (the FP64 solver's set),
        // and the best pressure seen by the running solve
        double** sd[10] = {&ctx->kp, &ctx->q,  &ctx->s,   &ctx->t,    &ctx->r0,
                           &ctx->pp, &ctx->ss, &ctx->tmp, &ctx->tmp2, &ctx->pbest};
        for (int q = 0; q < 10; ++q)
            if ((rc = dev_alloc(ctx, sd[q], L))) return fail(rc);
    } else {
        double** sc[10] = {&ctx->nut, &ctx->r0, &ctx->kp, &ctx->pp, &ctx->q, &ctx->s, &ctx->ss, &ctx->t,
                           &ctx->tmp, &ctx->tmp2};
        for (int q = 0; q < 10; ++q)
            if ((rc = dev_alloc(ctx, sc[q], L))) return fail(rc);
    }
    if ((rc = dev_alloc(ctx, &ctx->partials, 8 * (size_t)std::max(ctx->grid, ctx->sgrid))) ||
        (rc = dev_alloc(ctx, &ctx->counter, 16)) || (rc = dev_alloc(ctx, &ctx->solve, 2)) ||
        (rc = dev_alloc(ctx, &ctx->scal, 1)))
        return fail(rc);
    if (hipHostMalloc((void**)&ctx->h_solve, 2 * sizeof(SolveState)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_scal, sizeof(StepScalars)) != hipSuccess)
        return fail(DALM_HIP_ERROR);
    std::memset(ctx->h_solve, 0, 2 * sizeof(SolveState));
    std::memset(ctx->h_scal, 0, sizeof(StepScalars));
    if ((rc = build_operator(ctx))) return fail(rc);
    if (mixed && (rc = setup_mixed(ctx))) return fail(rc);
    *out = ctx;
    return DALM_OK;
}

int dalm_destroy(dalm_ctx* ctx) {
    if (!ctx) return DALM_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    adjoint_free(ctx);
    free_diag_buffers(ctx);
    for (hipEvent_t e : ctx->probe.ev) (void)hipEventDestroy(e);
    double* ptrs[] = {ctx->x, ctx->y, ctx->z, ctx->dx, ctx->dy, ctx->dz, ctx->p, ctx->nut, ctx->div,
                      ctx->b, ctx->solid, ctx->r, ctx->r0, ctx->kp, ctx->pp, ctx->q, ctx->s, ctx->ss,
                      ctx->t, ctx->tmp, ctx->tmp2, ctx->pbest, ctx->partials, ctx->pc.ae_n, ctx->pc.aw_n,
                      ctx->pc.an_n, ctx->pc.as_n, ctx->pc.at_n, ctx->pc.ab_n, ctx->pc.acx,
                      ctx->pc.an_u, ctx->pc.as_u, ctx->pc.at_u, ctx->pc.ab_u, ctx->alm.wt, ctx->alm.ap,
                      ctx->alm.cd, ctx->alm.cl, ctx->alm.atk, ctx->inflow_modes, ctx->inflow};
    for (double* p : ptrs)
        if (p) (void)hipFree(p);
    for (int q = 0; q < kXGroup - 1; ++q) {
        if (ctx->ppx[q]) (void)hipFree(ctx->ppx[q]);
        if (ctx->ssx[q]) (void)hipFree(ctx->ssx[q]);
    }
    for (int q = 0; q < kXGroupF - 1; ++q) {
        if (ctx->fzx[q]) (void)hipFree(ctx->fzx[q]);
        if (ctx->fz2x[q]) (void)hipFree(ctx->fz2x[q]);
    }
    for (int q = 0; q < 3; ++q) {
        if (ctx->U[q]) (void)hipFree(ctx->U[q]);
        if (ctx->Uold[q]) (void)hipFree(ctx->Uold[q]);
        if (ctx->JU[q]) (void)hipFree(ctx->JU[q]);
        if (ctx->f[q]) (void)hipFree(ctx->f[q]);
    }
    for (int q = 0; q < 3; ++q) {
        if (ctx->U32[q]) (void)hipFree(ctx->U32[q]);
        if (ctx->Uold32[q]) (void)hipFree(ctx->Uold32[q]);
        if (ctx->f32[q]) (void)hipFree(ctx->f32[q]);
    }
    float* fptrs[] = {ctx->nut32, ctx->fr,   ctx->fr0,  ctx->fp,  ctx->fq,   ctx->fs,  ctx->ft,
                      ctx->fz,    ctx->fz2,  ctx->ftmp, ctx->ftmp2};
    for (float* p : fptrs)
        if (p) (void)hipFree(p);
    for (void* p : ctx->owned_f) (void)hipFree(p);
    if (ctx->alm.box_dev) (void)hipFree(ctx->alm.box_dev);
    if (ctx->counter) (void)hipFree(ctx->counter);
    if (ctx->solve) (void)hipFree(ctx->solve);
    if (ctx->scal) (void)hipFree(ctx->scal);
    if (ctx->h_solve) (void)hipHostFree(ctx->h_solve);
    if (ctx->h_scal) (void)hipHostFree(ctx->h_scal);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DALM_OK;
}

int dalm_upload(dalm_ctx* ctx, int64_t field, const double* host) {
    CHECK_CTX();
    if (!host) return DALM_BAD_ARG;
    void* c[3];
    int nc;
    bool f32;
    TRY(field_ptrs(ctx, field, c, &nc, &f32));
    std::vector<double> tmp(ctx->L);
    std::vector<float> tmpf(f32 ? ctx->L : 0);
    for (int q = 0; q < nc; ++q) {
        if (f32) {
            for (long long e = 0; e < ctx->L; ++e) tmpf[e] = (float)host[e * nc + q];
            HIP_TRY(hipMemcpy(c[q], tmpf.data(), ctx->L * sizeof(float), hipMemcpyHostToDevice));
        } else {
            for (long long e = 0; e < ctx->L; ++e) tmp[e] = host[e * nc + q];
            HIP_TRY(hipMemcpy(c[q], tmp.data(), ctx->L * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    if (field == DALM_F_F) ctx->alm.f_cleared = false;
    if (field == DALM_F_SOLID) {
        ctx->has_solid = false;
        for (long long e = 0; e < ctx->L; ++e)
            if (host[e] != 0) {
                ctx->has_solid = true;
                break;
            }
    }
    return DALM_OK;
}

int dalm_download(dalm_ctx* ctx, int64_t field, double* host) {
    CHECK_CTX();
    if (!host) return DALM_BAD_ARG;
    void* c[3];
    int nc;
    bool f32;
    TRY(field_ptrs(ctx, field, c, &nc, &f32));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<double> tmp(f32 ? 0 : ctx->L);
    std::vector<float> tmpf(f32 ? ctx->L : 0);
    for (int q = 0; q < nc; ++q) {
        if (f32) {
            HIP_TRY(hipMemcpy(tmpf.data(), c[q], ctx->L * sizeof(float), hipMemcpyDeviceToHost));
            for (long long e = 0; e < ctx->L; ++e) host[e * nc + q] = tmpf[e];
        } else {
            HIP_TRY(hipMemcpy(tmp.data(), c[q], ctx->L * sizeof(double), hipMemcpyDeviceToHost));
            for (long long e = 0; e < ctx->L; ++e) host[e * nc + q] = tmp[e];
        }
    }
    return DALM_OK;
}

int dalm_set_params(dalm_ctx* ctx, const dalm_params* prm) {
    CHECK_CTX();
    if (!prm || !(prm->dt > 0) || !(prm->Re > 0) || prm->max_it < 1 ||
        (prm->method != DALM_SOR && prm->method != DALM_BICGSTAB) ||
        (prm->pc_kind != DALM_PC_JACOBI && prm->pc_kind != DALM_PC_SOR) || prm->pc_it < 0)
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "invalid dalm_params");
    ctx->prm = *prm;
    ctx->have_prm = true;
    return DALM_OK;
}

// ---- per-kernel entry points --------------------------------------------------
int dalm_copy_u_to_uold(dalm_ctx* ctx) {
    CHECK_FP64();
    TRY(launch_copy_uold(ctx));
    return sync_stream(ctx);
}
int dalm_predictor(dalm_ctx* ctx, double Re, double dt) {
    CHECK_FP64();
    TRY(launch_predictor(ctx, Re, dt));
    return sync_stream(ctx);
}
int dalm_interpolate_JU(dalm_ctx* ctx) {
    CHECK_FP64();
    TRY(launch_flux(ctx, nullptr, false));
    return sync_stream(ctx);
}
int dalm_apply_JUbc(dalm_ctx* ctx, const double Uin[3]) {
    CHECK_FP64();
    if (!Uin) return DALM_BAD_ARG;
    TRY(launch_jubc(ctx, Uin));
    return sync_stream(ctx);
}
int dalm_poisson_rhs(dalm_ctx* ctx, double dt, double scale) {
    CHECK_FP64();
    TRY(launch_rhs(ctx, dt, scale));
    return sync_stream(ctx);
}
int dalm_build_A(dalm_ctx* ctx, double* max_diag) {
    CHECK_CTX();
    if (!ctx->have_A) TRY(build_operator(ctx));
    if (max_diag) *max_diag = ctx->pc.max_diag;
    return DALM_OK;
}
int dalm_run_sor(dalm_ctx* ctx, double omega, int64_t max_it, double tol, int64_t* it, double* err) {
    CHECK_FP64();
    if (max_it < 1) return DALM_BAD_ARG;
    long long i = 0;
    double e = 0;
    int rc = run_sor(ctx, omega, max_it, tol, &i, &e);
    if (it) *it = i;
    if (err) *err = e;
    return rc;
}
int dalm_run_bicgstab(dalm_ctx* ctx, int64_t pc_kind, double pc_omega, int64_t pc_it, int64_t max_it,
                      double tol, int64_t* it, double* err) {
    CHECK_FP64();
    if (max_it < 1 || pc_it < 0 || (pc_kind != DALM_PC_JACOBI && pc_kind != DALM_PC_SOR))
        return DALM_BAD_ARG;
    long long i = 0;
    double e = 0;
    int rc = run_bicgstab(ctx, (int)pc_kind, pc_omega, (int)pc_it, max_it, tol, &i, &e);
    if (it) *it = i;
    if (err) *err = e;
    return rc;
}
int dalm_residual(dalm_ctx* ctx, double* out) {
    CHECK_FP64();
    TRY(launch_residual(ctx, ctx->p, ctx->b, ctx->r));
    TRY(sync_stream(ctx));
    if (out) {
        std::vector<double> tmp(ctx->L);
        HIP_TRY(hipMemcpy(tmp.data(), ctx->r, ctx->L * sizeof(double), hipMemcpyDeviceToHost));
        std::memcpy(out, tmp.data(), ctx->L * sizeof(double));
    }
    return DALM_OK;
}
int dalm_apply_pbc(dalm_ctx* ctx) {
    CHECK_FP64();
    TRY(launch_pbc(ctx));
    return sync_stream(ctx);
}
int dalm_project_p(dalm_ctx* ctx, double dt) {
    CHECK_FP64();
    TRY(launch_project(ctx, dt, nullptr));
    return sync_stream(ctx);
}
int dalm_set_solid_U(dalm_ctx* ctx) {
    CHECK_FP64();
    TRY(launch_solid(ctx));
    return sync_stream(ctx);
}
int dalm_apply_Ubc(dalm_ctx* ctx, const double Uin[3], double dt, int64_t retracted_floor,
                   double floor_start_x) {
    CHECK_FP64();
    if (!Uin) return DALM_BAD_ARG;
    TRY(launch_ubc(ctx, Uin, dt, (int)retracted_floor, floor_start_x));
    return sync_stream(ctx);
}
int dalm_calc_nut(dalm_ctx* ctx, double Cs) {
    CHECK_FP64();
    TRY(launch_nut(ctx, Cs));
    return sync_stream(ctx);
}
int dalm_calc_divergence(dalm_ctx* ctx) {
    CHECK_CTX();
    // the precision form's face fluxes are FP64 too: same arithmetic, store only
    TRY(ctx->precision == DALM_MIXED ? launch_div_store(ctx) : launch_divergence(ctx));
    return sync_stream(ctx);
}
int dalm_l2_norm(dalm_ctx* ctx, int64_t field, double* out) {
    CHECK_FP64();
    void* c[3];
    int nc;
    bool f32;
    TRY(field_ptrs(ctx, field, c, &nc, &f32));
    if (nc != 1 || f32 || !out) return DALM_BAD_ARG;
    return launch_l2(ctx, (const double*)c[0], out);
}
int dalm_max_cfl(dalm_ctx* ctx, double dt, double* out) {
    CHECK_FP64();
    if (!out) return DALM_BAD_ARG;
    return launch_max_cfl(ctx, dt, out);
}

// ---- ALM --------------------------------------------------------------------------
int dalm_alm_setup(dalm_ctx* ctx, const double* wt, int64_t wt_count, const double* ap,
                   int64_t ap_count, int64_t blade_per_wt, int64_t ap_per_blade, const double* cd,
                   const double* cl, const double* atk, int64_t atk_count, double width) {
    CHECK_CTX();
    // each factor is bounded before any product is formed
    if (wt_count < 0 || wt_count > DALM_MAX_TURBINES || blade_per_wt < 1 || ap_per_blade < 1 ||
        blade_per_wt > 1024 || ap_per_blade > 1024 || blade_per_wt * ap_per_blade > 1024 ||
        ap_count != wt_count * blade_per_wt * ap_per_blade ||
        (ap_count > 0 && (atk_count < 1 || atk_count > (1LL << 24) || !(width > 0) || !wt || !ap || !cd ||
                          !cl || !atk)))
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "invalid ALM set-up");
    AlmTables& A = ctx->alm;
    A.wt_count = (int)wt_count;
    A.ap_count = (int)ap_count;
    A.blade_per_wt = (int)blade_per_wt;
    A.ap_per_blade = (int)ap_per_blade;
    A.atk_count = (int)atk_count;
    A.width = width;
    if (ap_count == 0) return DALM_OK;
    A.h_wt.assign(wt, wt + DALM_WT_STRIDE * wt_count);
    A.h_rmax.assign(wt_count, 0.0);
    int per_wt = (int)(blade_per_wt * ap_per_blade);
    for (int64_t a = 0; a < ap_count; ++a) {
        double r = fabs(ap[DALM_AP_STRIDE * a + 3]);
        double& m = A.h_rmax[a / per_wt];
        m = r > m ? r : m;
    }
    int rc;
    if (!A.wt && ((rc = dev_alloc(ctx, &A.wt, DALM_WT_STRIDE * DALM_MAX_TURBINES)) ||
                  (rc = dev_alloc(ctx, &A.box_dev, 6 * DALM_MAX_TURBINES))))
        return rc;
    if (A.ap) { (void)hipFree(A.ap); (void)hipFree(A.cd); (void)hipFree(A.cl); (void)hipFree(A.atk); }
    if ((rc = dev_alloc(ctx, &A.ap, DALM_AP_STRIDE * ap_count)) ||
        (rc = dev_alloc(ctx, &A.cd, ap_count * atk_count)) ||
        (rc = dev_alloc(ctx, &A.cl, ap_count * atk_count)) || (rc = dev_alloc(ctx, &A.atk, atk_count)))
        return rc;
    HIP_TRY(hipMemcpy(A.wt, wt, sizeof(double) * DALM_WT_STRIDE * wt_count, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A.ap, ap, sizeof(double) * DALM_AP_STRIDE * ap_count, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A.cd, cd, sizeof(double) * ap_count * atk_count, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A.cl, cl, sizeof(double) * ap_count * atk_count, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A.atk, atk, sizeof(double) * atk_count, hipMemcpyHostToDevice));
    return DALM_OK;
}

int dalm_run_alm(dalm_ctx* ctx, double t, double* torque, double* thrust) {
    CHECK_CTX();
    TRY(alm_run(ctx, t, true));
    for (int w = 0; w < ctx->alm.wt_count; ++w) {
        if (torque) torque[w] = ctx->h_scal->torque[w];
        if (thrust) thrust[w] = ctx->h_scal->thrust[w];
    }
    return DALM_OK;
}

int dalm_alm_download(dalm_ctx* ctx, double* ap, double* wt) {
    CHECK_CTX();
    AlmTables& A = ctx->alm;
    if (A.ap_count == 0) return DALM_OK;
    if (ap)
        HIP_TRY(hipMemcpy(ap, A.ap, sizeof(double) * DALM_AP_STRIDE * A.ap_count, hipMemcpyDeviceToHost));
    if (wt) {
        HIP_TRY(hipMemcpy(wt, A.wt, sizeof(double) * DALM_WT_STRIDE * A.wt_count, hipMemcpyDeviceToHost));
        for (int w = 0; w < A.wt_count; ++w) {
            wt[DALM_WT_STRIDE * w + 14] = ctx->h_scal->torque[w];
            wt[DALM_WT_STRIDE * w + 15] = ctx->h_scal->thrust[w];
        }
    }
    return DALM_OK;
}

// ---- turbulent inflow ---------------------------------------------------------------
int dalm_set_inflow(dalm_ctx* ctx, const double* modes, int64_t n_modes, double speed) {
    CHECK_CTX();
    if (n_modes < 0 || n_modes > 4096 || (n_modes > 0 && !modes) || !std::isfinite(speed))
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "invalid inflow modes");
    for (int64_t q = 0; q < 7 * n_modes; ++q)
        if (!std::isfinite(modes[q])) return dalm_fail_msg(ctx, DALM_BAD_ARG, "non-finite inflow mode");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->inflow_modes) (void)hipFree(ctx->inflow_modes);
    if (ctx->inflow) (void)hipFree(ctx->inflow);
    ctx->inflow_modes = ctx->inflow = nullptr;
    ctx->inflow_n = 0;
    ctx->inflow_speed = 0;
    if (n_modes == 0) return DALM_OK;
    const Dims& d = ctx->d;
    const size_t cells = (3 * (size_t)d.gc + 1) * (size_t)d.ny * (size_t)d.nz;
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->inflow_modes, 7 * (size_t)n_modes)) || (rc = dev_alloc(ctx, &ctx->inflow, cells)))
        return rc;
    HIP_TRY(hipMemcpy(ctx->inflow_modes, modes, sizeof(double) * 7 * (size_t)n_modes, hipMemcpyHostToDevice));
    ctx->inflow_n = (int)n_modes;
    ctx->inflow_speed = speed;
    return DALM_OK;
}

int dalm_eval_inflow(dalm_ctx* ctx, double t) {
    CHECK_CTX();
    if (!ctx->have_prm) return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_set_params not called");
    TRY(launch_inflow(ctx, t));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

// ---- whole step -----------------------------------------------------------------------
int dalm_initialize(dalm_ctx* ctx, dalm_step_stats* stats) {
    CHECK_CTX();
    if (!ctx->have_prm) return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_set_params not called");
    const dalm_params& P = ctx->prm;
    ctx->step = 0;
    if (!ctx->have_A) TRY(build_operator(ctx));
    TRY(launch_inflow(ctx, 0.0));
    if (ctx->precision == DALM_MIXED) {
        if (ctx->has_solid) TRY(launch_solid_m(ctx));
        TRY(launch_ubc_m(ctx, P.Uin, P.dt, (int)P.retracted_floor, P.floor_start_x));
        TRY(launch_flux_rhs_m(ctx, P.Uin, P.dt, ctx->pc.max_diag, false));
        TRY(launch_post_m(ctx, P.Cs, P.dt, true));
    } else {
        if (ctx->has_solid) TRY(launch_solid(ctx));
        TRY(launch_ubc(ctx, P.Uin, P.dt, (int)P.retracted_floor, P.floor_start_x));
        TRY(launch_flux(ctx, P.Uin, true));
        TRY(launch_post(ctx, P.Cs, P.dt, true));
    }
    TRY(alm_run(ctx, 0.0, false));
    TRY(read_scalars(ctx));
    dalm_step_stats st;
    fill_stats(ctx, &st, 0, 0.0, DALM_OK);
    if (stats) *stats = st;
    return DALM_OK;
}

// One time step of the precision form (same order as the FP64 step below).
static int step_mixed(dalm_ctx* ctx, double t, dalm_step_stats* stats) {
    const dalm_params& P = ctx->prm;
    TRY(launch_inflow(ctx, t));
    for (int q = 0; q < 3; ++q) std::swap(ctx->U32[q], ctx->Uold32[q]);
    TRY(launch_ghost_copy_m(ctx));
    TRY(launch_predictor_m(ctx, P.Re, P.dt));
    TRY(launch_flux_rhs_m(ctx, P.Uin, P.dt, ctx->pc.max_diag, true));
    long long it = 0;
    double err = 0;
    int status = P.method == DALM_SOR
                     ? run_sor_mixed(ctx, P.omega, P.max_it, P.tol, &it, &err)
                     : run_poisson_mixed(ctx, (int)P.pc_kind, P.pc_omega, (int)P.pc_it, P.max_it, P.tol, &it, &err);
    if (status == DALM_HIP_ERROR) return status;
    TRY(launch_pbc(ctx));
    TRY(launch_project_m(ctx, P.dt, P.Uin));
    if (ctx->has_solid) TRY(launch_solid_m(ctx));
    TRY(launch_ubc_m(ctx, P.Uin, P.dt, (int)P.retracted_floor, P.floor_start_x));
    TRY(launch_post_m(ctx, P.Cs, P.dt, true));
    TRY(alm_run(ctx, t, false));
    TRY(read_scalars(ctx));
    TRY(probe_collect(ctx));
    dalm_step_stats st;
    fill_stats(ctx, &st, it, err, status);
    if (stats) *stats = st;
    return status == DALM_NONCONVERGED ? DALM_OK : status;
}

int dalm_step(dalm_ctx* ctx, dalm_step_stats* stats) {
    CHECK_CTX();
    if (!ctx->have_prm) return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_set_params not called");
    if (ctx->precision == DALM_MIXED) {
        ctx->step += 1;
        const int rc = step_mixed(ctx, ctx->prm.dt * (double)ctx->step, stats);
        if (dalm_ctx* sink = ctx->adj.sink) {
            // recorded into an FP64 context (dalm_adjoint_begin_from), on the same terms
            if (rc == DALM_OK || rc == DALM_BREAKDOWN || rc == DALM_NONFINITE) {
                TRY(adjoint_record_from(ctx, rc));
            } else {
                sink->adj.tape_bad = true;
                sink->adj.taping = false;
                sink->adj.source = nullptr;
                ctx->adj.sink = nullptr;
            }
        }
        return rc;
    }
    if (ctx->adj.source)
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "this context is recording a precision-form context's steps");
    const int rc = step_fp64(ctx, stats, nullptr);
    if (ctx->adj.taping) {
        // a BiCGStab breakdown ends the solve and the step goes on, as in the
        // reference (src/eq.h:349-353): the step is recorded like any other.  A
        // non-finite step is recorded as unusable; a HIP error leaves the step
        // unfinished, and the recording with it.
        if (rc == DALM_OK || rc == DALM_BREAKDOWN || rc == DALM_NONFINITE) {
            TRY(adjoint_record(ctx, rc));
        } else {
            ctx->adj.tape_bad = true;
            ctx->adj.taping = false;
        }
    }
    return rc;
}

int dalm_last_solve_info(dalm_ctx* ctx, int64_t out[3]) {
    CHECK_CTX();
    if (!out) return DALM_BAD_ARG;
    const long long it = ctx->last.it;
    if (ctx->precision == DALM_MIXED) {
        out[0] = it - ctx->fp64_finish_it;
        out[1] = ctx->fp64_finish_it;
        out[2] = ctx->outer_it;
    } else {
        out[0] = 0;
        out[1] = it;
        out[2] = 0;
    }
    return DALM_OK;
}

int dalm_set_fp32_pass_cap(dalm_ctx* ctx, int64_t n) {
    CHECK_CTX();
    if (n < 0) return DALM_BAD_ARG;
    ctx->fp32_pass_cap = n;
    return DALM_OK;
}

int dalm_probe_kernel(dalm_ctx* ctx, int64_t kernel) {
    CHECK_CTX();
    if (kernel < 0 || kernel > DALM_K_HALF_S) return DALM_BAD_ARG;
    TRY(probe_collect(ctx));
    ctx->probe.kernel = (int)kernel;
    ctx->probe.total_ms = 0;
    ctx->probe.launches = 0;
    return DALM_OK;
}

int dalm_probe_read(dalm_ctx* ctx, int64_t* launches, double* total_ms) {
    CHECK_CTX();
    TRY(probe_collect(ctx));
    if (launches) *launches = ctx->probe.launches;
    if (total_ms) *total_ms = ctx->probe.total_ms;
    return DALM_OK;
}

int dalm_run(dalm_ctx* ctx, int64_t n, dalm_step_stats* stats) {
    CHECK_CTX();
    dalm_step_stats st{};
    // a breakdown ends that step's solve and the run goes on (src/eq.h:349-353)
    for (int64_t s = 0; s < n; ++s) {
        const int rc = dalm_step(ctx, &st);
        if (rc && rc != DALM_BREAKDOWN) return rc;
    }
    if (stats) *stats = st;
    return DALM_OK;
}

}  // extern "C"
