// reduce.hip -- the one-block finalize kernel that ends every reduction.
//
// A reducing kernel stores one partial per block (store_partials); this kernel
// then sums the partials in block order (fixed tree: bitwise reproducible) and
// performs the scalar algebra that the reference does on the host between kernel
// launches (src/eq.h:297-305 SOR stopping test; src/eq.h:349-401 BiCGStab
// alpha / omega / beta / rho, breakdown and stopping tests; src/solver.h:1184-1186
// avg_div and max CFL), writing the result into device state that the following
// kernels and the host read.  Compiled with -ffp-contract=off.
#include <float.h>

#include "ctx.h"

namespace {

struct FinArgs {
    int mode;
    int nb;                 // number of partials per column
    const double* partials;
    SolveState* st;         // primary state (FP64 solve, or the inner solve)
    SolveState* st2;        // outer state (precision form)
    StepScalars* sc;
    double* out;
    double eta;
    long long inner_max;
};

__global__ __launch_bounds__(BLOCK) void k_finalize(FinArgs a) {
    double s[2];
    reduce_partials<2>(a.partials, a.nb, s);
    double mx = 0;
    if (a.mode == FIN_POST) mx = reduce_partials_max(a.partials + a.nb, a.nb);
    if (threadIdx.x != 0) return;
    SolveState* st = a.st;
    switch (a.mode) {
    case FIN_POST:
        a.sc->avg_div_sq = s[0];
        a.sc->max_cfl = mx;
        break;
    case FIN_SUM:
        *a.out = s[0];
        break;
    case FIN_RESID:
        st->err = sqrt(s[0]) / st->root_n;
        st->aux0 = s[0];
        break;
    case FIN_SOR: {
        if (st->done) break;
        double err = sqrt(s[0]) / st->root_n;
        st->err = err;
        long long it = st->it + 1;
        st->it = it;
        if (!(err == err) || isinf(err)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (!(it < st->max_it && err > st->tol)) {
            st->done = 1;
            st->status = err > st->tol ? DALM_NONCONVERGED : DALM_OK;
        }
        break;
    }
    case FIN_BI_INIT:
        st->rho = s[0];
        if (fabs(s[0]) < (double)FLT_MIN) {
            st->err = s[0];
            st->done = 1;
            st->status = DALM_BREAKDOWN;
        }
        break;
    case FIN_ALPHA:
        if (!st->done) st->alpha = st->rho / s[0];
        break;
    case FIN_OMEGA:
        if (!st->done) st->omega = s[0] / s[1];
        break;
    case FIN_BI_XR: {
        if (st->done) break;
        double err = sqrt(s[0]) / st->root_n;
        long long it = st->it + 1;
        st->it = it;
        st->err = err;
        if (!(err == err) || isinf(err)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (it < st->max_it && err > st->tol) {
            double rho_new = s[1];
            if (fabs(rho_new) < (double)FLT_MIN) {
                st->err = rho_new;
                st->done = 1;
                st->status = DALM_BREAKDOWN;
            } else {
                st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
                st->rho_prev = st->rho;
                st->rho = rho_new;
            }
        } else {
            st->done = 1;
            st->status = err > st->tol ? DALM_NONCONVERGED : DALM_OK;
        }
        break;
    }
    case FIN_RES64: {
        // start of one refinement pass of the precision form
        SolveState* outer = a.st2;
        double err = sqrt(s[0]) / outer->root_n;
        outer->aux1 = outer->err;  // previous pass's residual
        outer->err = err;
        outer->it += 1;
        st->it = 0;
        st->rho = s[1];
        st->alpha = st->omega = 1.0;
        st->beta = 0;
        st->err = err;
        st->root_n = outer->root_n;
        st->tol = fmax(0.5 * outer->tol, a.eta * err);
        st->max_it = a.inner_max;
        st->status = DALM_OK;
        st->done = 0;
        if (!(err == err) || isinf(err)) {
            outer->status = DALM_NONFINITE;
            outer->done = 1;
            st->done = 1;
        } else if (err <= outer->tol || a.inner_max <= 0) {
            outer->done = 1;
            outer->status = err <= outer->tol ? DALM_OK : DALM_NONCONVERGED;
            st->done = 1;
        } else if (fabs(s[1]) < (double)FLT_MIN) {
            st->done = 1;
            st->status = DALM_BREAKDOWN;
        }
        break;
    }
    default:
        break;
    }
}

}  // namespace

int launch_finalize(dalm_ctx* ctx, int mode, int nb, SolveState* st, SolveState* st2, double* out, double eta,
                    long long inner_max) {
    FinArgs a;
    a.mode = mode;
    a.nb = nb;
    a.partials = ctx->partials;
    a.st = st;
    a.st2 = st2;
    a.sc = ctx->scal;
    a.out = out;
    a.eta = eta;
    a.inner_max = inner_max;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(BLOCK), 0, ctx->stream, a);
    HIP_TRY(hipGetLastError());
    return DALM_OK;
}
