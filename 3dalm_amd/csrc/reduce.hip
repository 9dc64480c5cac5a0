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
    SolveState* st;         // state of the running solve
    StepScalars* sc;
    double* out;
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
    case FIN_HX_INIT: {
        // start of a precision-form pass: s[0] = <r0, r>, s[1] = |r|^2 of the
        // true FP64 residual; a pass whose residual already meets the test ends
        // here (the restart after a converged pass)
        double err = sqrt(s[1]) / st->root_n;
        st->rho = s[0];
        st->err = err;
        if (!(err == err) || isinf(err)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (err <= st->tol) {
            st->done = 1;
            st->status = DALM_OK;
        } else if (fabs(s[0]) < (double)FLT_MIN) {
            st->err = s[0];
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

int launch_finalize(dalm_ctx* ctx, int mode, int nb, SolveState* st, double* out) {
    FinArgs a;
    a.mode = mode;
    a.nb = nb;
    a.partials = ctx->partials;
    a.st = st;
    a.sc = ctx->scal;
    a.out = out;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(BLOCK), 0, ctx->stream, a);
    HIP_TRY(hipGetLastError());
    return DALM_OK;
}
