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
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (a.mode == FIN_MX_OMEGA) {
        reduce_partials<5>(a.partials, a.nb, s);
    } else {
        double s2[2];
        reduce_partials<2>(a.partials, a.nb, s2);
        s[0] = s2[0];
        s[1] = s2[1];
    }
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
        if (!st->done) {
            st->alpha = st->rho / s[0];
            st->pending = 0;  // a fused p-half has formed the previous iteration's r
        }
        break;
    case FIN_OMEGA:
        if (!st->done) st->omega = s[0] / s[1];
        break;
    case FIN_BI_XR: {
        if (st->done) break;
        // the fused iteration's x updates are applied kXGroup at a time, after the last
        // of them, with their alpha and omega
        for (int q = 0; q + 1 < kXGroup; ++q) {
            st->alpha_h[q] = st->alpha_h[q + 1];
            st->omega_h[q] = st->omega_h[q + 1];
        }
        st->alpha_h[kXGroup - 1] = st->alpha;
        st->omega_h[kXGroup - 1] = st->omega;
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
    case FIN_MX_INIT: {
        // precision form, start of the solve: s[0] = <r0, r>, s[1] = |r|^2 of the
        // true FP64 residual; a residual that already meets the test ends the solve
        double rn = sqrt(s[1]);
        double err = rn / st->root_n;
        st->rho = s[0];
        st->err = err;
        st->maxr = rn;
        st->ru = 0;
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
    case FIN_MX_XR: {
        // precision form, end of an iteration on the FP32 recursive residual:
        // s[0] = |r|^2, s[1] = <r0, r>.  Meeting the stopping test, reaching the
        // iteration limit or falling below delta x the largest norm since the last
        // replacement asks for a residual replacement (FIN_MX_RU decides then).
        if (st->done) break;
        double rn = sqrt(s[0]);
        double err = rn / st->root_n;
        long long it = st->it + 1;
        st->it = it;
        st->err = err;
        if (!(err == err) || isinf(err)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (err <= st->tol || it >= st->max_it || rn < kMixedReliableDelta * st->maxr) {
            st->ru = 1;
        } else {
            if (rn > st->maxr) st->maxr = rn;
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
        }
        break;
    }
    case FIN_MX_OMEGA: {
        // precision form, fused iteration (the s-half's dots): s[0] = <t, s>,
        // s[1] = <t, t>, s[2] = <r0, s>, s[3] = <r0, t>, s[4] = |s|^2.  omega, and
        // the new residual r = s - omega t through its norm and <r0, r> expanded
        // in these dots (no pass over r); the update of x and r is pending until
        // the next p-half (or a replacement, or the flush at the end of the pass).
        // Then the tests of FIN_MX_XR.
        if (st->done) break;
        const double om = s[0] / s[1];
        st->omega = om;
        st->pending = 1;
        // this iteration's pressure update is outstanding (kg_apply_pending4 applies the
        // outstanding ones when kXGroupF have gathered or a replacement is due, and
        // FIN_MX_RU clears the count, so it is below kXGroupF here)
        if (st->xdef >= kXGroupF) {
            st->done = 1;
            st->status = DALM_NONFINITE;
            break;
        }
        st->alpha_d[st->xdef] = st->alpha;
        st->omega_d[st->xdef] = om;
        st->xdef += 1;
        double r2 = s[4] - 2.0 * om * s[0] + om * om * s[1];
        if (r2 < 0.0) r2 = 0.0;
        const double rn = sqrt(r2);
        const double err = rn / st->root_n;
        const long long it = st->it + 1;
        st->it = it;
        st->err = err;
        if (!(err == err) || isinf(err) || !(om == om) || isinf(om)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (err <= st->tol || it >= st->max_it || rn < kMixedReliableDelta * st->maxr) {
            st->ru = 1;
        } else {
            if (rn > st->maxr) st->maxr = rn;
            const double rho_new = s[2] - om * s[3];
            if (fabs(rho_new) < (double)FLT_MIN) {
                st->err = rho_new;
                st->done = 1;
                st->status = DALM_BREAKDOWN;
            } else {
                st->beta = (rho_new / st->rho) * (st->alpha / om);
                st->rho_prev = st->rho;
                st->rho = rho_new;
            }
        }
        break;
    }
    case FIN_MX_RU: {
        // precision form, residual replacement: s[0] = <r0, r>, s[1] = |b - A x|^2
        // in FP64.  The solve ends only on the true residual.
        if (st->done) break;
        if (st->ru || st->xdef >= kXGroupF) st->xdef = 0;  // kg_apply_pending4 applied them
        if (!st->ru) break;
        st->ru = 0;
        st->pending = 0;  // the replacement replaced r
        st->n_ru += 1;
        double rn = sqrt(s[1]);
        double err = rn / st->root_n;
        st->err = err;
        st->maxr = rn;
        if (!(err == err) || isinf(err)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (err <= st->tol) {
            st->done = 1;
            st->status = DALM_OK;
        } else if (st->it >= st->max_it) {
            st->done = 1;
            st->status = DALM_NONCONVERGED;
        } else {
            double rho_new = s[0];
            if (fabs(rho_new) < (double)FLT_MIN) {
                st->err = rho_new;
                st->done = 1;
                st->status = DALM_BREAKDOWN;
            } else {
                st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
                st->rho_prev = st->rho;
                st->rho = rho_new;
            }
        }
        break;
    }
    case FIN_MX_UP: {
        // precision form handing its iteration over to FP64: s[0] = <r0, r>,
        // s[1] = |b - A x|^2 with the true residual; the iteration goes on from the
        // current direction, as after a residual replacement
        if (st->done) break;
        st->ru = 0;
        double rn = sqrt(s[1]);
        double err = rn / st->root_n;
        st->err = err;
        if (!(err == err) || isinf(err)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (err <= st->tol) {
            st->done = 1;
            st->status = DALM_OK;
        } else if (st->it >= st->max_it) {
            st->done = 1;
            st->status = DALM_NONCONVERGED;
        } else {
            double rho_new = s[0];
            if (fabs(rho_new) < (double)FLT_MIN) {
                st->err = rho_new;
                st->done = 1;
                st->status = DALM_BREAKDOWN;
            } else {
                st->beta = (rho_new / st->rho_prev) * (st->alpha / st->omega);
                st->rho = rho_new;
            }
        }
        break;
    }
    case FIN_MX_SOR: {
        // precision form of run_sor: s[0] = |b - A x|^2 in FP64 after the last
        // refinement; aux1 holds the SOR iterations of one refinement.  At least one
        // refinement runs, as the reference's do-while does at least one iteration.
        if (st->done) break;
        double err = sqrt(s[0]) / st->root_n;
        st->err = err;
        if (!(err == err) || isinf(err)) {
            st->done = 1;
            st->status = DALM_NONFINITE;
        } else if (st->it > 0 && err <= st->tol) {
            st->done = 1;
            st->status = DALM_OK;
        } else if (st->it >= st->max_it) {
            st->done = 1;
            st->status = DALM_NONCONVERGED;
        } else {
            st->it += (long long)st->aux1;
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
