// poisson_cg.hip -- preconditioned conjugate gradients for the pressure Poisson
// equation, in FP64 or with FP32 vectors (the precision form).
//
// No counterpart in the reference, which solves with SOR or BiCGStab
// (src/eq.h:145-473).  The normalised operator A of build_A (src/eq.h:701-798)
// is symmetric in the cell-volume inner product <u, v>_V = sum V u v with
// V = dx_i dy_j dz_k (V A is a symmetric matrix: every face coefficient is the
// face area over the centre distance, divided by the cell volume), and it is
// negative definite (Neumann faces plus the Dirichlet x+ face).  The m-sweep
// Jacobi preconditioner started from zero, M^-1 = q(D^-1 A) D^-1 with
// q(t) = (1 - (1 - t)^m) / t, is a polynomial in D^-1 A, which is self-adjoint
// in the (V D) inner product, so M^-1 V^-1 is symmetric as well and CG applies to
// V A p = V b.  One CG iteration applies the operator and the preconditioner once
// each, against twice each for BiCGStab.
//
// Iteration (rho = <r, z>_V):
//   p = z + beta p;  q = A p;  alpha = rho / <p, q>_V
//   x += alpha p;  r -= alpha q;  z = M^-1 r;  beta = <r, z>_V / rho
// The p update is folded into the operator kernel (p at the neighbours recomputed
// from z and the previous p, written to a second buffer); the first Jacobi sweep
// z = r / a_c is folded into the x / r update.
//
// FP32 vectors (precision form): r, z, p, q are stored in FP32, every dot product
// and norm accumulates in FP64, A p is evaluated in FP64 and rounded once, and the
// pressure x stays FP64, updated with exactly the FP32 p that produced q.  The
// recursive FP32 residual drifts from b - A x by the rounding it accumulates, so
// it is replaced by the FP64 true residual (reliable updates, Sleijpen & van der
// Vorst 1996) whenever its norm falls below 0.1 of the largest norm since the last
// replacement, and whenever it meets the stopping test: the solve ends only on the
// true residual, |b - A x| / sqrt(N) <= tol (the reference's test,
// src/eq.h:399-401).  Unlike BiCGStab, CG keeps converging with FP32 vectors on
// these badly conditioned stretched-grid operators: its iteration count matches
// the FP64 CG's.  A watchdog (no 30% gain in 192 iterations, a rise above 1e3 x
// the best, or a breakdown) still finishes a stalled solve with the FP64 CG.
//
// Algorithmic bytes per iteration and cell, Jacobi m = 3, FP32 vectors: operator
// and p update 16, x / r update and first sweep 36, two sweeps 24 = 76
// (FP64 vectors: 128); a replacement costs 28.
#include <float.h>

#include <cstdio>
#include <cstdlib>
#include <utility>

#include "ctx.h"

namespace {

struct Vol {
    const double *dx, *dy, *dz;  // cell widths (padded, FP64)
};
__device__ __forceinline__ double vol(const Vol& V, int i, int j, int k) { return V.dx[i] * V.dy[j] * V.dz[k]; }

__device__ __forceinline__ float diag_of(const PCoefF& A, int i, int j, int k) {
    return (((A.acx[i] - A.anu[j]) - A.asu[j]) - A.atu[k]) - A.abu[k];
}
__device__ __forceinline__ double diag_of(const PCoef& A, int i, int j, int k) {
    return ((((A.acx[i] - A.anu[j]) - A.asu[j]) - A.atu[k]) - A.abu[k]) / A.max_diag;
}
// v / a_c in the storage precision's arithmetic
__device__ __forceinline__ float over(float v, float ac) { return v * __builtin_amdgcn_rcpf(ac); }
__device__ __forceinline__ double over(double v, double ac) { return v / ac; }

// Operator coefficients in the arithmetic of the Jacobi sweeps: FP32 pieces for
// FP32 vectors, the FP64 operator otherwise.
template <typename T>
struct SweepCoef;
template <>
struct SweepCoef<float> {
    using type = PCoefF;
};
template <>
struct SweepCoef<double> {
    using type = PCoef;
};

__global__ void kc_state_reset(SolveState* st, double tol, long long max_it, double root_n, int reliable) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->rho = st->rho_prev = st->alpha = st->omega = 1.0;
        st->beta = 0;
        st->err = 0;
        st->tol = tol;
        st->root_n = root_n;
        st->it = 0;
        st->max_it = max_it;
        st->done = 0;
        st->status = DALM_OK;
        st->ru = 0;
        st->reliable = reliable;
        st->n_ru = 0;
        st->maxr = 0;
    }
}

// r = b - A x in FP64 (stored in T), z = r / a_c; partials <r, z>_V, |b - A x|^2.
// Without INIT this is the residual replacement, which returns at once unless the
// device state asks for one.
template <typename T, bool INIT>
__global__ __launch_bounds__(BLOCK) void kc_resid(Dims d, PCoef A, typename SweepCoef<T>::type As, Vol V,
                                                  const double* __restrict__ x, const double* __restrict__ b,
                                                  T* __restrict__ r, T* __restrict__ z, double* partials,
                                                  const SolveState* st) {
    if (!INIT && (st->done || !st->ru)) return;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const int j = cl.j, k = cl.k;
        int c = d.at(cl.i0, j, k);
        double xm = x[c - d.si], x0 = x[c];
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            const double xp = x[c + d.si];
            const double ac = diag_of(A, i, j, k);
            const double s = ac * x0 + A.ae[i] * xp + A.aw[i] * xm + A.an[j] * x[c + d.sj] +
                             A.as[j] * x[c - d.sj] + A.at[k] * x[c + 1] + A.ab[k] * x[c - 1];
            const double rv = b[c] - s;
            const T rt = (T)rv;
            const T zt = over(rt, diag_of(As, i, j, k));
            r[c] = rt;
            z[c] = zt;
            v[0] += vol(V, i, j, k) * (double)rt * (double)zt;
            v[1] += rv * rv;
            xm = x0;
            x0 = xp;
        }
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials);
}

// p = z (first iteration) or z + beta p_old, written to p_new; q = A p (FP64
// arithmetic, stored in T); partial <p, q>_V of the stored values.  The p of the
// neighbours is recomputed with the same rounding as its owner stores.
template <typename T>
__global__ __launch_bounds__(BLOCK) void kc_pq(Dims d, PCoef A, Vol V, const T* __restrict__ z,
                                               const T* __restrict__ pold, T* __restrict__ pnew,
                                               T* __restrict__ q, double* partials, const SolveState* st) {
    if (st->done) return;
    const bool first = st->it == 0;
    const double beta = st->beta;
    const Col cl = col_unit(d);
    double acc = 0.0;
    if (cl.ok) {
        auto P = [&](int c) -> double {
            return first ? (double)z[c] : (double)(T)((double)z[c] + beta * (double)pold[c]);
        };
        const int j = cl.j, k = cl.k;
        int c = d.at(cl.i0, j, k);
        double xm = P(c - d.si), x0 = P(c);
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            const double xp = P(c + d.si);
            const double ac = diag_of(A, i, j, k);
            const double y = ac * x0 + A.ae[i] * xp + A.aw[i] * xm + A.an[j] * P(c + d.sj) +
                             A.as[j] * P(c - d.sj) + A.at[k] * P(c + 1) + A.ab[k] * P(c - 1);
            const T yt = (T)y;
            pnew[c] = (T)x0;
            q[c] = yt;
            acc += vol(V, i, j, k) * x0 * (double)yt;
            xm = x0;
            x0 = xp;
        }
    }
    double w[1] = {acc};
    block_sum<1>(w);
    store_partials<1>(w, partials);
}

// x += alpha p (FP64); r -= alpha q; z = r / a_c; partials <r, z>_V, |r|^2
template <typename T>
__global__ __launch_bounds__(BLOCK) void kc_xr(Dims d, typename SweepCoef<T>::type As, Vol V, double* __restrict__ x,
                                               const T* __restrict__ p, const T* __restrict__ q, T* __restrict__ r,
                                               T* __restrict__ z, double* partials, const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const int j = cl.j, k = cl.k;
        int c = d.at(cl.i0, j, k);
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            x[c] = x[c] + alpha * (double)p[c];
            const T rt = (T)((double)r[c] - alpha * (double)q[c]);
            const T zt = over(rt, diag_of(As, i, j, k));
            r[c] = rt;
            z[c] = zt;
            v[0] += vol(V, i, j, k) * (double)rt * (double)zt;
            v[1] += (double)rt * (double)rt;
        }
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials);
}

// One Jacobi sweep z_out = z + (v - A z) / a_c in the storage precision's
// arithmetic; DOT: partial <v, z_out>_V (the last sweep)
template <typename T, bool DOT>
__global__ __launch_bounds__(BLOCK) void kc_jacobi(Dims d, typename SweepCoef<T>::type As, Vol V,
                                                   const T* __restrict__ zin, T* __restrict__ zout,
                                                   const T* __restrict__ v, double* partials, const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    double acc = 0.0;
    if (cl.ok) {
        const int j = cl.j, k = cl.k;
        int c = d.at(cl.i0, j, k);
        T xm = zin[c - d.si], x0 = zin[c];
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            const T xp = zin[c + d.si];
            const T ac = diag_of(As, i, j, k);
            const T s = ac * x0 + As.ae[i] * xp + As.aw[i] * xm + As.an[j] * zin[c + d.sj] +
                        As.as[j] * zin[c - d.sj] + As.at[k] * zin[c + 1] + As.ab[k] * zin[c - 1];
            const T vv = v[c];
            const T zn = x0 + over(vv - s, ac);
            zout[c] = zn;
            if (DOT) acc += vol(V, i, j, k) * (double)vv * (double)zn;
            xm = x0;
            x0 = xp;
        }
    }
    if (DOT) {
        double w[1] = {acc};
        block_sum<1>(w);
        store_partials<1>(w, partials);
    }
}

template <typename T>
struct CgBufs {
    T *r, *z, *zt, *pa, *pb, *q;
};
CgBufs<float> bufs_of(dalm_ctx* ctx, float*) {
    return CgBufs<float>{ctx->fr, ctx->fz, ctx->ftmp, ctx->fp, ctx->fs, ctx->fq};
}
CgBufs<double> bufs_of(dalm_ctx* ctx, double*) {
    return CgBufs<double>{ctx->r, ctx->pp, ctx->tmp, ctx->kp, ctx->s, ctx->q};
}
PCoefF sweep_coef(dalm_ctx* ctx, float*) { return ctx->pcf; }
PCoef sweep_coef(dalm_ctx* ctx, double*) { return ctx->coef(); }

int read_states(dalm_ctx* ctx) {
    HIP_TRY(hipMemcpyAsync(ctx->h_solve, ctx->solve, sizeof(SolveState), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

long long cg_chunk(long long done_it, long long predicted, long long max_it) {
    long long left = predicted - done_it;
    long long c = left > 8 ? left : 8;
    if (c > 64) c = 64;
    if (done_it + c > max_it) c = max_it - done_it;
    return c < 1 ? 1 : c;
}

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

// One CG solve on x with right-hand side b.  With `reliable` (FP32 vectors) the
// residual replacements run and *stalled reports a watchdog stop.
template <typename T>
int run_cg_impl(dalm_ctx* ctx, double* x, const double* b, int slot, int pc_it, long long max_it, double tol,
                bool reliable, long long* it, double* err, bool* stalled) {
    const Dims& d = ctx->d;
    const PCoef A = ctx->coef();
    const typename SweepCoef<T>::type As = sweep_coef(ctx, (T*)nullptr);
    const Vol V{ctx->dx, ctx->dy, ctx->dz};
    CgBufs<T> B = bufs_of(ctx, (T*)nullptr);
    SolveState* st = ctx->solve;
    const int kjac = sizeof(T) == 4 ? DALM_K_JACOBI32 : DALM_K_JACOBI;
    static const bool trace = getenv("DALM_TRACE") != nullptr;
    const double root_n = sqrt((double)ctx->N);
    *stalled = false;
    // z = M^-1 r from the first sweep already in B.z; returns the buffer holding z
    auto sweeps = [&]() -> T* {
        T* cur = B.z;
        T* other = B.zt;
        for (int m = 1; m < pc_it; ++m) {
            probe_begin(ctx, kjac);
            if (m == pc_it - 1)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kc_jacobi<T, true>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream,
                                   d, As, V, cur, other, B.r, ctx->partials, (const SolveState*)st);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kc_jacobi<T, false>), dim3(ctx->sgrid), dim3(BLOCK), 0,
                                   ctx->stream, d, As, V, cur, other, B.r, ctx->partials, (const SolveState*)st);
            probe_end(ctx, kjac);
            std::swap(cur, other);
        }
        return cur;
    };
    hipLaunchKernelGGL(kc_state_reset, dim3(1), dim3(64), 0, ctx->stream, st, tol, max_it, root_n, reliable ? 1 : 0);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kc_resid<T, true>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, As, V,
                       (const double*)x, b, B.r, B.z, ctx->partials, (const SolveState*)st);
    LAUNCH_CHECK();
    if (int rc = launch_finalize(ctx, FIN_CG_INIT, ctx->sgrid, st, nullptr)) return rc;
    T* zf = sweeps();
    if (int rc = launch_finalize(ctx, FIN_CG_RHO, ctx->sgrid, st, nullptr)) return rc;
    LAUNCH_CHECK();
    if (int rc = read_states(ctx)) return rc;
    const double start_err = ctx->h_solve->err;
    if (trace) fprintf(stderr, "[cg%d] residual %.3e at the start\n", (int)(8 * sizeof(T)), start_err);
    double best = start_err, anchor = start_err;
    long long anchor_it = 0, launched = 0;
    T* pold = B.pa;
    T* pnew = B.pb;
    probe_solve_begin(ctx);
    while (!ctx->h_solve->done && launched < max_it) {
        const long long chunk = cg_chunk(launched, ctx->prev_it[slot], max_it);
        for (long long c = 0; c < chunk; ++c) {
            probe_iteration(ctx, launched + c);
            probe_begin(ctx, DALM_K_CG_SPMV);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(kc_pq<T>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, V,
                               (const T*)zf, (const T*)pold, pnew, B.q, ctx->partials, (const SolveState*)st);
            probe_end(ctx, DALM_K_CG_SPMV);
            if (int rc = launch_finalize(ctx, FIN_CG_ALPHA, ctx->sgrid, st, nullptr)) return rc;
            probe_begin(ctx, DALM_K_CG_UPDATE);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(kc_xr<T>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, As, V, x,
                               (const T*)pnew, (const T*)B.q, B.r, B.z, ctx->partials, (const SolveState*)st);
            probe_end(ctx, DALM_K_CG_UPDATE);
            if (int rc = launch_finalize(ctx, FIN_CG_XR, ctx->sgrid, st, nullptr)) return rc;
            if (reliable) {
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kc_resid<T, false>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream,
                                   d, A, As, V, (const double*)x, b, B.r, B.z, ctx->partials, (const SolveState*)st);
                if (int rc = launch_finalize(ctx, FIN_CG_RU, ctx->sgrid, st, nullptr)) return rc;
            }
            zf = sweeps();
            if (int rc = launch_finalize(ctx, FIN_CG_RHO, ctx->sgrid, st, nullptr)) return rc;
            std::swap(pold, pnew);
        }
        LAUNCH_CHECK();
        launched += chunk;
        if (int rc = read_states(ctx)) return rc;
        const SolveState& h = *ctx->h_solve;
        if (trace)
            fprintf(stderr, "[cg%d]   it %lld: residual %.3e (%lld replacements)\n", (int)(8 * sizeof(T)), h.it, h.err,
                    h.n_ru);
        if (h.done || !reliable) continue;
        const double e = h.err;
        if (e < best) best = e;
        if (e < 0.7 * anchor) {
            anchor = e;
            anchor_it = h.it;
        }
        if (!(e == e) || e > 1e3 * best || h.it - anchor_it >= 192) {
            *stalled = true;
            break;
        }
    }
    const SolveState& h = *ctx->h_solve;
    probe_solve_end(ctx, h.it);
    *it = h.it;
    *err = h.err;
    ctx->prev_it[slot] = h.it;
    ctx->outer_it = h.n_ru;
    if (!h.done && !*stalled) return DALM_NONCONVERGED;
    return h.status;
}

}  // namespace

int run_cg_on(dalm_ctx* ctx, double* x, const double* b, int slot, int pc_it, long long max_it, double tol,
              long long* it, double* err) {
    bool stalled = false;
    return run_cg_impl<double>(ctx, x, b, slot, pc_it, max_it, tol, false, it, err, &stalled);
}

// Precision form: CG with FP32 vectors; a stalled or broken-down solve is
// finished by the FP64 CG from the better of the pressures before and after.
int run_cg_mixed(dalm_ctx* ctx, int pc_it, long long max_it, double tol, long long* it, double* err) {
    static const bool trace = getenv("DALM_TRACE") != nullptr;
    ctx->fp64_finish_it = 0;
    HIP_TRY(hipMemcpyAsync(ctx->tmp2, ctx->p, sizeof(double) * ctx->L, hipMemcpyDeviceToDevice, ctx->stream));
    bool stalled = false;
    long long it32 = 0;
    double e32 = 0;
    int status = run_cg_impl<float>(ctx, ctx->p, ctx->b, 1, pc_it, max_it, tol, true, &it32, &e32, &stalled);
    if (status == DALM_HIP_ERROR || status == DALM_BAD_ARG) return status;
    *it = it32;
    *err = e32;
    if (!(stalled || status == DALM_BREAKDOWN || status == DALM_NONFINITE)) return status;
    // true residual at the end against the start (both FP64)
    const Dims& d = ctx->d;
    const double root_n = sqrt((double)ctx->N);
    double e_end = 0, e_start = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const double* xv = pass == 0 ? ctx->p : ctx->tmp2;
        hipLaunchKernelGGL(kc_state_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->solve, tol, 1, root_n, 0);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kc_resid<double, true>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d,
                           ctx->coef(), ctx->coef(), Vol{ctx->dx, ctx->dy, ctx->dz}, xv, (const double*)ctx->b, ctx->r,
                           ctx->pp, ctx->partials, (const SolveState*)ctx->solve);
        HIP_TRY(hipGetLastError());
        if (int rc = launch_finalize(ctx, FIN_CG_INIT, ctx->sgrid, ctx->solve, nullptr)) return rc;
        if (int rc = read_states(ctx)) return rc;
        (pass == 0 ? e_end : e_start) = ctx->h_solve->err;
    }
    if (!(e_end <= e_start))
        HIP_TRY(hipMemcpyAsync(ctx->p, ctx->tmp2, sizeof(double) * ctx->L, hipMemcpyDeviceToDevice, ctx->stream));
    if (trace) fprintf(stderr, "[cg32] stalled at %.3e (start %.3e): FP64 finish\n", e_end, e_start);
    if (it32 >= max_it) return DALM_NONCONVERGED;
    long long it64 = 0;
    status = run_cg_on(ctx, ctx->p, ctx->b, 2, pc_it, max_it - it32, tol, &it64, err);
    ctx->fp64_finish_it = it64;
    *it = it32 + it64;
    return status;
}
