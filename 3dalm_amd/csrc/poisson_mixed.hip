// poisson_mixed.hip -- Poisson solve of the precision form.
//
// Preconditioned BiCGStab (the recurrences of src/eq.h:333-473) whose iterate,
// residual and direction vectors (x = p, r, p, s, q = A z, t = A z') and every dot
// product stay FP64, while the preconditioner -- m Jacobi sweeps or SOR double
// sweeps per application, more than half of the solver's memory traffic -- runs on
// FP32 vectors: FP32 copies of the vector being preconditioned and FP32 iterates z.
// The operator products A z are evaluated in FP64 from the FP32 z, so the residual
// recurrence stays consistent with the FP64 iterate; the shadow vector r0 (only
// ever used in dot products) is FP32.  A solve ends with an FP64 true-residual
// check |b - A p| / sqrt(N) <= tol (the reference's test, src/eq.h:399-401); if the
// recursive residual drifted above it, the solve restarts from the true residual
// (outer iterative refinement in FP64).
//
// The limit of FP32 here: the normalised operator of a stretched grid is badly
// conditioned (about 2.5e6 at 256x128x128 with tanh beta = 1.5).  Rounding a
// preconditioned direction z to FP32 adds high-frequency noise of relative size
// u32, which A amplifies by up to cond(A) against the smooth part of A z -- so
// once the residual is dominated by smooth modes (about 4e-7 at that size) the
// FP32-preconditioned iteration stagnates or diverges.  A watchdog on the
// recursive residual detects that; the pass is then abandoned (the pressure
// restored to its start if the pass made it worse) and the solve is finished by
// the FP64 BiCGStab, so every step meets the same tolerance as the FP64 form.
//
// Algorithmic bytes per iteration (m = 3): 208 per cell against 280 for the FP64
// solver; iterations reported = BiCGStab iterations over all passes.
#include <float.h>

#include <cstdio>
#include <cstdlib>

#include "ctx.h"

namespace {

__device__ __forceinline__ float diag_f(const PCoefF& A, int i, int j, int k) {
    return (((A.acx[i] - A.anu[j]) - A.asu[j]) - A.atu[k]) - A.abu[k];
}
// 7-point product in FP32 in the reference's order, the i-neighbours from
// registers (column streaming)
__device__ __forceinline__ float apply7fs(const Dims& d, const PCoefF& A, const float* x, int c, int i, int j,
                                          int k, float ac, float xc, float xp, float xm) {
    return ac * xc + A.ae[i] * xp + A.aw[i] * xm + A.an[j] * x[c + d.sj] + A.as[j] * x[c - d.sj] +
           A.at[k] * x[c + 1] + A.ab[k] * x[c - 1];
}
__device__ __forceinline__ double diag64(const PCoef& A, int i, int j, int k) {
    return ((((A.acx[i] - A.anu[j]) - A.asu[j]) - A.atu[k]) - A.abu[k]) / A.max_diag;
}

__global__ __launch_bounds__(BLOCK) void kf_jacobi(Dims d, PCoefF A, const float* __restrict__ zin,
                                                   float* __restrict__ zout, const float* __restrict__ v,
                                                   const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const int j = cl.j, k = cl.k;
    int c = d.at(cl.i0, j, k);
    float xm = zin[c - d.si], x0 = zin[c];
    for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
        const float xp = zin[c + d.si];
        float ac = diag_f(A, i, j, k);
        zout[c] = x0 + (v[c] - apply7fs(d, A, zin, c, i, j, k, ac, x0, xp, xm)) * __builtin_amdgcn_rcpf(ac);
        xm = x0;
        x0 = xp;
    }
}

__global__ __launch_bounds__(BLOCK) void kf_sor(Dims d, PCoefF A, float* __restrict__ x,
                                                const float* __restrict__ b, float omega, int color,
                                                const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const int j = cl.j, k = cl.k;
    int c = d.at(cl.i0, j, k);
    float xm = x[c - d.si], x0 = x[c];
    for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
        const float xp = x[c + d.si];
        if (((i + j + k) & 1) == color) {
            float ac = diag_f(A, i, j, k);
            x[c] = x0 + omega * (b[c] - apply7fs(d, A, x, c, i, j, k, ac, x0, xp, xm)) * __builtin_amdgcn_rcpf(ac);
        }
        xm = x0;
        x0 = xp;
    }
}


// A z in FP64 arithmetic from an FP32 vector, in the FP64 form's summation order
__device__ __forceinline__ double apply7h(const Dims& d, const PCoef& A, const float* z, int c, int i, int j, int k,
                                          double ac) {
    return ac * (double)z[c] + A.ae[i] * (double)z[c + d.si] + A.aw[i] * (double)z[c - d.si] +
           A.an[j] * (double)z[c + d.sj] + A.as[j] * (double)z[c - d.sj] + A.at[k] * (double)z[c + 1] +
           A.ab[k] * (double)z[c - 1];
}

__global__ void kh_state_reset(SolveState* st, double tol, long long max_it, double root_n) {
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
    }
}

// start of a pass: r = b - A x (FP64), r0 = (float) r; partials <r0, r>, |r|^2
__global__ __launch_bounds__(BLOCK) void kh_init(Dims d, PCoef A, const double* __restrict__ x,
                                                 const double* __restrict__ b, double* __restrict__ r,
                                                 float* __restrict__ r0, double* partials) {
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const int j = cl.j, k = cl.k;
        int c = d.at(cl.i0, j, k);
        double xm = x[c - d.si], x0 = x[c];
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            const double xp = x[c + d.si];
            double ac = diag64(A, i, j, k);
            double s = ac * x0 + A.ae[i] * xp + A.aw[i] * xm + A.an[j] * x[c + d.sj] + A.as[j] * x[c - d.sj] +
                       A.at[k] * x[c + 1] + A.ab[k] * x[c - 1];
            double rv = b[c] - s;
            float rf = (float)rv;
            r[c] = rv;
            r0[c] = rf;
            v[0] += (double)rf * rv;
            v[1] += rv * rv;
            xm = x0;
            x0 = xp;
        }
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials);
}

// p = r (first iteration) or r + beta (p - omega q) in FP64; its FP32 copy, and
// the first preconditioner sweep from zero (Jacobi: z = p / a_c; SOR: z = 0)
__global__ __launch_bounds__(BLOCK) void kh_p(Dims d, PCoefF Af, const double* __restrict__ r,
                                              double* __restrict__ p, const double* __restrict__ q,
                                              float* __restrict__ p32, float* __restrict__ z, int pc_first,
                                              const SolveState* st) {
    if (st->done) return;
    const bool first = st->it == 0;
    const double beta = st->beta, om = st->omega;
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        int c = d.at(cc.i, cc.j, cc.k);
        double pv = first ? r[c] : r[c] + beta * (p[c] - om * q[c]);
        p[c] = pv;
        float pf = (float)pv;
        p32[c] = pf;
        z[c] = pc_first ? pf * __builtin_amdgcn_rcpf(diag_f(Af, cc.i, cc.j, cc.k)) : 0.f;
    }
}

// s = r - alpha q in FP64; its FP32 copy and the first preconditioner sweep
__global__ __launch_bounds__(BLOCK) void kh_s(Dims d, PCoefF Af, const double* __restrict__ r,
                                              const double* __restrict__ q, double* __restrict__ s,
                                              float* __restrict__ s32, float* __restrict__ z, int pc_first,
                                              const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha;
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        int c = d.at(cc.i, cc.j, cc.k);
        double sv = r[c] - alpha * q[c];
        s[c] = sv;
        float sf = (float)sv;
        s32[c] = sf;
        z[c] = pc_first ? sf * __builtin_amdgcn_rcpf(diag_f(Af, cc.i, cc.j, cc.k)) : 0.f;
    }
}

// y = A z (FP64 from FP32 z) with <a1, y> (WHICH 0: a1 = r0) or <y, a1>, <y, y>
// (WHICH 1: a1 = s)
template <int WHICH, typename TA>
__global__ __launch_bounds__(BLOCK) void kh_spmv(Dims d, PCoef A, const float* __restrict__ z,
                                                 double* __restrict__ y, const TA* __restrict__ a1,
                                                 double* partials, const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const int j = cl.j, k = cl.k;
        int c = d.at(cl.i0, j, k);
        double xm = z[c - d.si], x0 = z[c];
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            const double xp = z[c + d.si];
            double ac = diag64(A, i, j, k);
            double yv = ac * x0 + A.ae[i] * xp + A.aw[i] * xm + A.an[j] * (double)z[c + d.sj] +
                        A.as[j] * (double)z[c - d.sj] + A.at[k] * (double)z[c + 1] + A.ab[k] * (double)z[c - 1];
            y[c] = yv;
            if (WHICH == 0) {
                v[0] += (double)a1[c] * yv;
            } else {
                v[0] += yv * (double)a1[c];
                v[1] += yv * yv;
            }
            xm = x0;
            x0 = xp;
        }
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials);
}

// x += alpha z_p + omega z_s; r = s - omega t; partials |r|^2, <r0, r>
__global__ __launch_bounds__(BLOCK) void kh_xr(Dims d, double* __restrict__ x, const float* __restrict__ pp,
                                               const float* __restrict__ ss, const double* __restrict__ s,
                                               const double* __restrict__ t_, double* __restrict__ r,
                                               const float* __restrict__ r0, double* partials,
                                               const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha, om = st->omega;
    TileRange tr = block_tiles(d);
    double v[2] = {0.0, 0.0};
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        int c = d.at(cc.i, cc.j, cc.k);
        x[c] = x[c] + alpha * (double)pp[c] + om * (double)ss[c];
        double rv = s[c] - om * t_[c];
        r[c] = rv;
        v[0] += rv * rv;
        v[1] += (double)r0[c] * rv;
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials);
}

int read_states(dalm_ctx* ctx) {
    HIP_TRY(hipMemcpyAsync(ctx->h_solve, ctx->solve, 2 * sizeof(SolveState), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

long long chunk_for(long long done_it, long long predicted, long long max_it) {
    long long left = predicted - done_it;
    long long c = left > 8 ? left : 8;
    if (c > 256) c = 256;
    if (done_it + c > max_it) c = max_it - done_it;
    return c < 1 ? 1 : c;
}

}  // namespace

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

int run_poisson_mixed(dalm_ctx* ctx, int pc_kind, double pc_omega, int pc_it, long long max_it, double tol,
                      long long* it, double* err) {
    const Dims& d = ctx->d;
    const PCoefF Af = ctx->pcf;
    const PCoef A = ctx->coef();
    SolveState* st = ctx->solve;
    const int jac = pc_kind == DALM_PC_JACOBI;
    const int pc_first = (jac && pc_it >= 1) ? 1 : 0;
    static const bool trace = getenv("DALM_TRACE") != nullptr;
    const double root_n = sqrt((double)ctx->N);
    // FP32 preconditioner application on v32 into z0 (spare: the other Jacobi buffer)
    auto precondition = [&](float* z0, float* spare, const float* v32) -> float* {
        float* cur = z0;
        if (jac) {
            float* other = spare;
            for (int m = 1; m < pc_it; ++m) {
                probe_begin(ctx, DALM_K_JACOBI32);
                hipLaunchKernelGGL(kf_jacobi, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, Af, cur, other, v32,
                                   (const SolveState*)st);
                probe_end(ctx, DALM_K_JACOBI32);
                float* tsw = cur;
                cur = other;
                other = tsw;
            }
        } else {
            for (int m = 0; m < pc_it; ++m)
                for (int color = 0; color < 2; ++color)
                    hipLaunchKernelGGL(kf_sor, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, Af, cur, v32,
                                       (float)pc_omega, color, (const SolveState*)st);
        }
        return cur;
    };
    long long total = 0;
    int status = DALM_NONCONVERGED;
    double e_out = 0;
    bool finish64 = false;
    ctx->outer_it = 0;
    ctx->fp64_finish_it = 0;
    for (int pass = 0; pass < 8; ++pass) {
        hipLaunchKernelGGL(kh_state_reset, dim3(1), dim3(64), 0, ctx->stream, st, tol, max_it - total, root_n);
        hipLaunchKernelGGL(kh_init, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, ctx->p, ctx->b, ctx->r,
                           ctx->fr0, ctx->partials);
        LAUNCH_CHECK();
        if (int rc = launch_finalize(ctx, FIN_HX_INIT, ctx->sgrid, st, nullptr)) return rc;
        int rc = read_states(ctx);
        if (rc) return rc;
        ctx->outer_it = pass + 1;
        e_out = ctx->h_solve[0].err;
        if (trace) fprintf(stderr, "[mixed] pass %d: fp64 residual %.3e after %lld iterations\n", pass, e_out, total);
        if (ctx->h_solve[0].done) {
            status = ctx->h_solve[0].status;
            break;
        }
        if (total >= max_it) break;
        const double start_err = e_out;
        // the pressure at the start of the pass, restored if the pass makes it worse
        HIP_TRY(hipMemcpyAsync(ctx->tmp2, ctx->p, sizeof(double) * ctx->L, hipMemcpyDeviceToDevice, ctx->stream));
        double best = start_err, anchor = start_err;  // lowest residual; last 30% milestone
        long long anchor_it = 0;
        bool stalled = false;
        long long launched = 0;
        while (!ctx->h_solve[0].done && launched < max_it - total) {
            long long chunk = chunk_for(launched, ctx->prev_it[1] - total, max_it - total);
            if (chunk > 32) chunk = 32;
            for (long long c = 0; c < chunk; ++c) {
                probe_begin(ctx, DALM_K_PUPDATE_MIXED);
                hipLaunchKernelGGL(kh_p, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, d, Af, ctx->r, ctx->kp,
                                   ctx->q, ctx->fp, ctx->fpp, pc_first, (const SolveState*)st);
                probe_end(ctx, DALM_K_PUPDATE_MIXED);
                float* ppz = precondition(ctx->fpp, ctx->ftmp, ctx->fp);
                probe_begin(ctx, DALM_K_SPMV_MIXED);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kh_spmv<0, float>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d,
                                   A, ppz, ctx->q, ctx->fr0, ctx->partials, (const SolveState*)st);
                probe_end(ctx, DALM_K_SPMV_MIXED);
                if ((rc = launch_finalize(ctx, FIN_ALPHA, ctx->sgrid, st, nullptr))) return rc;
                probe_begin(ctx, DALM_K_SUPDATE_MIXED);
                hipLaunchKernelGGL(kh_s, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, d, Af, ctx->r, ctx->q, ctx->s,
                                   ctx->fs, ctx->fss, pc_first, (const SolveState*)st);
                probe_end(ctx, DALM_K_SUPDATE_MIXED);
                float* ssz = precondition(ctx->fss, ctx->ftmp2, ctx->fs);
                probe_begin(ctx, DALM_K_SPMV_MIXED);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kh_spmv<1, double>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream,
                                   d, A, ssz, ctx->t, ctx->s, ctx->partials, (const SolveState*)st);
                probe_end(ctx, DALM_K_SPMV_MIXED);
                if ((rc = launch_finalize(ctx, FIN_OMEGA, ctx->sgrid, st, nullptr))) return rc;
                probe_begin(ctx, DALM_K_UPDATE_MIXED);
                hipLaunchKernelGGL(kh_xr, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, d, ctx->p, ppz, ssz, ctx->s,
                                   ctx->t, ctx->r, ctx->fr0, ctx->partials, (const SolveState*)st);
                probe_end(ctx, DALM_K_UPDATE_MIXED);
                if ((rc = launch_finalize(ctx, FIN_BI_XR, ctx->grid, st, nullptr))) return rc;
            }
            LAUNCH_CHECK();
            launched += chunk;
            rc = read_states(ctx);
            if (rc) return rc;
            if (trace)
                fprintf(stderr, "[mixed]   it %lld: recursive residual %.3e alpha %.3e omega %.3e\n",
                        ctx->h_solve[0].it, ctx->h_solve[0].err, ctx->h_solve[0].alpha, ctx->h_solve[0].omega);
            const double e = ctx->h_solve[0].err;
            const long long itn = ctx->h_solve[0].it;
            if (ctx->h_solve[0].done) break;
            if (e < best) best = e;
            if (e < 0.7 * anchor) {
                anchor = e;
                anchor_it = itn;
            }
            // divergence, or no 30% gain in 96 iterations
            if (!(e == e) || e > 100 * best || itn - anchor_it >= 96) {
                stalled = true;
                break;
            }
        }
        total += ctx->h_solve[0].it;
        status = ctx->h_solve[0].status;
        e_out = ctx->h_solve[0].err;
        if (stalled || status == DALM_BREAKDOWN || status == DALM_NONFINITE) {
            // true residual of the pass's end; keep the better of start and end
            hipLaunchKernelGGL(kh_state_reset, dim3(1), dim3(64), 0, ctx->stream, st, tol, 1, root_n);
            hipLaunchKernelGGL(kh_init, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, ctx->p, ctx->b, ctx->r,
                               ctx->fr0, ctx->partials);
            LAUNCH_CHECK();
            if ((rc = launch_finalize(ctx, FIN_HX_INIT, ctx->sgrid, st, nullptr))) return rc;
            if ((rc = read_states(ctx))) return rc;
            const double end_err = ctx->h_solve[0].err;
            if (!(end_err <= start_err))
                HIP_TRY(hipMemcpyAsync(ctx->p, ctx->tmp2, sizeof(double) * ctx->L, hipMemcpyDeviceToDevice,
                                       ctx->stream));
            if (trace)
                fprintf(stderr, "[mixed] FP32 preconditioning stalled at %.3e (pass start %.3e): FP64 finish\n",
                        end_err, start_err);
            finish64 = true;
            break;
        }
        if (status != DALM_OK) break;  // max_it: as the FP64 solver
        // converged by the recursive residual: the next pass checks the true one
    }
    ctx->prev_it[1] = total;
    if (finish64 && total < max_it) {
        long long it64 = 0;
        status = run_bicgstab_on(ctx, ctx->p, ctx->b, 2, pc_kind, pc_omega, pc_it, max_it - total, tol, &it64, &e_out);
        if (status == DALM_HIP_ERROR || status == DALM_BAD_ARG) return status;
        ctx->fp64_finish_it = it64;
        total += it64;
        if (trace) fprintf(stderr, "[mixed] FP64 finish: %lld iterations, residual %.3e\n", it64, e_out);
    } else if (finish64) {
        status = DALM_NONCONVERGED;
    }
    *it = total;
    *err = e_out;
    return status;
}
