// diag.hip -- the fields of the reference's output path (src/solver.h:1193-1250):
//   * the Q-criterion q = -1/2 (u_i,j u_j,i) by central differences over
//     x[i+1] - x[i-1] (calc_q, src/cfd.h:948-1077);
//   * the running time average Utavg = (1 - w) Utavg + w U over every padded cell
//     (calc_ax_plus_by on Cfd::Utavg, src/solver.h:1194-1204, src/mv.h:321-335);
//   * the divergence field div = sum of face fluxes / volume, stored for output
//     (calc_divergence, src/cfd.h:1079-1107) -- this file's kernel serves the
//     precision form, whose fused post pass keeps only the norm;
//   * U gathered at the probe cells (the per-probe `acc update host(U[id:1])`,
//     src/solver.h:1205-1216), one kernel and one copy for all probes.
//
// Arithmetic is FP64 in the reference's order and the file is compiled with
// -ffp-contract=off, so on the same U / JU the results match a g++ -O2 build of
// the reference bit for bit.  In a precision-form context U is FP32 and is widened
// before any arithmetic.
//
// All four are HBM-bound streams: calc_q reads U and writes q (32 B per interior
// cell in FP64), the time average reads U and Utavg and writes Utavg (72 B per
// padded cell).  calc_q and div use the column walk of the Poisson kernels (plane
// i+1 requested before plane i is processed).
#include "stencil.h"

namespace {

template <typename T>
struct QAux {
    T e[3], w[3], n[3], s[3], t[3], b[3];
};

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_calc_q(Dims d, const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ z, const T* __restrict__ u0,
                                                  const T* __restrict__ u1, const T* __restrict__ u2,
                                                  double* __restrict__ q) {
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const T* u[3] = {u0, u1, u2};
    const double dyns = y[cl.j + 1] - y[cl.j - 1];
    const double dztb = z[cl.k + 1] - z[cl.k - 1];
    walk_points<QAux<T>>(
        d, cl,
        [&](int c) {
            QAux<T> a;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                a.e[m] = u[m][c + d.si];
                a.w[m] = u[m][c - d.si];
                a.n[m] = u[m][c + d.sj];
                a.s[m] = u[m][c - d.sj];
                a.t[m] = u[m][c + 1];
                a.b[m] = u[m][c - 1];
            }
            return a;
        },
        [&](int i, int c, const QAux<T>& a) {
            const double dxew = x[i + 1] - x[i - 1];
            double gx[3], gy[3], gz[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                gx[m] = ((double)a.e[m] - (double)a.w[m]) / dxew;
                gy[m] = ((double)a.n[m] - (double)a.s[m]) / dyns;
                gz[m] = ((double)a.t[m] - (double)a.b[m]) / dztb;
            }
            const double dudx = gx[0], dudy = gy[0], dudz = gz[0];
            const double dvdx = gx[1], dvdy = gy[1], dvdz = gz[1];
            const double dwdx = gx[2], dwdy = gy[2], dwdz = gz[2];
            q[c] = -0.5 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz +
                           2 * (dudy * dvdx + dudz * dwdx + dvdz * dwdy));
        });
}

struct DivAux {
    double je, jw, jn, js, jt, jb;
};

__global__ __launch_bounds__(BLOCK) void k_div_store(Dims d, const double* __restrict__ dx,
                                                     const double* __restrict__ dy,
                                                     const double* __restrict__ dz,
                                                     const double* __restrict__ J0,
                                                     const double* __restrict__ J1,
                                                     const double* __restrict__ J2, double* __restrict__ div) {
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const double dy_j = dy[cl.j], dz_k = dz[cl.k];
    walk_points<DivAux>(
        d, cl,
        [&](int c) { return DivAux{J0[c], J0[c - d.si], J1[c], J1[c - d.sj], J2[c], J2[c - 1]}; },
        [&](int i, int c, const DivAux& a) {
            const double vol = dx[i] * dy_j * dz_k;
            div[c] = (a.je - a.jw + a.jn - a.js + a.jt - a.jb) / vol;
        });
}

// Utavg = a Utavg + b U over every padded cell (grid-stride loop).
template <typename T>
__global__ __launch_bounds__(256) void k_time_average(long long L, double a, double b, double* __restrict__ t0,
                                                      double* __restrict__ t1, double* __restrict__ t2,
                                                      const T* __restrict__ u0, const T* __restrict__ u1,
                                                      const T* __restrict__ u2) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < L; e += stride) {
        t0[e] = a * t0[e] + b * (double)u0[e];
        t1[e] = a * t1[e] + b * (double)u1[e];
        t2[e] = a * t2[e] + b * (double)u2[e];
    }
}

template <typename T>
__global__ void k_gather3(int n, const long long* __restrict__ idx, const T* __restrict__ u0,
                          const T* __restrict__ u1, const T* __restrict__ u2, double* __restrict__ out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const long long c = idx[m];
    out[3 * m] = (double)u0[c];
    out[3 * m + 1] = (double)u1[c];
    out[3 * m + 2] = (double)u2[c];
}

int alloc_zero(dalm_ctx* ctx, double** p, long long n) {
    HIP_TRY(hipMalloc((void**)p, n * sizeof(double)));
    HIP_TRY(hipMemset(*p, 0, n * sizeof(double)));
    return DALM_OK;
}

bool mixed(const dalm_ctx* ctx) { return ctx->precision == DALM_MIXED; }

}  // namespace

int ensure_diag_buffers(dalm_ctx* ctx) {
    if (ctx->qcrit) return DALM_OK;
    // the reference allocates q and Utavg zero-filled (`new Real[len]()`, src/solver.h:307-317)
    for (int m = 0; m < 3; ++m)
        if (int rc = alloc_zero(ctx, &ctx->Utavg[m], ctx->L)) return rc;
    return alloc_zero(ctx, &ctx->qcrit, ctx->L);
}

void free_diag_buffers(dalm_ctx* ctx) {
    for (int m = 0; m < 3; ++m)
        if (ctx->Utavg[m]) (void)hipFree(ctx->Utavg[m]);
    if (ctx->qcrit) (void)hipFree(ctx->qcrit);
    if (ctx->gather_idx) (void)hipFree(ctx->gather_idx);
    if (ctx->gather_out) (void)hipFree(ctx->gather_out);
}

int launch_calc_q(dalm_ctx* ctx) {
    if (int rc = ensure_diag_buffers(ctx)) return rc;
    probe_begin(ctx, DALM_K_CALC_Q);
    if (mixed(ctx))
        hipLaunchKernelGGL(k_calc_q<float>, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->x, ctx->y,
                           ctx->z, ctx->U32[0], ctx->U32[1], ctx->U32[2], ctx->qcrit);
    else
        hipLaunchKernelGGL(k_calc_q<double>, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->x,
                           ctx->y, ctx->z, ctx->U[0], ctx->U[1], ctx->U[2], ctx->qcrit);
    probe_end(ctx, DALM_K_CALC_Q);
    HIP_TRY(hipGetLastError());
    return DALM_OK;
}

int launch_div_store(dalm_ctx* ctx) {
    hipLaunchKernelGGL(k_div_store, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->dx, ctx->dy,
                       ctx->dz, ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->div);
    HIP_TRY(hipGetLastError());
    return DALM_OK;
}

int launch_time_average(dalm_ctx* ctx, double weight) {
    if (int rc = ensure_diag_buffers(ctx)) return rc;
    // the reference's calc_ax_plus_by(1. - weight, Utavg, weight, U, Utavg, len)
    const double a = 1. - weight, b = weight;
    long long blocks = (ctx->L + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    probe_begin(ctx, DALM_K_TIME_AVERAGE);
    if (mixed(ctx))
        hipLaunchKernelGGL(k_time_average<float>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, ctx->L, a, b,
                           ctx->Utavg[0], ctx->Utavg[1], ctx->Utavg[2], ctx->U32[0], ctx->U32[1], ctx->U32[2]);
    else
        hipLaunchKernelGGL(k_time_average<double>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, ctx->L, a,
                           b, ctx->Utavg[0], ctx->Utavg[1], ctx->Utavg[2], ctx->U[0], ctx->U[1], ctx->U[2]);
    probe_end(ctx, DALM_K_TIME_AVERAGE);
    HIP_TRY(hipGetLastError());
    return DALM_OK;
}

int launch_gather_U(dalm_ctx* ctx, const int64_t* cell, int64_t n, double* out) {
    for (int64_t m = 0; m < n; ++m)
        if (cell[m] < 0 || cell[m] >= ctx->L)
            return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_sample_U: cell index outside the padded box");
    if (n > ctx->gather_cap) {
        if (ctx->gather_idx) (void)hipFree(ctx->gather_idx);
        if (ctx->gather_out) (void)hipFree(ctx->gather_out);
        ctx->gather_idx = nullptr;
        ctx->gather_out = nullptr;
        ctx->gather_cap = 0;
        HIP_TRY(hipMalloc((void**)&ctx->gather_idx, n * sizeof(long long)));
        HIP_TRY(hipMalloc((void**)&ctx->gather_out, 3 * n * sizeof(double)));
        ctx->gather_cap = n;
    }
    static_assert(sizeof(long long) == sizeof(int64_t), "index width");
    HIP_TRY(hipMemcpyAsync(ctx->gather_idx, cell, n * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
    const int blocks = (int)((n + 255) / 256);
    if (mixed(ctx))
        hipLaunchKernelGGL(k_gather3<float>, dim3(blocks), dim3(256), 0, ctx->stream, (int)n, ctx->gather_idx,
                           ctx->U32[0], ctx->U32[1], ctx->U32[2], ctx->gather_out);
    else
        hipLaunchKernelGGL(k_gather3<double>, dim3(blocks), dim3(256), 0, ctx->stream, (int)n, ctx->gather_idx,
                           ctx->U[0], ctx->U[1], ctx->U[2], ctx->gather_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, ctx->gather_out, 3 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

// ---- C ABI ------------------------------------------------------------------------

int dalm_calc_q(dalm_ctx* ctx) {
    if (!ctx) return DALM_BAD_ARG;
    if (int rc = launch_calc_q(ctx)) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

int dalm_time_average(dalm_ctx* ctx, double weight) {
    if (!ctx) return DALM_BAD_ARG;
    if (!(weight > 0 && weight <= 1))
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_time_average: weight must lie in (0, 1]");
    if (int rc = launch_time_average(ctx, weight)) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

int dalm_sample_U(dalm_ctx* ctx, const int64_t* cell, int64_t n, double* out) {
    if (!ctx) return DALM_BAD_ARG;
    if (n < 0 || n > (1LL << 24) || (n > 0 && (!cell || !out)))
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_sample_U: bad probe list");
    if (n == 0) return DALM_OK;
    return launch_gather_U(ctx, cell, n, out);
}
