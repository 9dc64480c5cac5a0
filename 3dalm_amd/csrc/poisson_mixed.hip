// poisson_mixed.hip -- Poisson solve of the precision form.
//
// Preconditioned BiCGStab with the recurrences of src/eq.h:333-473, run on FP32
// vectors with FP64 dot products, FP64 norms and an FP64 iterate:
//
//   FP32 storage  r, r0, p, q = A z, s, t = A z', the preconditioner iterates z, z'
//                 (Jacobi sweeps or SOR double sweeps in FP32 arithmetic)
//   FP64          the pressure x (updated every iteration: x += alpha z + omega z',
//                 exactly the FP32 z that produced q and t), every dot product and
//                 norm (FP32 products are exact in FP64), the scalars alpha, omega,
//                 beta, rho, and the products A z (rounded to FP32 once)
//
// Reliable updates (residual replacement, Sleijpen & van der Vorst 1996; van der
// Vorst & Ye 2000): the FP32 recursive residual drifts from b - A x by the FP32
// rounding it accumulates, so whenever its norm has fallen below delta = 0.1 of
// the largest norm since the last replacement -- and whenever it meets the
// stopping test -- the true residual b - A x is computed in FP64 from the FP64
// pressure and replaces it (rounded to FP32); the search direction p is kept, so
// the Krylov process continues instead of restarting.  The solve ends only when
// the TRUE residual meets the reference's test |b - A x| / sqrt(N) <= tol
// (src/eq.h:399-401), so the pressure meets the same tolerance as the FP64 form.
// The decisions are taken on the device (SolveState.ru); the replacement kernel
// and its finalize are enqueued every iteration and return at once when no
// replacement is due.
//
// Passes: FP32 rounding of the directions perturbs BiCGStab's short recurrences,
// and on the stretched C2-size operators the FP32 iteration gains three to five
// decades quickly and then stagnates near a floor (about 1e-8 at 256x128x128,
// 2e-8 at 512x256x256) or wanders off.  A watchdog on the residual (no 30% gain
// in 48 iterations; near the tolerance none in 16; a rise above 100x the lowest)
// ends the pass, and the FP32 iteration restarts from the true residual with a
// new shadow residual and direction: the floor is relative to where a pass
// starts, and a restarted pass gets a longer window (128 iterations) because it
// starts on a smooth residual.  The best pressure seen is kept, and a pass that
// ended well above it restarts from it.  A restarted pass that stalls near the
// tolerance, gains less than a factor 2 or is the last allowed is handed over to
// FP64 without restarting: the true residual replaces r, r0, p and q are widened
// to FP64 (exactly), rho and beta are recomputed as after a replacement, and the
// FP64 BiCGStab continues from that state.  A breakdown or a non-finite residual
// restarts the FP64 solver from the better of the current and the best pressure.
// Either way every step meets the tolerance on the true residual.
//
// Algorithmic bytes per iteration, Jacobi m = 3, per cell: p update 20, Jacobi
// sweeps 4 x 12, q = A z 12, s update 16, t = A z' 12, x/r update 40 = 148,
// against 280 for the FP64 solver; a replacement costs 24.
#include <float.h>

#include <cstdio>
#include <cstdlib>
#include <utility>

#include "ctx.h"
#include "stencil.h"

namespace {

// Watchdog of the FP32 iteration: a rise above kWatchRise x the lowest residual,
// or no 30% gain within kWatchWindow iterations.  Left to run longer, the FP32
// iteration on the stretched C2-size operators mostly wanders or diverges rather
// than recovering (measured: windows of 192 and 400 iterations cost more FP64
// iterations than they saved; with 96 the FP32 phase idles near the tolerance).
constexpr double kWatchRise = 1e2;
constexpr long long kWatchWindow = 48;
constexpr long long kChunk = 16;  // iterations enqueued between watchdog checks
// A stalled FP32 pass is restarted from the true residual (a new shadow residual
// and direction) while each pass gains at least kPassGain.
constexpr double kPassGain = 2.0;
constexpr double kPassFloor = 30.0;
constexpr double kNearTol = 3.0;  // first pass: near the tolerance a stall is called after one chunk
// A restarted pass starts from a smooth residual and gains slowly at first (in
// FP64 as well): it gets a longer window.
constexpr long long kRestartWindow = 128;
constexpr double kRestoreRise = 4.0;  // a pass ending this far above its best restarts from the best
constexpr int kMaxPasses = 6;


__global__ __launch_bounds__(BLOCK) void kf_jacobi(Dims d, PCoefF A, const float* __restrict__ zin,
                                                   float* __restrict__ zout, const float* __restrict__ v,
                                                   const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const OpF op(A, cl.j, cl.k);
    walk<float, float>(d, cl, zin, [&](int c) { return v[c]; },
                       [&](int i, int c, float x0, float xp, float xm, const Nb4<float>& q, float vc) {
                           const float ac = op.diag(A, i);
                           zout[c] = x0 + (vc - op.apply(A, i, ac, x0, xp, xm, q)) * __builtin_amdgcn_rcpf(ac);
                       });
}

// The same sweep with four consecutive k per lane (one wavefront covers 256 k of a
// row), for rows of a multiple of 256 interior cells: each lane issues 16-byte loads
// and stores, which stream from HBM about 20 % faster than 4-byte ones at
// 512x256x256; k-neighbours inside the four come from registers.  Identical
// arithmetic per cell.  Launched over tj x (nz / 256) x nch units.  Stores are
// nontemporal, as in the FP64 four-cell kernels (3 % on the precision form's
// iteration at 512x256x256).
typedef float f4k __attribute__((ext_vector_type(kVecK), aligned(4)));

__global__ __launch_bounds__(BLOCK) void kf_jacobi4(Dims d, PCoefF A, const float* __restrict__ zin,
                                                    float* __restrict__ zout, const float* __restrict__ v,
                                                    const SolveState* st) {
    if (st->done) return;
    const int tkv = vec4_tiles(d);
    const int u = xcd_unit((int)blockIdx.x, (int)gridDim.x);
    const int jt = u % d.tj, rest = u / d.tj;
    const int kt = rest % tkv, ic = rest / tkv;
    const int lane = threadIdx.x % TK, row = threadIdx.x / TK;
    const int k0 = d.gc + (kt * TK + lane) * kVecK;
    const int j = d.gc + jt * TJ + row;
    if (j >= d.gc + d.ny || k0 >= d.gc + d.nz) return;
    const int own0 = d.gc + ic * d.ni;
    int i1 = own0 + d.ni;
    if (i1 > d.gc + d.nx) i1 = d.gc + d.nx;
    const int i0 = i1 - d.ni;
    const float an = A.an[j], as = A.as[j], anu = A.anu[j], asu = A.asu[j];
    float at[kVecK], ab[kVecK], atu[kVecK], abu[kVecK];
#pragma unroll
    for (int q = 0; q < kVecK; ++q) {
        at[q] = A.at[k0 + q];
        ab[q] = A.ab[k0 + q];
        atu[q] = A.atu[k0 + q];
        abu[q] = A.abu[k0 + q];
    }
    int c = d.at(i0, j, k0);
    f4k xm = *(const f4k*)(zin + c - d.si), x0 = *(const f4k*)(zin + c), xp = *(const f4k*)(zin + c + d.si);
    f4k xn = *(const f4k*)(zin + c + d.sj), xs = *(const f4k*)(zin + c - d.sj), vc = *(const f4k*)(v + c);
    float kb = zin[c - 1], kf = zin[c + kVecK];
    // plane i: OpF's diagonal and operator, cell by cell, in its order
    auto plane = [&](int i, int cc) {
        const float acx = A.acx[i], ae = A.ae[i], aw = A.aw[i];
        f4k out;
#pragma unroll
        for (int q = 0; q < kVecK; ++q) {
            const float ac = (((acx - anu) - asu) - atu[q]) - abu[q];
            const float t = q + 1 < kVecK ? x0[q + 1] : kf;
            const float b = q > 0 ? x0[q - 1] : kb;
            const float ax = ac * x0[q] + ae * xp[q] + aw * xm[q] + an * xn[q] + as * xs[q] + at[q] * t + ab[q] * b;
            out[q] = x0[q] + (vc[q] - ax) * __builtin_amdgcn_rcpf(ac);
        }
        __builtin_nontemporal_store(out, (f4k*)(zout + cc));
    };
    // every value of plane i+1 is requested before plane i is computed (the walk's
    // look-ahead; the last plane re-reads its own)
    auto step = [&](int i, int nxt) {
        const int cn = c + nxt;
        const f4k xpp = *(const f4k*)(zin + cn + d.si);
        const f4k xn1 = *(const f4k*)(zin + cn + d.sj), xs1 = *(const f4k*)(zin + cn - d.sj);
        const f4k vc1 = *(const f4k*)(v + cn);
        const float kb1 = zin[cn - 1], kf1 = zin[cn + kVecK];
        if (i >= own0) plane(i, c);
        xm = x0;
        x0 = xp;
        xp = xpp;
        xn = xn1;
        xs = xs1;
        vc = vc1;
        kb = kb1;
        kf = kf1;
        c += d.si;
    };
    if (d.ni == kMaxPlanes) {
#pragma unroll
        for (int w = 0; w < kMaxPlanes; ++w) step(i0 + w, w + 1 < kMaxPlanes ? d.si : 0);
    } else {
        for (int i = i0; i < i1; ++i) step(i, i + 1 < i1 ? d.si : 0);
    }
}

// kg_spmv with four cells per lane: y = A z in FP64 arithmetic (OpD's order), stored
// rounded, and the partial dot products of the stored values, cells in k order
template <int WHICH>
__global__ __launch_bounds__(BLOCK) void kg_spmv4(Dims d, PCoef A, const float* __restrict__ z,
                                                  float* __restrict__ y, const float* __restrict__ a1,
                                                  double* partials, const SolveState* st) {
    if (st->done) return;
    const Col4 cl = col_unit4(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const int j = cl.j, k0 = cl.k0;
        const double an = A.an[j], as = A.as[j], anu = A.anu[j], asu = A.asu[j];
        double at[kVecK], ab[kVecK], atu[kVecK], abu[kVecK];
#pragma unroll
        for (int q = 0; q < kVecK; ++q) {
            at[q] = A.at[k0 + q];
            ab[q] = A.ab[k0 + q];
            atu[q] = A.atu[k0 + q];
            abu[q] = A.abu[k0 + q];
        }
        int c = d.at(cl.i0, j, k0);
        f4k xm = *(const f4k*)(z + c - d.si), x0 = *(const f4k*)(z + c), xp = *(const f4k*)(z + c + d.si);
        f4k xn = *(const f4k*)(z + c + d.sj), xs = *(const f4k*)(z + c - d.sj), av = *(const f4k*)(a1 + c);
        float kb = z[c - 1], kf = z[c + kVecK];
        auto plane = [&](int i, int cc) {
            const double acx = A.acx[i], ae = A.ae[i], aw = A.aw[i];
            f4k out;
#pragma unroll
            for (int q = 0; q < kVecK; ++q) {
                const double ac = ((((acx - anu) - asu) - atu[q]) - abu[q]) / A.max_diag;
                const double t = q + 1 < kVecK ? (double)x0[q + 1] : (double)kf;
                const double b = q > 0 ? (double)x0[q - 1] : (double)kb;
                const double yv = ac * (double)x0[q] + ae * (double)xp[q] + aw * (double)xm[q] + an * (double)xn[q] +
                                  as * (double)xs[q] + at[q] * t + ab[q] * b;
                const float yf = (float)yv;
                out[q] = yf;
                const double ys = (double)yf;
                if (WHICH == 0) {
                    v[0] += (double)av[q] * ys;
                } else {
                    v[0] += ys * (double)av[q];
                    v[1] += ys * ys;
                }
            }
            __builtin_nontemporal_store(out, (f4k*)(y + cc));
        };
        auto step = [&](int i, int nxt) {
            const int cn = c + nxt;
            const f4k xpp = *(const f4k*)(z + cn + d.si);
            const f4k xn1 = *(const f4k*)(z + cn + d.sj), xs1 = *(const f4k*)(z + cn - d.sj);
            const f4k av1 = *(const f4k*)(a1 + cn);
            const float kb1 = z[cn - 1], kf1 = z[cn + kVecK];
            if (i >= cl.own0) plane(i, c);
            xm = x0;
            x0 = xp;
            xp = xpp;
            xn = xn1;
            xs = xs1;
            av = av1;
            kb = kb1;
            kf = kf1;
            c += d.si;
        };
        if (d.ni == kMaxPlanes) {
#pragma unroll
            for (int w = 0; w < kMaxPlanes; ++w) step(cl.i0 + w, w + 1 < kMaxPlanes ? d.si : 0);
        } else {
            for (int i = cl.i0; i < cl.i1; ++i) step(i, i + 1 < cl.i1 ? d.si : 0);
        }
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// kg_p and kg_s with four cells per lane (the same arithmetic per cell)
__global__ __launch_bounds__(BLOCK) void kg_p4(Dims d, PCoefF Af, const float* __restrict__ r,
                                               float* __restrict__ p, const float* __restrict__ q,
                                               float* __restrict__ z, int pc_first, const SolveState* st) {
    if (st->done) return;
    const bool first = st->it == 0;
    const double beta = st->beta, om = st->omega;
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    const float anu = Af.anu[cl.j], asu = Af.asu[cl.j];
    float atu[kVecK], abu[kVecK];
#pragma unroll
    for (int m = 0; m < kVecK; ++m) {
        atu[m] = Af.atu[cl.k0 + m];
        abu[m] = Af.abu[cl.k0 + m];
    }
    struct RPQ {
        f4k r, p, q;
    };
    walk_points4<RPQ>(d, cl, [&](int c) { return RPQ{*(const f4k*)(r + c), *(const f4k*)(p + c), *(const f4k*)(q + c)}; },
                      [&](int i, int c, const RPQ& v) {
                          const float acx = Af.acx[i];
                          f4k pv, zv;
#pragma unroll
                          for (int m = 0; m < kVecK; ++m) {
                              const float pf = first ? v.r[m]
                                                     : (float)((double)v.r[m] + beta * ((double)v.p[m] - om * (double)v.q[m]));
                              pv[m] = pf;
                              const float ac = (((acx - anu) - asu) - atu[m]) - abu[m];
                              zv[m] = pc_first ? pf * __builtin_amdgcn_rcpf(ac) : 0.f;
                          }
                          __builtin_nontemporal_store(pv, (f4k*)(p + c));
                          __builtin_nontemporal_store(zv, (f4k*)(z + c));
                      });
}

__global__ __launch_bounds__(BLOCK) void kg_s4(Dims d, PCoefF Af, const float* __restrict__ r,
                                               const float* __restrict__ q, float* __restrict__ s,
                                               float* __restrict__ z, int pc_first, const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha;
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    const float anu = Af.anu[cl.j], asu = Af.asu[cl.j];
    float atu[kVecK], abu[kVecK];
#pragma unroll
    for (int m = 0; m < kVecK; ++m) {
        atu[m] = Af.atu[cl.k0 + m];
        abu[m] = Af.abu[cl.k0 + m];
    }
    struct RQ {
        f4k r, q;
    };
    walk_points4<RQ>(d, cl, [&](int c) { return RQ{*(const f4k*)(r + c), *(const f4k*)(q + c)}; },
                     [&](int i, int c, const RQ& v) {
                         const float acx = Af.acx[i];
                         f4k sv, zv;
#pragma unroll
                         for (int m = 0; m < kVecK; ++m) {
                             const float sf = (float)((double)v.r[m] - alpha * (double)v.q[m]);
                             sv[m] = sf;
                             const float ac = (((acx - anu) - asu) - atu[m]) - abu[m];
                             zv[m] = pc_first ? sf * __builtin_amdgcn_rcpf(ac) : 0.f;
                         }
                         __builtin_nontemporal_store(sv, (f4k*)(s + c));
                         __builtin_nontemporal_store(zv, (f4k*)(z + c));
                     });
}

// kg_xr with four cells per lane: x (FP64) as two 16-byte halves, the FP32 vectors
// as one 16-byte load each; partial sums cell by cell in k order
typedef double d2k __attribute__((ext_vector_type(2), aligned(8)));
__global__ __launch_bounds__(BLOCK) void kg_xr4(Dims d, double* __restrict__ x, const float* __restrict__ zp,
                                                const float* __restrict__ zs, const float* __restrict__ s,
                                                const float* __restrict__ t_, float* __restrict__ r,
                                                const float* __restrict__ r0, double* partials,
                                                const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha, om = st->omega;
    const Col4 cl = col_unit4(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        struct XR {
            d2k xa, xb;
            f4k zp, zs, s, t, r0;
        };
        walk_points4<XR>(
            d, cl,
            [&](int c) {
                return XR{*(const d2k*)(x + c), *(const d2k*)(x + c + 2), *(const f4k*)(zp + c),
                          *(const f4k*)(zs + c), *(const f4k*)(s + c), *(const f4k*)(t_ + c), *(const f4k*)(r0 + c)};
            },
            [&](int, int c, const XR& a) {
                d2k xa, xb;
                f4k rv4;
#pragma unroll
                for (int m = 0; m < kVecK; ++m) {
                    const double xo = m < 2 ? a.xa[m] : a.xb[m - 2];
                    const double xn = xo + alpha * (double)a.zp[m] + om * (double)a.zs[m];
                    if (m < 2) xa[m] = xn;
                    else xb[m - 2] = xn;
                    const float rf = (float)((double)a.s[m] - om * (double)a.t[m]);
                    rv4[m] = rf;
                    const double rv = (double)rf;
                    v[0] += rv * rv;
                    v[1] += (double)a.r0[m] * rv;
                }
                __builtin_nontemporal_store(xa, (d2k*)(x + c));
                __builtin_nontemporal_store(xb, (d2k*)(x + c + 2));
                __builtin_nontemporal_store(rv4, (f4k*)(r + c));
            });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

__global__ __launch_bounds__(BLOCK) void kf_sor(Dims d, PCoefF A, float* __restrict__ x,
                                                const float* __restrict__ b, float omega, int color,
                                                const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const int j = cl.j, k = cl.k;
    const OpF op(A, j, k);
    // neighbours of a cell of this colour have the other colour: unchanged here
    walk<float, float>(d, cl, x, [&](int c) { return b[c]; },
                       [&](int i, int c, float x0, float xp, float xm, const Nb4<float>& q, float bc) {
                           if (((i + j + k) & 1) == color) {
                               const float ac = op.diag(A, i);
                               x[c] = x0 + omega * (bc - op.apply(A, i, ac, x0, xp, xm, q)) * __builtin_amdgcn_rcpf(ac);
                           }
                       });
}

__global__ void kg_state_reset(SolveState* st, double tol, long long max_it, double root_n) {
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
        st->pending = 0;
        st->xdef = 0;
        st->n_ru = 0;
        st->maxr = 0;
    }
}

// z_p and z_s of kXGroupF consecutive iterations of the fused iteration, oldest first
struct ZRing {
    const float* zp[kXGroupF];
    const float* zs[kXGroupF];
};

// kg_flush_pending4: the same with the values the host read from the state, ring
// positions [first, kXGroupF) with their alpha and omega, at the end of a pass
// (out == x) or into the best pressure's copy of x (out != x)
struct XFlush {
    double alpha[kXGroupF], omega[kXGroupF];
    int first;
};

// The fused iteration's outstanding pressure updates after iteration k (z: iterations
// k - kXGroupF + 1 .. k), the xdef of them in turn (alpha_d, omega_d, oldest first):
// when kXGroupF have gathered or a residual replacement is due, otherwise it returns
// at once (FIN_MX_RU then clears the count).  Rows of 256 cells, four cells per lane
// with 16-byte accesses: x and the pairs of z stream through once.
__global__ __launch_bounds__(BLOCK) void kg_apply_pending4(Dims d, double* __restrict__ x, ZRing z,
                                                           const SolveState* st) {
    if (st->done || !(st->ru || st->xdef >= kXGroupF) || st->xdef <= 0) return;
    const int first = kXGroupF - st->xdef;
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    int c = d.at(cl.own0, cl.j, cl.k0);
    for (int i = cl.own0; i < cl.i1; ++i, c += d.si) {
        d2k xa = *(const d2k*)(x + c), xb = *(const d2k*)(x + c + 2);
#pragma unroll
        for (int q = 0; q < kXGroupF; ++q) {
            if (q < first) continue;
            const f4k a = *(const f4k*)(z.zp[q] + c), b = *(const f4k*)(z.zs[q] + c);
            const double aq = st->alpha_d[q - first], oq = st->omega_d[q - first];
            xa[0] = x_update(xa[0], a[0], b[0], aq, oq);
            xa[1] = x_update(xa[1], a[1], b[1], aq, oq);
            xb[0] = x_update(xb[0], a[2], b[2], aq, oq);
            xb[1] = x_update(xb[1], a[3], b[3], aq, oq);
        }
        __builtin_nontemporal_store(xa, (d2k*)(x + c));
        __builtin_nontemporal_store(xb, (d2k*)(x + c + 2));
    }
}

__global__ __launch_bounds__(BLOCK) void kg_flush_pending4(Dims d, const double* x, ZRing z, double* out, XFlush f) {
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    int c = d.at(cl.own0, cl.j, cl.k0);
    for (int i = cl.own0; i < cl.i1; ++i, c += d.si) {
        d2k xa = *(const d2k*)(x + c), xb = *(const d2k*)(x + c + 2);
#pragma unroll
        for (int q = 0; q < kXGroupF; ++q) {
            if (q < f.first) continue;
            const f4k a = *(const f4k*)(z.zp[q] + c), b = *(const f4k*)(z.zs[q] + c);
            xa[0] = x_update(xa[0], a[0], b[0], f.alpha[q], f.omega[q]);
            xa[1] = x_update(xa[1], a[1], b[1], f.alpha[q], f.omega[q]);
            xb[0] = x_update(xb[0], a[2], b[2], f.alpha[q], f.omega[q]);
            xb[1] = x_update(xb[1], a[3], b[3], f.alpha[q], f.omega[q]);
        }
        *(d2k*)(out + c) = xa;
        *(d2k*)(out + c + 2) = xb;
    }
}

// True residual in FP64: r64 = b - A x, stored rounded as r32; INIT also sets the
// shadow r0 = r32.  Partials <r0, r32>, |r64|^2.  Without INIT this is the
// reliable update, which returns at once unless the device state asks for one.
template <bool INIT>
__global__ __launch_bounds__(BLOCK) void kg_resid(Dims d, PCoef A, const double* __restrict__ x,
                                                  const double* __restrict__ b, float* __restrict__ r,
                                                  float* __restrict__ r0, double* partials, const SolveState* st) {
    if (!INIT && (st->done || !st->ru)) return;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const OpD op(A, cl.j, cl.k);
        struct BR {
            double b;
            float r0;
        };
        walk<double, BR>(d, cl, x, [&](int c) { return BR{b[c], INIT ? 0.f : r0[c]}; },
                         [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, const BR& a) {
                             const double rv = a.b - op.apply(A, i, op.diag(A, i), x0, xp, xm, q);
                             const float rf = (float)rv;
                             r[c] = rf;
                             float r0v = a.r0;
                             if (INIT) {
                                 r0[c] = rf;
                                 r0v = rf;
                             }
                             v[0] += (double)r0v * (double)rf;
                             v[1] += rv * rv;
                         });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// p = r (first iteration) or r + beta (p - omega q), and the first preconditioner
// sweep from zero (Jacobi: z = p / a_c; SOR: z = 0)
__global__ __launch_bounds__(BLOCK) void kg_p(Dims d, PCoefF Af, const float* __restrict__ r,
                                              float* __restrict__ p, const float* __restrict__ q,
                                              float* __restrict__ z, int pc_first, const SolveState* st) {
    if (st->done) return;
    const bool first = st->it == 0;
    const double beta = st->beta, om = st->omega;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const OpF op(Af, cl.j, cl.k);
    struct RPQ {
        float r, p, q;
    };
    walk_points<RPQ>(d, cl, [&](int c) { return RPQ{r[c], p[c], q[c]}; },
                     [&](int i, int c, const RPQ& v) {
                         const float pf =
                             first ? v.r : (float)((double)v.r + beta * ((double)v.p - om * (double)v.q));
                         p[c] = pf;
                         z[c] = pc_first ? pf * __builtin_amdgcn_rcpf(op.diag(Af, i)) : 0.f;
                     });
}

// s = r - alpha q and the first preconditioner sweep
__global__ __launch_bounds__(BLOCK) void kg_s(Dims d, PCoefF Af, const float* __restrict__ r,
                                              const float* __restrict__ q, float* __restrict__ s,
                                              float* __restrict__ z, int pc_first, const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const OpF op(Af, cl.j, cl.k);
    struct RQ {
        float r, q;
    };
    walk_points<RQ>(d, cl, [&](int c) { return RQ{r[c], q[c]}; },
                    [&](int i, int c, const RQ& v) {
                        const float sf = (float)((double)v.r - alpha * (double)v.q);
                        s[c] = sf;
                        z[c] = pc_first ? sf * __builtin_amdgcn_rcpf(op.diag(Af, i)) : 0.f;
                    });
}

// y = A z in FP64 arithmetic, stored rounded to FP32, with <a1, y> (WHICH 0:
// a1 = r0) or <y, a1>, <y, y> (WHICH 1: a1 = s) of the stored values
template <int WHICH>
__global__ __launch_bounds__(BLOCK) void kg_spmv(Dims d, PCoef A, const float* __restrict__ z,
                                                 float* __restrict__ y, const float* __restrict__ a1,
                                                 double* partials, const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const OpD op(A, cl.j, cl.k);
        walk<float, float>(d, cl, z, [&](int c) { return a1[c]; },
                           [&](int i, int c, float x0, float xp, float xm, const Nb4<float>& q, float ac1) {
                               const double yv = op.apply(A, i, op.diag(A, i), (double)x0, (double)xp,
                                                          (double)xm, q);
                               const float yf = (float)yv;
                               y[c] = yf;
                               const double ys = (double)yf;
                               if (WHICH == 0) {
                                   v[0] += (double)ac1 * ys;
                               } else {
                                   v[0] += ys * (double)ac1;
                                   v[1] += ys * ys;
                               }
                           });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// x += alpha z + omega z' (FP64); r = s - omega t; partials |r|^2, <r0, r>
__global__ __launch_bounds__(BLOCK) void kg_xr(Dims d, double* __restrict__ x, const float* __restrict__ zp,
                                               const float* __restrict__ zs, const float* __restrict__ s,
                                               const float* __restrict__ t_, float* __restrict__ r,
                                               const float* __restrict__ r0, double* partials,
                                               const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha, om = st->omega;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        struct XR {
            double x;
            float zp, zs, s, t, r0;
        };
        walk_points<XR>(d, cl, [&](int c) { return XR{x[c], zp[c], zs[c], s[c], t_[c], r0[c]}; },
                        [&](int, int c, const XR& a) {
                            x[c] = a.x + alpha * (double)a.zp + om * (double)a.zs;
                            const float rf = (float)((double)a.s - om * (double)a.t);
                            r[c] = rf;
                            const double rv = (double)rf;
                            v[0] += rv * rv;
                            v[1] += (double)a.r0 * rv;
                        });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// Hand-over to FP64: r = b - A x (FP64), and r0, p, q widened to FP64 (exactly);
// partials <r0, r>, |r|^2
__global__ __launch_bounds__(BLOCK) void kg_upgrade(Dims d, PCoef A, const double* __restrict__ x,
                                                    const double* __restrict__ b, const float* __restrict__ r0f,
                                                    const float* __restrict__ pf, const float* __restrict__ qf,
                                                    double* __restrict__ r, double* __restrict__ r0,
                                                    double* __restrict__ p, double* __restrict__ q, double* partials) {
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const OpD op(A, cl.j, cl.k);
        struct U4 {
            double b;
            float r0, p, q;
        };
        walk<double, U4>(d, cl, x, [&](int c) { return U4{b[c], r0f[c], pf[c], qf[c]}; },
                         [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& nb, const U4& a) {
                             const double rv = a.b - op.apply(A, i, op.diag(A, i), x0, xp, xm, nb);
                             r[c] = rv;
                             r0[c] = (double)a.r0;
                             p[c] = (double)a.p;
                             q[c] = (double)a.q;
                             v[0] += (double)a.r0 * rv;
                             v[1] += rv * rv;
                         });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// ---- precision form of run_sor (src/eq.h:291-306) ------------------------------
// Defect correction: the FP64 residual r = b - A x of the FP64 pressure is rounded
// to FP32, kSorInner red-black SOR iterations solve A e = r from e = 0 in FP32, and
// x += e in FP64.  SOR is a stationary iteration, so this is the FP64 iteration
// with FP32 rounding confined to the correction e, which shrinks with the
// residual: the result meets the reference's test |b - A x| / sqrt(N) <= tol on
// the true FP64 residual.  The test runs once per refinement, so the iteration
// count is a multiple of kSorInner.  Bytes per cell and SOR iteration: 2 colour
// sweeps of 10 + (residual 24 + update 20) / kSorInner = 25.5 against 48 in FP64.
constexpr int kSorInner = 8;

__global__ void ks_state_reset(SolveState* st, double tol, long long max_it, double root_n, double inner) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->err = 0;
        st->tol = tol;
        st->root_n = root_n;
        st->it = 0;
        st->max_it = max_it;
        st->done = 0;
        st->status = DALM_OK;
        st->aux1 = inner;
    }
}

// r32 = b - A x (FP64 arithmetic), e = 0; partials |b - A x|^2
__global__ __launch_bounds__(BLOCK) void ks_resid(Dims d, PCoef A, const double* __restrict__ x,
                                                  const double* __restrict__ b, float* __restrict__ r,
                                                  float* __restrict__ e, double* partials, const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    double v[1] = {0.0};
    if (cl.ok) {
        const OpD op(A, cl.j, cl.k);
        walk<double, double>(d, cl, x, [&](int c) { return b[c]; },
                             [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, double bc) {
                                 const double rv = bc - op.apply(A, i, op.diag(A, i), x0, xp, xm, q);
                                 r[c] = (float)rv;
                                 e[c] = 0.f;
                                 v[0] += rv * rv;
                             });
    }
    block_sum<1>(v);
    double pv[1] = {v[0]};
    store_partials<1>(pv, partials, cl.slot);
}

// x += e (FP64)
__global__ __launch_bounds__(BLOCK) void ks_add(Dims d, double* __restrict__ x, const float* __restrict__ e,
                                                const SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    struct XE {
        double x;
        float e;
    };
    walk_points<XE>(d, cl, [&](int c) { return XE{x[c], e[c]}; },
                    [&](int, int c, const XE& a) { x[c] = a.x + (double)a.e; });
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
    if (c > kChunk) c = kChunk;
    if (done_it + c > max_it) c = max_it - done_it;
    return c < 1 ? 1 : c;
}

}  // namespace

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

int run_sor_mixed(dalm_ctx* ctx, double omega, long long max_it, double tol, long long* it, double* err) {
    const Dims& d = ctx->d;
    const PCoef A = ctx->coef();
    const PCoefF Af = ctx->pcf;
    float* r = ctx->fr;
    float* e = ctx->fz;
    ctx->outer_it = 0;
    ctx->fp64_finish_it = 0;
    hipLaunchKernelGGL(ks_state_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->solve, tol, max_it,
                       sqrt((double)ctx->N), (double)kSorInner);
    LAUNCH_CHECK();
    // refinements enqueued between host checks: the previous solve's count predicts this one
    const long long limit = max_it / kSorInner + 2;
    long long launched = 0;
    for (;;) {
        long long chunk = ctx->prev_it[0] / kSorInner + 1 - launched;
        if (chunk < 2) chunk = 2;
        if (chunk > 32) chunk = 32;
        for (long long c = 0; c < chunk; ++c) {
            hipLaunchKernelGGL(ks_resid, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, ctx->p, ctx->b, r, e,
                               ctx->partials, ctx->solve);
            if (int rc = launch_finalize(ctx, FIN_MX_SOR, ctx->sgrid, ctx->solve, nullptr)) return rc;
            for (int s = 0; s < kSorInner; ++s)
                for (int color = 0; color < 2; ++color)
                    hipLaunchKernelGGL(kf_sor, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, Af, e, r,
                                       (float)omega, color, ctx->solve);
            hipLaunchKernelGGL(ks_add, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, ctx->p, e, ctx->solve);
        }
        LAUNCH_CHECK();
        launched += chunk;
        if (int rc = read_states(ctx)) return rc;
        // the device test ends the solve once it >= max_it; `limit` only guards the loop
        if (ctx->h_solve->done || launched > limit) break;
    }
    *it = ctx->h_solve->it;
    *err = ctx->h_solve->err;
    ctx->prev_it[0] = *it;
    if (!ctx->h_solve->done) return dalm_fail_msg(ctx, DALM_HIP_ERROR, "precision-form SOR did not finish");
    return ctx->h_solve->status;
}

int run_poisson_mixed(dalm_ctx* ctx, int pc_kind, double pc_omega, int pc_it, long long max_it, double tol,
                      long long* it, double* err) {
    const Dims& d = ctx->d;
    const PCoefF Af = ctx->pcf;
    const PCoef A = ctx->coef();
    SolveState* st = ctx->solve;
    const int jac = pc_kind == DALM_PC_JACOBI;
    const int pc_first = (jac && pc_it >= 1) ? 1 : 0;
    static const bool trace = getenv("DALM_TRACE") != nullptr;
    // testing aid (dalm_set_fp32_pass_cap): every FP32 pass ends after this many
    // iterations as if it had stalled, so that small cases exercise the restarted
    // passes and the FP64 hand-over; 0 in production
    const long long test_cap = ctx->fp32_pass_cap;
    const double root_n = sqrt((double)ctx->N);
    // rows of a multiple of 256 interior cells: the four-per-lane Jacobi sweep
    const bool vec4 = vec4_applies(d);
    const int grid4 = d.tj * vec4_tiles(d) * d.nch;
    // rows of exactly 256 cells with three Jacobi sweeps: each half-iteration is one
    // fused kernel (fused_mixed.hip); p and q alternate between a buffer pair (the
    // Jacobi partners, which the fused path does not use)
    const bool fused = fused_half_applies(ctx, pc_kind, pc_it);
    int fni = 0;
    const int fgrid = fused ? fused_half_grid(ctx, &fni) : 0;
    if (fused) {
        if (const int rz = ensure_fused_z(ctx)) return rz;
    }
    // the fused iteration k writes z_p / z_s into buffer pair k mod kXGroupF
    float* ZP[kXGroupF] = {ctx->fz};
    float* ZS[kXGroupF] = {ctx->fz2};
    for (int q = 1; q < kXGroupF; ++q) {
        ZP[q] = ctx->fzx[q - 1];
        ZS[q] = ctx->fz2x[q - 1];
    }
    // z of iterations base - kXGroupF + 1 + q, q = 0 .. kXGroupF - 1
    auto ring = [&](long long base) {
        ZRing z;
        for (int q = 0; q < kXGroupF; ++q) {
            z.zp[q] = ZP[(base + 1 + q) % kXGroupF];
            z.zs[q] = ZS[(base + 1 + q) % kXGroupF];
        }
        return z;
    };
    // the outstanding x updates after iteration `last` of a pass (the last xdef
    // iterations') applied to x into out
    auto flush = [&](const SolveState& h, double* out) {
        const long long last = h.it - 1;
        XFlush f{};
        f.first = kXGroupF - h.xdef;
        for (int q = f.first; q < kXGroupF; ++q) {
            f.alpha[q] = h.alpha_d[q - f.first];
            f.omega[q] = h.omega_d[q - f.first];
        }
        hipLaunchKernelGGL(kg_flush_pending4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, (const double*)ctx->p,
                           ring(last), out, f);
    };
    // FP32 preconditioner application on v32 into z0 (spare: the other Jacobi buffer)
    auto precondition = [&](float* z0, float* spare, const float* v32) -> float* {
        float* cur = z0;
        if (jac) {
            float* other = spare;
            // the sweeps of one application run back to back: one event pair times them
            if (pc_it > 1) probe_begin(ctx, DALM_K_JACOBI32);
            for (int m = 1; m < pc_it; ++m) {
                if (vec4)
                    hipLaunchKernelGGL(kf_jacobi4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, Af, cur, other, v32,
                                       (const SolveState*)st);
                else
                    hipLaunchKernelGGL(kf_jacobi, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, Af, cur, other,
                                       v32, (const SolveState*)st);
                float* tsw = cur;
                cur = other;
                other = tsw;
            }
            if (pc_it > 1) probe_end(ctx, DALM_K_JACOBI32, pc_it - 1);
        } else {
            for (int m = 0; m < pc_it; ++m)
                for (int color = 0; color < 2; ++color)
                    hipLaunchKernelGGL(kf_sor, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, Af, cur, v32,
                                       (float)pc_omega, color, (const SolveState*)st);
        }
        return cur;
    };
    ctx->outer_it = 0;
    ctx->fp64_finish_it = 0;
    // pbest keeps the best pressure seen (by the running residual estimate): an FP32
    // pass that wanders off after a good stretch is restarted from it
    HIP_TRY(hipMemcpyAsync(ctx->pbest, ctx->p, sizeof(double) * ctx->L, hipMemcpyDeviceToDevice, ctx->stream));
    long long total = 0, n_ru = 0;
    double prev_pass = -1, e_out = 0, best_saved = -1;
    int status = DALM_NONCONVERGED;
    bool fp64 = false;
    int rc;
    // (re)start from the true residual of the pressure: r = b - A x, r0 = r
    auto restart = [&]() -> int {
        hipLaunchKernelGGL(kg_state_reset, dim3(1), dim3(64), 0, ctx->stream, st, tol, max_it - total, root_n);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_resid<true>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A,
                           ctx->p, ctx->b, ctx->fr, ctx->fr0, ctx->partials, (const SolveState*)st);
        HIP_TRY(hipGetLastError());
        if (int r = launch_finalize(ctx, FIN_MX_INIT, ctx->sgrid, st, nullptr)) return r;
        return read_states(ctx);
    };
    for (int pass = 0;; ++pass) {
        if ((rc = restart())) return rc;
        double e0 = ctx->h_solve[0].err;
        if (pass > 0 && e0 > kRestoreRise * best_saved) {
            // the last pass ended well above its best: restart from the best pressure
            // if it is really better (the saved value was a running estimate)
            std::swap(ctx->p, ctx->pbest);
            if ((rc = restart())) return rc;
            if (!(ctx->h_solve[0].err < e0)) {
                std::swap(ctx->p, ctx->pbest);
                if ((rc = restart())) return rc;
            } else if (trace) {
                fprintf(stderr, "[mixed] restored the best pressure (%.3e instead of %.3e)\n", ctx->h_solve[0].err, e0);
            }
            e0 = ctx->h_solve[0].err;
        }
        e_out = e0;
        if (pass == 0) best_saved = e0;
        if (trace) fprintf(stderr, "[mixed] pass %d: fp64 residual %.3e after %lld iterations\n", pass, e0, total);
        if (ctx->h_solve[0].done) {
            status = ctx->h_solve[0].status;
            fp64 = status == DALM_BREAKDOWN || status == DALM_NONFINITE;
            break;
        }
        if (total >= max_it) {
            status = DALM_NONCONVERGED;
            break;
        }
        // FP64 from here after a pass that gained less than kPassGain or after too
        // many passes
        if (pass > 0 && (!(e0 < prev_pass / kPassGain) || pass >= kMaxPasses)) {
            fp64 = true;
            break;
        }
        prev_pass = e0;
        double best = e0, anchor = e0;  // lowest residual; last 30% milestone
        long long anchor_it = 0, launched = 0;
        bool stalled = false;
        probe_solve_begin(ctx);
        while (launched < max_it - total) {
            const long long chunk = chunk_for(launched, ctx->prev_it[1], max_it - total);
            for (long long c = 0; c < chunk; ++c) {
                probe_iteration(ctx, launched + c);
                if (fused) {
                    // iteration k reads p, q from the pair's k-parity buffer and writes the
                    // other, and its z_p, z_s into pair k mod kXGroupF of ZP, ZS; the p-half
                    // forms the previous iteration's r, the s-half's dots give omega, |r|
                    // and <r0, r> (FIN_MX_OMEGA), and kg_apply_pending4 applies the
                    // outstanding x updates every kXGroupF-th iteration or ahead of a
                    // replacement
                    const long long k = launched + c;
                    float* P[2] = {ctx->fp, ctx->ftmp};
                    float* Q[2] = {ctx->fq, ctx->ftmp2};
                    probe_begin(ctx, DALM_K_HALF_P_MIXED);
                    if ((rc = launch_half_p(ctx, P[k & 1], Q[k & 1], P[(k + 1) & 1], Q[(k + 1) & 1],
                                            ZP[k % kXGroupF])))
                        return dalm_fail_msg(ctx, rc, "fused p half-iteration launch");
                    probe_end(ctx, DALM_K_HALF_P_MIXED);
                    if ((rc = launch_finalize(ctx, FIN_ALPHA, fgrid, st, nullptr))) return rc;
                    probe_begin(ctx, DALM_K_HALF_S_MIXED);
                    if ((rc = launch_half_s(ctx, Q[(k + 1) & 1], ZS[k % kXGroupF])))
                        return dalm_fail_msg(ctx, rc, "fused s half-iteration launch");
                    probe_end(ctx, DALM_K_HALF_S_MIXED);
                    if ((rc = launch_finalize(ctx, FIN_MX_OMEGA, fgrid, st, nullptr))) return rc;
                    hipLaunchKernelGGL(kg_apply_pending4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, ctx->p,
                                       ring(k), (const SolveState*)st);
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_resid<false>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream,
                                       d, A, ctx->p, ctx->b, ctx->fr, ctx->fr0, ctx->partials, (const SolveState*)st);
                    if ((rc = launch_finalize(ctx, FIN_MX_RU, ctx->sgrid, st, nullptr))) return rc;
                    continue;
                }
                probe_begin(ctx, DALM_K_PUPDATE_MIXED);
                if (vec4)
                    hipLaunchKernelGGL(kg_p4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, Af, ctx->fr, ctx->fp,
                                       ctx->fq, ctx->fz, pc_first, (const SolveState*)st);
                else
                    hipLaunchKernelGGL(kg_p, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, Af, ctx->fr, ctx->fp,
                                       ctx->fq, ctx->fz, pc_first, (const SolveState*)st);
                probe_end(ctx, DALM_K_PUPDATE_MIXED);
                float* zp = precondition(ctx->fz, ctx->ftmp, ctx->fp);
                probe_begin(ctx, DALM_K_SPMV_MIXED);
                if (vec4)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_spmv4<0>), dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A,
                                       zp, ctx->fq, ctx->fr0, ctx->partials, (const SolveState*)st);
                else
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_spmv<0>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d,
                                       A, zp, ctx->fq, ctx->fr0, ctx->partials, (const SolveState*)st);
                probe_end(ctx, DALM_K_SPMV_MIXED);
                if ((rc = launch_finalize(ctx, FIN_ALPHA, vec4 ? grid4 : ctx->sgrid, st, nullptr))) return rc;
                probe_begin(ctx, DALM_K_SUPDATE_MIXED);
                if (vec4)
                    hipLaunchKernelGGL(kg_s4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, Af, ctx->fr, ctx->fq,
                                       ctx->fs, ctx->fz2, pc_first, (const SolveState*)st);
                else
                    hipLaunchKernelGGL(kg_s, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, Af, ctx->fr, ctx->fq,
                                       ctx->fs, ctx->fz2, pc_first, (const SolveState*)st);
                probe_end(ctx, DALM_K_SUPDATE_MIXED);
                float* zs = precondition(ctx->fz2, ctx->ftmp2, ctx->fs);
                probe_begin(ctx, DALM_K_SPMV_MIXED);
                if (vec4)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_spmv4<1>), dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A,
                                       zs, ctx->ft, ctx->fs, ctx->partials, (const SolveState*)st);
                else
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_spmv<1>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d,
                                       A, zs, ctx->ft, ctx->fs, ctx->partials, (const SolveState*)st);
                probe_end(ctx, DALM_K_SPMV_MIXED);
                if ((rc = launch_finalize(ctx, FIN_OMEGA, vec4 ? grid4 : ctx->sgrid, st, nullptr))) return rc;
                probe_begin(ctx, DALM_K_UPDATE_MIXED);
                if (vec4)
                    hipLaunchKernelGGL(kg_xr4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, ctx->p, zp, zs, ctx->fs,
                                       ctx->ft, ctx->fr, ctx->fr0, ctx->partials, (const SolveState*)st);
                else
                    hipLaunchKernelGGL(kg_xr, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, ctx->p, zp, zs,
                                       ctx->fs, ctx->ft, ctx->fr, ctx->fr0, ctx->partials, (const SolveState*)st);
                probe_end(ctx, DALM_K_UPDATE_MIXED);
                if ((rc = launch_finalize(ctx, FIN_MX_XR, vec4 ? grid4 : ctx->sgrid, st, nullptr))) return rc;
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_resid<false>), dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d,
                                   A, ctx->p, ctx->b, ctx->fr, ctx->fr0, ctx->partials, (const SolveState*)st);
                if ((rc = launch_finalize(ctx, FIN_MX_RU, ctx->sgrid, st, nullptr))) return rc;
            }
            LAUNCH_CHECK();
            launched += chunk;
            if ((rc = read_states(ctx))) return rc;
            const SolveState& h = ctx->h_solve[0];
            if (trace)
                fprintf(stderr, "[mixed]   it %lld: residual %.3e (%lld replacements) alpha %.3e omega %.3e\n",
                        total + h.it, h.err, h.n_ru, h.alpha, h.omega);
            if (h.done) break;
            const double e = h.err;
            if (e < best) best = e;
            if (e < 0.5 * best_saved) {
                HIP_TRY(hipMemcpyAsync(ctx->pbest, ctx->p, sizeof(double) * ctx->L, hipMemcpyDeviceToDevice,
                                       ctx->stream));
                if (fused && h.xdef > 0)  // the pressure of the estimate e includes them
                    flush(h, ctx->pbest);
                best_saved = e;
            }
            if (e < 0.7 * anchor) {
                anchor = e;
                anchor_it = h.it;
            }
            const long long window = pass == 0 ? kWatchWindow : kRestartWindow;
            if (!(e == e) || e > kWatchRise * best || h.it - anchor_it >= window ||
                (pass == 0 && best < kNearTol * tol && h.it - anchor_it >= kChunk) ||
                (test_cap > 0 && h.it >= test_cap)) {
                stalled = true;
                break;
            }
        }
        const SolveState& h = ctx->h_solve[0];
        probe_solve_end(ctx, h.it);
        if (fused && h.xdef > 0) {
            // the pass's outstanding x updates
            flush(h, ctx->p);
            LAUNCH_CHECK();
        }
        if (fused && (h.it & 1)) {
            // after an odd number of iterations the current p and q are the partners
            std::swap(ctx->fp, ctx->ftmp);
            std::swap(ctx->fq, ctx->ftmp2);
        }
        total += h.it;
        n_ru += h.n_ru;
        status = h.status;
        e_out = h.err;
        if (h.done) {
            fp64 = status == DALM_BREAKDOWN || status == DALM_NONFINITE;
            break;  // met the test on the true residual, or reached max_it
        }
        if (!stalled) {
            status = DALM_NONCONVERGED;
            break;
        }
        // a restarted pass that stalled near the tolerance, gained too little or was
        // the last allowed: continue it in FP64; otherwise start another FP32 pass
        if (pass > 0 && (e_out < kPassFloor * tol || !(e_out < e0 / kPassGain) || pass + 1 >= kMaxPasses) &&
            e_out <= kRestoreRise * best_saved && total < max_it) {
            hipLaunchKernelGGL(kg_upgrade, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, (const double*)ctx->p,
                               (const double*)ctx->b, (const float*)ctx->fr0, (const float*)ctx->fp,
                               (const float*)ctx->fq, ctx->r, ctx->r0, ctx->kp, ctx->q, ctx->partials);
            LAUNCH_CHECK();
            if ((rc = launch_finalize(ctx, FIN_MX_UP, ctx->sgrid, st, nullptr))) return rc;
            if ((rc = read_states(ctx))) return rc;
            const long long it32 = ctx->h_solve[0].it;
            if (trace)
                fprintf(stderr, "[mixed] stalled near the tolerance (%.3e): FP64 continuation\n", ctx->h_solve[0].err);
            if (ctx->h_solve[0].done) {
                status = ctx->h_solve[0].status;
                e_out = ctx->h_solve[0].err;
                fp64 = status == DALM_BREAKDOWN || status == DALM_NONFINITE;
                break;
            }
            long long it_all = 0;
            status = run_bicgstab_on(ctx, ctx->p, ctx->b, 2, pc_kind, pc_omega, pc_it, max_it - total, tol, &it_all,
                                     &e_out, true);
            if (status == DALM_HIP_ERROR || status == DALM_BAD_ARG) return status;
            ctx->fp64_finish_it = it_all - it32;
            total += it_all - it32;
            if (trace)
                fprintf(stderr, "[mixed] FP64 continuation: %lld iterations, residual %.3e\n", ctx->fp64_finish_it,
                        e_out);
            fp64 = status == DALM_BREAKDOWN || status == DALM_NONFINITE;
            break;
        }
    }
    ctx->outer_it = n_ru;
    ctx->prev_it[1] = total;
    if (fp64 && total < max_it) {
        // the better (true residual) of the current and the best saved pressure,
        // then the FP64 BiCGStab from there
        if ((rc = restart())) return rc;
        const double end_err = ctx->h_solve[0].err;
        std::swap(ctx->p, ctx->pbest);
        if ((rc = restart())) return rc;
        double from = ctx->h_solve[0].err;
        if (!(from < end_err)) {
            std::swap(ctx->p, ctx->pbest);
            from = end_err;
        }
        long long it64 = 0;
        status = run_bicgstab_on(ctx, ctx->p, ctx->b, 2, pc_kind, pc_omega, pc_it, max_it - total, tol, &it64, &e_out);
        if (status == DALM_HIP_ERROR || status == DALM_BAD_ARG) return status;
        ctx->fp64_finish_it = it64;
        total += it64;
        if (trace)
            fprintf(stderr, "[mixed] FP64 finish from %.3e: %lld iterations, residual %.3e\n", from, it64, e_out);
    } else if (fp64) {
        status = DALM_NONCONVERGED;
    }
    *it = total;
    *err = e_out;
    return status;
}
