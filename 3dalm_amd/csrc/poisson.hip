// poisson.hip -- FP64 Poisson solvers of the drop-in form: red-black SOR
// (src/eq.h:145-306) and preconditioned BiCGStab with Jacobi or SOR
// preconditioning (src/eq.h:6-143, 308-473), on the matrix-free 7-point operator
// of build_A (src/eq.h:701-798).
//
// Design:
//  * the operator is never stored: off-diagonals are 1-D arrays, the diagonal is
//    rebuilt per cell from 1-D pieces in the reference's accumulation order, so
//    each sweep / SpMV reads only the iterate and the right-hand side;
//  * every scalar of the iteration (rho, alpha, omega, beta, err, it) lives in
//    device memory and is produced by the last block of the reduction that feeds
//    it (deterministic fixed-order tree); the stopping test runs on the device;
//  * the host enqueues iterations in chunks without waiting; every kernel of an
//    iteration returns at once when the device flag `done` is set, so the solve
//    stops at exactly the reference's iteration.  The host reads the flag once per
//    chunk.
// Compiled with -ffp-contract=off: sweep, residual and SpMV values are the
// reference's bit for bit; dot products differ only in summation order.
#include <float.h>

#include <utility>

#include "ctx.h"
#include "stencil.h"

namespace {

__device__ __forceinline__ bool finished(const SolveState* st) {
    return st != nullptr && st->done;
}

// ---------------------------------------------------------------------------- SOR
__global__ __launch_bounds__(BLOCK) void k_sor(Dims d, PCoef A, double* __restrict__ x,
                                               const double* __restrict__ b, double omega,
                                               int color, const SolveState* st) {
    if (finished(st)) return;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const int j = cl.j, k = cl.k;
    const OpD op(A, j, k);
    // every neighbour of a cell of this colour has the other colour, which this
    // sweep does not change, so values requested ahead of the updates stay valid
    walk<double, double>(d, cl, x, [&](int c) { return b[c]; },
                         [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, double bc) {
                             if (((i + j + k) & 1) == color) {
                                 const double ac = op.diag(A, i);
                                 const double rel = (bc - op.apply(A, i, ac, x0, xp, xm, q)) / ac;
                                 x[c] = x0 + omega * rel;
                             }
                         });
}

// r = b - A x (stored if r != null) and the stopping test of run_sor
// (src/eq.h:297-305): err = |r| / sqrt(N); it += 1; stop when !(it < max_it &&
// err > tol).
__global__ __launch_bounds__(BLOCK) void k_resnorm(Dims d, PCoef A, const double* __restrict__ x,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ r, double* partials,
                                                   unsigned* counter, SolveState* st, int count_it) {
    if (count_it && finished(st)) return;
    const Col cl = col_unit(d);
    double v[1] = {0.0};
    if (cl.ok) {
        const OpD op(A, cl.j, cl.k);
        walk<double, double>(d, cl, x, [&](int c) { return b[c]; },
                             [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, double bc) {
                                 const double rv = bc - op.apply(A, i, op.diag(A, i), x0, xp, xm, q);
                                 if (r) r[c] = rv;
                                 v[0] += rv * rv;
                             });
    }
    block_sum<1>(v);
    double pv[1] = {v[0]};
    store_partials<1>(pv, partials, cl.slot);
}

__global__ void k_state_reset(SolveState* st, double tol, long long max_it, double root_n) {
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
        st->pending = 0;
        st->xdef = 0;
    }
}

// -------------------------------------------------------------------- BiCGStab
// r = b - A x, r0 = r, rho = <r0, r> (src/eq.h:342-353)
__global__ __launch_bounds__(BLOCK) void k_bi_init(Dims d, PCoef A, const double* __restrict__ x,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ r, double* __restrict__ r0,
                                                   double* partials, unsigned* counter,
                                                   SolveState* st) {
    const Col cl = col_unit(d);
    double v[1] = {0.0};
    if (cl.ok) {
        const OpD op(A, cl.j, cl.k);
        walk<double, double>(d, cl, x, [&](int c) { return b[c]; },
                             [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, double bc) {
                                 const double rv = bc - op.apply(A, i, op.diag(A, i), x0, xp, xm, q);
                                 r[c] = rv;
                                 r0[c] = rv;
                                 v[0] += rv * rv;
                             });
    }
    block_sum<1>(v);
    double pv[1] = {v[0]};
    store_partials<1>(pv, partials, cl.slot);
}

// p = r (first iteration) or r + beta (p - omega q); then the first
// preconditioner sweep from zero: Jacobi gives z = p / a_c pointwise, SOR starts
// from z = 0 (src/eq.h:355-368, 437-439).
__global__ __launch_bounds__(BLOCK) void k_bi_p(Dims d, PCoef A, const double* __restrict__ r,
                                                double* __restrict__ p, const double* __restrict__ q,
                                                double* __restrict__ z, int pc_first,
                                                const SolveState* st) {
    if (st->done) return;
    const bool first = st->it == 0;
    const double beta = st->beta, om = st->omega;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const OpD op(A, cl.j, cl.k);
    struct RPQ {
        double r, p, q;
    };
    walk_points<RPQ>(d, cl, [&](int c) { return RPQ{r[c], p[c], q[c]}; },
                     [&](int i, int c, const RPQ& v) {
                         const double pv = first ? v.r : v.r + beta * (v.p - om * v.q);
                         p[c] = pv;
                         z[c] = pc_first ? pv / op.diag(A, i) : 0.0;
                     });
}

// One Jacobi sweep z_out = z_in + (v - A z_in) / a_c (src/eq.h:6-39)
__global__ __launch_bounds__(BLOCK) void k_jacobi(Dims d, PCoef A, const double* __restrict__ zin,
                                                  double* __restrict__ zout,
                                                  const double* __restrict__ v,
                                                  const SolveState* st) {
    if (finished(st)) return;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const OpD op(A, cl.j, cl.k);
    walk<double, double>(d, cl, zin, [&](int c) { return v[c]; },
                         [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, double vc) {
                             const double ac = op.diag(A, i);
                             const double rel = (vc - op.apply(A, i, ac, x0, xp, xm, q)) / ac;
                             zout[c] = x0 + rel;
                         });
}

// k_jacobi with four consecutive k per lane: the same per-cell arithmetic (OpD's
// diagonal and operator, in its order), so the sweep is bit-identical; each vector
// moves with two 16-byte accesses per lane and plane.  The four-cell kernels store
// nontemporally: a stored vector is next read one kernel or one iteration later,
// after more than the 256 MB last-level cache has streamed past, so keeping it
// there only evicts data that is still to be re-read.
typedef double d4k __attribute__((ext_vector_type(kVecK), aligned(8)));

__global__ __launch_bounds__(BLOCK) void k_jacobi4(Dims d, PCoef A, const double* __restrict__ zin,
                                                   double* __restrict__ zout, const double* __restrict__ v,
                                                   const SolveState* st) {
    if (finished(st)) return;
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    const double an = A.an[cl.j], as = A.as[cl.j], anu = A.anu[cl.j], asu = A.asu[cl.j];
    double at[kVecK], ab[kVecK], atu[kVecK], abu[kVecK];
#pragma unroll
    for (int q = 0; q < kVecK; ++q) {
        at[q] = A.at[cl.k0 + q];
        ab[q] = A.ab[cl.k0 + q];
        atu[q] = A.atu[cl.k0 + q];
        abu[q] = A.abu[cl.k0 + q];
    }
    int c = d.at(cl.i0, cl.j, cl.k0);
    d4k xm = *(const d4k*)(zin + c - d.si), x0 = *(const d4k*)(zin + c), xp = *(const d4k*)(zin + c + d.si);
    d4k xn = *(const d4k*)(zin + c + d.sj), xs = *(const d4k*)(zin + c - d.sj), vc = *(const d4k*)(v + c);
    double kb = zin[c - 1], kf = zin[c + kVecK];
    auto plane = [&](int i, int cc) {
        const double acx = A.acx[i], ae = A.ae[i], aw = A.aw[i];
        d4k out;
#pragma unroll
        for (int q = 0; q < kVecK; ++q) {
            const double ac = ((((acx - anu) - asu) - atu[q]) - abu[q]) / A.max_diag;
            const double t = q + 1 < kVecK ? x0[q + 1] : kf;
            const double b = q > 0 ? x0[q - 1] : kb;
            const double ax = ac * x0[q] + ae * xp[q] + aw * xm[q] + an * xn[q] + as * xs[q] + at[q] * t + ab[q] * b;
            out[q] = x0[q] + (vc[q] - ax) / ac;
        }
        __builtin_nontemporal_store(out, (d4k*)(zout + cc));
    };
    // every value of plane i+1 is requested before plane i is computed
    auto step = [&](int i, int nxt) {
        const int cn = c + nxt;
        const d4k xpp = *(const d4k*)(zin + cn + d.si);
        const d4k xn1 = *(const d4k*)(zin + cn + d.sj), xs1 = *(const d4k*)(zin + cn - d.sj);
        const d4k vc1 = *(const d4k*)(v + cn);
        const double kb1 = zin[cn - 1], kf1 = zin[cn + kVecK];
        if (i >= cl.own0) plane(i, c);
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
        for (int w = 0; w < kMaxPlanes; ++w) step(cl.i0 + w, w + 1 < kMaxPlanes ? d.si : 0);
    } else {
        for (int i = cl.i0; i < cl.i1; ++i) step(i, i + 1 < cl.i1 ? d.si : 0);
    }
}

// y = A z and the partial sums of <a1, y> (and <y, a2> if a2): with which = 0 the
// last block sets alpha = rho / <r0, q>; with which = 1 it sets
// omega = <t, s> / <t, t> (src/eq.h:369-383).
template <int WHICH>
__global__ __launch_bounds__(BLOCK) void k_bi_spmv(Dims d, PCoef A, const double* __restrict__ z,
                                                   double* __restrict__ y,
                                                   const double* __restrict__ a1, double* partials,
                                                   unsigned* counter, SolveState* st) {
    if (st->done) return;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const OpD op(A, cl.j, cl.k);
        walk<double, double>(d, cl, z, [&](int c) { return a1[c]; },
                             [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, double ac1) {
                                 const double yv = op.apply(A, i, op.diag(A, i), x0, xp, xm, q);
                                 y[c] = yv;
                                 if (WHICH == 0) {
                                     v[0] += ac1 * yv;  // <r0, q>
                                 } else {
                                     v[0] += yv * ac1;  // <t, s>
                                     v[1] += yv * yv;   // <t, t>
                                 }
                             });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// s = r - alpha q, then the first preconditioner sweep for ss (src/eq.h:373-380)
__global__ __launch_bounds__(BLOCK) void k_bi_s(Dims d, PCoef A, const double* __restrict__ r,
                                                const double* __restrict__ q, double* __restrict__ s,
                                                double* __restrict__ z, int pc_first,
                                                const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha;
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const OpD op(A, cl.j, cl.k);
    struct RQ {
        double r, q;
    };
    walk_points<RQ>(d, cl, [&](int c) { return RQ{r[c], q[c]}; },
                    [&](int i, int c, const RQ& v) {
                        const double sv = v.r - alpha * v.q;
                        s[c] = sv;
                        z[c] = pc_first ? sv / op.diag(A, i) : 0.0;
                    });
}

// x += alpha pp + omega ss; r = s - omega t; err = |r| / sqrt(N); next rho =
// <r0, r>; stopping test and breakdown test of the next iteration
// (src/eq.h:385-401, 349-353).
__global__ __launch_bounds__(BLOCK) void k_bi_xr(Dims d, double* __restrict__ x,
                                                 const double* __restrict__ pp,
                                                 const double* __restrict__ ss,
                                                 const double* __restrict__ s,
                                                 const double* __restrict__ t_,
                                                 double* __restrict__ r,
                                                 const double* __restrict__ r0, double* partials,
                                                 unsigned* counter, SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha, om = st->omega;
    const Col cl = col_unit(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        struct XR {
            double x, pp, ss, s, t, r0;
        };
        walk_points<XR>(d, cl, [&](int c) { return XR{x[c], pp[c], ss[c], s[c], t_[c], r0[c]}; },
                        [&](int, int c, const XR& a) {
                            x[c] = a.x + alpha * a.pp + om * a.ss;
                            const double rv = a.s - om * a.t;
                            r[c] = rv;
                            v[0] += rv * rv;
                            v[1] += a.r0 * rv;
                        });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// ---- the BiCGStab kernels with four cells per lane (rows of 256 interior cells):
// the per-cell arithmetic of the one-cell kernels above, so every stored vector is
// bit-identical; the dot products add the four cells of a lane in k order.
__device__ __forceinline__ d4k ld4(const double* a, int c) { return *(const d4k*)(a + c); }

// the diagonal of OpD for the four cells of a lane at plane i
struct Diag4 {
    double anu, asu, atu[kVecK], abu[kVecK];
    __device__ __forceinline__ Diag4(const PCoef& A, const Col4& cl) : anu(A.anu[cl.j]), asu(A.asu[cl.j]) {
#pragma unroll
        for (int m = 0; m < kVecK; ++m) {
            atu[m] = A.atu[cl.k0 + m];
            abu[m] = A.abu[cl.k0 + m];
        }
    }
    __device__ __forceinline__ double at(const PCoef& A, double acx, int m) const {
        return ((((acx - anu) - asu) - atu[m]) - abu[m]) / A.max_diag;
    }
};

__global__ __launch_bounds__(BLOCK) void k_bi_p4(Dims d, PCoef A, const double* __restrict__ r,
                                                 double* __restrict__ p, const double* __restrict__ q,
                                                 double* __restrict__ z, int pc_first, const SolveState* st) {
    if (st->done) return;
    const bool first = st->it == 0;
    const double beta = st->beta, om = st->omega;
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    const Diag4 dg(A, cl);
    struct RPQ {
        d4k r, p, q;
    };
    walk_points4<RPQ>(d, cl, [&](int c) { return RPQ{ld4(r, c), ld4(p, c), ld4(q, c)}; },
                      [&](int i, int c, const RPQ& v) {
                          const double acx = A.acx[i];
                          d4k pv, zv;
#pragma unroll
                          for (int m = 0; m < kVecK; ++m) {
                              pv[m] = first ? v.r[m] : v.r[m] + beta * (v.p[m] - om * v.q[m]);
                              zv[m] = pc_first ? pv[m] / dg.at(A, acx, m) : 0.0;
                          }
                          __builtin_nontemporal_store(pv, (d4k*)(p + c));
                          __builtin_nontemporal_store(zv, (d4k*)(z + c));
                      });
}

__global__ __launch_bounds__(BLOCK) void k_bi_s4(Dims d, PCoef A, const double* __restrict__ r,
                                                 const double* __restrict__ q, double* __restrict__ s,
                                                 double* __restrict__ z, int pc_first, const SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha;
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    const Diag4 dg(A, cl);
    struct RQ {
        d4k r, q;
    };
    walk_points4<RQ>(d, cl, [&](int c) { return RQ{ld4(r, c), ld4(q, c)}; },
                     [&](int i, int c, const RQ& v) {
                         const double acx = A.acx[i];
                         d4k sv, zv;
#pragma unroll
                         for (int m = 0; m < kVecK; ++m) {
                             sv[m] = v.r[m] - alpha * v.q[m];
                             zv[m] = pc_first ? sv[m] / dg.at(A, acx, m) : 0.0;
                         }
                         __builtin_nontemporal_store(sv, (d4k*)(s + c));
                         __builtin_nontemporal_store(zv, (d4k*)(z + c));
                     });
}

// z_p and z_s of kXGroup consecutive iterations, oldest first
struct PRing {
    const double* pp[kXGroup];
    const double* ss[kXGroup];
};

// XM 1: x += alpha pp + omega ss and r = s - omega t with the dots of k_bi_xr (pp, ss
// in the ring's last place).  XM 0 (the fused iteration) forms r only: its x updates
// are applied kXGroup at a time by k_x_group4.
template <int XM>
__global__ __launch_bounds__(BLOCK) void k_bi_xr4(Dims d, double* __restrict__ x, PRing z,
                                                  const double* __restrict__ s, const double* __restrict__ t_,
                                                  double* __restrict__ r, const double* __restrict__ r0,
                                                  double* partials, SolveState* st) {
    if (st->done) return;
    const double alpha = st->alpha, om = st->omega;
    const Col4 cl = col_unit4(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const double* __restrict__ pp = z.pp[kXGroup - 1];
        const double* __restrict__ ss = z.ss[kXGroup - 1];
        struct XR {
            d4k x, pp, ss, s, t, r0;
        };
        const d4k z4 = {0.0, 0.0, 0.0, 0.0};
        walk_points4<XR>(d, cl,
                         [&](int c) {
                             XR a{z4, z4, z4, ld4(s, c), ld4(t_, c), ld4(r0, c)};
                             if (XM == 1) {
                                 a.x = ld4(x, c);
                                 a.pp = ld4(pp, c);
                                 a.ss = ld4(ss, c);
                             }
                             return a;
                         },
                         [&](int, int c, const XR& a) {
                             d4k xv, rv;
#pragma unroll
                             for (int m = 0; m < kVecK; ++m) {
                                 xv[m] = a.x[m] + alpha * a.pp[m] + om * a.ss[m];
                                 rv[m] = a.s[m] - om * a.t[m];
                                 v[0] += rv[m] * rv[m];
                                 v[1] += a.r0[m] * rv[m];
                             }
                             if (XM == 1) __builtin_nontemporal_store(xv, (d4k*)(x + c));
                             __builtin_nontemporal_store(rv, (d4k*)(r + c));
                         });
    }
    block_sum<2>(v);
    double pv[2] = {v[0], v[1]};
    store_partials<2>(pv, partials, cl.slot);
}

// The fused iteration's x updates of the ring's iterations, in turn (x += alpha pp +
// omega ss with alpha_h, omega_h: the same two roundings per update as k_bi_xr4<1>, in
// the same order, so x is the same to the bit), after iteration `want` - 1 of the
// solve; enqueued every kXGroup iterations, it returns at once when the solve ended
// before (the end of the solve then applies the outstanding ones).  Rows of 256 cells,
// four cells per lane: x and the pairs of z stream through once.
__global__ __launch_bounds__(BLOCK) void k_x_group4(Dims d, double* __restrict__ x, PRing z, long long want,
                                                    const SolveState* st) {
    if (st->it != want) return;
    const Col4 cl = col_unit4(d);
    if (!cl.ok) return;
    int c = d.at(cl.own0, cl.j, cl.k0);
    for (int i = cl.own0; i < cl.i1; ++i, c += d.si) {
        d4k xv = ld4(x, c);
#pragma unroll
        for (int q = 0; q < kXGroup; ++q) {
            const d4k a = ld4(z.pp[q], c), b = ld4(z.ss[q], c);
            const double aq = st->alpha_h[q], oq = st->omega_h[q];
#pragma unroll
            for (int m = 0; m < kVecK; ++m) xv[m] = xv[m] + aq * a[m] + oq * b[m];
        }
        __builtin_nontemporal_store(xv, (d4k*)(x + c));
    }
}
// The updates outstanding at the end of a solve: x += alpha pp + omega ss for the
// ring's places [first, kXGroup), in turn, on the interior
struct XFlush64 {
    double alpha[kXGroup], omega[kXGroup];
    int first;
};
__global__ __launch_bounds__(BLOCK) void k_x_flush(Dims d, double* __restrict__ x, PRing z, XFlush64 f) {
    const TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        const Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int c = d.at(cc.i, cc.j, cc.k);
        double v = x[c];
#pragma unroll
        for (int q = 0; q < kXGroup; ++q)
            if (q >= f.first) v = v + f.alpha[q] * z.pp[q][c] + f.omega[q] * z.ss[q][c];
        x[c] = v;
    }
}

template <int WHICH>
__global__ __launch_bounds__(BLOCK) void k_bi_spmv4(Dims d, PCoef A, const double* __restrict__ z,
                                                    double* __restrict__ y, const double* __restrict__ a1,
                                                    double* partials, SolveState* st) {
    if (st->done) return;
    const Col4 cl = col_unit4(d);
    double v[2] = {0.0, 0.0};
    if (cl.ok) {
        const Diag4 dg(A, cl);
        const double an = A.an[cl.j], as = A.as[cl.j];
        double at[kVecK], ab[kVecK];
#pragma unroll
        for (int m = 0; m < kVecK; ++m) {
            at[m] = A.at[cl.k0 + m];
            ab[m] = A.ab[cl.k0 + m];
        }
        int c = d.at(cl.i0, cl.j, cl.k0);
        d4k xm = ld4(z, c - d.si), x0 = ld4(z, c), xp = ld4(z, c + d.si);
        d4k xn = ld4(z, c + d.sj), xs = ld4(z, c - d.sj), av = ld4(a1, c);
        double kb = z[c - 1], kf = z[c + kVecK];
        auto plane = [&](int i, int cc) {
            const double acx = A.acx[i], ae = A.ae[i], aw = A.aw[i];
            d4k out;
#pragma unroll
            for (int m = 0; m < kVecK; ++m) {
                const double ac = dg.at(A, acx, m);
                const double t = m + 1 < kVecK ? x0[m + 1] : kf;
                const double b = m > 0 ? x0[m - 1] : kb;
                const double yv = ac * x0[m] + ae * xp[m] + aw * xm[m] + an * xn[m] + as * xs[m] + at[m] * t + ab[m] * b;
                out[m] = yv;
                if (WHICH == 0) {
                    v[0] += av[m] * yv;  // <r0, q>
                } else {
                    v[0] += yv * av[m];  // <t, s>
                    v[1] += yv * yv;     // <t, t>
                }
            }
            __builtin_nontemporal_store(out, (d4k*)(y + cc));
        };
        auto step = [&](int i, int nxt) {
            const int cn = c + nxt;
            const d4k xpp = ld4(z, cn + d.si);
            const d4k xn1 = ld4(z, cn + d.sj), xs1 = ld4(z, cn - d.sj);
            const d4k av1 = ld4(a1, cn);
            const double kb1 = z[cn - 1], kf1 = z[cn + kVecK];
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

// ---- fused update + preconditioner (rows of exactly 256 cells, Jacobi x3) -----------
// The vector update of a BiCGStab half-iteration and the three Jacobi sweeps that
// follow it (src/eq.h:355-380, 6-38) as one kernel: p = r + beta (p - omega q) (or
// s = r - alpha q), z1 = v / a_c, z2 = J(z1; v), z3 = J(z2; v), writing only v and z3
// (40 / 32 B per cell instead of 40 + 48 / 32 + 48 for the update and two sweeps
// through HBM).  Temporal blocking along i as in the precision form's fused
// half-iteration, two stencil levels deep: a block of 12 wavefronts (one j row
// each, four k per lane; 12 rather than 16 so that a wavefront may hold 170 VGPRs
// of FP64 pipeline state) owns 8 rows and computes 2 / 1 halo rows of levels 0 / 1
// on either side; level 0 runs at plane t, the sweeps at t-1 and t-2; the new plane
// of each sweep's input goes through LDS (two copies alternate: read, compute, write,
// barrier).
// The SpMV that ends the half-iteration stays a separate kernel (k_bi_spmv4).
// Arithmetic per cell identical to k_bi_p4 / k_bi_s4 and k_jacobi4 (bit for bit);
// the iterates' ghost cells are zero, as in the reference.  p and q are read at
// halo cells owned by neighbouring blocks, so the p-half reads them from one buffer
// of a pair and writes the new p to the other.
constexpr int kFRows = 12;
constexpr int kFOwn = kFRows - 4;
constexpr int kFLag = 2;
constexpr int kFPad = 4;
constexpr int kFRowLen = 256 + 2 * kFPad;
typedef double d4a __attribute__((ext_vector_type(kVecK), aligned(32)));

struct FLds {
    // the new planes of z1 and z2, in two alternating copies: a step reads the
    // copy the previous step wrote and writes the other, so one barrier per step
    // orders both
    alignas(32) double lv[2][2][kFRows][kFRowLen];
    double at[256], ab[256], atu[256], abu[256];
};

template <int WHICH>
__global__ __launch_bounds__(64 * kFRows) void k_half_pc4(Dims d, PCoef A, const double* __restrict__ r,
                                                          const double* __restrict__ vin,
                                                          const double* __restrict__ q, double* __restrict__ vout,
                                                          double* __restrict__ zout, int ni, const SolveState* st) {
    if (st->done) return;
    __shared__ FLds sh;
    const int lane = threadIdx.x & 63, row = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = d.gc;
    // j tile fastest, a contiguous range of units per XCD (as kg_half4)
    const int tjf = (d.ny + kFOwn - 1) / kFOwn;
    const int u = xcd_unit((int)blockIdx.x, (int)gridDim.x);
    const int jt = u % tjf, ic = u / tjf;
    const int own0 = g + ic * ni;
    int i1 = own0 + ni;
    if (i1 > g + d.nx) i1 = g + d.nx;
    const int i0 = i1 - ni;
    const int j = g + jt * kFOwn + row - kFLag;
    const bool jin = j >= g && j < g + d.ny;
    const bool owned = jin && row >= kFLag && row < kFRows - kFLag;
    const bool l1 = jin && row >= 1 && row < kFRows - 1;
    const int k0 = g + 4 * lane, lk = 4 * lane, o = kFPad + lk;
    const int rr = row == 0 ? 1 : (row == kFRows - 1 ? kFRows - 2 : row);  // neighbour rows stay in range
    const double beta = st->beta, om = st->omega, alpha = st->alpha;
    const bool first = st->it == 0;
    const int jc = jin ? j : g;
    const double an = A.an[jc], as = A.as[jc], anu = A.anu[jc], asu = A.asu[jc];
    for (int t = threadIdx.x; t < 256; t += 64 * kFRows) {
        sh.at[t] = A.at[g + t];
        sh.ab[t] = A.ab[g + t];
        sh.atu[t] = A.atu[g + t];
        sh.abu[t] = A.abu[g + t];
    }
    for (int t = threadIdx.x; t < 4 * kFRows * kFRowLen; t += 64 * kFRows) (&sh.lv[0][0][0][0])[t] = 0.0;
    __syncthreads();

    const d4k zero = {0.0, 0.0, 0.0, 0.0};
    d4k z1a = zero, z1b = zero;  // z1 at t-1, t-2
    d4k va = zero, vb = zero;    // v at t-1, t-2
    d4k z2a = zero, z2b = zero;  // z2 at t-2, t-3
    struct In {
        d4k r, v, q;
    };
    auto load0 = [&](int i) {
        In x{zero, zero, zero};
        if (jin && i >= g && i < g + d.nx) {
            const int c = d.at(i, j, k0);
            x.r = ld4(r, c);
            x.q = ld4(q, c);
            if (WHICH == 0) x.v = ld4(vin, c);
        }
        return x;
    };
    struct Nb {
        d4k n, s;
        double kb, kf;
    };
    auto nbrs = [&](int cp, int lev) {
        Nb b;
        b.n = *(const d4a*)&sh.lv[cp][lev][rr + 1][o];
        b.s = *(const d4a*)&sh.lv[cp][lev][rr - 1][o];
        b.kb = sh.lv[cp][lev][rr][o - 1];
        b.kf = sh.lv[cp][lev][rr][o + 4];
        return b;
    };
    // one Jacobi sweep at plane ip (k_jacobi4's arithmetic).  The s-half takes the plane's
    // diagonal acv, computed by level 0 when it reached the plane and kept in registers
    // since (the same value, divided by max|a_c| once per cell instead of three times:
    // 236 against 258 us at 512x256x256); the p-half recomputes it, which measured
    // faster there (290 against 307 us with the carried diagonal)
    constexpr bool carry = WHICH == 1;
    auto sweep = [&](int ip, const Nb& nb, const d4k& xb, const d4k& xa, const d4k& xn, const d4k& rhs,
                     const d4k& acv) {
        const double acj = (A.acx[ip] - anu) - asu, ae = A.ae[ip], aw = A.aw[ip];
        d4k out;
#pragma unroll
        for (int m = 0; m < kVecK; ++m) {
            const double ac = carry ? acv[m] : ((acj - sh.atu[lk + m]) - sh.abu[lk + m]) / A.max_diag;
            const double tq = m + 1 < kVecK ? xa[m + 1] : nb.kf;
            const double bq = m > 0 ? xa[m - 1] : nb.kb;
            const double ax = ac * xa[m] + ae * xn[m] + aw * xb[m] + an * nb.n[m] + as * nb.s[m] +
                              sh.at[lk + m] * tq + sh.ab[lk + m] * bq;
            out[m] = xa[m] + (rhs[m] - ax) / ac;
        }
        return out;
    };
    In nx0 = load0(i0 - kFLag);
    d4k aca = zero, acb = zero;  // the diagonal (OpD's, divided by max|a_c|) at t-1, t-2
    for (int t = i0 - kFLag; t < i1 + kFLag; ++t) {
        // level 0 at plane t, then the next plane's inputs are requested
        d4k z1n = zero, vn = zero, acn = zero;
        if (jin && t >= g && t < g + d.nx) {
            const double acj = (A.acx[t] - anu) - asu;
#pragma unroll
            for (int m = 0; m < kVecK; ++m) {
                if (WHICH == 0)
                    vn[m] = first ? nx0.r[m] : nx0.r[m] + beta * (nx0.v[m] - om * nx0.q[m]);
                else
                    vn[m] = nx0.r[m] - alpha * nx0.q[m];
                acn[m] = ((acj - sh.atu[lk + m]) - sh.abu[lk + m]) / A.max_diag;
                z1n[m] = vn[m] / acn[m];
            }
            if (owned && t >= own0 && t < i1) __builtin_nontemporal_store(vn, (d4k*)(vout + d.at(t, j, k0)));
        }
        nx0 = load0(t + 1);
        // neighbours of the centre planes of both sweeps (written at the previous step,
        // into the other copy)
        const int cw = (t - i0 + kFLag) & 1;
        const int p1 = t - 1, p2 = t - 2;
        d4k z2n = zero;
        if (l1 && p1 >= g && p1 < g + d.nx && p1 >= i0 - 1) z2n = sweep(p1, nbrs(cw ^ 1, 0), z1b, z1a, z1n, va, aca);
        if (owned && p2 >= own0 && p2 < i1)
            __builtin_nontemporal_store(sweep(p2, nbrs(cw ^ 1, 1), z2b, z2a, z2n, vb, acb),
                                        (d4k*)(zout + d.at(p2, j, k0)));
        *(d4a*)&sh.lv[cw][0][row][o] = z1n;
        *(d4a*)&sh.lv[cw][1][row][o] = z2n;
        __syncthreads();
        z1b = z1a;
        z1a = z1n;
        vb = va;
        va = vn;
        z2b = z2a;
        z2a = z2n;
        if (carry) {
            acb = aca;
            aca = acn;
        }
    }
}

__global__ void k_zero(long long n, double* __restrict__ a) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x)
        a[e] = 0.0;
}

int read_state(dalm_ctx* ctx) {
    HIP_TRY(hipMemcpyAsync(ctx->h_solve, ctx->solve, sizeof(SolveState), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DALM_OK;
}

// Iterations enqueued before the host looks at the device flag again.
long long next_chunk(long long done_it, long long predicted, long long max_it) {
    long long left = predicted - done_it;
    long long c = left > 8 ? left : 8;
    if (c > 256) c = 256;
    if (done_it + c > max_it) c = max_it - done_it;
    return c < 1 ? 1 : c;
}

}  // namespace

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

int launch_sor_sweep(dalm_ctx* ctx, double* x, const double* b, double omega, int color) {
    hipLaunchKernelGGL(k_sor, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->coef(), x, b,
                       omega, color, (const SolveState*)nullptr);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_residual(dalm_ctx* ctx, const double* x, const double* b, double* r) {
    hipLaunchKernelGGL(k_resnorm, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->coef(), x,
                       b, r, ctx->partials, ctx->counter, ctx->solve, 0);
    LAUNCH_CHECK();
    return launch_finalize(ctx, FIN_RESID, ctx->sgrid, ctx->solve, nullptr);
}

__global__ void k_Ax_plain(Dims d, PCoef A, const double* __restrict__ z, double* __restrict__ y) {
    const Col cl = col_unit(d);
    if (!cl.ok) return;
    const OpD op(A, cl.j, cl.k);
    walk<double, NoAux>(d, cl, z, [&](int) { return NoAux{}; },
                        [&](int i, int c, double x0, double xp, double xm, const Nb4<double>& q, NoAux) {
                            y[c] = op.apply(A, i, op.diag(A, i), x0, xp, xm, q);
                        });
}

int launch_Ax(dalm_ctx* ctx, const double* x, double* y) {
    hipLaunchKernelGGL(k_Ax_plain, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->coef(),
                       x, y);
    LAUNCH_CHECK();
    return DALM_OK;
}

int run_sor_on(dalm_ctx* ctx, double* x, const double* b, int slot, double omega, long long max_it, double tol,
               long long* it, double* err) {
    const double root_n = sqrt((double)ctx->N);
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->solve, tol, max_it, root_n);
    LAUNCH_CHECK();
    PCoef A = ctx->coef();
    long long launched = 0;
    probe_solve_begin(ctx);
    for (;;) {
        long long chunk = next_chunk(launched, ctx->prev_it[slot], max_it);
        for (long long c = 0; c < chunk; ++c) {
            probe_iteration(ctx, launched + c);
            for (int color = 0; color < 2; ++color) {
                probe_begin(ctx, DALM_K_SOR);
                hipLaunchKernelGGL(k_sor, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, A, x, b, omega,
                                   color, ctx->solve);
                probe_end(ctx, DALM_K_SOR);
            }
            probe_begin(ctx, DALM_K_RESNORM);
            hipLaunchKernelGGL(k_resnorm, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, ctx->d, A, x, b,
                               ctx->r, ctx->partials, ctx->counter, ctx->solve, 1);
            probe_end(ctx, DALM_K_RESNORM);
            if (int rc = launch_finalize(ctx, FIN_SOR, ctx->sgrid, ctx->solve, nullptr)) return rc;
        }
        LAUNCH_CHECK();
        launched += chunk;
        int rc = read_state(ctx);
        if (rc) return rc;
        if (ctx->h_solve->done || launched >= max_it) break;
    }
    probe_solve_end(ctx, ctx->h_solve->it);
    *it = ctx->h_solve->it;
    *err = ctx->h_solve->err;
    ctx->prev_it[slot] = *it;
    return ctx->h_solve->status;
}

int run_sor(dalm_ctx* ctx, double omega, long long max_it, double tol, long long* it, double* err) {
    return run_sor_on(ctx, ctx->p, ctx->b, 0, omega, max_it, tol, it, err);
}

int run_bicgstab_on(dalm_ctx* ctx, double* x, const double* b, int slot, int pc_kind, double pc_omega, int pc_it,
                    long long max_it, double tol, long long* it, double* err, bool resume) {
    const double root_n = sqrt((double)ctx->N);
    PCoef A = ctx->coef();
    const Dims& d = ctx->d;
    if (!resume) {
        hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->solve, tol, max_it, root_n);
        hipLaunchKernelGGL(k_bi_init, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, x, b, ctx->r, ctx->r0,
                           ctx->partials, ctx->counter, ctx->solve);
        LAUNCH_CHECK();
        if (int rc0 = launch_finalize(ctx, FIN_BI_INIT, ctx->sgrid, ctx->solve, nullptr)) return rc0;
    }
    const int jac = pc_kind == DALM_PC_JACOBI;
    const int pc_first = (jac && pc_it >= 1) ? 1 : 0;
    // rows of a multiple of 256 interior cells: the four-per-lane kernels, whose
    // reductions leave grid4 partials
    const bool vec4 = vec4_applies(d);
    const int grid4 = d.tj * vec4_tiles(d) * d.nch;
    const int nb = vec4 ? grid4 : ctx->sgrid;
    // rows of exactly 256 cells with three Jacobi sweeps: the update and the sweeps of
    // each half-iteration are one kernel (k_half_pc4); p and q alternate between a
    // buffer pair (the Jacobi partners, unused on that path)
    const bool fused = jac && pc_it == 3 && d.nz == 256 && d.ny >= kFOwn;
    if (fused) {
        if (const int rz = ensure_fused_z64(ctx)) return rz;
    }
    // the fused iteration k writes z_p / z_s into buffer pair k mod kXGroup; its x updates
    // are applied kXGroup at a time, after iterations k = kXGroup - 1 mod kXGroup
    double* PP[kXGroup] = {ctx->pp};
    double* SS[kXGroup] = {ctx->ss};
    for (int q = 1; q < kXGroup; ++q) {
        PP[q] = ctx->ppx[q - 1];
        SS[q] = ctx->ssx[q - 1];
    }
    // z of iterations base - kXGroup + 1 + q, q = 0 .. kXGroup - 1
    auto pring = [&](long long base) {
        PRing z;
        for (int q = 0; q < kXGroup; ++q) {
            z.pp[q] = PP[(base + 1 + q) % kXGroup];
            z.ss[q] = SS[(base + 1 + q) % kXGroup];
        }
        return z;
    };
    const int ftiles = (d.ny + kFOwn - 1) / kFOwn;
    const int fni = fused_planes(ctx, ftiles);
    const int fgrid = ftiles * ((d.nx + fni - 1) / fni);
    // preconditioner application for one vector; returns the buffer holding z
    auto precondition = [&](double* z0, double* spare, const double* v) -> double* {
        double* cur = z0;
        if (jac) {
            double* other = spare;
            // the sweeps of one application run back to back: one event pair times them
            if (pc_it > 1) probe_begin(ctx, DALM_K_JACOBI);
            for (int m = 1; m < pc_it; ++m) {
                if (vec4)
                    hipLaunchKernelGGL(k_jacobi4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A, cur, other, v,
                                       (const SolveState*)ctx->solve);
                else
                    hipLaunchKernelGGL(k_jacobi, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, cur,
                                       other, v, (const SolveState*)ctx->solve);
                double* tsw = cur;
                cur = other;
                other = tsw;
            }
            if (pc_it > 1) probe_end(ctx, DALM_K_JACOBI, pc_it - 1);
        } else {
            for (int m = 0; m < pc_it; ++m) {
                hipLaunchKernelGGL(k_sor, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, cur, v,
                                   pc_omega, 0, (const SolveState*)ctx->solve);
                hipLaunchKernelGGL(k_sor, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, cur, v,
                                   pc_omega, 1, (const SolveState*)ctx->solve);
            }
        }
        return cur;
    };
    long long launched = 0;
    int rc = read_state(ctx);
    if (rc) return rc;
    const long long it_start = ctx->h_solve->it;
    probe_solve_begin(ctx);
    while (!ctx->h_solve->done && launched < max_it) {
        long long chunk = next_chunk(launched, ctx->prev_it[slot], max_it);
        for (long long c = 0; c < chunk; ++c) {
            probe_iteration(ctx, launched + c);
            if (fused) {
                // iteration k reads p, q from the pair's k-parity buffer and writes the other
                const long long k = launched + c;
                double* P[2] = {ctx->kp, ctx->tmp};
                double* Q[2] = {ctx->q, ctx->tmp2};
                probe_begin(ctx, DALM_K_HALF_P);
                hipLaunchKernelGGL(k_half_pc4<0>, dim3(fgrid), dim3(64 * kFRows), 0, ctx->stream, d, A,
                                   (const double*)ctx->r, (const double*)P[k & 1], (const double*)Q[k & 1],
                                   P[(k + 1) & 1], PP[k % kXGroup], fni, (const SolveState*)ctx->solve);
                probe_end(ctx, DALM_K_HALF_P);
                probe_begin(ctx, DALM_K_SPMV);
                hipLaunchKernelGGL(k_bi_spmv4<0>, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A, PP[k % kXGroup],
                                   Q[(k + 1) & 1], ctx->r0, ctx->partials, ctx->solve);
                probe_end(ctx, DALM_K_SPMV);
                if (int rc1 = launch_finalize(ctx, FIN_ALPHA, grid4, ctx->solve, nullptr)) return rc1;
                probe_begin(ctx, DALM_K_HALF_S);
                hipLaunchKernelGGL(k_half_pc4<1>, dim3(fgrid), dim3(64 * kFRows), 0, ctx->stream, d, A,
                                   (const double*)ctx->r, (const double*)nullptr, (const double*)Q[(k + 1) & 1],
                                   ctx->s, SS[k % kXGroup], fni, (const SolveState*)ctx->solve);
                probe_end(ctx, DALM_K_HALF_S);
                probe_begin(ctx, DALM_K_SPMV);
                hipLaunchKernelGGL(k_bi_spmv4<1>, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A, SS[k % kXGroup], ctx->t,
                                   ctx->s, ctx->partials, ctx->solve);
                probe_end(ctx, DALM_K_SPMV);
                if (int rc2 = launch_finalize(ctx, FIN_OMEGA, grid4, ctx->solve, nullptr)) return rc2;
                probe_begin(ctx, DALM_K_UPDATE);
                hipLaunchKernelGGL(k_bi_xr4<0>, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, x, pring(k),
                                   (const double*)ctx->s, (const double*)ctx->t, ctx->r, (const double*)ctx->r0,
                                   ctx->partials, ctx->solve);
                probe_end(ctx, DALM_K_UPDATE);
                if (int rc3 = launch_finalize(ctx, FIN_BI_XR, grid4, ctx->solve, nullptr)) return rc3;
                if (k % kXGroup == kXGroup - 1)
                    hipLaunchKernelGGL(k_x_group4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, x, pring(k),
                                       it_start + k + 1, (const SolveState*)ctx->solve);
                continue;
            }
            probe_begin(ctx, DALM_K_PUPDATE);
            if (vec4)
                hipLaunchKernelGGL(k_bi_p4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A, ctx->r, ctx->kp, ctx->q,
                                   ctx->pp, pc_first, (const SolveState*)ctx->solve);
            else
                hipLaunchKernelGGL(k_bi_p, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, ctx->r, ctx->kp,
                                   ctx->q, ctx->pp, pc_first, (const SolveState*)ctx->solve);
            probe_end(ctx, DALM_K_PUPDATE);
            double* ppz = precondition(ctx->pp, ctx->tmp, ctx->kp);
            probe_begin(ctx, DALM_K_SPMV);
            if (vec4)
                hipLaunchKernelGGL(k_bi_spmv4<0>, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A, ppz, ctx->q,
                                   ctx->r0, ctx->partials, ctx->solve);
            else
                hipLaunchKernelGGL(k_bi_spmv<0>, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, ppz,
                                   ctx->q, ctx->r0, ctx->partials, ctx->counter, ctx->solve);
            probe_end(ctx, DALM_K_SPMV);
            if (int rc1 = launch_finalize(ctx, FIN_ALPHA, nb, ctx->solve, nullptr)) return rc1;
            probe_begin(ctx, DALM_K_SUPDATE);
            if (vec4)
                hipLaunchKernelGGL(k_bi_s4, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A, ctx->r, ctx->q, ctx->s,
                                   ctx->ss, pc_first, (const SolveState*)ctx->solve);
            else
                hipLaunchKernelGGL(k_bi_s, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, ctx->r, ctx->q,
                                   ctx->s, ctx->ss, pc_first, (const SolveState*)ctx->solve);
            probe_end(ctx, DALM_K_SUPDATE);
            double* ssz = precondition(ctx->ss, ctx->tmp2, ctx->s);
            probe_begin(ctx, DALM_K_SPMV);
            if (vec4)
                hipLaunchKernelGGL(k_bi_spmv4<1>, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, A, ssz, ctx->t,
                                   ctx->s, ctx->partials, ctx->solve);
            else
                hipLaunchKernelGGL(k_bi_spmv<1>, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, A, ssz,
                                   ctx->t, ctx->s, ctx->partials, ctx->counter, ctx->solve);
            probe_end(ctx, DALM_K_SPMV);
            if (int rc2 = launch_finalize(ctx, FIN_OMEGA, nb, ctx->solve, nullptr)) return rc2;
            probe_begin(ctx, DALM_K_UPDATE);
            if (vec4) {
                PRing z{};
                z.pp[kXGroup - 1] = ppz;
                z.ss[kXGroup - 1] = ssz;
                hipLaunchKernelGGL(k_bi_xr4<1>, dim3(grid4), dim3(BLOCK), 0, ctx->stream, d, x, z,
                                   (const double*)ctx->s, (const double*)ctx->t, ctx->r, (const double*)ctx->r0,
                                   ctx->partials, ctx->solve);
            }
            else
                hipLaunchKernelGGL(k_bi_xr, dim3(ctx->sgrid), dim3(BLOCK), 0, ctx->stream, d, x, ppz, ssz,
                                   ctx->s, ctx->t, ctx->r, ctx->r0, ctx->partials, ctx->counter, ctx->solve);
            probe_end(ctx, DALM_K_UPDATE);
            if (int rc3 = launch_finalize(ctx, FIN_BI_XR, nb, ctx->solve, nullptr)) return rc3;
        }
        LAUNCH_CHECK();
        launched += chunk;
        rc = read_state(ctx);
        if (rc) return rc;
    }
    probe_solve_end(ctx, ctx->h_solve->it);
    const long long done_it = ctx->h_solve->it - it_start;
    const long long last = done_it - 1;
    if (fused && done_it > 0 && last % kXGroup != kXGroup - 1) {
        // the last iterations' x updates are outstanding: ring places [first, kXGroup)
        // hold them, with alpha and omega in the same places of alpha_h, omega_h (the
        // last place is iteration `last`, whose alpha FIN_BI_XR shifted in last)
        XFlush64 f{};
        f.first = kXGroup - 1 - (int)(last % kXGroup);
        for (int q = f.first; q < kXGroup; ++q) {
            f.alpha[q] = ctx->h_solve->alpha_h[q];
            f.omega[q] = ctx->h_solve->omega_h[q];
        }
        hipLaunchKernelGGL(k_x_flush, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, d, x, pring(last), f);
        LAUNCH_CHECK();
    }
    if (fused && (done_it & 1)) {
        // after an odd number of iterations the current p and q are the partners
        std::swap(ctx->kp, ctx->tmp);
        std::swap(ctx->q, ctx->tmp2);
    }
    *it = ctx->h_solve->it;
    *err = ctx->h_solve->err;
    ctx->prev_it[slot] = *it;
    return ctx->h_solve->status;
}

int run_bicgstab(dalm_ctx* ctx, int pc_kind, double pc_omega, int pc_it, long long max_it, double tol,
                 long long* it, double* err) {
    return run_bicgstab_on(ctx, ctx->p, ctx->b, 1, pc_kind, pc_omega, pc_it, max_it, tol, it, err);
}

int zero_device(dalm_ctx* ctx, double* a, long long n) {
    hipLaunchKernelGGL(k_zero, dim3(n > 256 * 4096 ? 4096 : (int)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, n, a);
    LAUNCH_CHECK();
    return DALM_OK;
}
