// stencil.h -- the matrix-free 7-point Poisson operator on one (j, k) column and
// the software-pipelined walk along i shared by the column-streaming kernels.
//
// A column unit (col_unit) owns one (j, k) line per thread and walks ni planes
// along i, keeping x[i-1], x[i], x[i+1] in registers.  Each plane needs the j and
// k neighbours of the streamed vector plus the values of the other vectors at the
// cell; walk() requests all of plane i+1's values (and x[i+2]) before plane i is
// processed, so every wavefront keeps two planes of loads in flight instead of
// waiting on each plane's loads in turn.
//
// The operator: a_c x_c + a_e x_e + a_w x_w + a_n x_n + a_s x_s + a_t x_t + a_b x_b,
// left to right (src/mv.h:87), with the diagonal rebuilt per cell in build_A's
// accumulation order (src/eq.h:701-798); the j and k coefficients of the column are
// loaded once.  OpD is the FP64 operator (exact: identical operands in identical
// order to the reference), OpF its FP32 pieces for the precision form's sweeps.
#pragma once
#include "ctx.h"

// j- and k-neighbours of one cell of a vector
template <typename T>
struct Nb4 {
    T n, s, t, b;
};
template <typename T>
__device__ __forceinline__ Nb4<T> nb4(const Dims& d, const T* x, int c) {
    return Nb4<T>{x[c + d.sj], x[c - d.sj], x[c + 1], x[c - 1]};
}

struct OpD {
    double an, as, at, ab, anu, asu, atu, abu;
    __device__ __forceinline__ OpD(const PCoef& A, int j, int k)
        : an(A.an[j]), as(A.as[j]), at(A.at[k]), ab(A.ab[k]), anu(A.anu[j]), asu(A.asu[j]), atu(A.atu[k]),
          abu(A.abu[k]) {}
    __device__ __forceinline__ double diag(const PCoef& A, int i) const {
        return ((((A.acx[i] - anu) - asu) - atu) - abu) / A.max_diag;
    }
    template <typename V>
    __device__ __forceinline__ double apply(const PCoef& A, int i, double ac, double x0, double xp, double xm,
                                            const Nb4<V>& q) const {
        return ac * x0 + A.ae[i] * xp + A.aw[i] * xm + an * (double)q.n + as * (double)q.s + at * (double)q.t +
               ab * (double)q.b;
    }
};

struct OpF {
    float an, as, at, ab, anu, asu, atu, abu;
    __device__ __forceinline__ OpF(const PCoefF& A, int j, int k)
        : an(A.an[j]), as(A.as[j]), at(A.at[k]), ab(A.ab[k]), anu(A.anu[j]), asu(A.asu[j]), atu(A.atu[k]),
          abu(A.abu[k]) {}
    __device__ __forceinline__ float diag(const PCoefF& A, int i) const {
        return (((A.acx[i] - anu) - asu) - atu) - abu;
    }
    __device__ __forceinline__ float apply(const PCoefF& A, int i, float ac, float x0, float xp, float xm,
                                           const Nb4<float>& q) const {
        return ac * x0 + A.ae[i] * xp + A.aw[i] * xm + an * q.n + as * q.s + at * q.t + ab * q.b;
    }
};

struct NoAux {};

// Walk the column of cl along i over the streamed vector x.  load(c) returns the
// other vectors' values at cell c (an Aux); f(i, c, x0, xp, xm, q, aux) processes
// plane i, for the planes the unit owns.  Every value of plane i+1 is requested
// before plane i is processed.  With the full kMaxPlanes planes per unit the walk
// is straight-line code (unrolled, no per-plane branches around the loads), so
// the compiler's wait counts leave the next plane's loads in flight; the look-ahead
// of the last plane re-reads the current plane (cache hits) instead of branching.
template <typename T, typename Aux, typename Load, typename F>
__device__ __forceinline__ void walk(const Dims& d, const Col& cl, const T* x, Load&& load, F&& f) {
    int c = d.at(cl.i0, cl.j, cl.k);
    T xm = x[c - d.si], x0 = x[c], xp = x[c + d.si];
    Nb4<T> q0 = nb4(d, x, c);
    Aux a0 = load(c);
    if (d.ni == kMaxPlanes) {
#pragma unroll
        for (int u = 0; u < kMaxPlanes; ++u) {
            const int nxt = u + 1 < kMaxPlanes ? d.si : 0;
            const T xpp = x[c + d.si + nxt];
            const Nb4<T> q1 = nb4(d, x, c + nxt);
            const Aux a1 = load(c + nxt);
            if (cl.i0 + u >= cl.own0) f(cl.i0 + u, c, x0, xp, xm, q0, a0);
            xm = x0;
            x0 = xp;
            xp = xpp;
            q0 = q1;
            a0 = a1;
            c += d.si;
        }
    } else {
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            const int nxt = i + 1 < cl.i1 ? d.si : 0;
            const T xpp = x[c + d.si + nxt];
            const Nb4<T> q1 = nb4(d, x, c + nxt);
            const Aux a1 = load(c + nxt);
            if (i >= cl.own0) f(i, c, x0, xp, xm, q0, a0);
            xm = x0;
            x0 = xp;
            xp = xpp;
            q0 = q1;
            a0 = a1;
        }
    }
}

// Pointwise kernels in the column layout: load(c) returns every value cell c
// needs, f(i, c, aux) processes it, for the planes the unit owns; plane i+1's
// values are requested before plane i is processed (the same straight-line walk).
template <typename Aux, typename Load, typename F>
__device__ __forceinline__ void walk_points(const Dims& d, const Col& cl, Load&& load, F&& f) {
    int c = d.at(cl.i0, cl.j, cl.k);
    Aux a0 = load(c);
    if (d.ni == kMaxPlanes) {
#pragma unroll
        for (int u = 0; u < kMaxPlanes; ++u) {
            const Aux a1 = load(c + (u + 1 < kMaxPlanes ? d.si : 0));
            if (cl.i0 + u >= cl.own0) f(cl.i0 + u, c, a0);
            a0 = a1;
            c += d.si;
        }
    } else {
        for (int i = cl.i0; i < cl.i1; ++i, c += d.si) {
            const Aux a1 = load(c + (i + 1 < cl.i1 ? d.si : 0));
            if (i >= cl.own0) f(i, c, a0);
            a0 = a1;
        }
    }
}

// ---- four cells per lane (rows of at least 256 interior cells, nz a multiple of 4) ----
// A wavefront covers 256 consecutive k of one row; each vector moves with one
// 16-byte (FP32) or two 16-byte (FP64) accesses per lane and plane.  The last tile
// of a row may be partly empty (nz = 384: 64 + 32 lanes).  Launch with
// tj * vec4_tiles(d) * nch blocks.
constexpr int kVecK = 4;
DALM_HD int vec4_tiles(const Dims& d) { return (d.nz + TK * kVecK - 1) / (TK * kVecK); }
// the four-cell kernels apply: nz a multiple of 4 and at least three quarters of the
// tiles' lanes in use (nz = 256 k, 384, 512, ...)
DALM_HD bool vec4_applies(const Dims& d) { return d.nz % kVecK == 0 && 4 * d.nz >= 3 * TK * kVecK * vec4_tiles(d); }

// Unit of a four-per-lane kernel: a lane owns k0 .. k0+3 of row j over the unit's
// planes [i0, i1), of which it writes [own0, i1).  ok is false on a ragged j row and
// on the empty lanes of a row's last, partial tile.  Units are numbered j tile
// fastest and dealt to the XCDs in contiguous ranges, so the j-neighbour tiles
// whose rows a unit's stencil reads run on the same XCD and share its L2 (the
// four-cell SpMV moved 1.18x its algorithmic bytes with channel-fastest numbering).
struct Col4 {
    int i0, i1, own0, j, k0;
    int slot;  // index of this unit's partial results (see col_unit4)
    bool ok;
};
__device__ __forceinline__ Col4 col_unit4(const Dims& d) {
    const int tkv = vec4_tiles(d);
    const int u = xcd_unit((int)blockIdx.x, (int)gridDim.x);
    const int jt = u % d.tj, rest = u / d.tj;
    const int kt = rest % tkv, ic = rest / tkv;
    Col4 r;
    r.k0 = d.gc + (kt * TK + (int)(threadIdx.x % TK)) * kVecK;
    r.j = d.gc + jt * TJ + (int)(threadIdx.x / TK);
    r.own0 = d.gc + ic * d.ni;
    r.i1 = r.own0 + d.ni;
    if (r.i1 > d.gc + d.nx) r.i1 = d.gc + d.nx;
    r.i0 = r.i1 - d.ni;
    r.ok = r.j < d.gc + d.ny && r.k0 < d.gc + d.nz;
    // The partial-result slot does not follow the launch numbering: it is the block
    // that held this unit when units were numbered channel fastest and dealt out by
    // virtual_block, so the finalize adds the dot products in that fixed order and
    // the iterates are the same bit for bit as with that numbering.
    r.slot = virtual_block_index(ic + d.nch * (kt + tkv * jt), (int)gridDim.x);
    return r;
}

// Pointwise walk of a four-per-lane unit: load(c) returns the vectors' values at
// c .. c+3, f(i, c, values) processes plane i for the owned planes; plane i+1 is
// requested before plane i is processed.
template <typename Aux, typename Load, typename F>
__device__ __forceinline__ void walk_points4(const Dims& d, const Col4& cl, Load&& load, F&& f) {
    int c = d.at(cl.i0, cl.j, cl.k0);
    Aux a0 = load(c);
    auto step = [&](int i, int nxt) {
        const Aux a1 = load(c + nxt);
        if (i >= cl.own0) f(i, c, a0);
        a0 = a1;
        c += d.si;
    };
    if (d.ni == kMaxPlanes) {
#pragma unroll
        for (int w = 0; w < kMaxPlanes; ++w) step(cl.i0 + w, w + 1 < kMaxPlanes ? d.si : 0);
    } else {
        for (int i = cl.i0; i < cl.i1; ++i) step(i, i + 1 < cl.i1 ? d.si : 0);
    }
}
