// star_walk.h -- column walk with an LDS-staged plane for the radius-2 star stencil
// of the predictor (MUSCL convection reads f(+-1), f(+-2) along each axis).
//
// A block of RJ wavefronts owns a TK (k) x RJ (j) tile and walks `ni` consecutive
// interior planes along i.  Each thread keeps its own (j, k) line of the three
// velocity components for planes i-2 .. i+2 in registers (the i neighbours; plane
// i+3 is requested while plane i is computed).  The centre plane of the tile, with
// two halo rows on either side along j and two halo cells on either side along k,
// goes through LDS: every thread writes its own centre value, the halo cells are
// loaded from memory, and after one barrier every thread reads its j and k
// neighbours there.  Two LDS planes alternate, so one barrier per plane suffices.
// Each value of the three components is then read from memory once per unit, plus
// 4 planes at the unit's ends and the halo ring, instead of up to 13 times through
// the caches.
//
// Units are numbered k tile fastest, then j tile, and each XCD receives a
// contiguous range (xcd_unit): the tiles that run at the same time on one XCD are
// k and j neighbours, so the halo cells of one (whole 128-byte lines fetched for a
// few cells) are the edge cells of another, served by the XCD's L2.
//
// The per-cell arithmetic is the caller's functor, given the same values a direct
// kernel would load, so results are bit-identical to it.
#pragma once

#include "common.h"

template <int RJ>
struct StarShape {
    static constexpr int W = TK + 4;                // LDS row: 64 k and two halo cells per side
    static constexpr int H = RJ + 4;                // LDS rows: RJ rows and two halo rows per side
    static constexpr int threads = TK * RJ;
    static constexpr int halo = 4 * W + 4 * RJ;     // halo cells of one plane
    static constexpr int per = (halo + threads - 1) / threads;  // halo cells per thread at most
};

template <typename T, int RJ>
struct StarLds {
    T v[2][3][StarShape<RJ>::H][StarShape<RJ>::W];
};

// the j and k neighbours of a thread's cell in the current LDS plane:
// y(q, s) = component q at j-2+s, z(q, s) at k-2+s
template <typename T, int RJ>
struct StarView {
    const T (*p)[StarShape<RJ>::H][StarShape<RJ>::W];
    int row, lane;
    __device__ __forceinline__ T y(int q, int s) const { return p[q][row + s][lane + 2]; }
    __device__ __forceinline__ T z(int q, int s) const { return p[q][row + 2][lane + s]; }
};

// planes per unit of the star walk (32, 16 or 8), or 0 where the direct kernel is
// the better choice: fewer units than compute units at 8 planes each (a unit adds
// 4 planes of its own lines at its ends, so shorter units would re-read too much)
template <int RJ>
inline int star_planes(const Dims& d) {
    const long long cols = (long long)((d.ny + RJ - 1) / RJ) * d.tk;
    for (int ni = 32; ni >= 8; ni >>= 1) {
        if (ni > d.nx) continue;
        if (cols * ((d.nx + ni - 1) / ni) >= 256) return ni;
    }
    return 0;
}
template <int RJ>
inline int star_grid(const Dims& d, int ni) {
    return ((d.ny + RJ - 1) / RJ) * d.tk * ((d.nx + ni - 1) / ni);
}

// f(i, j, k, c, x, v): x[q][0..4] = component q at i-2 .. i+2 (same j, k), v its j
// and k neighbours (StarView); called for the interior cells of the unit's owned
// planes.
template <int RJ, typename T, typename F>
__device__ __forceinline__ void star_walk(const Dims& d, const T* __restrict__ a0, const T* __restrict__ a1,
                                          const T* __restrict__ a2, int ni, StarLds<T, RJ>& sh, const F& f) {
    using S = StarShape<RJ>;
    const int lane = threadIdx.x & 63;
    const int row = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = d.gc;
    const int tjs = (d.ny + RJ - 1) / RJ;
    const int u = xcd_unit((int)blockIdx.x, (int)gridDim.x);
    const int kt = u % d.tk, rest = u / d.tk;
    const int jt = rest % tjs, ic = rest / tjs;
    const int own0 = g + ic * ni;
    int i1 = own0 + ni;
    if (i1 > g + d.nx) i1 = g + d.nx;
    const int i0 = i1 - ni;
    const int j0 = g + jt * RJ, k0 = g + kt * TK;
    const int j = j0 + row, k = k0 + lane;
    const bool inside = j < g + d.ny && k < g + d.nz;
    const bool live = j < d.n1 && k < d.n2;  // inside the padded box: may be loaded
    const T* src[3] = {a0, a1, a2};
    // this thread's halo cells of the LDS plane: cell number e = threadIdx.x + m *
    // threads; rows -2, -1, RJ, RJ + 1 (every column), then columns -2, -1, TK,
    // TK + 1 of the tile's rows
    int hr[S::per], hc[S::per], hl[S::per];
    bool hlive[S::per];
#pragma unroll
    for (int m = 0; m < S::per; ++m) {
        const int e = (int)threadIdx.x + m * S::threads;
        hr[m] = -1;
        hc[m] = 0;
        if (e < 4 * S::W) {
            const int rr = e / S::W;
            hr[m] = rr < 2 ? rr : RJ + rr;
            hc[m] = e % S::W;
        } else if (e < S::halo) {
            const int x = e - 4 * S::W, s = x & 3;
            hr[m] = 2 + (x >> 2);
            hc[m] = s < 2 ? s : TK + s;
        }
        const int hj = j0 + hr[m] - 2, hk = k0 + hc[m] - 2;
        hlive[m] = hr[m] >= 0 && hj < d.n1 && hk < d.n2;
        hl[m] = hlive[m] ? d.at(0, hj, hk) : 0;  // plane-0 offset of the halo line
    }
    const int cl = live ? d.at(0, j, k) : 0;      // plane-0 offset of the own line

    T w[3][5], pf[3], hcur[3][S::per], hpf[3][S::per];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int s = 0; s < 5; ++s) w[q][s] = live ? src[q][cl + (i0 - 2 + s) * d.si] : T(0);
#pragma unroll
        for (int m = 0; m < S::per; ++m) hcur[q][m] = hlive[m] ? src[q][hl[m] + i0 * d.si] : T(0);
    }
    for (int i = i0; i < i1; ++i) {
        const int b = (i - i0) & 1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            sh.v[b][q][row + 2][lane + 2] = w[q][2];
#pragma unroll
            for (int m = 0; m < S::per; ++m)
                if (hr[m] >= 0) sh.v[b][q][hr[m]][hc[m]] = hcur[q][m];
        }
        __syncthreads();
        // next plane's values, in flight while this one is computed
        const bool more = i + 1 < i1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pf[q] = (live && more) ? src[q][cl + (i + 3) * d.si] : T(0);
#pragma unroll
            for (int m = 0; m < S::per; ++m)
                hpf[q][m] = (hlive[m] && more) ? src[q][hl[m] + (i + 1) * d.si] : T(0);
        }
        if (inside && i >= own0) {
            const StarView<T, RJ> v{sh.v[b], row, lane};
            f(i, j, k, d.at(i, j, k), w, v);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
#pragma unroll
            for (int s = 0; s < 4; ++s) w[q][s] = w[q][s + 1];
            w[q][4] = pf[q];
#pragma unroll
            for (int m = 0; m < S::per; ++m) hcur[q][m] = hpf[q][m];
        }
    }
}
