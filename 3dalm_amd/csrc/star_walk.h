// star_walk.h -- column walk with an LDS-staged plane for the radius-2 star stencil
// of the predictor (MUSCL convection reads f(+-1), f(+-2) along each axis).
//
// A block of kSJ wavefronts owns a TK (k) x kSJ (j) tile and walks `ni` consecutive
// interior planes along i.  Each thread keeps its own (j, k) line of the three
// velocity components for planes i-2 .. i+2 in registers (the i neighbours; plane
// i+3 is requested while plane i is computed).  The centre plane of the tile, with
// two halo rows on either side along j and two halo cells on either side along k,
// goes through LDS: every thread writes its own centre value, the first threads
// also load the halo cells from memory, and after one barrier every thread reads
// its j and k neighbours there.  Two LDS planes alternate, so one barrier per plane
// suffices.  Each value of the three components is then read from memory once per
// unit, plus 4 planes at the unit's ends and the halo ring ((kSJ + 4) x (TK + 4)
// cells per kSJ x TK), instead of up to 13 times through the caches.
//
// Units are numbered j tile fastest and each XCD receives a contiguous range
// (xcd_unit): the tiles that run at the same time on one XCD are j neighbours, and
// the halo rows of one are the edge rows of the next, served by the XCD's L2.
//
// The per-cell arithmetic is the caller's functor, given the same values a direct
// kernel would load, so results are bit-identical to it.
#pragma once

#include "common.h"

constexpr int kSJ = 8;             // j rows of a tile (one wavefront each)
constexpr int kSW = TK + 4;        // LDS row: the tile's 64 k and two halo cells per side
constexpr int kSH = kSJ + 4;       // LDS rows: the tile's rows and two halo rows per side
constexpr int kSBlock = TK * kSJ;  // threads of a block
constexpr int kSHalo = 4 * kSW + 4 * kSJ;  // halo cells of one plane (<= kSBlock)
static_assert(kSHalo <= kSBlock, "one halo cell per thread at most");

template <typename T>
struct StarLds {
    T v[2][3][kSH][kSW];
};

// the j and k neighbours of a thread's cell in the current LDS plane:
// y(q, s) = component q at j-2+s, z(q, s) at k-2+s
template <typename T>
struct StarView {
    const T (*p)[kSH][kSW];
    int row, lane;
    __device__ __forceinline__ T y(int q, int s) const { return p[q][row + s][lane + 2]; }
    __device__ __forceinline__ T z(int q, int s) const { return p[q][row + 2][lane + s]; }
};

// planes per unit of the star walk (32, 16 or 8), or 0 where the direct kernel is
// the better choice: fewer units than compute units at 8 planes each (a unit adds
// 4 planes of its own lines at its ends, so shorter units would re-read too much)
inline int star_planes(const Dims& d) {
    const long long cols = (long long)((d.ny + kSJ - 1) / kSJ) * d.tk;
    for (int ni = 32; ni >= 8; ni >>= 1) {
        if (ni > d.nx) continue;
        if (cols * ((d.nx + ni - 1) / ni) >= 256) return ni;
    }
    return 0;
}
inline int star_grid(const Dims& d, int ni) {
    return ((d.ny + kSJ - 1) / kSJ) * d.tk * ((d.nx + ni - 1) / ni);
}

// f(i, j, k, c, x, v): x[q][0..4] = component q at i-2 .. i+2 (same j, k), v its j
// and k neighbours (StarView); called for the interior cells of the unit's owned
// planes.
template <typename T, typename F>
__device__ __forceinline__ void star_walk(const Dims& d, const T* __restrict__ a0, const T* __restrict__ a1,
                                          const T* __restrict__ a2, int ni, StarLds<T>& sh, const F& f) {
    const int lane = threadIdx.x & 63;
    const int row = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = d.gc;
    const int tjs = (d.ny + kSJ - 1) / kSJ;
    const int u = xcd_unit((int)blockIdx.x, (int)gridDim.x);
    const int jt = u % tjs, rest = u / tjs;
    const int kt = rest % d.tk, ic = rest / d.tk;
    const int own0 = g + ic * ni;
    int i1 = own0 + ni;
    if (i1 > g + d.nx) i1 = g + d.nx;
    const int i0 = i1 - ni;
    const int j0 = g + jt * kSJ, k0 = g + kt * TK;
    const int j = j0 + row, k = k0 + lane;
    const bool inside = j < g + d.ny && k < g + d.nz;
    const bool live = j < d.n1 && k < d.n2;  // inside the padded box: may be loaded
    const T* src[3] = {a0, a1, a2};
    // this thread's halo cell of the LDS plane, if any: rows -2, -1, kSJ, kSJ + 1
    // (every column), then columns -2, -1, TK, TK + 1 of the tile's rows
    int hr = -1, hc = 0;
    {
        const int t = (int)threadIdx.x;
        if (t < 4 * kSW) {
            const int rr = t / kSW;
            hr = rr < 2 ? rr : kSJ + rr;
            hc = t % kSW;
        } else if (t < kSHalo) {
            const int e = t - 4 * kSW, s = e & 3;
            hr = 2 + (e >> 2);
            hc = s < 2 ? s : TK + s;
        }
    }
    const int hj = j0 + hr - 2, hk = k0 + hc - 2;
    const bool hlive = hr >= 0 && hj < d.n1 && hk < d.n2;
    const int cl = d.at(0, j, k), hl = d.at(0, hj, hk);  // plane-0 offsets of the two lines

    T w[3][5], pf[3], hcur[3], hpf[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int s = 0; s < 5; ++s) w[q][s] = live ? src[q][cl + (i0 - 2 + s) * d.si] : T(0);
        hcur[q] = hlive ? src[q][hl + i0 * d.si] : T(0);
    }
    for (int i = i0; i < i1; ++i) {
        const int b = (i - i0) & 1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            sh.v[b][q][row + 2][lane + 2] = w[q][2];
            if (hr >= 0) sh.v[b][q][hr][hc] = hcur[q];
        }
        __syncthreads();
        // next plane's values, in flight while this one is computed
        const bool more = i + 1 < i1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pf[q] = (live && more) ? src[q][cl + (i + 3) * d.si] : T(0);
            hpf[q] = (hlive && more) ? src[q][hl + (i + 1) * d.si] : T(0);
        }
        if (inside && i >= own0) {
            const StarView<T> v{sh.v[b], row, lane};
            f(i, j, k, d.at(i, j, k), w, v);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
#pragma unroll
            for (int s = 0; s < 4; ++s) w[q][s] = w[q][s + 1];
            w[q][4] = pf[q];
            hcur[q] = hpf[q];
        }
    }
}
