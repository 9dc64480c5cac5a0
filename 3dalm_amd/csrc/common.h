// common.h -- shared device/host definitions for the 3dalm hot path on gfx950.
//
// Layout on the device (HBM): every field is stored structure-of-arrays, one FP64
// (or FP32 in the precision form) array per component, over the reference's padded
// box n[0] x n[1] x n[2] (gc ghost layers per side) with the reference's linear
// index i*n1*n2 + j*n2 + k (z fastest, src/util.h:10-13).  A wavefront covers 64
// consecutive k of one (i, j) row, so every stencil load is a coalesced 512-byte
// (FP64) row segment; the j and i neighbours of a block's tile are re-read from
// L2 / the 256 MB last-level cache rather than HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DALM_HD __host__ __device__ __forceinline__

// Block shape used by all cell kernels: 64 lanes along k (one wavefront row) x
// 4 rows along j = 256 threads = 4 wavefronts.
constexpr int TK = 64;
constexpr int TJ = 4;
constexpr int BLOCK = TK * TJ;

struct Dims {
    int n0, n1, n2;     // padded extents
    int gc;             // ghost layers
    int nx, ny, nz;     // interior counts
    int si, sj;         // strides of i and j (k stride is 1)
    int tk, tj;         // tiles along k and j: ceil(nz/TK), ceil(ny/TJ)
    int ntiles;         // nx * tj * tk  (one tile = 1 x TJ x TK interior cells)
    int ni;             // planes per column unit (column-streaming stencil kernels)
    int nch;            // column units along i: ceil(nx / ni)
    int nunits;         // tj * tk * nch
    DALM_HD int at(int i, int j, int k) const { return i * si + j * sj + k; }
};

// Grid-stride assignment of interior tiles to blocks.  Blocks are dispatched to
// the eight XCDs round-robin (block b -> XCD b % 8); the virtual block number
// below gives each XCD one contiguous range of tiles (i-major order), so halo
// planes shared by neighbouring tiles are mostly served by the same XCD's L2.
DALM_HD int virtual_block(int b, int nblocks) {
    if (nblocks % 8) return b;
    return (b & 7) * (nblocks >> 3) + (b >> 3);
}

// The same for any number of blocks: XCD x receives the contiguous unit range
// starting at x * (n / 8) + min(x, n % 8), one unit longer on the first n % 8 XCDs
// (the fused half-iterations launch one block per compute unit, rarely a multiple
// of eight).
DALM_HD int xcd_unit(int b, int n) {
    const int q = n >> 3, rem = n & 7, x = b & 7;
    return x * q + (x < rem ? x : rem) + (b >> 3);
}

// Inverse of virtual_block: the block that virtual_block maps to unit u.
DALM_HD int virtual_block_index(int u, int nblocks) {
    if (nblocks % 8) return u;
    const int q = nblocks >> 3;
    return (u % q) * 8 + u / q;
}

// Tile t -> base interior coordinates of this thread (may fall outside the
// interior along j or k on ragged edges; callers test `inside`).
struct Cell {
    int i, j, k;
    bool inside;
};
__device__ __forceinline__ Cell tile_cell(const Dims& d, int t) {
    int kt = t % d.tk;
    int r = t / d.tk;
    int jt = r % d.tj;
    int it = r / d.tj;
    Cell c;
    c.k = d.gc + kt * TK + (int)(threadIdx.x % TK);
    c.j = d.gc + jt * TJ + (int)(threadIdx.x / TK);
    c.i = d.gc + it;
    c.inside = (c.k < d.gc + d.nz) && (c.j < d.gc + d.ny);
    return c;
}

// Column unit of a stencil kernel that streams along i: one block owns a TK x TJ
// (k, j) tile and ni consecutive interior planes, so each thread walks its own
// (j, k) line and keeps x[i-1], x[i], x[i+1] in registers -- every value of the
// streamed vector is read once per unit (plus one halo plane on each end) instead
// of three times from the caches.  Units are numbered column-major over i so that
// the consecutive units a single XCD receives (virtual_block) share halo planes.
//
// Every unit walks exactly ni planes, [i0, i1): the last unit of a column starts
// inside its predecessor's range when ni does not divide nx, and owns (writes,
// accumulates) only the planes from own0 on.
constexpr int kMaxPlanes = 8;         // planes of the unrolled walk (units of other lengths loop)
constexpr int kMaxColumnPlanes = 32;  // planes per column unit at most
// Units are numbered j tile fastest, then k tile, then channel, and dealt to the
// XCDs in contiguous ranges, so the j-neighbour tiles whose rows a unit's stencil
// reads run on the same XCD and share its L2.  slot is the unit's index for its
// partial results: the block that held it when units were numbered channel fastest
// and dealt out by virtual_block, so the finalize adds the dot products in that
// fixed order whatever the launch numbering.
struct Col {
    int i0, i1, own0, j, k;
    int slot;
    bool ok;  // (j, k) inside the interior (ragged tiles have idle lanes)
};
__device__ __forceinline__ Col col_unit(const Dims& d) {
    const int u = xcd_unit((int)blockIdx.x, (int)gridDim.x);
    const int jt = u % d.tj, rest = u / d.tj;
    const int kt = rest % d.tk, ic = rest / d.tk;
    Col r;
    r.slot = virtual_block_index(ic + d.nch * (kt + d.tk * jt), (int)gridDim.x);
    r.k = d.gc + kt * TK + (int)(threadIdx.x % TK);
    r.j = d.gc + jt * TJ + (int)(threadIdx.x / TK);
    r.own0 = d.gc + ic * d.ni;
    r.i1 = r.own0 + d.ni;
    if (r.i1 > d.gc + d.nx) r.i1 = d.gc + d.nx;
    r.i0 = r.i1 - d.ni;
    r.ok = r.k < d.gc + d.nz && r.j < d.gc + d.ny;
    return r;
}

// Contiguous chunk of tiles for this block.
struct TileRange {
    int begin, end;
};
__device__ __forceinline__ TileRange block_tiles(const Dims& d) {
    int vb = virtual_block((int)blockIdx.x, (int)gridDim.x);
    int per = (d.ntiles + (int)gridDim.x - 1) / (int)gridDim.x;
    int b = vb * per;
    int e = b + per;
    if (e > d.ntiles) e = d.ntiles;
    if (b > d.ntiles) b = d.ntiles;
    TileRange r;
    r.begin = b;
    r.end = e;
    return r;
}

// ---------------------------------------------------------------------------
// Block reductions (wavefront shuffle, then LDS across the 4 wavefronts).
// Fixed trees: results are bitwise reproducible from run to run.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Sum NV values over the block; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[NV][BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q][wave] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = red[q][0];
            for (int w = 1; w < BLOCK / 64; ++w) s += red[q][w];
            v[q] = s;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ double block_max(double v) {
    __shared__ double red[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_max(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; ++w) v = red[w] > v ? red[w] : v;
    }
    __syncthreads();
    return v;
}

// Per-block partial results: thread 0 stores the block's NV values with plain
// stores; a separate one-block finalize kernel (reduce.hip) sums them in a fixed
// order after the producer kernel has ended.  No device-scope fence or atomic is
// needed inside the producer: with one L2 per XCD such a fence writes the L2 back,
// which made every reducing kernel several times slower than a plain sweep.
template <int NV>
__device__ __forceinline__ void store_partials(const double (&v)[NV], double* partials) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * gridDim.x + blockIdx.x] = v[q];
    }
}
// The same into slot `slot` of [0, gridDim.x) instead of the block's own index.
template <int NV>
__device__ __forceinline__ void store_partials(const double (&v)[NV], double* partials, int slot) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * gridDim.x + slot] = v[q];
    }
}

// Reduce partials[q*nb + b] over b in a fixed order (one block).  Result in all
// threads' out[] after the call.  Each thread requests its values in groups of
// eight independent loads (the partials were just written by other XCDs, so every
// load is a trip past the local L2) and sums them in a fixed tree.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* partials, int nb, double (&out)[NV]) {
    constexpr int U = 8;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = 0;
        for (int b0 = (int)threadIdx.x; b0 < nb; b0 += U * BLOCK) {
            double w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u * BLOCK;
                w[u] = b < nb ? partials[q * nb + b] : 0.0;
            }
            s += ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
        }
        v[q] = s;
    }
    block_sum<NV>(v);
    __shared__ double bc[NV];
    if (threadIdx.x == 0) {
        for (int q = 0; q < NV; ++q) bc[q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = bc[q];
}

__device__ __forceinline__ double reduce_partials_max(const double* partials, int nb) {
    double v = 0;
    for (int b = (int)threadIdx.x; b < nb; b += BLOCK) {
        double w = partials[b];
        v = w > v ? w : v;
    }
    v = block_max(v);
    __shared__ double bc;
    if (threadIdx.x == 0) bc = v;
    __syncthreads();
    return bc;
}
