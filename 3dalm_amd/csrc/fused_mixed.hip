// fused_mixed.hip -- the precision form's BiCGStab half-iterations as single kernels
// (rows of exactly 256 interior cells, three Jacobi sweeps as preconditioner).
//
// One half-iteration of the preconditioned BiCGStab (src/eq.h:355-385) is a
// pointwise vector update followed by the Jacobi preconditioner and the SpMV:
//
//   p-half   p = r + beta (p - omega q);  z1 = p / a_c;  z2 = J(z1; p);  z3 = J(z2; p);
//            q = A z3,  <r0, q>
//   s-half   s = r - alpha q;              z1 = s / a_c;  z2 = J(z1; s);  z3 = J(z2; s);
//            t = A z3,  <t, s>, <t, t>
//
// where J(z; v) = z + (v - A z) / a_c is one sweep (src/eq.h:6-38: the first sweep
// from zero is z1 = v / a_c).  Run as separate kernels these move 148 B per cell
// and iteration (every level is written to HBM and read back); fused, a block keeps
// the intermediate levels on chip and reads r, p, q, r0 and writes p, z3, q once:
// 28 B per cell for the p-half and 20 B for the s-half.
//
// Temporal blocking along i.  A block of 16 wavefronts owns 10 consecutive j rows
// (one wavefront per row, four consecutive k per lane: 256 k per row) of a chunk of
// `ni` planes, and computes the levels with a lag of one plane each: at step t,
// level 0 (the update) at plane t, the first stencil sweep at t-1, the second at
// t-2 and the SpMV at t-3.  Each stencil level needs its input on the neighbouring
// rows, so the block also computes 3 / 2 / 1 halo rows of levels 0 / 1 / 2 on
// either side (overlapped tiling: 16, 14, 12 and 10 rows per level); the i
// neighbours are the thread's own registers of the previous and next planes, the
// k neighbours its own four cells or the adjacent lanes' through LDS.  Each step
// exchanges the new plane of every stencil input through LDS (one row per
// wavefront; two copies alternate): read, compute, write, barrier.  A chunk starts 3 planes
// early and ends 3 planes late (the pipeline's fill and drain).
//
// The arithmetic per cell is that of the separate kernels (kg_p4 / kg_s4, kf_jacobi4,
// kg_spmv4): FP32 sweeps with FP32 coefficients, the SpMV in FP64 with the FP64
// operator and its result rounded once (its diagonal scaled by the reciprocal of
// max|a_c| instead of divided by it), dot products in FP64 of the stored values.
// Ghost cells of the preconditioner iterates are zero, as in the reference
// (fill_array(pp, 0) before the sweeps, which write the interior only).
//
// p and q are read at halo cells owned by neighbouring blocks, so the p-half reads
// the previous p and q from one buffer of a pair and writes the new ones to the
// other (the host alternates the pair every iteration).
//
// The end of an iteration, x += alpha z_p + omega z_s and r = s - omega t, has no
// pass of its own.  The s-half also returns <r0, s>, <r0, t> and |s|^2, from which
// the finalize expands |r|^2 and <r0, r> (FIN_MX_OMEGA); both updates are then
// pending (SolveState.pending) and the next p-half applies them: it forms r from s
// and t where it needs r (every row it computes) and stores r on its owned cells.
// Neither half updates the FP64 pressure x: the finalize records each iteration's
// alpha and omega (SolveState.xdef counts the outstanding updates), z_p and z_s rotate
// through kXGroupF buffer pairs with the iteration, and kg_apply_pending4 applies the
// outstanding updates, in turn, once kXGroupF have gathered or ahead of a residual
// replacement; the host applies the rest after the last iteration of a pass
// (kg_flush_pending4).  Each update is the same two roundings in the same order as one
// update per iteration, so x is the same to the bit.  Per iteration this moves 70 B
// per cell (the p-half 36, the s-half 24, the grouped x update (16 + 8 kXGroupF) /
// kXGroupF) instead of 88.
#include "ctx.h"

namespace {

constexpr int kRows = 16;           // rows (wavefronts) of a block
constexpr int kOwn = kRows - 6;     // rows a block owns
constexpr int kLag = 3;             // planes from level 0 to the SpMV
constexpr int kPad = 4;             // zero cells on either side of an LDS row
constexpr int kRowLen = 256 + 2 * kPad;

typedef float f4 __attribute__((ext_vector_type(4), aligned(16)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ f4 ld4(const float* p) { return *(const f4u*)p; }
// nontemporal: a stored vector is next read a kernel later, past the last-level cache
__device__ __forceinline__ void st4(float* p, f4 v) { __builtin_nontemporal_store((f4u)v, (f4u*)p); }

struct Lds {
    // the new plane of each stencil input (z1, z2, z3), in two alternating copies: a
    // step reads the copy the previous step wrote and writes the other, so one
    // barrier per step orders both (101 KB of the CU's 160 KB)
    float lv[2][3][kRows][kRowLen];
    double atd[256], abd[256], atud[256], abud[256];  // FP64 k coefficients (SpMV)
    f4 atf[64], abf[64], atuf[64], abuf[64];         // FP32 k coefficients per lane (sweeps)
};

// j and k neighbours of one level at this thread's four cells, from LDS
struct Nb {
    f4 n, s;
    float kb, kf;
};
__device__ __forceinline__ Nb nbrs(const float (*lv)[kRowLen], int row, int lane) {
    const int o = kPad + 4 * lane;
    Nb b;
    b.n = *(const f4*)&lv[row + 1][o];
    b.s = *(const f4*)&lv[row - 1][o];
    b.kb = lv[row][o - 1];
    b.kf = lv[row][o + 4];
    return b;
}

// 16-wavefront block sum in a fixed order (result in thread 0)
template <int NV>
__device__ __forceinline__ void block_sum16(double (&v)[NV]) {
    __shared__ double red[NV][kRows];
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
            for (int w = 1; w < kRows; ++w) s += red[q][w];
            v[q] = s;
        }
    }
}

// WHICH 0: the p-half (vin = previous p, qin = previous q; vout = new p, zout = new
//          z_p, yout = new q; r read, or formed from s_in, t_in and stored; dot
//          <r0, q>)
// WHICH 1: the s-half (r, qin = q; vout = s, zout = z_s, yout = t; dots <t, s>,
//          <t, t>, <r0, s>, <r0, t>, |s|^2)
struct HalfArgs {
    float* r;
    const float *s_in, *t_in;  // p-half: the previous iteration's s, t
    const float *vin, *qin;
    float *vout, *zout, *yout;
    const float* r0;
    double* partials;
};

template <int WHICH>
__global__ __launch_bounds__(64 * kRows) void kg_half4(Dims d, PCoefF Af, PCoef A, HalfArgs h, int ni,
                                                       const SolveState* st) {
    if (st->done) return;
    float* __restrict__ r = h.r;
    const float* __restrict__ vin = h.vin;
    const float* __restrict__ qin = h.qin;
    float* __restrict__ vout = h.vout;
    float* __restrict__ zout = h.zout;
    float* __restrict__ yout = h.yout;
    const float* __restrict__ r0 = h.r0;
    __shared__ Lds sh;
    // the row is uniform in a wavefront: its coefficients and tests live in scalars
    const int lane = threadIdx.x & 63, row = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = d.gc;
    // unit: j tile jt (rows jt*kOwn .. +kOwn), plane chunk ic; the last chunk of a
    // column starts inside its predecessor (it computes ni planes, owns from own0).
    // Units are numbered j tile fastest, and each XCD runs a contiguous range of
    // them: the halo rows that neighbouring tiles of a chunk both read are fetched
    // from HBM once and served to the second from the XCD's L2.
    const int tjf = (d.ny + kOwn - 1) / kOwn;
    const int u = xcd_unit((int)blockIdx.x, (int)gridDim.x);
    const int jt = u % tjf, ic = u / tjf;
    const int own0 = g + ic * ni;
    int i1 = own0 + ni;
    if (i1 > g + d.nx) i1 = g + d.nx;
    const int i0 = i1 - ni;
    const int j = g + jt * kOwn + row - kLag;
    const bool jin = j >= g && j < g + d.ny;         // interior row: holds values
    const bool owned = jin && row >= kLag && row < kRows - kLag;
    const bool l1 = jin && row >= 1 && row < kRows - 1;
    const bool l2 = jin && row >= 2 && row < kRows - 2;
    const int k0 = g + 4 * lane;
    const double beta = st->beta, om = st->omega, alpha = st->alpha;
    const bool first = st->it == 0;
    // p-half: the previous iteration's r is still to be formed (its omega is that of
    // st: FIN_ALPHA runs after this kernel)
    const bool pend = WHICH == 0 && st->pending;
    // the SpMV's diagonal is scaled by the reciprocal of max|a_c| (within an ulp of the
    // division: far below the FP32 rounding of the product it feeds)
    const double inv_md = 1.0 / A.max_diag;

    // coefficients: j (this row), k (this lane's four cells), FP32 and FP64
    const int jc = jin ? j : g;
    const float an = Af.an[jc], as = Af.as[jc], anu = Af.anu[jc], asu = Af.asu[jc];
    const double and_ = A.an[jc], asd = A.as[jc], anud = A.anu[jc], asud = A.asu[jc];
    for (int t = threadIdx.x; t < 256; t += 64 * kRows) {
        sh.atd[t] = A.at[g + t];
        sh.abd[t] = A.ab[g + t];
        sh.atud[t] = A.atu[g + t];
        sh.abud[t] = A.abu[g + t];
        sh.atf[t >> 2][t & 3] = Af.at[g + t];
        sh.abf[t >> 2][t & 3] = Af.ab[g + t];
        sh.atuf[t >> 2][t & 3] = Af.atu[g + t];
        sh.abuf[t >> 2][t & 3] = Af.abu[g + t];
    }
    for (int t = threadIdx.x; t < 6 * kRows * kRowLen; t += 64 * kRows) (&sh.lv[0][0][0][0])[t] = 0.f;
    __syncthreads();

    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    // pipeline registers: level values at the planes they are needed
    f4 z1a = zero, z1b = zero;  // z1 at t-1, t-2
    f4 va = zero, vb = zero, vc = zero;  // v (p or s) at t-1, t-2, t-3
    f4 z2a = zero, z2b = zero;  // z2 at t-2, t-3
    f4 z3a = zero, z3b = zero;  // z3 at t-3, t-4
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};

    // level-0 inputs of plane t (prefetched one step ahead); with a pending update
    // r is s - om t, formed from s (in .r) and t
    struct In {
        f4 r, v, q, t;
    };
    auto load0 = [&](int i) {
        In x{zero, zero, zero, zero};
        if (jin && i >= g && i < g + d.nx) {
            const int c = d.at(i, j, k0);
            x.r = ld4((pend ? h.s_in : r) + c);
            if (pend) x.t = ld4(h.t_in + c);
            x.q = ld4(qin + c);
            if (WHICH == 0) x.v = ld4(vin + c);
        }
        return x;
    };
    In nx0 = load0(i0 - kLag);
    const int rr = row == 0 ? 1 : (row == kRows - 1 ? kRows - 2 : row);  // neighbour rows stay in range
    f4 r0n = zero;  // r0 of the SpMV's plane

    for (int t = i0 - kLag; t < i1 + kLag; ++t) {
        const In in = nx0;
        nx0 = load0(t + 1);
        const int iy = t - kLag;  // SpMV plane
        const bool yplane = owned && iy >= own0 && iy < i1;
        f4 r0c = r0n;
        if (owned && iy + 1 >= own0 && iy + 1 < i1) r0n = ld4(r0 + d.at(iy + 1, j, k0));
        if (!(owned && iy >= own0 && iy < i1)) r0c = zero;
        // neighbours of the centre planes (written at the previous step, into the
        // other copy)
        const int cw = (t - i0 + kLag) & 1;
        const float (*prev)[kRows][kRowLen] = sh.lv[cw ^ 1];

        // this lane's k coefficients of the FP32 diagonal, read once for level 0 and both sweeps
        const f4 atu = sh.atuf[lane], abu = sh.abuf[lane];
        // level 0 at plane t
        f4 z1n = zero, vn = zero;
        if (jin && t >= g && t < g + d.nx) {
            const float acj = (Af.acx[t] - anu) - asu;  // the diagonal's (i, j) part, in its order
            const bool mine = owned && t >= own0 && t < i1;
            const int c = d.at(t, j, k0);
            f4 rv = in.r;
            if (pend) {
#pragma unroll
                for (int q = 0; q < 4; ++q) rv[q] = (float)((double)in.r[q] - om * (double)in.t[q]);
                if (mine) st4(r + c, rv);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float vf;
                if (WHICH == 0)
                    vf = first ? rv[q] : (float)((double)rv[q] + beta * ((double)in.v[q] - om * (double)in.q[q]));
                else
                    vf = (float)((double)rv[q] - alpha * (double)in.q[q]);
                vn[q] = vf;
                const float ac = (acj - atu[q]) - abu[q];
                z1n[q] = vf * __builtin_amdgcn_rcpf(ac);
            }
            if (mine) st4(vout + c, vn);
        }
        // a stencil sweep of level `x` (centre xa at plane ic, xb before, xn after)
        const f4 at = sh.atf[lane], ab = sh.abf[lane];
        auto sweep = [&](int ip, const f4& xb, const f4& xa, const f4& xn, const Nb& nb, const f4& rhs) {
            const float acj = (Af.acx[ip] - anu) - asu, ae = Af.ae[ip], aw = Af.aw[ip];
            f4 out;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float ac = (acj - atu[q]) - abu[q];
                const float tq = q + 1 < 4 ? xa[q + 1] : nb.kf;
                const float bq = q > 0 ? xa[q - 1] : nb.kb;
                const float ax = ac * xa[q] + ae * xn[q] + aw * xb[q] + an * nb.n[q] + as * nb.s[q] + at[q] * tq +
                                 ab[q] * bq;
                out[q] = xa[q] + (rhs[q] - ax) * __builtin_amdgcn_rcpf(ac);
            }
            return out;
        };
        // level 1 at plane t-1, level 2 at plane t-2
        const int p1 = t - 1, p2 = t - 2;
        f4 z2n = zero, z3n = zero;
        if (l1 && p1 >= g && p1 < g + d.nx && p1 >= i0 - 2) z2n = sweep(p1, z1b, z1a, z1n, nbrs(prev[0], rr, lane), va);
        if (l2 && p2 >= g && p2 < g + d.nx && p2 >= i0 - 1) z3n = sweep(p2, z2b, z2a, z2n, nbrs(prev[1], rr, lane), vb);
        // SpMV at plane t-3 (FP64 operator, FP64 arithmetic, rounded once)
        if (yplane) {
            const Nb n3 = nbrs(prev[2], rr, lane);
            const double acj = (A.acx[iy] - anud) - asud, ae = A.ae[iy], aw = A.aw[iy];
            f4 y;
            const int lk = 4 * lane;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double ac = ((acj - sh.atud[lk + q]) - sh.abud[lk + q]) * inv_md;
                const double tq = q + 1 < 4 ? (double)z3a[q + 1] : (double)n3.kf;
                const double bq = q > 0 ? (double)z3a[q - 1] : (double)n3.kb;
                const double yv = ac * (double)z3a[q] + ae * (double)z3n[q] + aw * (double)z3b[q] +
                                  and_ * (double)n3.n[q] + asd * (double)n3.s[q] + sh.atd[lk + q] * tq +
                                  sh.abd[lk + q] * bq;
                const float yf = (float)yv;
                y[q] = yf;
                const double ys = (double)yf;
                if (WHICH == 0) {
                    acc[0] += (double)r0c[q] * ys;
                } else {
                    const double sv = (double)vc[q], rz = (double)r0c[q];
                    acc[0] += ys * sv;
                    acc[1] += ys * ys;
                    acc[2] += rz * sv;
                    acc[3] += rz * ys;
                    acc[4] += sv * sv;
                }
            }
            const int c = d.at(iy, j, k0);
            st4(yout + c, y);
            st4(zout + c, z3a);
        }
        // the new planes become the neighbours of the next step
        const int o = kPad + 4 * lane;
        *(f4*)&sh.lv[cw][0][row][o] = z1n;
        *(f4*)&sh.lv[cw][1][row][o] = z2n;
        *(f4*)&sh.lv[cw][2][row][o] = z3n;
        __syncthreads();
        z1b = z1a;
        z1a = z1n;
        vc = vb;
        vb = va;
        va = vn;
        z2b = z2a;
        z2a = z2n;
        z3b = z3a;
        z3a = z3n;
    }
    block_sum16<5>(acc);
    if (threadIdx.x == 0) {
        h.partials[blockIdx.x] = acc[0];
        if (WHICH == 1)
            for (int q = 1; q < 5; ++q) h.partials[q * gridDim.x + blockIdx.x] = acc[q];
    }
}

}  // namespace

// Rows of exactly 256 interior cells and three Jacobi sweeps: the fused path applies.
bool fused_half_applies(const dalm_ctx* ctx, int pc_kind, int pc_it) {
    return ctx->d.nz == 256 && pc_kind == DALM_PC_JACOBI && pc_it == 3 && ctx->d.ny >= kOwn;
}

int fused_planes(const dalm_ctx* ctx, int tiles) {
    const int nx = ctx->d.nx;
    int nch = ctx->ncu / (tiles > 0 ? tiles : 1);
    const int most = (nx + 15) / 16;
    if (nch > most) nch = most;
    if (nch < 1) nch = 1;
    return (nx + nch - 1) / nch;
}

// planes per unit and the grid of the fused half-iterations: one block per compute
// unit (its 16 wavefronts fill the CU's VGPRs), in a single round when the units
// allow it -- a second, partial round would leave most CUs idle
int fused_half_grid(const dalm_ctx* ctx, int* ni) {
    const Dims& d = ctx->d;
    const int tjf = (d.ny + kOwn - 1) / kOwn;
    const int n = fused_planes(ctx, tjf);
    *ni = n;
    return tjf * ((d.nx + n - 1) / n);
}

int launch_half_p(dalm_ctx* ctx, const float* p_in, const float* q_in, float* p_out, float* q_out, float* zp_out) {
    int ni;
    const int grid = fused_half_grid(ctx, &ni);
    HalfArgs h{};
    h.r = ctx->fr;
    h.s_in = ctx->fs;
    h.t_in = ctx->ft;
    h.vin = p_in;
    h.qin = q_in;
    h.vout = p_out;
    h.zout = zp_out;
    h.yout = q_out;
    h.r0 = ctx->fr0;
    h.partials = ctx->partials;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_half4<0>), dim3(grid), dim3(64 * kRows), 0, ctx->stream, ctx->d, ctx->pcf,
                       ctx->coef(), h, ni, (const SolveState*)ctx->solve);
    return hipGetLastError() == hipSuccess ? DALM_OK : DALM_HIP_ERROR;
}

int launch_half_s(dalm_ctx* ctx, const float* q, float* zs_out) {
    int ni;
    const int grid = fused_half_grid(ctx, &ni);
    HalfArgs h{};
    h.r = ctx->fr;
    h.qin = q;
    h.vout = ctx->fs;
    h.zout = zs_out;
    h.yout = ctx->ft;
    h.r0 = ctx->fr0;
    h.partials = ctx->partials;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kg_half4<1>), dim3(grid), dim3(64 * kRows), 0, ctx->stream, ctx->d, ctx->pcf,
                       ctx->coef(), h, ni, (const SolveState*)ctx->solve);
    return hipGetLastError() == hipSuccess ? DALM_OK : DALM_HIP_ERROR;
}
