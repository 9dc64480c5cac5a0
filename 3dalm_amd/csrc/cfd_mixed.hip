// cfd_mixed.hip -- stencil kernels of the precision form (DALM_MIXED).
//
// Storage: cell velocity U / Uold, eddy viscosity and ALM force in FP32; face
// fluxes JU, pressure p, right-hand side b and divergence in FP64.  The face
// fluxes stay FP64 because after projection div(JU) = dt * max_diag * (b - A p)
// (src/cfd.h:716-740 against src/eq.h:741-767): FP32 fluxes would leave a rounding
// divergence of ~2.4 u32 U / dx, above the 1e-6 target on fine grids.
//
// Arithmetic: FP32 with FMA contraction and precomputed reciprocal metrics (no
// divisions in the cell loops) for the predictor and nu_t; FP64 for everything
// that writes JU, b or div.  Same HBM-bound streaming structure as the FP64 form,
// with about half the bytes per cell (compulsory reads + writes per interior cell:
// predictor 64 B, flux + rhs 44 B, projection 56 B, post 40 B).
#include <math.h>

#include "ctx.h"
#include "star_walk.h"

namespace {

struct M1 {
    const double *x, *y, *z, *dx, *dy, *dz;
};
M1 mesh(const dalm_ctx* c) { return M1{c->x, c->y, c->z, c->dx, c->dy, c->dz}; }

__device__ __forceinline__ float sgnf(float v) { return (float)((v > 0.0f) - (v < 0.0f)); }

struct FluxPairF {
    float p, m;
};
// MUSCL kappa = 1/3, compression-limited minmod (src/cfd.h:37-60), FP32.
__device__ __forceinline__ FluxPairF muscl_faces_f(float vm2, float vm1, float v0, float vp1, float vp2,
                                                   float Jp, float Jm) {
    const float kap = 1.f / 3.f;
    const float bb = (3.f - kap) / (1.f - kap);
    const float q = 0.25f;
    const float cm = q * (1.f - kap), cp = q * (1.f + kap);
    float a1 = vm1 - vm2, a2 = v0 - vm1, a3 = vp1 - v0, a4 = vp2 - vp1;
    float t1 = sgnf(a1), t2 = sgnf(a2), t3 = sgnf(a3), t4 = sgnf(a4);
    float h6 = t4 * fmaxf(0.f, fminf(fabsf(a4), t4 * bb * a3));
    float h5 = t3 * fmaxf(0.f, fminf(fabsf(a3), t3 * bb * a4));
    float h4 = t3 * fmaxf(0.f, fminf(fabsf(a3), t3 * bb * a2));
    float h3 = t2 * fmaxf(0.f, fminf(fabsf(a2), t2 * bb * a3));
    float h2 = t2 * fmaxf(0.f, fminf(fabsf(a2), t2 * bb * a1));
    float h1 = t1 * fmaxf(0.f, fminf(fabsf(a1), t1 * bb * a2));
    float pr = vp1 - (cm * h6 + cp * h5);
    float pl = v0 + (cm * h3 + cp * h4);
    float mr = v0 - (cm * h4 + cp * h3);
    float ml = vm1 + (cm * h1 + cp * h2);
    FluxPairF r;
    r.p = 0.5f * (Jp * (pr + pl) - fabsf(Jp) * (pr - pl));
    r.m = 0.5f * (Jm * (mr + ml) - fabsf(Jm) * (mr - ml));
    return r;
}

// predictor (src/cfd.h:175-242) in FP32 with reciprocal metrics

// the cell's metrics, face fluxes (rounded to FP32) and viscosity
struct PredCellF {
    float Je, Jw, Jn, Js, Jt, Jb;
    float rdx, rdy, rdz, rvol, rxp, rxm, ryp, rym, rzp, rzm, nu;
};
__device__ __forceinline__ PredCellF pred_cell_f(const Dims& d, const MetricF& m, const double* __restrict__ J0,
                                                 const double* __restrict__ J1, const double* __restrict__ J2,
                                                 const float* __restrict__ nut, float inv_re, int i, int j, int k,
                                                 int c) {
    PredCellF P;
    P.Je = (float)J0[c], P.Jw = (float)J0[c - d.si], P.Jn = (float)J1[c], P.Js = (float)J1[c - d.sj];
    P.Jt = (float)J2[c], P.Jb = (float)J2[c - 1];
    P.rdx = m.rw[0][i], P.rdy = m.rw[1][j], P.rdz = m.rw[2][k];
    P.rvol = P.rdx * P.rdy * P.rdz;
    P.rxp = m.rp[0][i], P.rxm = m.rm[0][i], P.ryp = m.rp[1][j], P.rym = m.rm[1][j];
    P.rzp = m.rp[2][k], P.rzm = m.rm[2][k];
    P.nu = inv_re + nut[c];
    return P;
}
__device__ __forceinline__ float pred_value_f(const PredCellF& P, float vc, float xww, float xw, float xe, float xee,
                                              float yss, float ys, float yn, float ynn, float zbb, float zb, float zt,
                                              float ztt, float fq, float dt) {
    FluxPairF fx = muscl_faces_f(xww, xw, vc, xe, xee, P.Je, P.Jw);
    FluxPairF fy = muscl_faces_f(yss, ys, vc, yn, ynn, P.Jn, P.Js);
    FluxPairF fz = muscl_faces_f(zbb, zb, vc, zt, ztt, P.Jt, P.Jb);
    float conv = ((fx.p - fx.m) + (fy.p - fy.m) + (fz.p - fz.m)) * P.rvol;
    float diff = P.nu * (((xe - vc) * P.rxp - (vc - xw) * P.rxm) * P.rdx +
                         ((yn - vc) * P.ryp - (vc - ys) * P.rym) * P.rdy +
                         ((zt - vc) * P.rzp - (vc - zb) * P.rzm) * P.rdz);
    return vc + dt * (diff - conv + fq);
}

// direct form (grids too small for the walk)
__global__ __launch_bounds__(BLOCK) void km_predictor(
    Dims d, MetricF m, const float* __restrict__ o0, const float* __restrict__ o1,
    const float* __restrict__ o2, const double* __restrict__ J0, const double* __restrict__ J1,
    const double* __restrict__ J2, const float* __restrict__ nut, const float* __restrict__ f0,
    const float* __restrict__ f1, const float* __restrict__ f2, float* __restrict__ u0,
    float* __restrict__ u1, float* __restrict__ u2, float inv_re, float dt) {
    TileRange tr = block_tiles(d);
    const int si = d.si, sj = d.sj;
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int i = cc.i, j = cc.j, k = cc.k;
        const int c = d.at(i, j, k);
        const PredCellF P = pred_cell_f(d, m, J0, J1, J2, nut, inv_re, i, j, k, c);
        const float* src[3] = {o0, o1, o2};
        const float* frc[3] = {f0, f1, f2};
        float* dst[3] = {u0, u1, u2};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float* a = src[q];
            dst[q][c] = pred_value_f(P, a[c], a[c - 2 * si], a[c - si], a[c + si], a[c + 2 * si], a[c - 2 * sj],
                                     a[c - sj], a[c + sj], a[c + 2 * sj], a[c - 2], a[c - 1], a[c + 1], a[c + 2],
                                     frc[q][c], dt);
        }
    }
}

// column walk with the centre plane staged through LDS (star_walk.h), kPredRowsF
// wavefronts per block
constexpr int kPredRowsF = 8;
__global__ __launch_bounds__(TK * kPredRowsF) void km_predictor_star(
    Dims d, MetricF m, const float* __restrict__ o0, const float* __restrict__ o1,
    const float* __restrict__ o2, const double* __restrict__ J0, const double* __restrict__ J1,
    const double* __restrict__ J2, const float* __restrict__ nut, const float* __restrict__ f0,
    const float* __restrict__ f1, const float* __restrict__ f2, float* __restrict__ u0,
    float* __restrict__ u1, float* __restrict__ u2, float inv_re, float dt, int ni) {
    __shared__ StarLds<float, kPredRowsF> sh;
    star_walk<kPredRowsF>(d, o0, o1, o2, ni, sh,
              [&](int i, int j, int k, int c, const float (&x)[3][5], const StarView<float, kPredRowsF>& v) {
                  const PredCellF P = pred_cell_f(d, m, J0, J1, J2, nut, inv_re, i, j, k, c);
                  const float* frc[3] = {f0, f1, f2};
                  float* dst[3] = {u0, u1, u2};
#pragma unroll
                  for (int q = 0; q < 3; ++q)
                      dst[q][c] = pred_value_f(P, x[q][2], x[q][0], x[q][1], x[q][3], x[q][4], v.y(q, 0),
                                               v.y(q, 1), v.y(q, 3), v.y(q, 4), v.z(q, 0), v.z(q, 1), v.z(q, 3),
                                               v.z(q, 4), frc[q][c], dt);
              });
}

// Face fluxes (FP64) with their boundary values, and -- for interior cells -- the
// Poisson right-hand side from the six faces of the cell (the west/south/bottom
// faces are recomputed here rather than re-read).  src/cfd.h:394-598, src/bc.h:235-302.
// The x- face takes the constant inflow uin0, or the turbulent inflow's face value
// (face != nullptr, [ny][nz]).
__device__ __forceinline__ double face_x(const Dims& d, const M1& m, const float* u0, int i, int j, int k,
                                         double uin0, const double* face) {
    const int g = d.gc;
    const int c = d.at(i, j, k);
    double a = m.dy[j] * m.dz[k];
    if (i == g - 1 && j >= g && j < d.n1 - g && k >= g && k < d.n2 - g)
        return a * (face ? face[(j - g) * d.nz + (k - g)] : uin0);
    return 0.5 * ((double)u0[c] + (double)u0[c + d.si]) * a;
}
__device__ __forceinline__ double face_y(const Dims& d, const M1& m, const float* u1, int i, int j, int k) {
    const int g = d.gc;
    if ((j == g - 1 || j == d.n1 - g - 1) && i >= g && i < d.n0 - g && k >= g && k < d.n2 - g) return 0.0;
    const int c = d.at(i, j, k);
    return 0.5 * ((double)u1[c] + (double)u1[c + d.sj]) * (m.dx[i] * m.dz[k]);
}
__device__ __forceinline__ double face_z(const Dims& d, const M1& m, const float* u2, int i, int j, int k) {
    const int g = d.gc;
    if ((k == g - 1 || k == d.n2 - g - 1) && i >= g && i < d.n0 - g && j >= g && j < d.n1 - g) return 0.0;
    const int c = d.at(i, j, k);
    return 0.5 * ((double)u2[c] + (double)u2[c + 1]) * (m.dx[i] * m.dy[j]);
}

__global__ void km_flux_rhs(Dims d, M1 m, const float* __restrict__ u0, const float* __restrict__ u1,
                            const float* __restrict__ u2, double* __restrict__ J0, double* __restrict__ J1,
                            double* __restrict__ J2, double* __restrict__ b, double uin0, double inv_dts,
                            int rhs, const double* __restrict__ face) {
    long long n = (long long)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    const int g = d.gc;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int k = g - 1 + (int)(t % (d.nz + 1));
        long long r = t / (d.nz + 1);
        int j = g - 1 + (int)(r % (d.ny + 1));
        int i = g - 1 + (int)(r / (d.ny + 1));
        const int c = d.at(i, j, k);
        double fe = face_x(d, m, u0, i, j, k, uin0, face);
        double fn = face_y(d, m, u1, i, j, k);
        double ft = face_z(d, m, u2, i, j, k);
        J0[c] = fe;
        J1[c] = fn;
        J2[c] = ft;
        if (rhs && i >= g && j >= g && k >= g) {
            double fw = face_x(d, m, u0, i - 1, j, k, uin0, face);
            double fs = face_y(d, m, u1, i, j - 1, k);
            double fb = face_z(d, m, u2, i, j, k - 1);
            double vol = m.dx[i] * m.dy[j] * m.dz[k];
            b[c] = ((fe - fw) + (fn - fs) + (ft - fb)) / vol * inv_dts;
        }
    }
}

// Projection (src/cfd.h:600-783) + face BCs; U in FP32 updated from FP64 math.
__global__ void km_project(Dims d, M1 m, MetricF mf, const double* __restrict__ p, float* __restrict__ u0,
                           float* __restrict__ u1, float* __restrict__ u2, double* __restrict__ J0,
                           double* __restrict__ J1, double* __restrict__ J2, double dt, double uin0,
                           const double* __restrict__ face) {
    long long n = (long long)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    const int g = d.gc;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int k = g - 1 + (int)(t % (d.nz + 1));
        long long r = t / (d.nz + 1);
        int j = g - 1 + (int)(r % (d.ny + 1));
        int i = g - 1 + (int)(r / (d.ny + 1));
        const int c = d.at(i, j, k);
        const double pc = p[c], pe = p[c + d.si], pn = p[c + d.sj], pt = p[c + 1];
        const bool ii = i >= g, ji = j >= g, ki = k >= g;
        if (ii && ji && ki) {
            u0[c] = (float)((double)u0[c] - dt * (pe - p[c - d.si]) / (m.x[i + 1] - m.x[i - 1]));
            u1[c] = (float)((double)u1[c] - dt * (pn - p[c - d.sj]) / (m.y[j + 1] - m.y[j - 1]));
            u2[c] = (float)((double)u2[c] - dt * (pt - p[c - 1]) / (m.z[k + 1] - m.z[k - 1]));
        }
        const bool jint = ji && j < d.n1 - g, kint = ki && k < d.n2 - g, iint = ii && i < d.n0 - g;
        double ayz = m.dy[j] * m.dz[k];
        double vx = J0[c] - dt * ayz * (pe - pc) / (m.x[i + 1] - m.x[i]);
        if (i == g - 1 && jint && kint) vx = ayz * (face ? face[(j - g) * d.nz + (k - g)] : uin0);
        J0[c] = vx;
        double vy = J1[c] - dt * (m.dx[i] * m.dz[k]) * (pn - pc) / (m.y[j + 1] - m.y[j]);
        if (iint && kint && (j == g - 1 || j == d.n1 - g - 1)) vy = 0;
        J1[c] = vy;
        double vz = J2[c] - dt * (m.dx[i] * m.dy[j]) * (pt - pc) / (m.z[k + 1] - m.z[k]);
        if (iint && jint && (k == g - 1 || k == d.n2 - g - 1)) vz = 0;
        J2[c] = vz;
    }
}

__device__ __forceinline__ double quad3(double x0, double x1, double x2, double f0, double f1, double f2,
                                        double x3) {
    double l0 = ((x3 - x1) * (x3 - x2)) / ((x0 - x1) * (x0 - x2));
    double l1 = ((x3 - x0) * (x3 - x2)) / ((x1 - x0) * (x1 - x2));
    double l2 = ((x3 - x0) * (x3 - x1)) / ((x2 - x0) * (x2 - x1));
    return f0 * l0 + f1 * l1 + f2 * l2;
}

// velocity ghosts (src/bc.h:5-233) on FP32 storage, FP64 arithmetic
__global__ void km_ubc(Dims d, M1 m, float* __restrict__ u0, float* __restrict__ u1, float* __restrict__ u2,
                       const float* __restrict__ o0, const float* __restrict__ o1,
                       const float* __restrict__ o2, double uin0, double uin1, double uin2, double dt,
                       int retracted, double floor_x, const double* __restrict__ cells) {
    const int g = d.gc;
    long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz, nxy = (long long)d.nx * d.ny;
    long long n = 2 * nyz + 2 * nxz + 2 * nxy;
    float* U[3] = {u0, u1, u2};
    const float* O[3] = {o0, o1, o2};
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long s = t;
        if (s < 2 * nyz) {
            int side = (int)(s / nyz);
            s %= nyz;
            int j = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            if (!side) {
                for (int i = 0; i < g; ++i) {
                    int c = d.at(i, j, k);
                    if (cells) {
                        // turbulent inflow: [3][gc][ny][nz]
                        const long long o = (long long)i * nyz + (long long)(j - g) * d.nz + (k - g);
                        u0[c] = (float)cells[o];
                        u1[c] = (float)cells[g * nyz + o];
                        u2[c] = (float)cells[2 * g * nyz + o];
                    } else {
                        u0[c] = (float)uin0;
                        u1[c] = (float)uin1;
                        u2[c] = (float)uin2;
                    }
                }
            } else {
                for (int i = d.n0 - g; i < d.n0; ++i) {
                    int c = d.at(i, j, k);
                    double uo = u0[c];
                    for (int q = 0; q < 3; ++q) {
                        double a0 = O[q][c], a1 = O[q][c - d.si], a2 = O[q][c - 2 * d.si];
                        U[q][c] = (float)(a0 - uo * dt * (3 * a0 - 4 * a1 + a2) / 2 / m.dx[i]);
                    }
                }
            }
            continue;
        }
        s -= 2 * nyz;
        if (s < 2 * nxz) {
            int side = (int)(s / nxz);
            s %= nxz;
            int i = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            int jin2 = side ? d.n1 - g - 2 : g + 1, jin1 = side ? d.n1 - g - 1 : g;
            int jg1 = side ? d.n1 - g : g - 1, jg2 = side ? d.n1 - g + 1 : g - 2;
            double yw = side ? m.y[jin1] + 0.5 * m.dy[jin1] : m.y[jin1] - 0.5 * m.dy[jin1];
            int c1 = d.at(i, jin1, k), c2 = d.at(i, jin2, k);
            double wall[3] = {u0[c1], 0, u2[c1]};
            for (int q = 0; q < 3; ++q) {
                double f0 = U[q][c2], f1 = U[q][c1];
                U[q][d.at(i, jg1, k)] = (float)quad3(m.y[jin2], m.y[jin1], yw, f0, f1, wall[q], m.y[jg1]);
                U[q][d.at(i, jg2, k)] = (float)quad3(m.y[jin2], m.y[jin1], yw, f0, f1, wall[q], m.y[jg2]);
            }
            continue;
        }
        s -= 2 * nxz;
        int side = (int)(s / nxy);
        s %= nxy;
        int i = g + (int)(s / d.ny), j = g + (int)(s % d.ny);
        int kin2 = side ? d.n2 - g - 2 : g + 1, kin1 = side ? d.n2 - g - 1 : g;
        int kg1 = side ? d.n2 - g : g - 1, kg2 = side ? d.n2 - g + 1 : g - 2;
        double zw = side ? m.z[kin1] + 0.5 * m.dz[kin1] : m.z[kin1] - 0.5 * m.dz[kin1];
        int c1 = d.at(i, j, kin1), c2 = d.at(i, j, kin2);
        double wall[3] = {0, 0, 0};
        if (side || (retracted && m.x[i] < floor_x)) {
            wall[0] = u0[c1];
            wall[1] = u1[c1];
        }
        for (int q = 0; q < 3; ++q) {
            double f0 = U[q][c2], f1 = U[q][c1];
            U[q][d.at(i, j, kg1)] = (float)quad3(m.z[kin2], m.z[kin1], zw, f0, f1, wall[q], m.z[kg1]);
            U[q][d.at(i, j, kg2)] = (float)quad3(m.z[kin2], m.z[kin1], zw, f0, f1, wall[q], m.z[kg2]);
        }
    }
}

__device__ __forceinline__ long long shell_count(const Dims& d) {
    return 2LL * d.gc * d.n1 * d.n2 + 2LL * d.gc * d.nx * d.n2 + 2LL * d.gc * d.nx * d.ny;
}
__device__ __forceinline__ int shell_index(const Dims& d, long long t) {
    const int g = d.gc;
    long long A = 2LL * g * d.n1 * d.n2;
    if (t < A) {
        int pl = (int)(t / ((long long)d.n1 * d.n2));
        int r = (int)(t % ((long long)d.n1 * d.n2));
        int i = pl < g ? pl : d.n0 - 2 * g + pl;
        return i * d.si + r;
    }
    t -= A;
    long long B = 2LL * g * d.nx * d.n2;
    if (t < B) {
        int k = (int)(t % d.n2);
        long long r = t / d.n2;
        int jj = (int)(r % (2 * g));
        int i = g + (int)(r / (2 * g));
        int j = jj < g ? jj : d.n1 - 2 * g + jj;
        return d.at(i, j, k);
    }
    t -= B;
    int kk = (int)(t % (2 * g));
    long long r = t / (2 * g);
    int j = g + (int)(r % d.ny);
    int i = g + (int)(r / d.ny);
    int k = kk < g ? kk : d.n2 - 2 * g + kk;
    return d.at(i, j, k);
}

__global__ void km_ghost_copy(Dims d, const float* __restrict__ a0, const float* __restrict__ a1,
                              const float* __restrict__ a2, float* __restrict__ b0, float* __restrict__ b1,
                              float* __restrict__ b2) {
    long long n = shell_count(d);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int e = shell_index(d, t);
        b0[e] = a0[e];
        b1[e] = a1[e];
        b2[e] = a2[e];
    }
}

__global__ __launch_bounds__(BLOCK) void km_solid(Dims d, const double* __restrict__ solid,
                                                  float* __restrict__ u0, float* __restrict__ u1,
                                                  float* __restrict__ u2) {
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        int c = d.at(cc.i, cc.j, cc.k);
        float w = (float)(1. - solid[c]);
        u0[c] *= w;
        u1[c] *= w;
        u2[c] *= w;
    }
}

// nu_t (FP32), divergence (FP64, stored on request), sum div^2 and max CFL
// nu_t in FP32 from the central differences of U (e, w, n, s, t, b: the three
// components at i+1, i-1, j+1, j-1, k+1, k-1), and the FP64 CFL number of a cell
__device__ __forceinline__ float nutf_from(const MetricF& mf, float Cs, int i, int j, int k, const float (&e)[3],
                                           const float (&w)[3], const float (&n)[3], const float (&s)[3],
                                           const float (&t)[3], const float (&b)[3]) {
    const float rx = mf.r2[0][i], ry = mf.r2[1][j], rz = mf.r2[2][k];
    float dudx = (e[0] - w[0]) * rx, dudy = (n[0] - s[0]) * ry;
    float dudz = (t[0] - b[0]) * rz;
    float dvdx = (e[1] - w[1]) * rx, dvdy = (n[1] - s[1]) * ry;
    float dvdz = (t[1] - b[1]) * rz;
    float dwdx = (e[2] - w[2]) * rx, dwdy = (n[2] - s[2]) * ry;
    float dwdz = (t[2] - b[2]) * rz;
    float a = dudy + dvdx, bb = dudz + dwdx, ee = dvdz + dwdy;
    float S = sqrtf(2.f * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + a * a + bb * bb + ee * ee);
    float l = Cs * cbrtf(mf.w[0][i] * mf.w[1][j] * mf.w[2][k]);
    return l * l * S;
}
__device__ __forceinline__ double cflm_cell(const M1& m, int i, int j, int k, float u0, float u1, float u2,
                                            double dt) {
    return dt * (fabs((double)u0) / m.dx[i] + fabs((double)u1) / m.dy[j] + fabs((double)u2) / m.dz[k]);
}
__device__ __forceinline__ double divm_cell(const Dims& d, const M1& m, const double* __restrict__ J0,
                                            const double* __restrict__ J1, const double* __restrict__ J2, int i,
                                            int j, int k, int c) {
    double vol = m.dx[i] * m.dy[j] * m.dz[k];
    return ((J0[c] - J0[c - d.si]) + (J1[c] - J1[c - d.sj]) + (J2[c] - J2[c - 1])) / vol;
}

__global__ __launch_bounds__(BLOCK) void km_post(Dims d, M1 m, MetricF mf, const float* __restrict__ u0,
                                                 const float* __restrict__ u1, const float* __restrict__ u2,
                                                 const double* __restrict__ J0, const double* __restrict__ J1,
                                                 const double* __restrict__ J2, float* __restrict__ nut,
                                                 double* __restrict__ div, float Cs, double dt,
                                                 double* partials, unsigned* counter, StepScalars* out) {
    TileRange tr = block_tiles(d);
    double v[1] = {0.0};
    double cfl = 0.0;
    const float* u[3] = {u0, u1, u2};
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int i = cc.i, j = cc.j, k = cc.k;
        const int c = d.at(i, j, k);
        float e[3], w[3], n[3], s[3], tt[3], b[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            e[q] = u[q][c + d.si], w[q] = u[q][c - d.si], n[q] = u[q][c + d.sj], s[q] = u[q][c - d.sj];
            tt[q] = u[q][c + 1], b[q] = u[q][c - 1];
        }
        nut[c] = nutf_from(mf, Cs, i, j, k, e, w, n, s, tt, b);
        const double dv = divm_cell(d, m, J0, J1, J2, i, j, k, c);
        if (div) div[c] = dv;
        v[0] += dv * dv;
        const double lc = cflm_cell(m, i, j, k, u0[c], u1[c], u2[c], dt);
        cfl = lc > cfl ? lc : cfl;
    }
    block_sum<1>(v);
    double bm = block_max(cfl);
    double pv[2] = {v[0], bm};
    store_partials<2>(pv, partials);
}

// the same as a column walk with the centre plane of U staged through LDS
// (star_walk.h, 4-row blocks of BLOCK threads)
__global__ __launch_bounds__(BLOCK) void km_post_star(Dims d, M1 m, MetricF mf, const float* __restrict__ u0,
                                                      const float* __restrict__ u1, const float* __restrict__ u2,
                                                      const double* __restrict__ J0, const double* __restrict__ J1,
                                                      const double* __restrict__ J2, float* __restrict__ nut,
                                                      double* __restrict__ div, float Cs, double dt, double* partials,
                                                      int ni) {
    __shared__ StarLds<float, TJ> sh;
    double v[1] = {0.0};
    double cfl = 0.0;
    star_walk<TJ>(d, u0, u1, u2, ni, sh,
                  [&](int i, int j, int k, int c, const float (&x)[3][5], const StarView<float, TJ>& a) {
                      float e[3], w[3], n[3], s[3], tt[3], b[3];
#pragma unroll
                      for (int q = 0; q < 3; ++q) {
                          e[q] = x[q][3], w[q] = x[q][1], n[q] = a.y(q, 3), s[q] = a.y(q, 1);
                          tt[q] = a.z(q, 3), b[q] = a.z(q, 1);
                      }
                      nut[c] = nutf_from(mf, Cs, i, j, k, e, w, n, s, tt, b);
                      const double dv = divm_cell(d, m, J0, J1, J2, i, j, k, c);
                      if (div) div[c] = dv;
                      v[0] += dv * dv;
                      const double lc = cflm_cell(m, i, j, k, x[0][2], x[1][2], x[2][2], dt);
                      cfl = lc > cfl ? lc : cfl;
                  });
    block_sum<1>(v);
    double bm = block_max(cfl);
    double pv[2] = {v[0], bm};
    store_partials<2>(pv, partials);
}

int grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

template <typename T>
int dev_upload(dalm_ctx* ctx, const std::vector<T>& h, const T** out) {
    T* p = nullptr;
    HIP_TRY(hipMalloc((void**)&p, h.size() * sizeof(T) + 256));
    HIP_TRY(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    ctx->owned_f.push_back(p);
    *out = p;
    return DALM_OK;
}

}  // namespace

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

// FP32 metric reciprocals and operator pieces, built once from the FP64 mesh and
// the FP64 operator (which must already exist).
int setup_mixed(dalm_ctx* ctx) {
    const std::vector<double>* cs[3] = {&ctx->hx, &ctx->hy, &ctx->hz};
    const std::vector<double>* ws[3] = {&ctx->hdx, &ctx->hdy, &ctx->hdz};
    int rc;
    for (int a = 0; a < 3; ++a) {
        const std::vector<double>& c = *cs[a];
        const std::vector<double>& w = *ws[a];
        size_t n = c.size();
        std::vector<float> rw(n), rp(n, 0.f), rm(n, 0.f), r2(n, 0.f), wf(n);
        for (size_t i = 0; i < n; ++i) {
            rw[i] = (float)(1.0 / w[i]);
            wf[i] = (float)w[i];
            if (i + 1 < n) rp[i] = (float)(1.0 / (c[i + 1] - c[i]));
            if (i > 0) rm[i] = (float)(1.0 / (c[i] - c[i - 1]));
            if (i > 0 && i + 1 < n) r2[i] = (float)(1.0 / (c[i + 1] - c[i - 1]));
        }
        if ((rc = dev_upload(ctx, rw, &ctx->mf.rw[a])) || (rc = dev_upload(ctx, rp, &ctx->mf.rp[a])) ||
            (rc = dev_upload(ctx, rm, &ctx->mf.rm[a])) || (rc = dev_upload(ctx, r2, &ctx->mf.r2[a])) ||
            (rc = dev_upload(ctx, wf, &ctx->mf.w[a])))
            return rc;
    }
    // operator pieces: copy the FP64 1-D arrays back and round them to FP32
    const Dims& d = ctx->d;
    const PoissonCoef& P = ctx->pc;
    auto down = [&](const double* src, int n, const float** dst, bool normalise) -> int {
        std::vector<double> h(n);
        HIP_TRY(hipMemcpy(h.data(), src, n * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<float> f(n);
        for (int q = 0; q < n; ++q) f[q] = (float)(normalise ? h[q] / P.max_diag : h[q]);
        return dev_upload(ctx, f, dst);
    };
    if ((rc = down(P.ae_n, d.n0, &ctx->pcf.ae, false)) || (rc = down(P.aw_n, d.n0, &ctx->pcf.aw, false)) ||
        (rc = down(P.an_n, d.n1, &ctx->pcf.an, false)) || (rc = down(P.as_n, d.n1, &ctx->pcf.as, false)) ||
        (rc = down(P.at_n, d.n2, &ctx->pcf.at, false)) || (rc = down(P.ab_n, d.n2, &ctx->pcf.ab, false)) ||
        (rc = down(P.acx, d.n0, &ctx->pcf.acx, true)) || (rc = down(P.an_u, d.n1, &ctx->pcf.anu, true)) ||
        (rc = down(P.as_u, d.n1, &ctx->pcf.asu, true)) || (rc = down(P.at_u, d.n2, &ctx->pcf.atu, true)) ||
        (rc = down(P.ab_u, d.n2, &ctx->pcf.abu, true)))
        return rc;
    return DALM_OK;
}

int launch_ghost_copy_m(dalm_ctx* ctx) {
    long long n = 2LL * ctx->d.gc * ctx->d.n1 * ctx->d.n2 + 2LL * ctx->d.gc * ctx->d.nx * ctx->d.n2 +
                  2LL * ctx->d.gc * ctx->d.nx * ctx->d.ny;
    hipLaunchKernelGGL(km_ghost_copy, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, ctx->Uold32[0],
                       ctx->Uold32[1], ctx->Uold32[2], ctx->U32[0], ctx->U32[1], ctx->U32[2]);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_predictor_m(dalm_ctx* ctx, double Re, double dt) {
    probe_begin(ctx, DALM_K_PREDICTOR);
    const int ni = ctx->d.gc >= 2 ? star_planes<kPredRowsF>(ctx->d) : 0;
    if (ni > 0)
        hipLaunchKernelGGL(km_predictor_star, dim3(star_grid<kPredRowsF>(ctx->d, ni)), dim3(TK * kPredRowsF), 0,
                           ctx->stream, ctx->d,
                           ctx->mf, ctx->Uold32[0], ctx->Uold32[1], ctx->Uold32[2], ctx->JU[0], ctx->JU[1],
                           ctx->JU[2], ctx->nut32, ctx->f32[0], ctx->f32[1], ctx->f32[2], ctx->U32[0], ctx->U32[1],
                           ctx->U32[2], (float)(1 / Re), (float)dt, ni);
    else
        hipLaunchKernelGGL(km_predictor, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->mf,
                           ctx->Uold32[0], ctx->Uold32[1], ctx->Uold32[2], ctx->JU[0], ctx->JU[1], ctx->JU[2],
                           ctx->nut32, ctx->f32[0], ctx->f32[1], ctx->f32[2], ctx->U32[0], ctx->U32[1],
                           ctx->U32[2], (float)(1 / Re), (float)dt);
    probe_end(ctx, DALM_K_PREDICTOR);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_flux_rhs_m(dalm_ctx* ctx, const double Uin[3], double dt, double scale, bool rhs) {
    long long n = (long long)(ctx->d.nx + 1) * (ctx->d.ny + 1) * (ctx->d.nz + 1);
    hipLaunchKernelGGL(km_flux_rhs, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->U32[0], ctx->U32[1], ctx->U32[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->b, Uin[0],
                       rhs ? 1.0 / (dt * scale) : 0.0, rhs ? 1 : 0, ctx->inflow_face());
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_project_m(dalm_ctx* ctx, double dt, const double Uin[3]) {
    long long n = (long long)(ctx->d.nx + 1) * (ctx->d.ny + 1) * (ctx->d.nz + 1);
    hipLaunchKernelGGL(km_project, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx), ctx->mf,
                       ctx->p, ctx->U32[0], ctx->U32[1], ctx->U32[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], dt,
                       Uin[0], ctx->inflow_face());
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_ubc_m(dalm_ctx* ctx, const double Uin[3], double dt, int retracted, double floor_x) {
    const Dims& d = ctx->d;
    long long n = 2LL * d.ny * d.nz + 2LL * d.nx * d.nz + 2LL * d.nx * d.ny;
    hipLaunchKernelGGL(km_ubc, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx), ctx->U32[0],
                       ctx->U32[1], ctx->U32[2], ctx->Uold32[0], ctx->Uold32[1], ctx->Uold32[2], Uin[0], Uin[1],
                       Uin[2], dt, retracted, floor_x, ctx->inflow_cells());
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_solid_m(dalm_ctx* ctx) {
    hipLaunchKernelGGL(km_solid, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->solid,
                       ctx->U32[0], ctx->U32[1], ctx->U32[2]);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_post_m(dalm_ctx* ctx, double Cs, double dt, bool store_div) {
    int ni = ctx->d.gc >= 2 ? star_planes<TJ>(ctx->d) : 0;
    const long long cap = 8LL * (ctx->grid > ctx->sgrid ? ctx->grid : ctx->sgrid);
    if (ni > 0 && 2LL * star_grid<TJ>(ctx->d, ni) > cap) ni = 0;
    if (ni > 0) {
        const int grid = star_grid<TJ>(ctx->d, ni);
        hipLaunchKernelGGL(km_post_star, dim3(grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), ctx->mf,
                           ctx->U32[0], ctx->U32[1], ctx->U32[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->nut32,
                           store_div ? ctx->div : nullptr, (float)Cs, dt, ctx->partials, ni);
        LAUNCH_CHECK();
        return launch_finalize(ctx, FIN_POST, grid, nullptr, nullptr);
    }
    hipLaunchKernelGGL(km_post, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), ctx->mf,
                       ctx->U32[0], ctx->U32[1], ctx->U32[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->nut32,
                       store_div ? ctx->div : nullptr, (float)Cs, dt, ctx->partials, ctx->counter, ctx->scal);
    LAUNCH_CHECK();
    return launch_finalize(ctx, FIN_POST, ctx->grid, nullptr, nullptr);
}
