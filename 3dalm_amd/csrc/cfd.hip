// cfd.hip -- FP64 stencil kernels of the predictor / projection / post-processing
// chain (drop-in form).  Arithmetic follows the reference's evaluation order and
// this file is compiled with -ffp-contract=off, so interior results match a
// g++ -O2 x86-64 build of the reference bit for bit (cbrt aside).
//
// All of these kernels are HBM-bound streaming stencils (no MFMA: nothing here is
// a dense contraction).  Loads are 64-wide coalesced k-rows; neighbours along j
// and i are re-reads served by L2 / the last-level cache.
#include "ctx.h"
#include "star_walk.h"

namespace {

struct M1 {
    const double *x, *y, *z, *dx, *dy, *dz;
};
M1 mesh(const dalm_ctx* c) { return M1{c->x, c->y, c->z, c->dx, c->dy, c->dz}; }

__device__ __forceinline__ double sgn(double v) { return (double)((v > 0.0) - (v < 0.0)); }

// Face fluxes on both sides of a cell along one direction (MUSCL kappa = 1/3 with
// the compression-limited minmod of src/cfd.h:37-60).  v: f(-2) .. f(+2).
struct FluxPair {
    double p, m;
};
__device__ __forceinline__ FluxPair muscl_faces(double vm2, double vm1, double v0, double vp1,
                                                double vp2, double Jp, double Jm) {
    const double kap = 1. / 3.;
    const double bb = (3. - kap) / (1. - kap);
    const double q = 1. / 4.;
    double a1 = vm1 - vm2, a2 = v0 - vm1, a3 = vp1 - v0, a4 = vp2 - vp1;
    double t1 = sgn(a1), t2 = sgn(a2), t3 = sgn(a3), t4 = sgn(a4);
    double h6 = t4 * fmax(0., fmin(fabs(a4), t4 * bb * a3));
    double h5 = t3 * fmax(0., fmin(fabs(a3), t3 * bb * a4));
    double h4 = t3 * fmax(0., fmin(fabs(a3), t3 * bb * a2));
    double h3 = t2 * fmax(0., fmin(fabs(a2), t2 * bb * a3));
    double h2 = t2 * fmax(0., fmin(fabs(a2), t2 * bb * a1));
    double h1 = t1 * fmax(0., fmin(fabs(a1), t1 * bb * a2));
    double pr = vp1 - q * ((1. - kap) * h6 + (1. + kap) * h5);
    double pl = v0 + q * ((1. - kap) * h3 + (1. + kap) * h4);
    double mr = v0 - q * ((1. - kap) * h4 + (1. + kap) * h3);
    double ml = vm1 + q * ((1. - kap) * h1 + (1. + kap) * h2);
    FluxPair r;
    r.p = 0.5 * (Jp * (pr + pl) - fabs(Jp) * (pr - pl));
    r.m = 0.5 * (Jm * (mr + ml) - fabs(Jm) * (mr - ml));
    return r;
}

// ---------------------------------------------------------------------------
// predictor: U = Uold + dt (-conv + diff + f), src/cfd.h:175-242

// the cell's metrics, face fluxes and viscosity
struct PredCell {
    double Je, Jw, Jn, Js, Jt, Jb;
    double dxc, dyc, dzc, vol, hxp, hxm, hyp, hym, hzp, hzm, nu;
};
__device__ __forceinline__ PredCell pred_cell(const Dims& d, const M1& m, const double* __restrict__ J0,
                                              const double* __restrict__ J1, const double* __restrict__ J2,
                                              const double* __restrict__ nut, double inv_re, int i, int j, int k,
                                              int c) {
    PredCell P;
    P.Je = J0[c], P.Jw = J0[c - d.si], P.Jn = J1[c], P.Js = J1[c - d.sj], P.Jt = J2[c], P.Jb = J2[c - 1];
    P.dxc = m.dx[i], P.dyc = m.dy[j], P.dzc = m.dz[k];
    P.vol = P.dxc * P.dyc * P.dzc;
    P.hxp = m.x[i + 1] - m.x[i], P.hxm = m.x[i] - m.x[i - 1];
    P.hyp = m.y[j + 1] - m.y[j], P.hym = m.y[j] - m.y[j - 1];
    P.hzp = m.z[k + 1] - m.z[k], P.hzm = m.z[k] - m.z[k - 1];
    P.nu = inv_re + nut[c];
    return P;
}
// one component from its 13-point star (x*, y*, z*: the neighbours along i, j, k)
__device__ __forceinline__ double pred_value(const PredCell& P, double vc, double xww, double xw, double xe,
                                             double xee, double yss, double ys, double yn, double ynn, double zbb,
                                             double zb, double zt, double ztt, double fq, double dt) {
    FluxPair fx = muscl_faces(xww, xw, vc, xe, xee, P.Je, P.Jw);
    FluxPair fy = muscl_faces(yss, ys, vc, yn, ynn, P.Jn, P.Js);
    FluxPair fz = muscl_faces(zbb, zb, vc, zt, ztt, P.Jt, P.Jb);
    double conv = (fx.p - fx.m + fy.p - fy.m + fz.p - fz.m) / P.vol;
    double diff = P.nu * (((xe - vc) / P.hxp - (vc - xw) / P.hxm) / P.dxc +
                          ((yn - vc) / P.hyp - (vc - ys) / P.hym) / P.dyc +
                          ((zt - vc) / P.hzp - (vc - zb) / P.hzm) / P.dzc);
    return vc + dt * (-conv + diff + fq);
}

// direct form: every value through the caches (grids too small for the walk)
__global__ __launch_bounds__(BLOCK) void k_predictor(
    Dims d, M1 m, const double* __restrict__ o0, const double* __restrict__ o1,
    const double* __restrict__ o2, const double* __restrict__ J0, const double* __restrict__ J1,
    const double* __restrict__ J2, const double* __restrict__ nut, const double* __restrict__ f0,
    const double* __restrict__ f1, const double* __restrict__ f2, double* __restrict__ u0,
    double* __restrict__ u1, double* __restrict__ u2, double inv_re, double dt) {
    TileRange tr = block_tiles(d);
    const int si = d.si, sj = d.sj;
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int i = cc.i, j = cc.j, k = cc.k;
        const int c = d.at(i, j, k);
        const PredCell P = pred_cell(d, m, J0, J1, J2, nut, inv_re, i, j, k, c);
        const double* src[3] = {o0, o1, o2};
        const double* frc[3] = {f0, f1, f2};
        double* dst[3] = {u0, u1, u2};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double* a = src[q];
            dst[q][c] = pred_value(P, a[c], a[c - 2 * si], a[c - si], a[c + si], a[c + 2 * si], a[c - 2 * sj],
                                   a[c - sj], a[c + sj], a[c + 2 * sj], a[c - 2], a[c - 1], a[c + 1], a[c + 2],
                                   frc[q][c], dt);
        }
    }
}

// column walk with the centre plane staged through LDS (star_walk.h); same
// arithmetic per cell.  kPredRows wavefronts per block.
constexpr int kPredRows = 4;
__global__ __launch_bounds__(TK * kPredRows) void k_predictor_star(
    Dims d, M1 m, const double* __restrict__ o0, const double* __restrict__ o1,
    const double* __restrict__ o2, const double* __restrict__ J0, const double* __restrict__ J1,
    const double* __restrict__ J2, const double* __restrict__ nut, const double* __restrict__ f0,
    const double* __restrict__ f1, const double* __restrict__ f2, double* __restrict__ u0,
    double* __restrict__ u1, double* __restrict__ u2, double inv_re, double dt, int ni) {
    __shared__ StarLds<double, kPredRows> sh;
    star_walk<kPredRows>(d, o0, o1, o2, ni, sh,
              [&](int i, int j, int k, int c, const double (&x)[3][5], const StarView<double, kPredRows>& v) {
                  const PredCell P = pred_cell(d, m, J0, J1, J2, nut, inv_re, i, j, k, c);
                  const double* frc[3] = {f0, f1, f2};
                  double* dst[3] = {u0, u1, u2};
#pragma unroll
                  for (int q = 0; q < 3; ++q)
                      dst[q][c] = pred_value(P, x[q][2], x[q][0], x[q][1], x[q][3], x[q][4], v.y(q, 0), v.y(q, 1),
                                             v.y(q, 3), v.y(q, 4), v.z(q, 0), v.z(q, 1), v.z(q, 3), v.z(q, 4),
                                             frc[q][c], dt);
              });
}

// Full-length copy of three arrays (cpy_array on a [L][3] field).
__global__ void k_copy3(long long L, const double* __restrict__ a0, const double* __restrict__ a1,
                        const double* __restrict__ a2, double* __restrict__ b0,
                        double* __restrict__ b1, double* __restrict__ b2) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < L;
         e += (long long)gridDim.x * blockDim.x) {
        b0[e] = a0[e];
        b1[e] = a1[e];
        b2[e] = a2[e];
    }
}

// Ghost shell enumeration: i-planes, then j-rows of interior i, then k-cells of
// interior (i, j).
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
__host__ __device__ inline long long shell_count(const Dims& d) {
    return 2LL * d.gc * d.n1 * d.n2 + 2LL * d.gc * d.nx * d.n2 + 2LL * d.gc * d.nx * d.ny;
}

__global__ void k_ghost_copy(Dims d, const double* __restrict__ a0, const double* __restrict__ a1,
                             const double* __restrict__ a2, double* __restrict__ b0,
                             double* __restrict__ b1, double* __restrict__ b2) {
    long long n = shell_count(d);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int e = shell_index(d, t);
        b0[e] = a0[e];
        b1[e] = a1[e];
        b2[e] = a2[e];
    }
}

// Face box [gc-1, n-gc) in each direction, i.e. (nx+1)(ny+1)(nz+1) faces.
struct FaceIdx {
    int i, j, k;
};
__device__ __forceinline__ FaceIdx face_idx(const Dims& d, long long t) {
    FaceIdx f;
    int ez = d.nz + 1, ey = d.ny + 1;
    f.k = d.gc - 1 + (int)(t % ez);
    long long r = t / ez;
    f.j = d.gc - 1 + (int)(r % ey);
    f.i = d.gc - 1 + (int)(r / ey);
    return f;
}

// The x- inflow velocity of face (j, k): the constant uin0 (src/bc.h:243-257), or
// the turbulent inflow's face value when one is set (face != nullptr).
__device__ __forceinline__ double inflow_u(const Dims& d, const double* face, int j, int k, double uin0) {
    return face ? face[(j - d.gc) * d.nz + (k - d.gc)] : uin0;
}

// Turbulent inflow (not in the reference): the x- ghost cells and the x- face
// velocity of frozen turbulence convected by `shift` = speed * t,
// Uin + sum_n a_n cos(kx (x - shift) + ky y + kz z + psi_n), summed in mode order;
// modes[n] = (kx, ky, kz, psi, ax, ay, az).  One thread per (j, k) boundary cell;
// out: ghost cells [3][gc][ny][nz], then the face [ny][nz] at (x[gc-1] + x[gc]) / 2.
__global__ void k_inflow(Dims d, M1 m, const double* __restrict__ modes, int n_modes, double shift,
                         double uin0, double uin1, double uin2, double* __restrict__ out) {
    const int g = d.gc;
    const long long nyz = (long long)d.ny * d.nz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nyz;
         t += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(t / d.nz), k = (int)(t % d.nz);
        const double Y = m.y[g + j], Z = m.z[g + k];
        const double uin[3] = {uin0, uin1, uin2};
        for (int i = 0; i <= g; ++i) {
            const double X = (i < g ? m.x[i] : 0.5 * (m.x[g - 1] + m.x[g])) - shift;
            double s0 = 0, s1 = 0, s2 = 0;
            for (int n = 0; n < n_modes; ++n) {
                const double* M = modes + 7 * n;
                const double c = cos(M[0] * X + M[1] * Y + M[2] * Z + M[3]);
                s0 += M[4] * c;
                s1 += M[5] * c;
                s2 += M[6] * c;
            }
            if (i < g) {
                const double s[3] = {s0, s1, s2};
                for (int q = 0; q < 3; ++q) out[(q * g + i) * nyz + t] = uin[q] + s[q];
            } else {
                out[3 * g * nyz + t] = uin0 + s0;
            }
        }
    }
}

// Face fluxes (interpolate_JU, src/cfd.h:394-420) with, optionally, the face BCs
// of apply_JUbc (src/bc.h:235-302) written in place of the interpolated value.
__global__ void k_flux(Dims d, M1 m, const double* __restrict__ u0, const double* __restrict__ u1,
                       const double* __restrict__ u2, double* __restrict__ J0, double* __restrict__ J1,
                       double* __restrict__ J2, int apply_bc, double uin0, const double* __restrict__ face) {
    long long n = (long long)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    const int g = d.gc;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        FaceIdx f = face_idx(d, t);
        const int i = f.i, j = f.j, k = f.k;
        const int c = d.at(i, j, k);
        const bool ji = j >= g && j < d.n1 - g, ki = k >= g && k < d.n2 - g, ii = i >= g && i < d.n0 - g;
        double ayz = m.dy[j] * m.dz[k];
        double vx = 0.5 * (u0[c] * ayz + u0[c + d.si] * ayz);
        if (apply_bc && i == g - 1 && ji && ki) vx = ayz * inflow_u(d, face, j, k, uin0);
        J0[c] = vx;
        double axz = m.dx[i] * m.dz[k];
        double vy = 0.5 * (u1[c] * axz + u1[c + d.sj] * axz);
        if (apply_bc && ii && ki && (j == g - 1 || j == d.n1 - g - 1)) vy = 0;
        J1[c] = vy;
        double axy = m.dx[i] * m.dy[j];
        double vz = 0.5 * (u2[c] * axy + u2[c + 1] * axy);
        if (apply_bc && ii && ji && (k == g - 1 || k == d.n2 - g - 1)) vz = 0;
        J2[c] = vz;
    }
}

// apply_JUbc alone (src/bc.h:235-302): planes x-, y+-, z+-.
__global__ void k_jubc(Dims d, M1 m, double* __restrict__ J0, double* __restrict__ J1,
                       double* __restrict__ J2, double uin0, const double* __restrict__ face) {
    const int g = d.gc;
    long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz, nxy = (long long)d.nx * d.ny;
    long long n = nyz + 2 * nxz + 2 * nxy;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long s = t;
        if (s < nyz) {
            int j = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            J0[d.at(g - 1, j, k)] = m.dy[j] * m.dz[k] * inflow_u(d, face, j, k, uin0);
            continue;
        }
        s -= nyz;
        if (s < 2 * nxz) {
            int side = (int)(s / nxz);
            s %= nxz;
            int i = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            J1[d.at(i, side ? d.n1 - g - 1 : g - 1, k)] = 0;
            continue;
        }
        s -= 2 * nxz;
        int side = (int)(s / nxy);
        s %= nxy;
        int i = g + (int)(s / d.ny), j = g + (int)(s % d.ny);
        J2[d.at(i, j, side ? d.n2 - g - 1 : g - 1)] = 0;
    }
}

// divergence of the face fluxes at an interior cell (src/cfd.h:584-595)
__device__ __forceinline__ double flux_div(const Dims& d, const M1& m, const double* J0,
                                           const double* J1, const double* J2, int c, int i,
                                           int j, int k) {
    double vol = m.dx[i] * m.dy[j] * m.dz[k];
    return (J0[c] - J0[c - d.si] + J1[c] - J1[c - d.sj] + J2[c] - J2[c - 1]) / vol;
}

// Poisson RHS b = div(JU) / (dt * max_diag), src/cfd.h:569-598
__global__ __launch_bounds__(BLOCK) void k_rhs(Dims d, M1 m, const double* __restrict__ J0,
                                               const double* __restrict__ J1,
                                               const double* __restrict__ J2, double* __restrict__ b,
                                               double dts) {
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        int c = d.at(cc.i, cc.j, cc.k);
        b[c] = flux_div(d, m, J0, J1, J2, c, cc.i, cc.j, cc.k) / dts;
    }
}

// Pressure ghosts (src/bc.h:304-411): Neumann copies on x-, y+-, z+-; the x+
// Dirichlet image p_o = 0 - (p_i - 0) * (h_o / h_i).
__global__ void k_pbc(Dims d, M1 m, double* __restrict__ p) {
    const int g = d.gc;
    long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz, nxy = (long long)d.nx * d.ny;
    long long n = 2 * nyz + 2 * nxz + 2 * nxy;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long s = t;
        if (s < 2 * nyz) {
            int side = (int)(s / nyz);
            s %= nyz;
            int j = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            if (!side) {
                p[d.at(g - 1, j, k)] = p[d.at(g, j, k)];
            } else {
                int ii = d.n0 - g - 1, io = d.n0 - g;
                double hi = 0.5 * m.dx[ii], ho = 0.5 * m.dx[io], pb = 0;
                p[d.at(io, j, k)] = pb - (p[d.at(ii, j, k)] - pb) * (ho / hi);
            }
            continue;
        }
        s -= 2 * nyz;
        if (s < 2 * nxz) {
            int side = (int)(s / nxz);
            s %= nxz;
            int i = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            if (!side) p[d.at(i, g - 1, k)] = p[d.at(i, g, k)];
            else p[d.at(i, d.n1 - g, k)] = p[d.at(i, d.n1 - g - 1, k)];
            continue;
        }
        s -= 2 * nxz;
        int side = (int)(s / nxy);
        s %= nxy;
        int i = g + (int)(s / d.ny), j = g + (int)(s % d.ny);
        if (!side) p[d.at(i, j, g - 1)] = p[d.at(i, j, g)];
        else p[d.at(i, j, d.n2 - g)] = p[d.at(i, j, d.n2 - g - 1)];
    }
}

// Projection (src/cfd.h:600-783) over the face box; interior cells also update
// U.  With apply_bc the face BCs of apply_JUbc replace the projected value (the
// reference re-applies them right after, src/solver.h:1148-1153).
__global__ void k_project(Dims d, M1 m, const double* __restrict__ p, double* __restrict__ u0,
                          double* __restrict__ u1, double* __restrict__ u2, double* __restrict__ J0,
                          double* __restrict__ J1, double* __restrict__ J2, double dt, int apply_bc,
                          double uin0, const double* __restrict__ face) {
    long long n = (long long)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    const int g = d.gc;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        FaceIdx f = face_idx(d, t);
        const int i = f.i, j = f.j, k = f.k;
        const int c = d.at(i, j, k);
        const bool ii = i >= g, ji = j >= g, ki = k >= g;
        const double pc = p[c], pe = p[c + d.si], pn = p[c + d.sj], pt = p[c + 1];
        if (ii && ji && ki) {
            u0[c] -= dt * (pe - p[c - d.si]) / (m.x[i + 1] - m.x[i - 1]);
            u1[c] -= dt * (pn - p[c - d.sj]) / (m.y[j + 1] - m.y[j - 1]);
            u2[c] -= dt * (pt - p[c - 1]) / (m.z[k + 1] - m.z[k - 1]);
        }
        const bool jint = ji && j < d.n1 - g, kint = ki && k < d.n2 - g, iint = ii && i < d.n0 - g;
        double gx = (pe - pc) / (m.x[i + 1] - m.x[i]);
        double vx = J0[c] - dt * (m.dy[j] * m.dz[k]) * gx;
        if (apply_bc && i == g - 1 && jint && kint) vx = m.dy[j] * m.dz[k] * inflow_u(d, face, j, k, uin0);
        J0[c] = vx;
        double gy = (pn - pc) / (m.y[j + 1] - m.y[j]);
        double vy = J1[c] - dt * (m.dx[i] * m.dz[k]) * gy;
        if (apply_bc && iint && kint && (j == g - 1 || j == d.n1 - g - 1)) vy = 0;
        J1[c] = vy;
        double gz = (pt - pc) / (m.z[k + 1] - m.z[k]);
        double vz = J2[c] - dt * (m.dx[i] * m.dy[j]) * gz;
        if (apply_bc && iint && jint && (k == g - 1 || k == d.n2 - g - 1)) vz = 0;
        J2[c] = vz;
    }
}

// set_solid_U (src/cfd.h:1143-1160)
__global__ __launch_bounds__(BLOCK) void k_solid(Dims d, const double* __restrict__ solid,
                                                 double* __restrict__ u0, double* __restrict__ u1,
                                                 double* __restrict__ u2) {
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        int c = d.at(cc.i, cc.j, cc.k);
        double w = 1. - solid[c];
        u0[c] = w * u0[c];
        u1[c] = w * u1[c];
        u2[c] = w * u2[c];
    }
}

__device__ __forceinline__ double quad3(double x0, double x1, double x2, double f0, double f1,
                                        double f2, double x3) {
    double l0 = ((x3 - x1) * (x3 - x2)) / ((x0 - x1) * (x0 - x2));
    double l1 = ((x3 - x0) * (x3 - x2)) / ((x1 - x0) * (x1 - x2));
    double l2 = ((x3 - x0) * (x3 - x1)) / ((x2 - x0) * (x2 - x1));
    return f0 * l0 + f1 * l1 + f2 * l2;
}

// Velocity ghosts (src/bc.h:5-233): x- inflow, x+ convective outflow (advected
// by the current ghost u, gradient of Uold), y+- slip, z- no-slip (slip upstream
// of a retracted floor), z+ slip.  Each work item fills both ghost layers of one
// wall cell column, so items never touch each other's cells.
__global__ void k_ubc(Dims d, M1 m, double* __restrict__ u0, double* __restrict__ u1,
                      double* __restrict__ u2, const double* __restrict__ o0,
                      const double* __restrict__ o1, const double* __restrict__ o2, double uin0,
                      double uin1, double uin2, double dt, int retracted, double floor_x,
                      const double* __restrict__ cells) {
    const int g = d.gc;
    long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz, nxy = (long long)d.nx * d.ny;
    long long n = 2 * nyz + 2 * nxz + 2 * nxy;
    double* U[3] = {u0, u1, u2};
    const double* O[3] = {o0, o1, o2};
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
                        u0[c] = cells[o];
                        u1[c] = cells[g * nyz + o];
                        u2[c] = cells[2 * g * nyz + o];
                    } else {
                        u0[c] = uin0;
                        u1[c] = uin1;
                        u2[c] = uin2;
                    }
                }
            } else {
                for (int i = d.n0 - g; i < d.n0; ++i) {
                    int c = d.at(i, j, k);
                    double uo = u0[c];
                    for (int q = 0; q < 3; ++q) {
                        double a0 = O[q][c], a1 = O[q][c - d.si], a2 = O[q][c - 2 * d.si];
                        double gr = (3 * a0 - 4 * a1 + a2) / 2;
                        U[q][c] = a0 - uo * dt * gr / m.dx[i];
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
                U[q][d.at(i, jg1, k)] = quad3(m.y[jin2], m.y[jin1], yw, f0, f1, wall[q], m.y[jg1]);
                U[q][d.at(i, jg2, k)] = quad3(m.y[jin2], m.y[jin1], yw, f0, f1, wall[q], m.y[jg2]);
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
            U[q][d.at(i, j, kg1)] = quad3(m.z[kin2], m.z[kin1], zw, f0, f1, wall[q], m.z[kg1]);
            U[q][d.at(i, j, kg2)] = quad3(m.z[kin2], m.z[kin1], zw, f0, f1, wall[q], m.z[kg2]);
        }
    }
}

// Smagorinsky nu_t at one cell (src/cfd.h:785-826)
// Smagorinsky nu_t from the central differences of U (src/cfd.h:785-826): e, w, n,
// s, t, b hold the three components at i+1, i-1, j+1, j-1, k+1, k-1
__device__ __forceinline__ double nut_from(const M1& m, double Cs, int i, int j, int k, const double (&e)[3],
                                           const double (&w)[3], const double (&n)[3], const double (&s)[3],
                                           const double (&t)[3], const double (&b)[3]) {
    double hx = m.x[i + 1] - m.x[i - 1], hy = m.y[j + 1] - m.y[j - 1], hz = m.z[k + 1] - m.z[k - 1];
    double vol = m.dx[i] * m.dy[j] * m.dz[k];
    double dudx = (e[0] - w[0]) / hx;
    double dudy = (n[0] - s[0]) / hy;
    double dudz = (t[0] - b[0]) / hz;
    double dvdx = (e[1] - w[1]) / hx;
    double dvdy = (n[1] - s[1]) / hy;
    double dvdz = (t[1] - b[1]) / hz;
    double dwdx = (e[2] - w[2]) / hx;
    double dwdy = (n[2] - s[2]) / hy;
    double dwdz = (t[2] - b[2]) / hz;
    double S = sqrt(2 * (dudx * dudx) + 2 * (dvdy * dvdy) + 2 * (dwdz * dwdz) +
                    (dudy + dvdx) * (dudy + dvdx) + (dudz + dwdx) * (dudz + dwdx) +
                    (dvdz + dwdy) * (dvdz + dwdy));
    double l = Cs * cbrt(vol);
    return l * l * S;
}
__device__ __forceinline__ double nut_cell(const Dims& d, const M1& m, const double* u0,
                                           const double* u1, const double* u2, double Cs, int c,
                                           int i, int j, int k) {
    const double* u[3] = {u0, u1, u2};
    double e[3], w[3], n[3], s[3], t[3], b[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        e[q] = u[q][c + d.si], w[q] = u[q][c - d.si], n[q] = u[q][c + d.sj], s[q] = u[q][c - d.sj];
        t[q] = u[q][c + 1], b[q] = u[q][c - 1];
    }
    return nut_from(m, Cs, i, j, k, e, w, n, s, t, b);
}
// the CFL number of a cell (src/cfd.h:1100-1141)
__device__ __forceinline__ double cfl_cell(const M1& m, int i, int j, int k, double u0, double u1, double u2,
                                           double dt) {
    return fabs(u0) * dt / m.dx[i] + fabs(u1) * dt / m.dy[j] + fabs(u2) * dt / m.dz[k];
}

// Post pass fused over the interior: nu_t, divergence (stored on request),
// sum of div^2 and max CFL; the last block writes both to the step scalars.
__global__ __launch_bounds__(BLOCK) void k_post(Dims d, M1 m, const double* __restrict__ u0,
                                                const double* __restrict__ u1,
                                                const double* __restrict__ u2,
                                                const double* __restrict__ J0,
                                                const double* __restrict__ J1,
                                                const double* __restrict__ J2, double* __restrict__ nut,
                                                double* __restrict__ div, double Cs, double dt,
                                                int do_nut, double* partials, unsigned* counter,
                                                StepScalars* out) {
    TileRange tr = block_tiles(d);
    double v[1] = {0.0};
    double cfl = 0.0;
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int i = cc.i, j = cc.j, k = cc.k;
        int c = d.at(i, j, k);
        if (do_nut) nut[c] = nut_cell(d, m, u0, u1, u2, Cs, c, i, j, k);
        double dv = flux_div(d, m, J0, J1, J2, c, i, j, k);
        if (div) div[c] = dv;
        v[0] += dv * dv;
        double lc = cfl_cell(m, i, j, k, u0[c], u1[c], u2[c], dt);
        cfl = lc > cfl ? lc : cfl;
    }
    block_sum<1>(v);
    double bm = block_max(cfl);
    double pv[2] = {v[0], bm};
    store_partials<2>(pv, partials);
}

// the post pass as a column walk with the centre plane of U staged through LDS
// (star_walk.h, 4-row blocks of BLOCK threads); same arithmetic per cell, the
// reductions summed per block of the walk's grid
__global__ __launch_bounds__(BLOCK) void k_post_star(Dims d, M1 m, const double* __restrict__ u0,
                                                     const double* __restrict__ u1, const double* __restrict__ u2,
                                                     const double* __restrict__ J0, const double* __restrict__ J1,
                                                     const double* __restrict__ J2, double* __restrict__ nut,
                                                     double* __restrict__ div, double Cs, double dt, int do_nut,
                                                     double* partials, int ni) {
    __shared__ StarLds<double, TJ> sh;
    double v[1] = {0.0};
    double cfl = 0.0;
    star_walk<TJ>(d, u0, u1, u2, ni, sh,
                  [&](int i, int j, int k, int c, const double (&x)[3][5], const StarView<double, TJ>& a) {
                      if (do_nut) {
                          double e[3], w[3], n[3], s[3], t[3], b[3];
#pragma unroll
                          for (int q = 0; q < 3; ++q) {
                              e[q] = x[q][3], w[q] = x[q][1], n[q] = a.y(q, 3), s[q] = a.y(q, 1);
                              t[q] = a.z(q, 3), b[q] = a.z(q, 1);
                          }
                          nut[c] = nut_from(m, Cs, i, j, k, e, w, n, s, t, b);
                      }
                      const double dv = flux_div(d, m, J0, J1, J2, c, i, j, k);
                      if (div) div[c] = dv;
                      v[0] += dv * dv;
                      const double lc = cfl_cell(m, i, j, k, x[0][2], x[1][2], x[2][2], dt);
                      cfl = lc > cfl ? lc : cfl;
                  });
    block_sum<1>(v);
    double bm = block_max(cfl);
    double pv[2] = {v[0], bm};
    store_partials<2>(pv, partials);
}

__global__ void k_scalar_sum(Dims d, const double* __restrict__ a, const double* __restrict__ b,
                             double* partials, unsigned* counter, double* result) {
    TileRange tr = block_tiles(d);
    double v[1] = {0.0};
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        int c = d.at(cc.i, cc.j, cc.k);
        v[0] += a[c] * b[c];
    }
    block_sum<1>(v);
    double pv[1] = {v[0]};
    store_partials<1>(pv, partials);
}

int grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

}  // namespace

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

int launch_predictor(dalm_ctx* ctx, double Re, double dt) {
    probe_begin(ctx, DALM_K_PREDICTOR);
    const int ni = ctx->d.gc >= 2 ? star_planes<kPredRows>(ctx->d) : 0;
    if (ni > 0)
        hipLaunchKernelGGL(k_predictor_star, dim3(star_grid<kPredRows>(ctx->d, ni)), dim3(TK * kPredRows), 0,
                           ctx->stream, ctx->d,
                           mesh(ctx), ctx->Uold[0], ctx->Uold[1], ctx->Uold[2], ctx->JU[0], ctx->JU[1], ctx->JU[2],
                           ctx->nut, ctx->f[0], ctx->f[1], ctx->f[2], ctx->U[0], ctx->U[1], ctx->U[2], 1 / Re, dt,
                           ni);
    else
        hipLaunchKernelGGL(k_predictor, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx),
                           ctx->Uold[0], ctx->Uold[1], ctx->Uold[2], ctx->JU[0], ctx->JU[1], ctx->JU[2],
                           ctx->nut, ctx->f[0], ctx->f[1], ctx->f[2], ctx->U[0], ctx->U[1], ctx->U[2],
                           1 / Re, dt);
    probe_end(ctx, DALM_K_PREDICTOR);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_copy_uold(dalm_ctx* ctx) {
    hipLaunchKernelGGL(k_copy3, dim3(grid_for(ctx->L)), dim3(256), 0, ctx->stream, ctx->L, ctx->U[0],
                       ctx->U[1], ctx->U[2], ctx->Uold[0], ctx->Uold[1], ctx->Uold[2]);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_ghost_copy(dalm_ctx* ctx) {
    hipLaunchKernelGGL(k_ghost_copy, dim3(grid_for(shell_count(ctx->d))), dim3(256), 0, ctx->stream,
                       ctx->d, ctx->Uold[0], ctx->Uold[1], ctx->Uold[2], ctx->U[0], ctx->U[1],
                       ctx->U[2]);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_flux(dalm_ctx* ctx, const double Uin[3], bool apply_bc) {
    long long n = (long long)(ctx->d.nx + 1) * (ctx->d.ny + 1) * (ctx->d.nz + 1);
    hipLaunchKernelGGL(k_flux, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->U[0], ctx->U[1], ctx->U[2], ctx->JU[0], ctx->JU[1], ctx->JU[2],
                       apply_bc ? 1 : 0, Uin ? Uin[0] : 0.0, ctx->inflow_face());
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_inflow(dalm_ctx* ctx, double t) {
    if (!ctx->inflow_n) return DALM_OK;
    const dalm_params& P = ctx->prm;
    hipLaunchKernelGGL(k_inflow, dim3(grid_for((long long)ctx->d.ny * ctx->d.nz)), dim3(256), 0, ctx->stream,
                       ctx->d, mesh(ctx), ctx->inflow_modes, ctx->inflow_n, ctx->inflow_speed * t, P.Uin[0],
                       P.Uin[1], P.Uin[2], ctx->inflow);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_jubc(dalm_ctx* ctx, const double Uin[3]) {
    const Dims& d = ctx->d;
    long long n = (long long)d.ny * d.nz + 2LL * d.nx * d.nz + 2LL * d.nx * d.ny;
    hipLaunchKernelGGL(k_jubc, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->JU[0], ctx->JU[1], ctx->JU[2], Uin[0], ctx->inflow_face());
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_rhs(dalm_ctx* ctx, double dt, double scale) {
    hipLaunchKernelGGL(k_rhs, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->b, dt * scale);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_pbc(dalm_ctx* ctx) {
    const Dims& d = ctx->d;
    long long n = 2LL * d.ny * d.nz + 2LL * d.nx * d.nz + 2LL * d.nx * d.ny;
    hipLaunchKernelGGL(k_pbc, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx), ctx->p);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_project(dalm_ctx* ctx, double dt, const double* Uin_bc) {
    long long n = (long long)(ctx->d.nx + 1) * (ctx->d.ny + 1) * (ctx->d.nz + 1);
    hipLaunchKernelGGL(k_project, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->p, ctx->U[0], ctx->U[1], ctx->U[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], dt,
                       Uin_bc ? 1 : 0, Uin_bc ? Uin_bc[0] : 0.0, ctx->inflow_face());
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_solid(dalm_ctx* ctx) {
    hipLaunchKernelGGL(k_solid, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->solid,
                       ctx->U[0], ctx->U[1], ctx->U[2]);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_ubc(dalm_ctx* ctx, const double Uin[3], double dt, int retracted, double floor_x) {
    const Dims& d = ctx->d;
    long long n = 2LL * d.ny * d.nz + 2LL * d.nx * d.nz + 2LL * d.nx * d.ny;
    hipLaunchKernelGGL(k_ubc, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->U[0], ctx->U[1], ctx->U[2], ctx->Uold[0], ctx->Uold[1], ctx->Uold[2],
                       Uin[0], Uin[1], Uin[2], dt, retracted, floor_x, ctx->inflow_cells());
    LAUNCH_CHECK();
    return DALM_OK;
}

// planes per unit of the post pass's column walk, or 0 for the direct kernel (too
// few units, or more partials than the partials buffer holds)
static int star_post_planes(const dalm_ctx* ctx) {
    if (ctx->d.gc < 2) return 0;
    const int ni = star_planes<TJ>(ctx->d);
    if (ni == 0) return 0;
    const long long cap = 8LL * (ctx->grid > ctx->sgrid ? ctx->grid : ctx->sgrid);
    return 2LL * star_grid<TJ>(ctx->d, ni) <= cap ? ni : 0;
}

int launch_post(dalm_ctx* ctx, double Cs, double dt, bool store_div) {
    const int ni = star_post_planes(ctx);
    if (ni > 0) {
        const int grid = star_grid<TJ>(ctx->d, ni);
        hipLaunchKernelGGL(k_post_star, dim3(grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), ctx->U[0],
                           ctx->U[1], ctx->U[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->nut,
                           store_div ? ctx->div : nullptr, Cs, dt, 1, ctx->partials, ni);
        LAUNCH_CHECK();
        return launch_finalize(ctx, FIN_POST, grid, nullptr, nullptr);
    }
    hipLaunchKernelGGL(k_post, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->U[0], ctx->U[1], ctx->U[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->nut,
                       store_div ? ctx->div : nullptr, Cs, dt, 1, ctx->partials, ctx->counter, ctx->scal);
    LAUNCH_CHECK();
    return launch_finalize(ctx, FIN_POST, ctx->grid, nullptr, nullptr);
}

int launch_nut(dalm_ctx* ctx, double Cs) {
    // nu_t only: reuse the post pass and ignore its reductions
    return launch_post(ctx, Cs, 0.0, false);
}

int launch_divergence(dalm_ctx* ctx) {
    hipLaunchKernelGGL(k_post, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->U[0], ctx->U[1], ctx->U[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->nut,
                       ctx->div, 0.0, 0.0, 0, ctx->partials, ctx->counter, ctx->scal);
    LAUNCH_CHECK();
    return launch_finalize(ctx, FIN_POST, ctx->grid, nullptr, nullptr);
}

int launch_l2(dalm_ctx* ctx, const double* v, double* out_host) {
    double* res = &ctx->solve->aux0;
    hipLaunchKernelGGL(k_scalar_sum, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, v, v,
                       ctx->partials, ctx->counter, res);
    LAUNCH_CHECK();
    if (int rc = launch_finalize(ctx, FIN_SUM, ctx->grid, nullptr, res)) return rc;
    double h = 0;
    HIP_TRY(hipMemcpyAsync(&h, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out_host = sqrt(h);
    return DALM_OK;
}

int launch_max_cfl(dalm_ctx* ctx, double dt, double* out_host) {
    hipLaunchKernelGGL(k_post, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx),
                       ctx->U[0], ctx->U[1], ctx->U[2], ctx->JU[0], ctx->JU[1], ctx->JU[2], ctx->nut,
                       nullptr, 0.0, dt, 0, ctx->partials, ctx->counter, ctx->scal);
    LAUNCH_CHECK();
    if (int rc = launch_finalize(ctx, FIN_POST, ctx->grid, nullptr, nullptr)) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->h_scal, ctx->scal, sizeof(StepScalars), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out_host = ctx->h_scal->max_cfl;
    return DALM_OK;
}
