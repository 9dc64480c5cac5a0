// adjoint.hip -- discrete adjoint (reverse mode) of the FP64 time step, and the
// checkpointed reverse sweep over a recorded run.
//
// The reference has no adjoint.  Each kernel below is the transpose of the
// Jacobian of one forward kernel (cfd.hip, poisson.hip, alm.hip), written as a
// gather: every output cell collects the contributions of the forward outputs that
// read it, so no two threads write one cell and the result does not depend on the
// launch order (the actuator-point scatter onto its 8 interpolation corners is the
// one exception and uses FP64 atomics).  Non-differentiable points take the branch
// the forward took.  The Poisson solve is differentiated as the exact solve
// p = A^-1 b; since V A is symmetric (V the cell volumes), A^-T = V A^-1 V^-1, so the
// adjoint solve reuses the forward solvers on V^-1 p_b.
//
// Reverse step n (state S_{n-1} -> S_n of src/solver.h:1042-1186, in reverse):
//   ALM forces (recomputed from U_n) -> projection bars -> AP force bars
//   -> nu_t bars -> face-BC bars -> velocity-BC bars -> projection bars (p_b)
//   -> pressure-BC fold and V^-1 -> adjoint Poisson solve -> RHS bars -> face BCs
//   -> face-flux bars -> predictor bars (U_{n-1}, JU_{n-1}, nu_t_{n-1}, f_{n-1}).
// Compiled with -ffp-contract=off.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "ctx.h"

namespace {

constexpr double PI_ = 3.14159265358979323846;
constexpr double CUT = 8.0;  // projection cutoff in units of eps (as alm.hip)

struct M1 {
    const double *x, *y, *z, *dx, *dy, *dz;
};
M1 mesh(const dalm_ctx* c) { return M1{c->x, c->y, c->z, c->dx, c->dy, c->dz}; }

struct C3 {
    const double* c[3];
};
struct V3 {
    double* c[3];
};
C3 c3(double* const a[3]) { return C3{{a[0], a[1], a[2]}}; }
V3 v3(double* const a[3]) { return V3{{a[0], a[1], a[2]}}; }

__device__ __forceinline__ double sgn(double v) { return (double)((v > 0.0) - (v < 0.0)); }

__device__ __forceinline__ void cell_of(const Dims& d, long long e, int& i, int& j, int& k) {
    i = (int)(e / d.si);
    int r = (int)(e - (long long)i * d.si);
    j = r / d.sj;
    k = r - j * d.sj;
}

// Partials of the MUSCL limiter h = t max(0, min(|a|, t bb b)), t = sgn(a).
__device__ __forceinline__ void lim_d(double a, double b, double bb, double& da, double& db) {
    da = 0;
    db = 0;
    double t = sgn(a);
    if (t == 0) return;
    double inner = fmin(fabs(a), t * bb * b);
    if (!(inner > 0)) return;
    if (fabs(a) <= t * bb * b) da = 1;  // h = a
    else db = bb;                       // h = bb b
}

// Reverse of one upwind MUSCL face flux (src/cfd.h:37-60): the face between w1
// and w2 with the 4-point line w0..w3 and face flux J.  Returns the bars of the
// four line values for a flux bar Fb, and the bar of J.
__device__ __forceinline__ void face_adj(double w0, double w1, double w2, double w3, double J, double Fb,
                                         double wb[4], double& Jb) {
    const double kap = 1. / 3.;
    const double bb = (3. - kap) / (1. - kap);
    const double q = 1. / 4.;
    double a2 = w1 - w0, a3 = w2 - w1, a4 = w3 - w2;
    double t2 = sgn(a2), t3 = sgn(a3), t4 = sgn(a4);
    double h6 = t4 * fmax(0., fmin(fabs(a4), t4 * bb * a3));
    double h5 = t3 * fmax(0., fmin(fabs(a3), t3 * bb * a4));
    double h4 = t3 * fmax(0., fmin(fabs(a3), t3 * bb * a2));
    double h3 = t2 * fmax(0., fmin(fabs(a2), t2 * bb * a3));
    double pr = w2 - q * ((1. - kap) * h6 + (1. + kap) * h5);
    double pl = w1 + q * ((1. - kap) * h3 + (1. + kap) * h4);
    Jb = Fb * 0.5 * ((pr + pl) - sgn(J) * (pr - pl));
    double prb = Fb * 0.5 * (J - fabs(J)), plb = Fb * 0.5 * (J + fabs(J));
    double h6b = -q * (1. - kap) * prb, h5b = -q * (1. + kap) * prb;
    double h3b = q * (1. - kap) * plb, h4b = q * (1. + kap) * plb;
    double a2b = 0, a3b = 0, a4b = 0, da, db;
    lim_d(a4, a3, bb, da, db); a4b += h6b * da; a3b += h6b * db;
    lim_d(a3, a4, bb, da, db); a3b += h5b * da; a4b += h5b * db;
    lim_d(a2, a3, bb, da, db); a2b += h3b * da; a3b += h3b * db;
    lim_d(a3, a2, bb, da, db); a3b += h4b * da; a2b += h4b * db;
    wb[0] = -a2b;
    wb[1] = plb + a2b - a3b;
    wb[2] = prb + a3b - a4b;
    wb[3] = a4b;
}

// ---------------------------------------------------------------------------
// Predictor (src/cfd.h:175-242) in reverse.  In: ub = bar of U* on the interior.
// Out (overwritten, every cell): un = bar of Uold (plus ub itself on ghosts when
// fold_ghost: U* shares Uold's ghost shell), jn = bar of JU on the live faces,
// nutn = bar of nu_t, fn = bar of f.  xp (optional) adds the x+ outflow bars.
__global__ __launch_bounds__(256) void ka_predictor(Dims d, M1 m, C3 ub, C3 uo, C3 ju,
                                                    const double* __restrict__ nut, double inv_re, double dt,
                                                    V3 un, V3 jn, double* __restrict__ nutn, V3 fn, int fold_ghost,
                                                    const double* __restrict__ xp) {
    const int g = d.gc;
    const int n[3] = {d.n0, d.n1, d.n2};
    const int st[3] = {d.si, d.sj, 1};
    const double* X[3] = {m.x, m.y, m.z};
    const double* W[3] = {m.dx, m.dy, m.dz};
    const long long L = (long long)d.n0 * d.si;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < L;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)e;
        int co[3];
        cell_of(d, e, co[0], co[1], co[2]);
        bool in[3];
        for (int a = 0; a < 3; ++a) in[a] = co[a] >= g && co[a] < n[a] - g;
        const int nout = (!in[0]) + (!in[1]) + (!in[2]);
        if (nout >= 2) {  // edges and corners: never read by the step's stencils
            for (int q = 0; q < 3; ++q) {
                un.c[q][c] = 0;
                jn.c[q][c] = 0;
                fn.c[q][c] = 0;
            }
            nutn[c] = 0;
            continue;
        }
        const bool ci = nout == 0;
        double acc[3], ubc[3];
        for (int q = 0; q < 3; ++q) {
            ubc[q] = ub.c[q][c];
            acc[q] = (ci || fold_ghost) ? ubc[q] : 0.0;
        }
        double jb[3] = {0, 0, 0};
        double nb = 0;
        for (int a = 0; a < 3; ++a) {
            const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
            if (!(in[b1] && in[b2])) continue;
            const int s_ = st[a], ic = co[a];
            const double wo = W[b1][co[b1]] * W[b2][co[b2]];
            // convection: the four faces whose 4-point line contains this cell
            for (int sh = -1; sh <= 2; ++sh) {
                const int l = ic - sh;  // face between cells l and l+1 along a
                const bool lown = l >= g && l < n[a] - g;
                const bool rown = l + 1 >= g && l + 1 < n[a] - g;
                if (!lown && !rown) continue;
                const int Lc = c - sh * s_;
                const double vl = wo * W[a][l], vr = wo * W[a][l + 1];
                const double Jf = ju.c[a][Lc];
                for (int q = 0; q < 3; ++q) {
                    double Fb = 0;
                    if (lown) Fb -= dt * ub.c[q][Lc] / vl;
                    if (rown) Fb += dt * ub.c[q][Lc + s_] / vr;
                    const double* o = uo.c[q];
                    double wb[4], Jbq;
                    face_adj(o[Lc - s_], o[Lc], o[Lc + s_], o[Lc + 2 * s_], Jf, Fb, wb, Jbq);
                    acc[q] += wb[sh + 1];
                    if (sh == 0) jb[a] += Jbq;
                }
            }
            // diffusion (src/cfd.h:144-173): this cell as the + neighbour of l = ic-1,
            // as the - neighbour of l = ic+1, and as the centre
            const double* x = X[a];
            const double* w = W[a];
            if (ic - 1 >= g && ic - 1 < n[a] - g) {
                const int l = ic - 1, cm = c - s_;
                const double coef = dt * (inv_re + nut[cm]) / ((x[l + 1] - x[l]) * w[l]);
                for (int q = 0; q < 3; ++q) acc[q] += coef * ub.c[q][cm];
            }
            if (ic + 1 >= g && ic + 1 < n[a] - g) {
                const int l = ic + 1, cp = c + s_;
                const double coef = dt * (inv_re + nut[cp]) / ((x[l] - x[l - 1]) * w[l]);
                for (int q = 0; q < 3; ++q) acc[q] += coef * ub.c[q][cp];
            }
            if (ci) {
                const double hp = x[ic + 1] - x[ic], hm = x[ic] - x[ic - 1];
                const double coef = dt * (inv_re + nut[c]) * (1 / hp + 1 / hm) / w[ic];
                for (int q = 0; q < 3; ++q) {
                    acc[q] -= coef * ubc[q];
                    const double* o = uo.c[q];
                    const double vc = o[c];
                    nb += dt * ubc[q] * (((o[c + s_] - vc) / hp - (vc - o[c - s_]) / hm) / w[ic]);
                }
            }
        }
        if (xp && in[1] && in[2] && co[0] >= d.n0 - g - 2) {
            const int plane = co[0] - (d.n0 - g - 2);
            const long long pl = (long long)d.ny * d.nz, o = (long long)(co[1] - g) * d.nz + (co[2] - g);
            for (int q = 0; q < 3; ++q) acc[q] += xp[(plane * 3 + q) * pl + o];
        }
        for (int q = 0; q < 3; ++q) {
            un.c[q][c] = acc[q];
            jn.c[q][c] = jb[q];
            fn.c[q][c] = ci ? dt * ubc[q] : 0.0;
        }
        nutn[c] = ci ? nb : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Smagorinsky nu_t (src/cfd.h:785-826) in reverse, two passes: per interior cell
// the 9 coefficients Q[q][a] = nut_b K dS/dG_qa / h_a, then every cell gathers
// U_b[q] += Q[q][a](c - e_a) - Q[q][a](c + e_a).
struct Q9 {
    double* q[9];
};
__global__ __launch_bounds__(BLOCK) void ka_nut_q(Dims d, M1 m, C3 u, const double* __restrict__ nutb, double Cs,
                                                  Q9 Q) {
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int i = cc.i, j = cc.j, k = cc.k;
        const int c = d.at(i, j, k);
        const double h[3] = {m.x[i + 1] - m.x[i - 1], m.y[j + 1] - m.y[j - 1], m.z[k + 1] - m.z[k - 1]};
        const int st[3] = {d.si, d.sj, 1};
        double G[3][3];
        for (int q = 0; q < 3; ++q)
            for (int a = 0; a < 3; ++a) G[q][a] = (u.c[q][c + st[a]] - u.c[q][c - st[a]]) / h[a];
        double S = sqrt(2 * (G[0][0] * G[0][0]) + 2 * (G[1][1] * G[1][1]) + 2 * (G[2][2] * G[2][2]) +
                        (G[0][1] + G[1][0]) * (G[0][1] + G[1][0]) + (G[0][2] + G[2][0]) * (G[0][2] + G[2][0]) +
                        (G[1][2] + G[2][1]) * (G[1][2] + G[2][1]));
        double l = Cs * cbrt(m.dx[i] * m.dy[j] * m.dz[k]);
        double nb = S > 0 ? nutb[c] * (l * l) / S : 0.0;
        for (int q = 0; q < 3; ++q)
            for (int a = 0; a < 3; ++a) {
                double dS = (q == a) ? 2 * G[q][q] : (G[q][a] + G[a][q]);
                Q.q[q * 3 + a][c] = nb * dS / h[a];
            }
    }
}

__global__ __launch_bounds__(256) void ka_nut_gather(Dims d, Q9 Q, V3 ub) {
    const int g = d.gc;
    const int n[3] = {d.n0, d.n1, d.n2};
    const int st[3] = {d.si, d.sj, 1};
    const long long L = (long long)d.n0 * d.si;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < L;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)e;
        int co[3];
        cell_of(d, e, co[0], co[1], co[2]);
        bool in[3];
        for (int a = 0; a < 3; ++a) in[a] = co[a] >= g && co[a] < n[a] - g;
        double acc[3] = {0, 0, 0};
        bool any = false;
        for (int a = 0; a < 3; ++a) {
            if (!(in[(a + 1) % 3] && in[(a + 2) % 3])) continue;
            const int ic = co[a];
            if (ic - 1 >= g && ic - 1 < n[a] - g) {
                for (int q = 0; q < 3; ++q) acc[q] += Q.q[q * 3 + a][c - st[a]];
                any = true;
            }
            if (ic + 1 >= g && ic + 1 < n[a] - g) {
                for (int q = 0; q < 3; ++q) acc[q] -= Q.q[q * 3 + a][c + st[a]];
                any = true;
            }
        }
        if (any)
            for (int q = 0; q < 3; ++q) ub.c[q][c] += acc[q];
    }
}

// ---------------------------------------------------------------------------
// Velocity ghosts (src/bc.h:5-233) in reverse, in place on U_b.  x+ columns: the
// ghost's bar becomes the bar of its own advecting velocity (U_pre.x), and the
// outflow's dependence on Uold goes to xp (4 planes i = n0-g-2 .. n0-1, written in
// full); y walls (both sides per (i, k) item) fold onto the two interior cells;
// x- ghosts are cleared.  z walls follow in ka_ubc_z.
__device__ __forceinline__ void quad3_coef(double x0, double x1, double x2, double x3, double l[3]) {
    l[0] = ((x3 - x1) * (x3 - x2)) / ((x0 - x1) * (x0 - x2));
    l[1] = ((x3 - x0) * (x3 - x2)) / ((x1 - x0) * (x1 - x2));
    l[2] = ((x3 - x0) * (x3 - x1)) / ((x2 - x0) * (x2 - x1));
}

__global__ void ka_ubc_xy(Dims d, M1 m, V3 ub, const double* __restrict__ upre_x, C3 uold, double dt,
                          double* __restrict__ xp) {
    const int g = d.gc;
    const long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz;
    const long long n = 2 * nyz + nxz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long s = t;
        if (s < nyz) {  // x+ column
            const int j = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            double acc[4][3] = {};
            for (int i = d.n0 - g; i < d.n0; ++i) {
                const int c = d.at(i, j, k);
                const double uo = upre_x[c];
                const int p0 = i - (d.n0 - g - 2);
                double uob = 0;
                for (int q = 0; q < 3; ++q) {
                    const double* O = uold.c[q];
                    const double a0 = O[c], a1 = O[c - d.si], a2 = O[c - 2 * d.si];
                    const double gr = (3 * a0 - 4 * a1 + a2) / 2;
                    const double b = ub.c[q][c];
                    acc[p0][q] += b;
                    uob -= b * dt * gr / m.dx[i];
                    const double grb = -b * uo * dt / m.dx[i];
                    acc[p0][q] += 1.5 * grb;
                    acc[p0 - 1][q] -= 2.0 * grb;
                    acc[p0 - 2][q] += 0.5 * grb;
                }
                ub.c[0][c] = uob;
                ub.c[1][c] = 0;
                ub.c[2][c] = 0;
            }
            const long long o = (long long)(j - g) * d.nz + (k - g);
            for (int p = 0; p < 4; ++p)
                for (int q = 0; q < 3; ++q) xp[(p * 3 + q) * nyz + o] = acc[p][q];
            continue;
        }
        s -= nyz;
        if (s < nxz) {  // y walls of one (i, k)
            const int i = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
            for (int side = 0; side < 2; ++side) {
                const int jin2 = side ? d.n1 - g - 2 : g + 1, jin1 = side ? d.n1 - g - 1 : g;
                const int jg1 = side ? d.n1 - g : g - 1, jg2 = side ? d.n1 - g + 1 : g - 2;
                const double yw = side ? m.y[jin1] + 0.5 * m.dy[jin1] : m.y[jin1] - 0.5 * m.dy[jin1];
                const int c1 = d.at(i, jin1, k), c2 = d.at(i, jin2, k);
                const int gh[2] = {d.at(i, jg1, k), d.at(i, jg2, k)};
                const double yg[2] = {m.y[jg1], m.y[jg2]};
                for (int o = 0; o < 2; ++o) {
                    double l[3];
                    quad3_coef(m.y[jin2], m.y[jin1], yw, yg[o], l);
                    for (int q = 0; q < 3; ++q) {
                        const double b = ub.c[q][gh[o]];
                        ub.c[q][c2] += b * l[0];
                        ub.c[q][c1] += b * (q != 1 ? l[1] + l[2] : l[1]);
                        ub.c[q][gh[o]] = 0;
                    }
                }
            }
            continue;
        }
        s -= nxz;  // x- ghosts of one (j, k)
        const int j = g + (int)(s / d.nz), k = g + (int)(s % d.nz);
        for (int i = 0; i < g; ++i) {
            const int c = d.at(i, j, k);
            ub.c[0][c] = 0;
            ub.c[1][c] = 0;
            ub.c[2][c] = 0;
        }
    }
}

__global__ void ka_ubc_z(Dims d, M1 m, V3 ub, int retracted, double floor_x) {
    const int g = d.gc;
    const long long n = (long long)d.nx * d.ny;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = g + (int)(t / d.ny), j = g + (int)(t % d.ny);
        for (int side = 0; side < 2; ++side) {
            const int kin2 = side ? d.n2 - g - 2 : g + 1, kin1 = side ? d.n2 - g - 1 : g;
            const int kg1 = side ? d.n2 - g : g - 1, kg2 = side ? d.n2 - g + 1 : g - 2;
            const double zw = side ? m.z[kin1] + 0.5 * m.dz[kin1] : m.z[kin1] - 0.5 * m.dz[kin1];
            const int c1 = d.at(i, j, kin1), c2 = d.at(i, j, kin2);
            const bool slip = side || (retracted && m.x[i] < floor_x);
            const int gh[2] = {d.at(i, j, kg1), d.at(i, j, kg2)};
            const double zg[2] = {m.z[kg1], m.z[kg2]};
            for (int o = 0; o < 2; ++o) {
                double l[3];
                quad3_coef(m.z[kin2], m.z[kin1], zw, zg[o], l);
                for (int q = 0; q < 3; ++q) {
                    const double b = ub.c[q][gh[o]];
                    ub.c[q][c2] += b * l[0];
                    ub.c[q][c1] += b * ((slip && q < 2) ? l[1] + l[2] : l[1]);
                    ub.c[q][gh[o]] = 0;
                }
            }
        }
    }
}

// add the x+ outflow bars (4 planes) into a bar field
__global__ void ka_xp_add(Dims d, const double* __restrict__ xp, V3 out) {
    const long long nyz = (long long)d.ny * d.nz;
    const long long n = 4 * nyz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int p = (int)(t / nyz);
        const long long o = t % nyz;
        const int j = d.gc + (int)(o / d.nz), k = d.gc + (int)(o % d.nz);
        const int c = d.at(d.n0 - d.gc - 2 + p, j, k);
        for (int q = 0; q < 3; ++q) out.c[q][c] += xp[(p * 3 + q) * nyz + o];
    }
}

// set_solid_U (src/cfd.h:1143-1160) in reverse
__global__ __launch_bounds__(BLOCK) void ka_solid(Dims d, const double* __restrict__ solid, V3 ub) {
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int c = d.at(cc.i, cc.j, cc.k);
        const double w = 1. - solid[c];
        for (int q = 0; q < 3; ++q) ub.c[q][c] *= w;
    }
}

// ---------------------------------------------------------------------------
// Face-flux boundary conditions (src/bc.h:235-302) in reverse: zero the bars of
// the faces they overwrite.
__device__ __forceinline__ bool bc_face(const Dims& d, int a, int i, int j, int k) {
    const int g = d.gc;
    const bool ii = i >= g && i < d.n0 - g, ji = j >= g && j < d.n1 - g, ki = k >= g && k < d.n2 - g;
    if (a == 0) return i == g - 1 && ji && ki;
    if (a == 1) return (j == g - 1 || j == d.n1 - g - 1) && ii && ki;
    return (k == g - 1 || k == d.n2 - g - 1) && ii && ji;
}

__global__ void ka_jubc(Dims d, V3 jb) {
    const int g = d.gc;
    const long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz, nxy = (long long)d.nx * d.ny;
    const long long n = nyz + 2 * nxz + 2 * nxy;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long s = t;
        if (s < nyz) {
            jb.c[0][d.at(g - 1, g + (int)(s / d.nz), g + (int)(s % d.nz))] = 0;
            continue;
        }
        s -= nyz;
        if (s < 2 * nxz) {
            const int side = (int)(s / nxz);
            s %= nxz;
            jb.c[1][d.at(g + (int)(s / d.nz), side ? d.n1 - g - 1 : g - 1, g + (int)(s % d.nz))] = 0;
            continue;
        }
        s -= 2 * nxz;
        const int side = (int)(s / nxy);
        s %= nxy;
        jb.c[2][d.at(g + (int)(s / d.ny), g + (int)(s % d.ny), side ? d.n2 - g - 1 : g - 1)] = 0;
    }
}

// Projection (src/cfd.h:600-783) in reverse: p_b = -dt (Gc^T U_b + Gf^T JU_b) on
// the cells [g-1, n-g+1) in each direction (interior plus one ghost layer);
// overwritten there.
__global__ void ka_project(Dims d, M1 m, C3 ub, C3 jb, double dt, double* __restrict__ pb) {
    const int g = d.gc;
    const int n[3] = {d.n0, d.n1, d.n2};
    const int st[3] = {d.si, d.sj, 1};
    const double* X[3] = {m.x, m.y, m.z};
    const double* W[3] = {m.dx, m.dy, m.dz};
    const int e0 = d.nx + 2, e1 = d.ny + 2, e2 = d.nz + 2;
    const long long total = (long long)e0 * e1 * e2;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int co[3];
        co[2] = g - 1 + (int)(t % e2);
        long long r = t / e2;
        co[1] = g - 1 + (int)(r % e1);
        co[0] = g - 1 + (int)(r / e1);
        const int c = d.at(co[0], co[1], co[2]);
        bool in[3], fb[3];
        for (int a = 0; a < 3; ++a) {
            in[a] = co[a] >= g && co[a] < n[a] - g;
            fb[a] = co[a] >= g - 1 && co[a] < n[a] - g;  // face-box range
        }
        double v = 0;
        for (int a = 0; a < 3; ++a) {
            const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
            const int ic = co[a];
            const double* x = X[a];
            if (in[b1] && in[b2]) {
                if (ic + 1 >= g && ic + 1 < n[a] - g) v += dt * ub.c[a][c + st[a]] / (x[ic + 2] - x[ic]);
                if (ic - 1 >= g && ic - 1 < n[a] - g) v -= dt * ub.c[a][c - st[a]] / (x[ic] - x[ic - 2]);
            }
            if (fb[b1] && fb[b2]) {
                const double area = W[b1][co[b1]] * W[b2][co[b2]];
                if (fb[a]) v += dt * area * jb.c[a][c] / (x[ic + 1] - x[ic]);
                if (ic - 1 >= g - 1 && ic - 1 < n[a] - g) v -= dt * area * jb.c[a][c - st[a]] / (x[ic] - x[ic - 1]);
            }
        }
        pb[c] = v;
    }
}

// Pressure ghosts (src/bc.h:304-411) in reverse, fused with the volume scaling of
// the adjoint solve: out = (p_b + folded ghost bars) / V on the interior, and the
// sum of out^2 per block.  fold = 0 skips the ghost bars, scale_v = 0 the volume.
// Each thread reads only its own interior cell and ghosts, so out may be pb.
__global__ __launch_bounds__(BLOCK) void ka_pbc_fold(Dims d, M1 m, const double* pb, double* out, int fold,
                                                     int scale_v, double* partials) {
    const int g = d.gc;
    TileRange tr = block_tiles(d);
    double v2[1] = {0.0};
    const double ratio = (0.5 * m.dx[d.n0 - g]) / (0.5 * m.dx[d.n0 - g - 1]);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int i = cc.i, j = cc.j, k = cc.k;
        const int c = d.at(i, j, k);
        double v = pb[c];
        if (fold) {
            if (i == g) v += pb[c - d.si];
            if (i == d.n0 - g - 1) v -= pb[c + d.si] * ratio;
            if (j == g) v += pb[c - d.sj];
            if (j == d.n1 - g - 1) v += pb[c + d.sj];
            if (k == g) v += pb[c - 1];
            if (k == d.n2 - g - 1) v += pb[c + 1];
        }
        if (scale_v) v /= m.dx[i] * m.dy[j] * m.dz[k];
        out[c] = v;
        v2[0] += v * v;
    }
    block_sum<1>(v2);
    double pv[1] = {v2[0]};
    store_partials<1>(pv, partials);
}

// clear the ghosts of a folded p_b (one layer of each face)
__global__ void ka_pbc_clear(Dims d, double* __restrict__ pb) {
    const int g = d.gc;
    const long long nyz = (long long)d.ny * d.nz, nxz = (long long)d.nx * d.nz, nxy = (long long)d.nx * d.ny;
    const long long n = 2 * nyz + 2 * nxz + 2 * nxy;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long s = t;
        if (s < 2 * nyz) {
            const int side = (int)(s / nyz);
            s %= nyz;
            pb[d.at(side ? d.n0 - g : g - 1, g + (int)(s / d.nz), g + (int)(s % d.nz))] = 0;
            continue;
        }
        s -= 2 * nyz;
        if (s < 2 * nxz) {
            const int side = (int)(s / nxz);
            s %= nxz;
            pb[d.at(g + (int)(s / d.nz), side ? d.n1 - g : g - 1, g + (int)(s % d.nz))] = 0;
            continue;
        }
        s -= 2 * nxz;
        const int side = (int)(s / nxy);
        s %= nxy;
        pb[d.at(g + (int)(s / d.ny), g + (int)(s % d.ny), side ? d.n2 - g : g - 1)] = 0;
    }
}

// V y (interior) after the adjoint solve: the bar of b
__global__ __launch_bounds__(BLOCK) void ka_scale_v(Dims d, M1 m, const double* __restrict__ y,
                                                    double* __restrict__ out) {
    TileRange tr = block_tiles(d);
    for (int t = tr.begin; t < tr.end; ++t) {
        Cell cc = tile_cell(d, t);
        if (!cc.inside) continue;
        const int c = d.at(cc.i, cc.j, cc.k);
        out[c] = y[c] * (m.dx[cc.i] * m.dy[cc.j] * m.dz[cc.k]);
    }
}

// Poisson RHS (src/cfd.h:569-598) in reverse over the face box:
// JU_b[a](c) += (s(c) - s(c + e_a)) / (dt max_diag), s = b_b / V on the interior
// (or s = y directly when y_given), zero outside; zero_bc then clears the
// boundary faces (the face BCs that precede the RHS).
__global__ void ka_rhs(Dims d, M1 m, const double* __restrict__ s_in, int y_given, double inv_dts, int zero_bc,
                       V3 jb) {
    const int g = d.gc;
    const int n[3] = {d.n0, d.n1, d.n2};
    const int st[3] = {d.si, d.sj, 1};
    const long long total = (long long)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int co[3];
        const int ez = d.nz + 1, ey = d.ny + 1;
        co[2] = g - 1 + (int)(t % ez);
        long long r = t / ez;
        co[1] = g - 1 + (int)(r % ey);
        co[0] = g - 1 + (int)(r / ey);
        const int c = d.at(co[0], co[1], co[2]);
        bool in[3];
        for (int a = 0; a < 3; ++a) in[a] = co[a] >= g && co[a] < n[a] - g;
        const bool ci = in[0] && in[1] && in[2];
        double sc = 0;
        if (ci) sc = y_given ? s_in[c] : s_in[c] / (m.dx[co[0]] * m.dy[co[1]] * m.dz[co[2]]);
        for (int a = 0; a < 3; ++a) {
            int cn[3] = {co[0], co[1], co[2]};
            cn[a] += 1;
            const bool ni = cn[a] >= g && cn[a] < n[a] - g && in[(a + 1) % 3] && in[(a + 2) % 3];
            double sn = 0;
            if (ni) sn = y_given ? s_in[c + st[a]] : s_in[c + st[a]] / (m.dx[cn[0]] * m.dy[cn[1]] * m.dz[cn[2]]);
            double v = jb.c[a][c] + (sc - sn) * inv_dts;
            if (zero_bc && bc_face(d, a, co[0], co[1], co[2])) v = 0;
            jb.c[a][c] = v;
        }
    }
}

// Face fluxes (interpolate_JU, src/cfd.h:394-420) in reverse on the cells
// [g-1, n-g+1): U_b[a] += A/2 (JU_b[a](c) + JU_b[a](c - e_a)) over face-box faces.
__global__ void ka_interp(Dims d, M1 m, C3 jb, V3 ub) {
    const int g = d.gc;
    const int n[3] = {d.n0, d.n1, d.n2};
    const int st[3] = {d.si, d.sj, 1};
    const double* W[3] = {m.dx, m.dy, m.dz};
    const int e0 = d.nx + 2, e1 = d.ny + 2, e2 = d.nz + 2;
    const long long total = (long long)e0 * e1 * e2;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int co[3];
        co[2] = g - 1 + (int)(t % e2);
        long long r = t / e2;
        co[1] = g - 1 + (int)(r % e1);
        co[0] = g - 1 + (int)(r / e1);
        const int c = d.at(co[0], co[1], co[2]);
        bool fb[3];
        for (int a = 0; a < 3; ++a) fb[a] = co[a] >= g - 1 && co[a] < n[a] - g;
        for (int a = 0; a < 3; ++a) {
            const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
            if (!(fb[b1] && fb[b2])) continue;
            const double half_area = 0.5 * W[b1][co[b1]] * W[b2][co[b2]];
            double v = 0;
            if (fb[a]) v += jb.c[a][c];
            if (co[a] - 1 >= g - 1 && co[a] - 1 < n[a] - g) v += jb.c[a][c - st[a]];
            ub.c[a][c] += half_area * v;
        }
    }
}

// ---------------------------------------------------------------------------
// ALM in reverse.  Rotation matrix of alm.hip's rotate() and its derivative.
__device__ __forceinline__ void rot_mat(double ang, int type, double M[3][3], double dM[3][3]) {
    const double c = cos(ang), s = sin(ang);
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) M[r][q] = dM[r][q] = 0;
    if (type == 0) {
        M[0][0] = 1; M[1][1] = c; M[1][2] = s; M[2][1] = -s; M[2][2] = c;
        dM[1][1] = -s; dM[1][2] = c; dM[2][1] = -c; dM[2][2] = -s;
    } else if (type == 1) {
        M[0][0] = c; M[0][2] = -s; M[1][1] = 1; M[2][0] = -s; M[2][2] = c;
        dM[0][0] = -s; dM[0][2] = -c; dM[2][0] = -c; dM[2][2] = -s;
    } else if (type == 2) {
        M[0][0] = c; M[0][1] = s; M[1][0] = -s; M[1][1] = c; M[2][2] = 1;
        dM[0][0] = -s; dM[0][1] = c; dM[1][0] = -c; dM[1][1] = -s;
    } else {
        M[0][0] = M[1][1] = M[2][2] = 1;
    }
}

// Jacobian of alm.hip's rotate_U (quirks included) in position, velocity, angle.
__device__ __forceinline__ void rotate_U_jac(const double xyz[3], const double U[3], double ang, double ow,
                                             int type, double dX[3][3], double dU[3][3], double dA[3]) {
    const double c = cos(ang), s = sin(ang);
    const double x = xyz[0], y = xyz[1], z = xyz[2];
    const double u = U[0], v = U[1], w = U[2];
    for (int r = 0; r < 3; ++r) {
        dA[r] = 0;
        for (int q = 0; q < 3; ++q) dX[r][q] = dU[r][q] = 0;
    }
    if (type == 0) {
        dU[0][0] = 1;
        dX[1][1] = -ow * s; dX[1][2] = ow * c; dU[1][1] = c; dU[1][2] = s;
        dA[1] = ow * (-c * y - s * z) + (-s * v + c * w);
        dX[2][1] = -ow * c; dX[2][2] = -ow * s; dU[2][1] = -s + w; dU[2][2] = v;
        dA[2] = ow * (s * y - c * z) + (-c * v);
    } else if (type == 1) {
        dX[0][0] = -ow * s; dX[0][2] = -ow * c; dU[0][0] = c; dU[0][2] = -s;
        dA[0] = ow * (-c * x + s * z) + (-s * u - c * w);
        dU[1][1] = 1;
        dX[2][0] = ow * c; dX[2][2] = -ow * s; dU[2][0] = s; dU[2][2] = c;
        dA[2] = ow * (-s * x - c * z) + (c * u - s * w);
    } else if (type == 2) {
        dX[0][0] = -ow * s; dX[0][1] = ow * c; dU[0][0] = c; dU[0][1] = s;
        dA[0] = ow * (-c * x - s * y) + (-s * u + c * v);
        dX[1][1] = -ow * s; dU[1][0] = -s; dU[1][1] = c;
        dA[1] = ow * (s * s - c * c - c * y) + (-c * u - s * v);
        dU[2][2] = 1;
    } else {
        dU[0][0] = dU[1][1] = dU[2][2] = 1;
    }
}

__device__ __forceinline__ void rotate_U(const double xyz[3], const double U[3], double out[3], double ang,
                                         double ow, int type) {
    const double c = cos(ang), s = sin(ang);
    const double x = xyz[0], y = xyz[1], z = xyz[2];
    const double u = U[0], v = U[1], w = U[2];
    if (type == 0) {
        out[0] = 0. + u;
        out[1] = ow * (-s * y + c * z) + (c * v + s * w);
        out[2] = ow * (-c * y - s * z) + (-s * v + v * w);
    } else if (type == 1) {
        out[0] = ow * (-s * x - c * z) + (c * u - s * w);
        out[1] = 0. + v;
        out[2] = ow * (c * x - s * z) + (s * u + c * w);
    } else if (type == 2) {
        out[0] = ow * (-s * x + c * y) + (c * u + s * v);
        out[1] = ow * (-c * s - s * y) + (-s * u + c * v);
        out[2] = 0. + w;
    } else {
        out[0] = u; out[1] = v; out[2] = w;
    }
}

__device__ __forceinline__ int floor_index(const double* a, double v, int n) {
    if (v < a[0]) return -1;
    if (v >= a[n - 1]) return n - 1;
    int lo = 0, hi = n - 1, hit = -1;
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        if (a[mid] <= v) {
            hit = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    return hit;
}

// Projection (src/alm.h:416-448) in reverse, one block per AP over the cells
// within the cutoff: red[6a..6a+2] = sum w f_b, red[6a+3..6a+5] = sum w (2/eps^2)
// (x_c - x_a) (f_b . F_a).
__global__ __launch_bounds__(BLOCK) void ka_ap_reduce(Dims d, M1 m, const double* __restrict__ ap, C3 fb,
                                                      double norm, double w2, double reach, double cut2,
                                                      double* __restrict__ red) {
    const int a = blockIdx.x;
    const double* P = ap + DALM_AP_STRIDE * (long long)a;
    const double xa = P[0], ya = P[1], za = P[2], F0 = P[8], F1 = P[9], F2 = P[10];
    const int g = d.gc;
    const int i0 = max(g, floor_index(m.x, xa - reach, d.n0) + 1), i1 = min(d.n0 - g, floor_index(m.x, xa + reach, d.n0) + 1);
    const int j0 = max(g, floor_index(m.y, ya - reach, d.n1) + 1), j1 = min(d.n1 - g, floor_index(m.y, ya + reach, d.n1) + 1);
    const int k0 = max(g, floor_index(m.z, za - reach, d.n2) + 1), k1 = min(d.n2 - g, floor_index(m.z, za + reach, d.n2) + 1);
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i1 > i0 && j1 > j0 && k1 > k0) {
        const long long nk = k1 - k0, nj = j1 - j0, n = (long long)(i1 - i0) * nj * nk;
        for (long long t = threadIdx.x; t < n; t += blockDim.x) {
            const int k = k0 + (int)(t % nk);
            const long long r = t / nk;
            const int j = j0 + (int)(r % nj);
            const int i = i0 + (int)(r / nj);
            const double ex = m.x[i] - xa, ey = m.y[j] - ya, ez = m.z[k] - za;
            const double r2 = ex * ex + ey * ey + ez * ez;
            if (r2 > cut2) continue;
            const int c = d.at(i, j, k);
            const double f0 = fb.c[0][c], f1 = fb.c[1][c], f2 = fb.c[2][c];
            const double wgt = norm * exp(-r2 / w2);
            v[0] += wgt * f0;
            v[1] += wgt * f1;
            v[2] += wgt * f2;
            const double s = wgt * (f0 * F0 + f1 * F1 + f2 * F2);
            v[3] += s * ex;
            v[4] += s * ey;
            v[5] += s * ez;
        }
    }
    block_sum<6>(v);
    if (threadIdx.x == 0) {
        double* o = red + 6LL * a;
        o[0] = v[0];
        o[1] = v[1];
        o[2] = v[2];
        o[3] = v[3] * 2 / w2;
        o[4] = v[4] * 2 / w2;
        o[5] = v[5] * 2 / w2;
    }
}

// AP positions and forces (src/alm.h:247-371) in reverse, one thread per AP:
// velocity bars go to the 8 interpolation corners (atomics: APs may share
// corners), parameter bars to apg[4a..4a+3] (turbine angle, pitch, chord, twist).
__global__ void ka_ap_force(Dims d, M1 m, C3 U, const double* __restrict__ wt, const double* __restrict__ ap, int n,
                            int blade_per_wt, int ap_per_blade, const double* __restrict__ cdt,
                            const double* __restrict__ clt, const double* __restrict__ atk, int natk,
                            const double* __restrict__ red, const double* __restrict__ tqw, V3 ub,
                            double* __restrict__ apg) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int per_wt = ap_per_blade * blade_per_wt;
    const double* T = wt + DALM_WT_STRIDE * (a / per_wt);
    const double* P = ap + DALM_AP_STRIDE * (long long)a;
    const int type = (int)T[6];
    const int g = d.gc;
    double Pb[3] = {red[6LL * a + 3], red[6LL * a + 4], red[6LL * a + 5]};
    double angb = 0, pitchb = 0, chordb = 0, twistb = 0;
    const int i = floor_index(m.x, P[0], d.n0);
    const int j = floor_index(m.y, P[1], d.n1);
    const int k = floor_index(m.z, P[2], d.n2);
    if (i >= g && i < d.n0 - g && j >= g && j < d.n1 - g && k >= g && k < d.n2 - g) {
        const double hx = m.x[i + 1] - m.x[i], hy = m.y[j + 1] - m.y[j], hz = m.z[k + 1] - m.z[k];
        const double fx = (P[0] - m.x[i]) / hx, fy = (P[1] - m.y[j]) / hy, fz = (P[2] - m.z[k]) / hz;
        int cell[8];
        for (int o = 0; o < 8; ++o) cell[o] = d.at(i + (o & 1), j + ((o >> 1) & 1), k + ((o >> 2) & 1));
        double v[3][8], Uap[3];
        for (int q = 0; q < 3; ++q) {
            for (int o = 0; o < 8; ++o) v[q][o] = U.c[q][cell[o]];
            double v01 = v[q][0] * (1 - fx) + v[q][1] * fx, v23 = v[q][2] * (1 - fx) + v[q][3] * fx;
            double v45 = v[q][4] * (1 - fx) + v[q][5] * fx, v67 = v[q][6] * (1 - fx) + v[q][7] * fx;
            double lo = v01 * (1 - fy) + v23 * fy, hi = v45 * (1 - fy) + v67 * fy;
            Uap[q] = lo * (1 - fz) + hi * fz;
        }
        const double rel[3] = {P[0] - T[0], P[1] - T[1], P[2] - T[2]};
        double Ur[3];
        rotate_U(rel, Uap, Ur, T[7], T[8], type);
        const double sth = sin(P[6]), cth = cos(P[6]);
        const double ux = Ur[0];
        const double ut = P[3] * T[13] + Ur[1] * sth - Ur[2] * cth;
        const double V2 = ux * ux + ut * ut;
        const double phi = atan(ux / ut);
        const double aoa = phi * 180 / PI_ - P[5] - P[13];
        const double* CD = cdt + (long long)natk * a;
        const double* CL = clt + (long long)natk * a;
        const int s = floor_index(atk, aoa, natk);
        double cd, cl, scd = 0, scl = 0;
        if (s < 0) {
            cd = CD[0]; cl = CL[0];
        } else if (s >= natk - 1) {
            cd = CD[natk - 1]; cl = CL[natk - 1];
        } else {
            const double h = atk[s + 1] - atk[s];
            const double wv = (aoa - atk[s]) / h;
            cd = CD[s] * (1 - wv) + CD[s + 1] * wv;
            cl = CL[s] * (1 - wv) + CL[s + 1] * wv;
            scd = (CD[s + 1] - CD[s]) / h;
            scl = (CL[s + 1] - CL[s]) / h;
        }
        const double K = 0.5 * V2 * P[4] * P[7];
        const double Fd = cd * K, Fl = cl * K;
        const double cphi = cos(phi), sphi = sin(phi);
        const double Fx = Fl * cphi + Fd * sphi;
        const double sg = sgn(T[13]);
        const double Ft = (Fl * sphi - Fd * cphi) * sg;
        double M[3][3], dM[3][3];
        rot_mat(-T[7], type, M, dM);
        const double loc[3] = {-Fx, Ft * sth, -Ft * cth};
        const double Fbar[3] = {red[6LL * a], red[6LL * a + 1], red[6LL * a + 2]};
        double locb[3];
        for (int r = 0; r < 3; ++r) {
            locb[r] = M[0][r] * Fbar[0] + M[1][r] * Fbar[1] + M[2][r] * Fbar[2];
            for (int q = 0; q < 3; ++q) angb -= Fbar[q] * dM[q][r] * loc[r];
        }
        const double Fxb = -locb[0];
        const double Ftb = locb[1] * sth - locb[2] * cth + tqw[a / per_wt] * sgn(Ft) * P[3];
        double Flb = Ftb * sg * sphi + Fxb * cphi;
        double Fdb = -Ftb * sg * cphi + Fxb * sphi;
        double phib = Ftb * sg * (Fl * cphi + Fd * sphi) + Fxb * (-Fl * sphi + Fd * cphi);
        const double cdb = Fdb * K, clb = Flb * K, Kb = Fdb * cd + Flb * cl;
        const double V2b = Kb * 0.5 * P[4] * P[7];
        chordb = Kb * 0.5 * V2 * P[7];
        const double aoab = cdb * scd + clb * scl;
        phib += aoab * 180 / PI_;
        twistb = -aoab;
        pitchb = -aoab;
        const double uxb = 2 * ux * V2b + phib * ut / V2;
        const double utb = 2 * ut * V2b - phib * ux / V2;
        const double Urb[3] = {uxb, utb * sth, -utb * cth};
        double dX[3][3], dU[3][3], dA[3];
        rotate_U_jac(rel, Uap, T[7], T[8], type, dX, dU, dA);
        double Uapb[3] = {0, 0, 0};
        for (int r = 0; r < 3; ++r) {
            angb += Urb[r] * dA[r];
            for (int q = 0; q < 3; ++q) {
                Uapb[q] += Urb[r] * dU[r][q];
                Pb[q] += Urb[r] * dX[r][q];
            }
        }
        double fxb = 0, fyb = 0, fzb = 0;
        for (int q = 0; q < 3; ++q) {
            const double* vv = v[q];
            const double b = Uapb[q];
            for (int o = 0; o < 8; ++o) {
                const double wx = (o & 1) ? fx : 1 - fx, wy = ((o >> 1) & 1) ? fy : 1 - fy;
                const double wz = ((o >> 2) & 1) ? fz : 1 - fz;
                atomicAdd(&ub.c[q][cell[o]], b * wx * wy * wz);
            }
            fxb += b * (((vv[1] - vv[0]) * (1 - fy) + (vv[3] - vv[2]) * fy) * (1 - fz) +
                        ((vv[5] - vv[4]) * (1 - fy) + (vv[7] - vv[6]) * fy) * fz);
            const double lx0 = vv[0] * (1 - fx) + vv[1] * fx, lx1 = vv[2] * (1 - fx) + vv[3] * fx;
            const double hx0 = vv[4] * (1 - fx) + vv[5] * fx, hx1 = vv[6] * (1 - fx) + vv[7] * fx;
            fyb += b * ((lx1 - lx0) * (1 - fz) + (hx1 - hx0) * fz);
            fzb += b * ((hx0 * (1 - fy) + hx1 * fy) - (lx0 * (1 - fy) + lx1 * fy));
        }
        Pb[0] += fxb / hx;
        Pb[1] += fyb / hy;
        Pb[2] += fzb / hz;
    }
    // position: P = base + rotate(loc_p, -angle)
    {
        const double lp[3] = {T[3], T[4] + P[3] * cos(P[6]), T[5] + P[3] * sin(P[6])};
        double M[3][3], dM[3][3];
        rot_mat(-T[7], type, M, dM);
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) angb -= Pb[r] * dM[r][q] * lp[q];
    }
    double* G = apg + 4LL * a;
    G[0] += angb;
    G[1] += pitchb;
    G[2] += chordb;
    G[3] += twistb;
}

int grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

Q9 q9(dalm_ctx* ctx) {
    Q9 Q;
    double* v[9] = {ctx->r, ctx->r0, ctx->kp, ctx->pp, ctx->q, ctx->s, ctx->ss, ctx->t, ctx->tmp};
    for (int q = 0; q < 9; ++q) Q.q[q] = v[q];
    return Q;
}

}  // namespace

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

// ---------------------------------------------------------------------------
// launchers
namespace {

int alloc_dev(dalm_ctx* ctx, double** p, size_t n) {
    HIP_TRY(hipMalloc((void**)p, n * sizeof(double) + 256));
    HIP_TRY(hipMemsetAsync(*p, 0, n * sizeof(double) + 256, ctx->stream));
    return DALM_OK;
}

}  // namespace

int ensure_adjoint_buffers(dalm_ctx* ctx) {
    AdjointState& S = ctx->adj;
    if (S.have_buffers) return DALM_OK;
    if (ctx->precision != DALM_FP64) return dalm_fail_msg(ctx, DALM_BAD_ARG, "the adjoint needs an FP64 context");
    if (ctx->d.gc != 2) return dalm_fail_msg(ctx, DALM_BAD_ARG, "the adjoint supports gc == 2");
    const size_t L = (size_t)ctx->L;
    for (int q = 0; q < 3; ++q) {
        double** v[6] = {&S.Ub[q], &S.Ubn[q], &S.JUb[q], &S.JUbn[q], &S.fb[q], &S.fbn[q]};
        for (double** p : v)
            if (int rc = alloc_dev(ctx, p, L)) return rc;
    }
    double** s1[5] = {&S.nutb, &S.nutbn, &S.pb, &S.ab, &S.ya};
    for (double** p : s1)
        if (int rc = alloc_dev(ctx, p, L)) return rc;
    if (int rc = alloc_dev(ctx, &S.xp, 12 * (size_t)ctx->d.ny * ctx->d.nz)) return rc;
    if (int rc = alloc_dev(ctx, &S.tqw, DALM_MAX_TURBINES)) return rc;
    S.have_buffers = true;
    return DALM_OK;
}

// per-AP scratch, sized for the current actuator-line setup
static int ensure_ap_buffers(dalm_ctx* ctx) {
    AdjointState& S = ctx->adj;
    const int na = ctx->alm.ap_count > 0 ? ctx->alm.ap_count : 1;
    if (na <= S.ap_cap) return DALM_OK;
    if (S.apred) (void)hipFree(S.apred);
    if (S.apg) (void)hipFree(S.apg);
    S.apred = S.apg = nullptr;
    S.ap_cap = 0;
    if (int rc = alloc_dev(ctx, &S.apred, 6 * (size_t)na)) return rc;
    if (int rc = alloc_dev(ctx, &S.apg, 4 * (size_t)na)) return rc;
    S.ap_cap = na;
    return DALM_OK;
}

namespace {

int zero(dalm_ctx* ctx, double* a, size_t n) {
    HIP_TRY(hipMemsetAsync(a, 0, n * sizeof(double), ctx->stream));
    return DALM_OK;
}

int launch_predictor_adj(dalm_ctx* ctx, double* const Uold[3], double* const JU[3], const double* nut, double Re,
                         double dt, int fold_ghost, const double* xp) {
    AdjointState& S = ctx->adj;
    hipLaunchKernelGGL(ka_predictor, dim3(grid_for(ctx->L)), dim3(256), 0, ctx->stream, ctx->d, mesh(ctx), c3(S.Ub),
                       c3(Uold), c3(JU), nut, 1 / Re, dt, v3(S.Ubn), v3(S.JUbn), S.nutbn, v3(S.fbn), fold_ghost, xp);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_nut_adj(dalm_ctx* ctx, double* const U[3], double Cs) {
    AdjointState& S = ctx->adj;
    Q9 Q = q9(ctx);
    hipLaunchKernelGGL(ka_nut_q, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), c3(U), S.nutb, Cs,
                       Q);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(ka_nut_gather, dim3(grid_for(ctx->L)), dim3(256), 0, ctx->stream, ctx->d, Q, v3(S.Ub));
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_ubc_adj(dalm_ctx* ctx, const double* upre_x, double* const Uold[3], double dt, int retracted,
                   double floor_x) {
    AdjointState& S = ctx->adj;
    const Dims& d = ctx->d;
    long long n = 2LL * d.ny * d.nz + (long long)d.nx * d.nz;
    hipLaunchKernelGGL(ka_ubc_xy, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, mesh(ctx), v3(S.Ub), upre_x,
                       c3(Uold), dt, S.xp);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(ka_ubc_z, dim3(grid_for((long long)d.nx * d.ny)), dim3(256), 0, ctx->stream, d, mesh(ctx),
                       v3(S.Ub), retracted, floor_x);
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_jubc_adj(dalm_ctx* ctx, double* const jb[3]) {
    const Dims& d = ctx->d;
    long long n = (long long)d.ny * d.nz + 2LL * d.nx * d.nz + 2LL * d.nx * d.ny;
    hipLaunchKernelGGL(ka_jubc, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, v3(jb));
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_project_adj(dalm_ctx* ctx, double dt) {
    AdjointState& S = ctx->adj;
    const Dims& d = ctx->d;
    if (ctx->has_solid) {
        hipLaunchKernelGGL(ka_solid, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, d, ctx->solid, v3(S.Ub));
        LAUNCH_CHECK();
    }
    long long n = (long long)(d.nx + 2) * (d.ny + 2) * (d.nz + 2);
    hipLaunchKernelGGL(ka_project, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, mesh(ctx), c3(S.Ub), c3(S.JUb),
                       dt, S.pb);
    LAUNCH_CHECK();
    return DALM_OK;
}

// b_b = A^-T p_b via A y = V^-1 p_b (fold_pbc: fold p_b's ghosts first).  The
// solve uses the forward's method and preconditioner, warm-started from the last
// adjoint solution, to a tolerance relative to |V^-1 p_b|.
int adjoint_solve(dalm_ctx* ctx, int fold_pbc, double tol_rel, long long max_it, long long* it_out,
                  double* err_out) {
    AdjointState& S = ctx->adj;
    const dalm_params& P = ctx->prm;
    hipLaunchKernelGGL(ka_pbc_fold, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), S.pb, S.ab,
                       fold_pbc, 1, ctx->partials);
    LAUNCH_CHECK();
    double* res = &ctx->solve->aux0;
    if (int rc = launch_finalize(ctx, FIN_SUM, ctx->grid, nullptr, res)) return rc;
    double h = 0;
    HIP_TRY(hipMemcpyAsync(&h, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *it_out = 0;
    *err_out = 0;
    if (!(h > 0)) {
        if (int rc = zero(ctx, S.ya, (size_t)ctx->L)) return rc;
        return DALM_OK;
    }
    const double tol = tol_rel * sqrt(h) / sqrt((double)ctx->N);
    int st;
    if (S.from_mixed && P.method == DALM_BICGSTAB) {
        // a precision-form recording: the precision form's solver (FP32 Krylov
        // vectors and sweeps, ended on the true FP64 residual), on y and the
        // right-hand side in place of the context's p and b
        std::swap(ctx->p, S.ya);
        std::swap(ctx->b, S.ab);
        st = run_poisson_mixed(ctx, (int)P.pc_kind, P.pc_omega, (int)P.pc_it, max_it, tol, it_out, err_out);
        std::swap(ctx->p, S.ya);
        std::swap(ctx->b, S.ab);
    } else if (P.method == DALM_BICGSTAB)
        st = run_bicgstab_on(ctx, S.ya, S.ab, 3, (int)P.pc_kind, P.pc_omega, (int)P.pc_it, max_it, tol, it_out,
                             err_out);
    else
        st = run_sor_on(ctx, S.ya, S.ab, 3, P.omega, max_it, tol, it_out, err_out);
    if (st == DALM_HIP_ERROR || st == DALM_BAD_ARG) return st;
    return DALM_OK;
}

int launch_rhs_adj(dalm_ctx* ctx, const double* s_in, int y_given, double dt, double scale, int zero_bc) {
    AdjointState& S = ctx->adj;
    const Dims& d = ctx->d;
    long long n = (long long)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    hipLaunchKernelGGL(ka_rhs, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, mesh(ctx), s_in, y_given,
                       1.0 / (dt * scale), zero_bc, v3(S.JUb));
    LAUNCH_CHECK();
    return DALM_OK;
}

int launch_interp_adj(dalm_ctx* ctx) {
    AdjointState& S = ctx->adj;
    const Dims& d = ctx->d;
    long long n = (long long)(d.nx + 2) * (d.ny + 2) * (d.nz + 2);
    hipLaunchKernelGGL(ka_interp, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, mesh(ctx), c3(S.JUb), v3(S.Ub));
    LAUNCH_CHECK();
    return DALM_OK;
}

// ALM at time t in reverse: forces recomputed from U, then projection and force
// bars with the objective weights already in S.tqw.
int launch_alm_adj(dalm_ctx* ctx, double t, double* const U[3]) {
    AdjointState& S = ctx->adj;
    AlmTables& A = ctx->alm;
    if (A.ap_count <= 0) return DALM_OK;
    if (int rc = alm_forces(ctx, t, U)) return rc;
    const double reach = A.width * CUT;
    const double sp = A.width * sqrt(PI_);
    hipLaunchKernelGGL(ka_ap_reduce, dim3(A.ap_count), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), A.ap,
                       c3(S.fb), 1 / (sp * sp * sp), A.width * A.width, reach, reach * reach, S.apred);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(ka_ap_force, dim3((A.ap_count + 63) / 64), dim3(64), 0, ctx->stream, ctx->d, mesh(ctx), c3(U),
                       A.wt, A.ap, A.ap_count, A.blade_per_wt, A.ap_per_blade, A.cd, A.cl, A.atk, A.atk_count,
                       S.apred, S.tqw, v3(S.Ub), S.apg);
    LAUNCH_CHECK();
    return DALM_OK;
}

// ---- recorded runs ----------------------------------------------------------------
constexpr int CK_FIELDS = 11;   // U(3), JU(3), p, nut, f(3)
constexpr int SEG_FIELDS = 7;   // U(3), JU(3), nut

void state_ptrs(dalm_ctx* ctx, double* v[CK_FIELDS]) {
    for (int q = 0; q < 3; ++q) {
        v[q] = ctx->U[q];
        v[3 + q] = ctx->JU[q];
        v[8 + q] = ctx->f[q];
    }
    v[6] = ctx->p;
    v[7] = ctx->nut;
}

int save_state(dalm_ctx* ctx, double* dst) {
    double* v[CK_FIELDS];
    state_ptrs(ctx, v);
    const size_t L = (size_t)ctx->L;
    for (int q = 0; q < CK_FIELDS; ++q)
        HIP_TRY(hipMemcpyAsync(dst + q * L, v[q], L * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return DALM_OK;
}

__global__ void ka_widen(long long n, const float* __restrict__ a, double* __restrict__ b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

// The precision form's state (src, DALM_MIXED) as an FP64 checkpoint in
// state_ptrs' order: its FP32 U, nu_t and f widened exactly, its FP64 JU and p
// copied.  Enqueued on src's stream.
int save_state_from(dalm_ctx* src, double* dst) {
    dalm_ctx* ctx = src;  // for HIP_TRY's error report
    const size_t L = (size_t)src->L;
    const int grid = (int)std::min<long long>(4096, ((long long)L + 255) / 256);
    const float* w[5] = {src->U32[0], src->U32[1], src->U32[2], src->nut32, nullptr};
    const int wslot[4] = {0, 1, 2, 7};
    for (int q = 0; q < 4; ++q)
        hipLaunchKernelGGL(ka_widen, dim3(grid), dim3(256), 0, src->stream, (long long)L, w[q], dst + wslot[q] * L);
    for (int q = 0; q < 3; ++q) {
        hipLaunchKernelGGL(ka_widen, dim3(grid), dim3(256), 0, src->stream, (long long)L, (const float*)src->f32[q],
                           dst + (8 + q) * L);
        HIP_TRY(hipMemcpyAsync(dst + (3 + q) * L, src->JU[q], L * sizeof(double), hipMemcpyDeviceToDevice,
                               src->stream));
    }
    HIP_TRY(hipMemcpyAsync(dst + 6 * L, src->p, L * sizeof(double), hipMemcpyDeviceToDevice, src->stream));
    HIP_TRY(hipGetLastError());
    return DALM_OK;
}

int load_state(dalm_ctx* ctx, const double* src) {
    double* v[CK_FIELDS];
    state_ptrs(ctx, v);
    const size_t L = (size_t)ctx->L;
    for (int q = 0; q < CK_FIELDS; ++q)
        HIP_TRY(hipMemcpyAsync(v[q], src + q * L, L * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    // the force field is complete: let the next projection clear the interior once
    ctx->alm.f_cleared = false;
    ctx->alm.prev_box.clear();
    return DALM_OK;
}

struct Slot {
    double* U[3];
    double* JU[3];
    double* nut;
};
Slot slot(dalm_ctx* ctx, long long s) {
    const size_t L = (size_t)ctx->L;
    double* b = ctx->adj.seg + (size_t)s * SEG_FIELDS * L;
    Slot r;
    for (int q = 0; q < 3; ++q) {
        r.U[q] = b + q * L;
        r.JU[q] = b + (3 + q) * L;
    }
    r.nut = b + 6 * L;
    return r;
}

int save_slot(dalm_ctx* ctx, long long s) {
    Slot r = slot(ctx, s);
    const size_t L = (size_t)ctx->L, B = L * sizeof(double);
    for (int q = 0; q < 3; ++q) {
        HIP_TRY(hipMemcpyAsync(r.U[q], ctx->U[q], B, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(r.JU[q], ctx->JU[q], B, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(r.nut, ctx->nut, B, hipMemcpyDeviceToDevice, ctx->stream));
    return DALM_OK;
}

// Reverse of step n: (Ub, JUb, nutb, fb) = bars of S_n on entry, of S_{n-1} on exit.
int reverse_step(dalm_ctx* ctx, long long n, const Slot& cur, const Slot& prev, double tol_rel) {
    AdjointState& S = ctx->adj;
    const dalm_params& P = ctx->prm;
    const double t = P.dt * (double)n;
    if (int rc = launch_alm_adj(ctx, t, cur.U)) return rc;
    if (int rc = launch_nut_adj(ctx, cur.U, P.Cs)) return rc;
    if (int rc = launch_jubc_adj(ctx, S.JUb)) return rc;
    // U before the boundary pass shares its ghost shell (and so its advecting x+
    // velocity) with U_{n-1}
    if (int rc = launch_ubc_adj(ctx, prev.U[0], prev.U, P.dt, (int)P.retracted_floor, P.floor_start_x)) return rc;
    if (int rc = launch_project_adj(ctx, P.dt)) return rc;
    long long it = 0;
    double err = 0;
    if (int rc = adjoint_solve(ctx, 1, tol_rel, P.max_it > 0 ? 4 * P.max_it : 4000, &it, &err)) return rc;
    S.adj_it += it;
    if (int rc = launch_rhs_adj(ctx, S.ya, 1, P.dt, ctx->pc.max_diag, 1)) return rc;
    if (int rc = launch_interp_adj(ctx)) return rc;
    if (int rc = launch_predictor_adj(ctx, prev.U, prev.JU, prev.nut, P.Re, P.dt, 1, S.xp)) return rc;
    for (int q = 0; q < 3; ++q) {
        std::swap(S.Ub[q], S.Ubn[q]);
        std::swap(S.JUb[q], S.JUbn[q]);
        std::swap(S.fb[q], S.fbn[q]);
    }
    std::swap(S.nutb, S.nutbn);
    return DALM_OK;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// record one finished forward step (called by dalm_step while taping)
int adjoint_record(dalm_ctx* ctx, int status) {
    AdjointState& S = ctx->adj;
    const long long n = ctx->step - S.start_step;
    if (n < 1 || n > S.steps || n != S.recorded + 1) {
        S.taping = false;
        S.tape_bad = true;
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "recorded steps out of sequence or beyond the recording's length");
    }
    if (status == DALM_NONFINITE) S.tape_bad = true;
    const size_t L = (size_t)ctx->L;
    if (S.keep_p)
        HIP_TRY(hipMemcpyAsync(S.ptape + (size_t)(n - 1) * L, ctx->p, L * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream));
    if (n % S.every == 0)
        if (int rc = save_state(ctx, S.ck + (size_t)(n / S.every) * CK_FIELDS * L)) return rc;
    for (int w = 0; w < ctx->alm.wt_count; ++w) S.J += S.h_tqw[w] * ctx->h_scal->torque[w];
    S.recorded += 1;
    if (S.recorded == S.steps) S.taping = false;
    return DALM_OK;
}

// One dalm_step of a precision-form context recording into its FP64 sink
// (dalm_adjoint_begin_from): the same bookkeeping as adjoint_record, with the
// checkpoints widened from the precision form's state.
int adjoint_record_from(dalm_ctx* src, int status) {
    dalm_ctx* sink = src->adj.sink;
    if (!sink) return DALM_OK;
    AdjointState& S = sink->adj;
    dalm_ctx* ctx = src;  // errors are reported on the stepping context
    const long long n = src->step - S.start_step;
    if (n < 1 || n > S.steps || n != S.recorded + 1) {
        S.taping = false;
        S.tape_bad = true;
        src->adj.sink = nullptr;
        S.source = nullptr;
        return dalm_fail_msg(src, DALM_BAD_ARG, "recorded steps out of sequence or beyond the recording's length");
    }
    if (status == DALM_NONFINITE) S.tape_bad = true;
    const size_t L = (size_t)src->L;
    if (S.keep_p)
        HIP_TRY(hipMemcpyAsync(S.ptape + (size_t)(n - 1) * L, src->p, L * sizeof(double), hipMemcpyDeviceToDevice,
                               src->stream));
    if (n % S.every == 0)
        if (int rc = save_state_from(src, S.ck + (size_t)(n / S.every) * CK_FIELDS * L)) return rc;
    HIP_TRY(hipStreamSynchronize(src->stream));
    for (int w = 0; w < src->alm.wt_count; ++w) S.J += S.h_tqw[w] * src->h_scal->torque[w];
    S.recorded += 1;
    sink->step = src->step;
    if (S.recorded == S.steps) {
        S.taping = false;
        S.source = nullptr;
        src->adj.sink = nullptr;
    }
    return DALM_OK;
}

void adjoint_free(dalm_ctx* ctx) {
    AdjointState& S = ctx->adj;
    // a recording between two contexts ends with either of them
    if (S.sink) {
        S.sink->adj.source = nullptr;
        if (S.sink->adj.taping) {
            S.sink->adj.taping = false;
            S.sink->adj.tape_bad = true;
        }
    }
    if (S.source) S.source->adj.sink = nullptr;
    double** big[4] = {&S.ck, &S.ptape, &S.seg, &S.fin};
    for (double** p : big)
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    if (S.have_buffers) {
        for (int q = 0; q < 3; ++q) {
            double* v[6] = {S.Ub[q], S.Ubn[q], S.JUb[q], S.JUbn[q], S.fb[q], S.fbn[q]};
            for (double* p : v) (void)hipFree(p);
        }
        double* s1[9] = {S.nutb, S.nutbn, S.pb, S.ab, S.ya, S.xp, S.apred, S.apg, S.tqw};
        for (double* p : s1) (void)hipFree(p);
    }
    S = AdjointState();
}

#define CHECK_ADJ()                                                            \
    do {                                                                       \
        if (!ctx) return DALM_BAD_ARG;                                         \
        if (int rc_ = ensure_adjoint_buffers(ctx)) return rc_;                 \
        if (int rc_ = ensure_ap_buffers(ctx)) return rc_;                      \
    } while (0)
#define TRY(x)              \
    do {                    \
        int rc_ = (x);      \
        if (rc_) return rc_; \
    } while (0)

extern "C" {

// ---- per-kernel adjoints (bars live in the context: DALM_F_*_B fields) ---------------
int dalm_predictor_adj(dalm_ctx* ctx, double Re, double dt) {
    CHECK_ADJ();
    TRY(launch_predictor_adj(ctx, ctx->Uold, ctx->JU, ctx->nut, Re, dt, 0, nullptr));
    return sync_stream(ctx);
}

int dalm_interpolate_JU_adj(dalm_ctx* ctx) {
    CHECK_ADJ();
    TRY(launch_interp_adj(ctx));
    return sync_stream(ctx);
}

int dalm_apply_JUbc_adj(dalm_ctx* ctx) {
    CHECK_ADJ();
    TRY(launch_jubc_adj(ctx, ctx->adj.JUb));
    return sync_stream(ctx);
}

int dalm_poisson_rhs_adj(dalm_ctx* ctx, double dt, double scale) {
    CHECK_ADJ();
    TRY(launch_rhs_adj(ctx, ctx->adj.ab, 0, dt, scale, 0));
    return sync_stream(ctx);
}

int dalm_apply_pbc_adj(dalm_ctx* ctx) {
    CHECK_ADJ();
    AdjointState& S = ctx->adj;
    hipLaunchKernelGGL(ka_pbc_fold, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), S.pb, S.pb, 1, 0,
                       ctx->partials);
    HIP_TRY(hipGetLastError());
    const Dims& d = ctx->d;
    long long n = 2LL * d.ny * d.nz + 2LL * d.nx * d.nz + 2LL * d.nx * d.ny;
    hipLaunchKernelGGL(ka_pbc_clear, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, S.pb);
    HIP_TRY(hipGetLastError());
    return sync_stream(ctx);
}

int dalm_project_p_adj(dalm_ctx* ctx, double dt) {
    CHECK_ADJ();
    TRY(launch_project_adj(ctx, dt));
    return sync_stream(ctx);
}

int dalm_set_solid_U_adj(dalm_ctx* ctx) {
    CHECK_ADJ();
    hipLaunchKernelGGL(ka_solid, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, ctx->solid, v3(ctx->adj.Ub));
    HIP_TRY(hipGetLastError());
    return sync_stream(ctx);
}

int dalm_apply_Ubc_adj(dalm_ctx* ctx, double dt, int64_t retracted_floor, double floor_start_x) {
    CHECK_ADJ();
    TRY(launch_ubc_adj(ctx, ctx->U[0], ctx->Uold, dt, (int)retracted_floor, floor_start_x));
    const Dims& d = ctx->d;
    hipLaunchKernelGGL(ka_xp_add, dim3(grid_for(4LL * d.ny * d.nz)), dim3(256), 0, ctx->stream, d, ctx->adj.xp,
                       v3(ctx->adj.Ubn));
    HIP_TRY(hipGetLastError());
    return sync_stream(ctx);
}

int dalm_calc_nut_adj(dalm_ctx* ctx, double Cs) {
    CHECK_ADJ();
    TRY(launch_nut_adj(ctx, ctx->U, Cs));
    return sync_stream(ctx);
}

int dalm_poisson_adj(dalm_ctx* ctx, double tol_rel, int64_t max_it, int64_t* it, double* err) {
    CHECK_ADJ();
    if (!ctx->have_prm) return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_set_params not called");
    AdjointState& S = ctx->adj;
    long long its = 0;
    double e = 0;
    TRY(adjoint_solve(ctx, 0, tol_rel, max_it, &its, &e));
    hipLaunchKernelGGL(ka_scale_v, dim3(ctx->grid), dim3(BLOCK), 0, ctx->stream, ctx->d, mesh(ctx), S.ya, S.ab);
    HIP_TRY(hipGetLastError());
    if (it) *it = its;
    if (err) *err = e;
    return sync_stream(ctx);
}

int dalm_run_alm_adj(dalm_ctx* ctx, double t, const double* tq_w, double* ap_grad) {
    CHECK_ADJ();
    AdjointState& S = ctx->adj;
    AlmTables& A = ctx->alm;
    if (A.ap_count <= 0) return DALM_OK;
    std::vector<double> w(DALM_MAX_TURBINES, 0.0);
    for (int q = 0; q < A.wt_count && tq_w; ++q) w[q] = tq_w[q];
    HIP_TRY(hipMemcpyAsync(S.tqw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    TRY(zero(ctx, S.apg, 4 * (size_t)A.ap_count));
    TRY(launch_alm_adj(ctx, t, ctx->U));
    if (ap_grad) {
        std::vector<double> g(4 * (size_t)A.ap_count);
        HIP_TRY(hipMemcpyAsync(g.data(), S.apg, g.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (size_t q = 0; q < g.size(); ++q) ap_grad[q] += g[q];
    }
    return sync_stream(ctx);
}

// ---- recorded run and reverse sweep --------------------------------------------------
}  // extern "C"

namespace {

// The tape of a recording of `steps` steps from step `start` (checkpoint every
// `every` steps, pressure per step with DALM_ADJ_KEEP_PRESSURE), allocated in ctx
// after a check against the free device memory.
int begin_tape(dalm_ctx* ctx, int64_t steps, int64_t every, const double* tq_w, int64_t flags, long long start) {
    // bounded so that the recording's byte counts below cannot overflow
    if (steps < 1 || every < 1 || steps > (1LL << 24)) return DALM_BAD_ARG;
    if (!ctx->have_prm) return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_set_params not called");
    AdjointState& S = ctx->adj;
    if (S.source) S.source->adj.sink = nullptr;
    S.source = nullptr;
    S.from_mixed = false;
    double** big[4] = {&S.ck, &S.ptape, &S.seg, &S.fin};
    for (double** p : big)
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    if (every > steps) every = steps;
    S.steps = steps;
    S.every = every;
    S.keep_p = (flags & DALM_ADJ_KEEP_PRESSURE) != 0;
    S.start_step = start;
    S.recorded = 0;
    S.tape_bad = false;
    S.J = 0;
    S.h_tqw.assign(DALM_MAX_TURBINES, 0.0);
    for (int q = 0; q < ctx->alm.wt_count && tq_w; ++q) S.h_tqw[q] = tq_w[q];
    const size_t L = (size_t)ctx->L;
    const size_t nck = (size_t)(steps / every) + 1;
    size_t need = (nck * CK_FIELDS + (size_t)(every + 1) * SEG_FIELDS + CK_FIELDS) * L * sizeof(double);
    if (S.keep_p) need += (size_t)steps * L * sizeof(double);
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    if (need + (size_t)(1ull << 30) > free_b)
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "not enough device memory for the recording (" +
                                                    std::to_string(need >> 20) + " MiB needed, " +
                                                    std::to_string(free_b >> 20) + " MiB free)");
    HIP_TRY(hipMalloc((void**)&S.ck, nck * CK_FIELDS * L * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&S.seg, (size_t)(every + 1) * SEG_FIELDS * L * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&S.fin, CK_FIELDS * L * sizeof(double)));
    if (S.keep_p) HIP_TRY(hipMalloc((void**)&S.ptape, (size_t)steps * L * sizeof(double)));
    return DALM_OK;
}

}  // namespace

extern "C" {

int dalm_adjoint_begin(dalm_ctx* ctx, int64_t steps, int64_t every, const double* tq_w, int64_t flags) {
    CHECK_ADJ();
    TRY(begin_tape(ctx, steps, every, tq_w, flags, ctx->step));
    AdjointState& S = ctx->adj;
    TRY(save_state(ctx, S.ck));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    S.taping = true;
    return DALM_OK;
}

int dalm_adjoint_begin_from(dalm_ctx* ctx, dalm_ctx* src, int64_t steps, int64_t every, const double* tq_w,
                            int64_t flags) {
    CHECK_ADJ();
    if (!src || src == ctx || src->precision != DALM_MIXED)
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_adjoint_begin_from records a DALM_MIXED context");
    const Dims &a = ctx->d, &b = src->d;
    if (a.n0 != b.n0 || a.n1 != b.n1 || a.n2 != b.n2 || a.gc != b.gc || ctx->alm.wt_count != src->alm.wt_count ||
        ctx->alm.ap_count != src->alm.ap_count)
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "the recorded context has another grid or actuator-line set-up");
    if (!src->have_prm) return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_set_params not called on the recorded context");
    if (src->adj.sink) return dalm_fail_msg(ctx, DALM_BAD_ARG, "the context is already being recorded");
    TRY(ensure_mixed_solver(ctx));
    TRY(begin_tape(ctx, steps, every, tq_w, flags, src->step));
    AdjointState& S = ctx->adj;
    TRY(save_state_from(src, S.ck));
    HIP_TRY(hipStreamSynchronize(src->stream));
    ctx->step = src->step;
    S.source = src;
    S.from_mixed = true;
    src->adj.sink = ctx;
    S.taping = true;
    return DALM_OK;
}

int dalm_adjoint_run(dalm_ctx* ctx, double tol_rel, double* J, double* ap_grad, double* wt_grad,
                     dalm_adjoint_stats* stats) {
    CHECK_ADJ();
    AdjointState& S = ctx->adj;
    if (!S.ck || S.recorded != S.steps)
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_adjoint_run needs a complete recording (dalm_adjoint_begin "
                                                "then that many dalm_step calls)");
    if (S.tape_bad)
        return dalm_fail_msg(ctx, DALM_NONFINITE, "the recording holds a step that did not finish or left a "
                                                  "non-finite state; it cannot be differentiated");
    const size_t L = (size_t)ctx->L;
    AlmTables& A = ctx->alm;
    const long long end_step = ctx->step;
    TRY(save_state(ctx, S.fin));
    for (int q = 0; q < 3; ++q) {
        TRY(zero(ctx, S.Ub[q], L));
        TRY(zero(ctx, S.JUb[q], L));
        TRY(zero(ctx, S.fb[q], L));
    }
    TRY(zero(ctx, S.nutb, L));
    TRY(zero(ctx, S.ya, L));
    const int na = A.ap_count > 0 ? A.ap_count : 1;
    TRY(zero(ctx, S.apg, 4 * (size_t)na));
    HIP_TRY(hipMemcpyAsync(S.tqw, S.h_tqw.data(), DALM_MAX_TURBINES * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream));
    S.adj_it = 0;
    S.replay_ms = S.reverse_ms = 0;
    const long long nseg = (S.steps + S.every - 1) / S.every;
    static const bool trace = getenv("DALM_TRACE") != nullptr;
    for (long long sg = nseg - 1; sg >= 0; --sg) {
        const long long s0 = sg * S.every;
        const long long s1 = std::min(s0 + S.every, S.steps);
        double t0 = now_ms();
        TRY(load_state(ctx, S.ck + (size_t)sg * CK_FIELDS * L));
        ctx->step = S.start_step + s0;
        TRY(save_slot(ctx, 0));
        S.replaying = true;
        for (long long n = s0 + 1; n <= s1; ++n) {
            const double* pg = S.keep_p ? S.ptape + (size_t)(n - 1) * L : nullptr;
            int rc = step_fp64(ctx, nullptr, pg);
            if (rc && rc != DALM_NONCONVERGED && rc != DALM_BREAKDOWN) {
                S.replaying = false;
                return rc;
            }
            if (int rc2 = save_slot(ctx, n - s0)) {
                S.replaying = false;
                return rc2;
            }
        }
        S.replaying = false;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        double t1 = now_ms();
        for (long long n = s1; n > s0; --n)
            TRY(reverse_step(ctx, S.start_step + n, slot(ctx, n - s0), slot(ctx, n - s0 - 1), tol_rel));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        double t2 = now_ms();
        S.replay_ms += t1 - t0;
        S.reverse_ms += t2 - t1;
        if (trace)
            fprintf(stderr, "[adjoint] segment %lld of %lld (steps %lld-%lld): replay %.1f ms, reverse %.1f ms, "
                            "%lld adjoint iterations so far\n",
                    nseg - sg, nseg, s0 + 1, s1, t1 - t0, t2 - t1, S.adj_it);
    }
    // the force of the initialisation pass (t of the recording's start) drives the
    // first step; its torque is not part of J
    if (A.ap_count > 0) {
        std::vector<double> z(DALM_MAX_TURBINES, 0.0);
        HIP_TRY(hipMemcpyAsync(S.tqw, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        TRY(launch_alm_adj(ctx, ctx->prm.dt * (double)S.start_step, slot(ctx, 0).U));
    }
    std::vector<double> g(4 * (size_t)na);
    HIP_TRY(hipMemcpyAsync(g.data(), S.apg, g.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TRY(load_state(ctx, S.fin));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->step = end_step;
    if (J) *J = S.J;
    if (ap_grad)
        for (int a = 0; a < A.ap_count; ++a)
            for (int q = 0; q < 4; ++q) ap_grad[4 * a + q] = g[4 * (size_t)a + q];
    if (wt_grad) {
        const int per_wt = A.blade_per_wt * A.ap_per_blade;
        for (int w = 0; w < A.wt_count; ++w) {
            double ya = 0, pa = 0;
            for (int a = w * per_wt; a < (w + 1) * per_wt; ++a) {
                ya += g[4 * (size_t)a];
                pa += g[4 * (size_t)a + 1];
            }
            wt_grad[2 * w] = ya;
            wt_grad[2 * w + 1] = pa;
        }
    }
    if (stats) {
        stats->steps = S.steps;
        stats->every = S.every;
        stats->adj_it = S.adj_it;
        stats->replay_ms = S.replay_ms;
        stats->reverse_ms = S.reverse_ms;
    }
    return DALM_OK;
}

int dalm_adjoint_end(dalm_ctx* ctx) {
    if (!ctx) return DALM_BAD_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    adjoint_free(ctx);
    return DALM_OK;
}

}  // extern "C"
