/*
 * oracle.c -- CPU restatement of the 3dalm per-step hot path (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / CPU baseline, never as product.
 *
 * How each function is pinned (tests/test_oracle.py, tests/test_oracle_ref.py):
 *   - every function is compared against the reference's own code compiled from
 *     its headers (oracle/_ref, built by oracle/Makefile) on seeded inputs:
 *     bit-exact except libm-based routines (cbrt/exp/sin/cos/atan), which agree to
 *     <= 1e-14 relative;
 *   - committed golden vectors from that build (tests/golden, npz files) are checked
 *     here without the reference present;
 *   - closed forms: MUSCL on constant/linear fields, diffusion of x^2 = 2*nu,
 *     Smagorinsky of a shear flow, V*A symmetric, A*1 zero except at x+, quadratic
 *     extrapolation exact for quadratics, projection identity
 *     avg_div = dt*max_diag*err, single-AP force in closed form, Gaussian weights
 *     summing to ~1, and the reference's AP-property fixture test/wtprop.json.
 *
 * Output path: calc_q and the time average of U (src/cfd.h:948-1077,
 *   src/solver.h:1194-1204), pinned the same way plus closed forms (solid-body
 *   rotation q = omega^2, pure strain q = -a^2, shear q = 0; the average of a
 *   constant is that constant, of n fields their mean).
 *
 * Reference files restated: src/cfd.h, src/eq.h, src/mv.h, src/bc.h, src/alm.h,
 * src/euler_angle.h, src/util.h, src/io.h:98-164, src/solver.h:727-833,1034-1190.
 */
#define _DEFAULT_SOURCE
#include "oracle.h"
#include "oracle_util.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif




/* ========================================================================== */
/* mesh widths: src/io.h:115-125 */
void or_mesh_widths(const double *c, i64 n, double *d) {
    for (i64 i = 1; i < n - 1; ++i) d[i] = 0.5 * (c[i + 1] - c[i - 1]);
    d[0] = -1.5 * c[0] + 2.0 * c[1] - 0.5 * c[2];
    d[n - 1] = 1.5 * c[n - 1] - 2.0 * c[n - 2] + 0.5 * c[n - 3];
}

/* ========================================================================== */
/* Convection of one component at one cell, src/cfd.h:8-109. */
static double muscl_cell(const double vx[5], const double vy[5], const double vz[5],
                         const double J[6], double dxc, double dyc, double dzc) {
    const double kap = 1. / 3.;
    const double bb = (3. - kap) / (1. - kap);
    const double q = 1. / 4.;
    double F[6];
    const double *vs[3] = {vx, vy, vz};
    for (int d = 0; d < 3; ++d) {
        const double *v = vs[d];
        double a1 = v[1] - v[0], a2 = v[2] - v[1], a3 = v[3] - v[2], a4 = v[4] - v[3];
        double t1 = sgn(a1), t2 = sgn(a2), t3 = sgn(a3), t4 = sgn(a4);
        double h6 = t4 * fmax(0., fmin(fabs(a4), t4 * bb * a3));
        double h5 = t3 * fmax(0., fmin(fabs(a3), t3 * bb * a4));
        double h4 = t3 * fmax(0., fmin(fabs(a3), t3 * bb * a2));
        double h3 = t2 * fmax(0., fmin(fabs(a2), t2 * bb * a3));
        double h2 = t2 * fmax(0., fmin(fabs(a2), t2 * bb * a1));
        double h1 = t1 * fmax(0., fmin(fabs(a1), t1 * bb * a2));
        double pr = v[3] - q * ((1. - kap) * h6 + (1. + kap) * h5);
        double pl = v[2] + q * ((1. - kap) * h3 + (1. + kap) * h4);
        double mr = v[2] - q * ((1. - kap) * h4 + (1. + kap) * h3);
        double ml = v[1] + q * ((1. - kap) * h1 + (1. + kap) * h2);
        double Jp = J[2 * d], Jm = J[2 * d + 1];
        F[2 * d] = 0.5 * (Jp * (pr + pl) - fabs(Jp) * (pr - pl));
        F[2 * d + 1] = 0.5 * (Jm * (mr + ml) - fabs(Jm) * (mr - ml));
    }
    double vol = dxc * dyc * dzc;
    return (F[0] - F[1] + F[2] - F[3] + F[4] - F[5]) / vol;
}

/* predictor: src/cfd.h:175-242 (cell), 244-392 (interior loop, single rank) */
void or_predictor(const or_grid *g, double *U, const double *Uold, const double *JU,
                  const double *nut, const double *f, double Re, double dt) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    PAR_INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double J[6] = {JU[3 * c + 0], JU[3 * (c - si) + 0], JU[3 * c + 1],
                       JU[3 * (c - sj) + 1], JU[3 * c + 2], JU[3 * (c - 1) + 2]};
        double nu = 1 / Re + nut[c];
        for (int m = 0; m < 3; ++m) {
            double vx[5], vy[5], vz[5];
            for (int o = -2; o <= 2; ++o) {
                vx[o + 2] = Uold[3 * (c + o * si) + m];
                vy[o + 2] = Uold[3 * (c + o * sj) + m];
                vz[o + 2] = Uold[3 * (c + o) + m];
            }
            double conv = muscl_cell(vx, vy, vz, J, DX[i], DY[j], DZ[k]);
            /* src/cfd.h:144-173 */
            double uc = vx[2];
            double dif = nu * (((vx[3] - uc) / (X[i + 1] - X[i]) - (uc - vx[1]) / (X[i] - X[i - 1])) / DX[i] +
                               ((vy[3] - uc) / (Y[j + 1] - Y[j]) - (uc - vy[1]) / (Y[j] - Y[j - 1])) / DY[j] +
                               ((vz[3] - uc) / (Z[k + 1] - Z[k]) - (uc - vz[1]) / (Z[k] - Z[k - 1])) / DZ[k]);
            U[3 * c + m] = Uold[3 * c + m] + dt * (-conv + dif + f[3 * c + m]);
        }
    }
}

/* face fluxes: src/cfd.h:394-420 (cell), 452-465 (range [gc-1, n-gc)) */
void or_interpolate_JU(const or_grid *g, const double *U, double *JU) {
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    OMP_FOR
    for (i64 i = g->gc - 1; i < g->n[0] - g->gc; ++i)
        for (i64 j = g->gc - 1; j < g->n[1] - g->gc; ++j)
            for (i64 k = g->gc - 1; k < g->n[2] - g->gc; ++k) {
                i64 c = lin(g, i, j, k);
                double ayz = DY[j] * DZ[k];
                JU[3 * c + 0] = 0.5 * (U[3 * c + 0] * ayz + U[3 * (c + si) + 0] * ayz);
                double axz = DX[i] * DZ[k];
                JU[3 * c + 1] = 0.5 * (U[3 * c + 1] * axz + U[3 * (c + sj) + 1] * axz);
                double axy = DX[i] * DY[j];
                JU[3 * c + 2] = 0.5 * (U[3 * c + 2] * axy + U[3 * (c + 1) + 2] * axy);
            }
}

/* face-flux boundary conditions: src/bc.h:235-302 */
void or_apply_JUbc(const or_grid *g, double *JU, const double Uin[3]) {
    const double *DY = g->d[1], *DZ = g->d[2];
    const i64 G = g->gc, NX = g->n[0], NY = g->n[1], NZ = g->n[2];
    for (i64 j = G; j < NY - G; ++j)
        for (i64 k = G; k < NZ - G; ++k) JU[3 * lin(g, G - 1, j, k) + 0] = DY[j] * DZ[k] * Uin[0];
    for (i64 i = G; i < NX - G; ++i)
        for (i64 k = G; k < NZ - G; ++k) {
            JU[3 * lin(g, i, G - 1, k) + 1] = 0;
            JU[3 * lin(g, i, NY - G - 1, k) + 1] = 0;
        }
    for (i64 i = G; i < NX - G; ++i)
        for (i64 j = G; j < NY - G; ++j) {
            JU[3 * lin(g, i, j, G - 1) + 2] = 0;
            JU[3 * lin(g, i, j, NZ - G - 1) + 2] = 0;
        }
}

static inline double flux_div(const or_grid *g, const double *JU, i64 i, i64 j, i64 k) {
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    i64 c = lin(g, i, j, k);
    double vol = g->d[0][i] * g->d[1][j] * g->d[2][k];
    return (JU[3 * c + 0] - JU[3 * (c - si) + 0] + JU[3 * c + 1] - JU[3 * (c - sj) + 1] +
            JU[3 * c + 2] - JU[3 * (c - 1) + 2]) / vol;
}

/* Poisson right-hand side: src/cfd.h:569-598 */
void or_poisson_rhs(const or_grid *g, const double *JU, double *b, double dt, double scale) {
    PAR_INTERIOR(g, i, j, k) b[lin(g, i, j, k)] = flux_div(g, JU, i, j, k) / (dt * scale);
}

/* divergence: src/cfd.h:1079-1107 */
void or_divergence(const or_grid *g, const double *JU, double *div) {
    PAR_INTERIOR(g, i, j, k) div[lin(g, i, j, k)] = flux_div(g, JU, i, j, k);
}

/* ========================================================================== */
/* operator: src/eq.h:701-798 (single domain: gsize = n, offset = 0) */
double or_build_A(const or_grid *g, double *A) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 G = g->gc;
    double mx = 0;
    PAR_INTERIOR_MAX(g, i, j, k) {
        double *a = A + 7 * lin(g, i, j, k);
        double e = 0, w = 0, nn = 0, s = 0, t = 0, bo = 0, c = 0;
        if (i < g->n[0] - G - 1) {
            e = 1 / (DX[i] * (X[i + 1] - X[i]));
            c -= e;
        } else {
            c -= 2 / (DX[i] * DX[i]);
        }
        if (i > G) { w = 1 / (DX[i] * (X[i] - X[i - 1])); c -= w; }
        if (j < g->n[1] - G - 1) { nn = 1 / (DY[j] * (Y[j + 1] - Y[j])); c -= nn; }
        if (j > G) { s = 1 / (DY[j] * (Y[j] - Y[j - 1])); c -= s; }
        if (k < g->n[2] - G - 1) { t = 1 / (DZ[k] * (Z[k + 1] - Z[k])); c -= t; }
        if (k > G) { bo = 1 / (DZ[k] * (Z[k] - Z[k - 1])); c -= bo; }
        a[0] = c; a[1] = e; a[2] = w; a[3] = nn; a[4] = s; a[5] = t; a[6] = bo;
        if (fabs(c) > mx) mx = fabs(c);
    }
    PAR_INTERIOR(g, i, j, k) {
        double *a = A + 7 * lin(g, i, j, k);
        for (int m = 0; m < 7; ++m) a[m] /= mx;
    }
    return mx;
}

/* sum of a[c] * b[c] over the interior.  Serial build: the reference's loop order.
 * OpenMP build: each thread sums its static share of the i planes in that order,
 * and the per-thread sums are added in thread order, so a given thread count always
 * gives the same result (an OpenMP reduction clause adds them in arrival order). */
static double interior_dot(const or_grid *g, const double *a, const double *b) {
#ifdef _OPENMP
    enum { MAX_THREADS = 512 };
    double part[MAX_THREADS];
    int nt = omp_get_max_threads();
    if (nt > MAX_THREADS) nt = MAX_THREADS;
#pragma omp parallel num_threads(nt)
    {
        double s = 0;
#pragma omp for schedule(static)
        for (i64 i = g->gc; i < g->n[0] - g->gc; ++i)
            for (i64 j = g->gc; j < g->n[1] - g->gc; ++j)
                for (i64 k = g->gc; k < g->n[2] - g->gc; ++k) {
                    i64 c = lin(g, i, j, k);
                    s += a[c] * b[c];
                }
        part[omp_get_thread_num()] = s;
    }
    double s = 0;
    for (int t = 0; t < nt; ++t) s += part[t];
    return s;
#else
    double s = 0;
    INTERIOR(g, i, j, k) { i64 c = lin(g, i, j, k); s += a[c] * b[c]; }
    return s;
#endif
}

/* src/mv.h:8-31 */
double or_dot(const or_grid *g, const double *a, const double *b) { return interior_dot(g, a, b); }
/* src/mv.h:33-55 (sq(v) = v * v) */
double or_l2_norm(const or_grid *g, const double *v) { return sqrt(interior_dot(g, v, v)); }

/* 7-point product at one cell in the reference's summation order (src/mv.h:87) */
static inline double stencil7(const or_grid *g, const double *A, const double *x, i64 c) {
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    const double *a = A + 7 * c;
    return a[0] * x[c] + a[1] * x[c + si] + a[2] * x[c - si] + a[3] * x[c + sj] +
           a[4] * x[c - sj] + a[5] * x[c + 1] + a[6] * x[c - 1];
}

/* src/mv.h:57-182 */
void or_residual(const or_grid *g, const double *A, const double *x, const double *b, double *r) {
    PAR_INTERIOR(g, i, j, k) { i64 c = lin(g, i, j, k); r[c] = b[c] - stencil7(g, A, x, c); }
}
/* src/mv.h:184-309 */
void or_Ax(const or_grid *g, const double *A, const double *x, double *y) {
    PAR_INTERIOR(g, i, j, k) { i64 c = lin(g, i, j, k); y[c] = stencil7(g, A, x, c); }
}
/* src/eq.h:145-289 (offset = 0) */
void or_sor_sweep(const or_grid *g, const double *A, double *x, const double *b,
                  double omega, i64 color) {
    PAR_INTERIOR(g, i, j, k) {
        if ((i + j + k) % 2 != color) continue;
        i64 c = lin(g, i, j, k);
        double rel = (b[c] - stencil7(g, A, x, c)) / A[7 * c];
        x[c] = x[c] + omega * rel;
    }
}
/* src/eq.h:6-143 */
void or_jacobi_sweep(const or_grid *g, const double *A, double *x, const double *xold,
                     const double *b) {
    PAR_INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double rel = (b[c] - stencil7(g, A, xold, c)) / A[7 * c];
        x[c] = xold[c] + rel;
    }
}


/* src/eq.h:291-306 */
void or_run_sor(const or_grid *g, const double *A, double *x, const double *b, double *r,
                double omega, i64 *it, i64 max_it, double *err, double tol) {
    double rootN = sqrt((double)n_interior(g));
    *it = 0;
    do {
        or_sor_sweep(g, A, x, b, omega, 0);
        or_sor_sweep(g, A, x, b, omega, 1);
        or_residual(g, A, x, b, r);
        *err = or_l2_norm(g, r) / rootN;
        ++*it;
    } while (*it < max_it && *err > tol);
}

/* preconditioner application z ~ A^-1 v starting from z = 0:
 *   Jacobi: src/eq.h:320-331; SOR: src/eq.h:308-318 */
static void apply_pc(const or_grid *g, const double *A, double *z, const double *v,
                     double *tmp, i64 kind, double pc_omega, i64 pc_it, i64 L) {
    OMP_FOR
    for (i64 c = 0; c < L; ++c) z[c] = 0;
    for (i64 m = 0; m < pc_it; ++m) {
        if (kind == 0) {
            OMP_FOR
            for (i64 c = 0; c < L; ++c) tmp[c] = z[c];
            or_jacobi_sweep(g, A, z, tmp, v);
        } else {
            or_sor_sweep(g, A, z, v, pc_omega, 0);
            or_sor_sweep(g, A, z, v, pc_omega, 1);
        }
    }
}

/* preconditioned BiCGStab: src/eq.h:333-402 (Jacobi PC) and 404-473 (SOR PC).
 * The reference updates p, s, x and r over the padded length; ghosts of every
 * Krylov vector stay zero (they start zero and only receive zero updates), so the
 * padded-length loops below reproduce it exactly. */
void or_run_bicgstab(const or_grid *g, const double *A, double *x, const double *b,
                     double *work, i64 pc_kind, double pc_omega, i64 pc_it, i64 *it,
                     i64 max_it, double *err, double tol) {
    const i64 L = g->n[0] * g->n[1] * g->n[2];
    double *r = work, *r0 = work + L, *p = work + 2 * L, *pp = work + 3 * L, *q = work + 4 * L;
    double *s = work + 5 * L, *ss = work + 6 * L, *t = work + 7 * L, *tmp = work + 8 * L;
    double rootN = sqrt((double)n_interior(g));

    or_residual(g, A, x, b, r);
    memcpy(r0, r, sizeof(double) * (size_t)L);
    double alpha = 1, beta, omega = 1, rho, rho_prev = 1;
    memset(q, 0, sizeof(double) * (size_t)L);
    *it = 0;
    do {
        rho = or_dot(g, r0, r);
        if (fabs(rho) < FLT_MIN) { *err = rho; break; }
        if (*it == 0) {
            memcpy(p, r, sizeof(double) * (size_t)L);
        } else {
            beta = (rho / rho_prev) * (alpha / omega);
            OMP_FOR
            for (i64 c = 0; c < L; ++c) p[c] = r[c] + beta * (p[c] - omega * q[c]);
        }
        apply_pc(g, A, pp, p, tmp, pc_kind, pc_omega, pc_it, L);
        or_Ax(g, A, pp, q);
        alpha = rho / or_dot(g, r0, q);
        OMP_FOR
        for (i64 c = 0; c < L; ++c) s[c] = r[c] - alpha * q[c];
        apply_pc(g, A, ss, s, tmp, pc_kind, pc_omega, pc_it, L);
        or_Ax(g, A, ss, t);
        omega = or_dot(g, t, s) / or_dot(g, t, t);
        OMP_FOR
        for (i64 c = 0; c < L; ++c) x[c] = x[c] + alpha * pp[c] + omega * ss[c];
        OMP_FOR
        for (i64 c = 0; c < L; ++c) r[c] = s[c] - omega * t[c];
        rho_prev = rho;
        *err = or_l2_norm(g, r) / rootN;
        ++*it;
    } while (*it < max_it && *err > tol);
}

/* ========================================================================== */
/* pressure ghosts: src/bc.h:304-411 */
void or_apply_pbc(const or_grid *g, double *p) {
    const i64 G = g->gc, NX = g->n[0], NY = g->n[1], NZ = g->n[2];
    const double *DX = g->d[0];
    for (i64 j = G; j < NY - G; ++j)
        for (i64 k = G; k < NZ - G; ++k) {
            p[lin(g, G - 1, j, k)] = p[lin(g, G, j, k)];
            i64 ii = NX - G - 1, io = NX - G;
            double hi = 0.5 * DX[ii], ho = 0.5 * DX[io], pb = 0;
            p[lin(g, io, j, k)] = pb - (p[lin(g, ii, j, k)] - pb) * (ho / hi);
        }
    for (i64 i = G; i < NX - G; ++i)
        for (i64 k = G; k < NZ - G; ++k) {
            p[lin(g, i, G - 1, k)] = p[lin(g, i, G, k)];
            p[lin(g, i, NY - G, k)] = p[lin(g, i, NY - G - 1, k)];
        }
    for (i64 i = G; i < NX - G; ++i)
        for (i64 j = G; j < NY - G; ++j) {
            p[lin(g, i, j, G - 1)] = p[lin(g, i, j, G)];
            p[lin(g, i, j, NZ - G)] = p[lin(g, i, j, NZ - G - 1)];
        }
}

/* projection: src/cfd.h:600-618 (cell velocity), 716-740 (face flux) */
void or_project(const or_grid *g, double *U, double *JU, const double *p, double dt) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    PAR_INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        U[3 * c + 0] -= dt * (p[c + si] - p[c - si]) / (X[i + 1] - X[i - 1]);
        U[3 * c + 1] -= dt * (p[c + sj] - p[c - sj]) / (Y[j + 1] - Y[j - 1]);
        U[3 * c + 2] -= dt * (p[c + 1] - p[c - 1]) / (Z[k + 1] - Z[k - 1]);
    }
    OMP_FOR
    for (i64 i = g->gc - 1; i < g->n[0] - g->gc; ++i)
        for (i64 j = g->gc - 1; j < g->n[1] - g->gc; ++j)
            for (i64 k = g->gc - 1; k < g->n[2] - g->gc; ++k) {
                i64 c = lin(g, i, j, k);
                double gx = (p[c + si] - p[c]) / (X[i + 1] - X[i]);
                JU[3 * c + 0] -= dt * (DY[j] * DZ[k]) * gx;
                double gy = (p[c + sj] - p[c]) / (Y[j + 1] - Y[j]);
                JU[3 * c + 1] -= dt * (DX[i] * DZ[k]) * gy;
                double gz = (p[c + 1] - p[c]) / (Z[k + 1] - Z[k]);
                JU[3 * c + 2] -= dt * (DX[i] * DY[j]) * gz;
            }
}

/* src/cfd.h:1143-1160 */
void or_set_solid(const or_grid *g, double *U, const double *solid) {
    PAR_INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        for (int m = 0; m < 3; ++m) U[3 * c + m] = (1. - solid[c]) * U[3 * c + m];
    }
}


/* turbulent inflow: not in the reference (its x- inflow is the constant Uin,
 * src/bc.h:16-30 and 243-257).  Frozen turbulence convected at `speed`: the x-
 * ghost cells and the x- face flux take Uin + u'(x - speed t, y, z) with
 * u'(x) = sum_n a_n cos(k_n . x + psi_n), modes[n] = (kx, ky, kz, psi, ax, ay, az),
 * summed in mode order.  cells: [gc][ny][nz][3] for ghost planes i = 0..gc-1 at the
 * cell centres; face: [ny][nz] the face velocity at x_f = (x[gc-1] + x[gc]) / 2.
 * Pinned by closed forms: each mode divergence-free, the shift property
 * u(x - s, t) = u(x, t + s / speed), the single-mode value. */
static void inflow_point(const double *modes, i64 n_modes, double X, double Y, double Z, double out[3]) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (i64 n = 0; n < n_modes; ++n) {
        const double *M = modes + 7 * n;
        const double c = cos(M[0] * X + M[1] * Y + M[2] * Z + M[3]);
        s0 += M[4] * c;
        s1 += M[5] * c;
        s2 += M[6] * c;
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
}

void or_inflow_eval(const or_grid *g, const double *modes, i64 n_modes, const double Uin[3], double shift,
                    double *cells, double *face) {
    const i64 G = g->gc, NY = g->n[1], NZ = g->n[2], ny = NY - 2 * G, nz = NZ - 2 * G;
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double xf = 0.5 * (X[G - 1] + X[G]);
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static)
#endif
    for (i64 j = 0; j < ny; ++j)
        for (i64 k = 0; k < nz; ++k) {
            double u[3];
            for (i64 i = 0; i < G; ++i) {
                inflow_point(modes, n_modes, X[i] - shift, Y[G + j], Z[G + k], u);
                double *c = cells + 3 * ((i * ny + j) * nz + k);
                for (int m = 0; m < 3; ++m) c[m] = Uin[m] + u[m];
            }
            inflow_point(modes, n_modes, xf - shift, Y[G + j], Z[G + k], u);
            face[j * nz + k] = Uin[0] + u[0];
        }
}

/* the x- ghost cells of U from or_inflow_eval's cells (in place of the constant) */
void or_apply_inflow_U(const or_grid *g, double *U, const double *cells) {
    const i64 G = g->gc, NY = g->n[1], NZ = g->n[2], ny = NY - 2 * G, nz = NZ - 2 * G;
    for (i64 i = 0; i < G; ++i)
        for (i64 j = 0; j < ny; ++j)
            for (i64 k = 0; k < nz; ++k)
                for (int m = 0; m < 3; ++m)
                    U[3 * lin(g, i, G + j, G + k) + m] = cells[3 * ((i * ny + j) * nz + k) + m];
}

/* the x- face flux JU = dy dz u_face (in place of dy dz Uin[0]) */
void or_apply_inflow_JU(const or_grid *g, double *JU, const double *face) {
    const double *DY = g->d[1], *DZ = g->d[2];
    const i64 G = g->gc, NY = g->n[1], NZ = g->n[2], ny = NY - 2 * G, nz = NZ - 2 * G;
    for (i64 j = 0; j < ny; ++j)
        for (i64 k = 0; k < nz; ++k) JU[3 * lin(g, G - 1, G + j, G + k) + 0] = DY[G + j] * DZ[G + k] * face[j * nz + k];
}

/* velocity ghosts: src/bc.h:5-233 */
void or_apply_Ubc(const or_grid *g, double *U, const double *Uold, const double Uin[3],
                  double dt, i64 retracted_floor, double floor_start_x) {
    const i64 G = g->gc, NX = g->n[0], NY = g->n[1], NZ = g->n[2];
    const i64 si = NY * NZ, sj = NZ;
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    /* x-: fixed inflow */
    for (i64 i = 0; i < G; ++i)
        for (i64 j = G; j < NY - G; ++j)
            for (i64 k = G; k < NZ - G; ++k)
                for (int m = 0; m < 3; ++m) U[3 * lin(g, i, j, k) + m] = Uin[m];
    /* x+: convective outflow advected by the current ghost u */
    for (i64 i = NX - G; i < NX; ++i)
        for (i64 j = G; j < NY - G; ++j)
            for (i64 k = G; k < NZ - G; ++k) {
                i64 c = lin(g, i, j, k);
                double uo = U[3 * c + 0];
                for (int m = 0; m < 3; ++m) {
                    double a0 = Uold[3 * c + m], a1 = Uold[3 * (c - si) + m], a2 = Uold[3 * (c - 2 * si) + m];
                    double gr = (3 * a0 - 4 * a1 + a2) / 2;
                    U[3 * c + m] = a0 - uo * dt * gr / DX[i];
                }
            }
    /* y- and y+: slip walls (v = 0 at the wall, u and w mirrored) */
    for (i64 i = G; i < NX - G; ++i)
        for (i64 k = G; k < NZ - G; ++k) {
            for (int side = 0; side < 2; ++side) {
                i64 jin2 = side ? NY - G - 2 : G + 1, jin1 = side ? NY - G - 1 : G;
                i64 jg1 = side ? NY - G : G - 1, jg2 = side ? NY - G + 1 : G - 2;
                double yw = side ? Y[jin1] + 0.5 * DY[jin1] : Y[jin1] - 0.5 * DY[jin1];
                i64 c1 = lin(g, i, jin1, k), c2 = lin(g, i, jin2, k);
                double wall[3] = {U[3 * c1 + 0], 0, U[3 * c1 + 2]};
                for (int m = 0; m < 3; ++m) {
                    double f0 = U[3 * c2 + m], f1 = U[3 * c1 + m];
                    U[3 * lin(g, i, jg1, k) + m] = quad3(Y[jin2], Y[jin1], yw, f0, f1, wall[m], Y[jg1]);
                    U[3 * lin(g, i, jg2, k) + m] = quad3(Y[jin2], Y[jin1], yw, f0, f1, wall[m], Y[jg2]);
                }
            }
        }
    (void)sj;
    /* z-: no-slip floor (slip upstream of a retracted floor); z+: slip */
    for (i64 i = G; i < NX - G; ++i)
        for (i64 j = G; j < NY - G; ++j) {
            for (int side = 0; side < 2; ++side) {
                i64 kin2 = side ? NZ - G - 2 : G + 1, kin1 = side ? NZ - G - 1 : G;
                i64 kg1 = side ? NZ - G : G - 1, kg2 = side ? NZ - G + 1 : G - 2;
                double zw = side ? Z[kin1] + 0.5 * DZ[kin1] : Z[kin1] - 0.5 * DZ[kin1];
                i64 c1 = lin(g, i, j, kin1), c2 = lin(g, i, j, kin2);
                double wall[3] = {0, 0, 0};
                if (side == 1 || (retracted_floor && X[i] < floor_start_x)) {
                    wall[0] = U[3 * c1 + 0];
                    wall[1] = U[3 * c1 + 1];
                }
                for (int m = 0; m < 3; ++m) {
                    double f0 = U[3 * c2 + m], f1 = U[3 * c1 + m];
                    U[3 * lin(g, i, j, kg1) + m] = quad3(Z[kin2], Z[kin1], zw, f0, f1, wall[m], Z[kg1]);
                    U[3 * lin(g, i, j, kg2) + m] = quad3(Z[kin2], Z[kin1], zw, f0, f1, wall[m], Z[kg2]);
                }
            }
        }
}

/* Smagorinsky eddy viscosity: src/cfd.h:785-826 */
void or_nut(const or_grid *g, const double *U, double *nut, double Cs) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    PAR_INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double hx = X[i + 1] - X[i - 1], hy = Y[j + 1] - Y[j - 1], hz = Z[k + 1] - Z[k - 1];
        double vol = DX[i] * DY[j] * DZ[k];
        double G3[3][3];
        for (int m = 0; m < 3; ++m) {
            G3[m][0] = (U[3 * (c + si) + m] - U[3 * (c - si) + m]) / hx;
            G3[m][1] = (U[3 * (c + sj) + m] - U[3 * (c - sj) + m]) / hy;
            G3[m][2] = (U[3 * (c + 1) + m] - U[3 * (c - 1) + m]) / hz;
        }
        double S = sqrt(2 * sq(G3[0][0]) + 2 * sq(G3[1][1]) + 2 * sq(G3[2][2]) +
                        sq(G3[0][1] + G3[1][0]) + sq(G3[0][2] + G3[2][0]) + sq(G3[1][2] + G3[2][1]));
        nut[c] = sq(Cs * cbrt(vol)) * S;
    }
}

/* Solid fraction of the turbulence-generating grid: src/solver.h:52-159, with
 * get_intersection (src/util.h:84-90) as the overlap of two intervals. */
static double overlap1(double h1, double t1, double h2, double t2) {
    double l1 = t1 - h1, l2 = t2 - h2, x1 = t1 - h2, x2 = t2 - h1;
    return fmax(0., fmin(l1, fmin(l2, fmin(x1, x2))));
}
void or_tgg_solid(const or_grid *g, double *solid, double thick, double mesh, double bar, const double plc[3]) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    INTERIOR(g, i, j, k) {
        double xc = X[i], yc = Y[j], zc = Z[k], hx = DX[i], hy = DY[j], hz = DZ[k];
        double xv = overlap1(xc - 0.5 * hx, xc + 0.5 * hx, plc[0] - thick, plc[0]);
        i64 nj = (i64)round((yc - plc[1]) / mesh);
        double by = nj * mesh + plc[1];
        double yi = overlap1(yc - 0.5 * hy, yc + 0.5 * hy, by - 0.5 * bar, by + 0.5 * bar);
        double xh = overlap1(xc - 0.5 * hx, xc + 0.5 * hx, plc[0] - 2 * thick, plc[0] - thick);
        i64 nk = (i64)round((zc - plc[2]) / mesh);
        double bz = nk * mesh + plc[2];
        double zi = overlap1(zc - 0.5 * hz, zc + 0.5 * hz, bz - 0.5 * bar, bz + 0.5 * bar);
        solid[lin(g, i, j, k)] = (xv * yi * hz + xh * zi * hy) / (hx * hy * hz);
    }
}

/* Q-criterion: src/cfd.h:948-976 (calc_q_at_cell), over the interior as calc_q
 * does on one rank (src/cfd.h:978-1077); other cells are left as they are. */
void or_calc_q(const or_grid *g, const double *U, double *q) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double hx = X[i + 1] - X[i - 1], hy = Y[j + 1] - Y[j - 1], hz = Z[k + 1] - Z[k - 1];
        double G3[3][3]; /* G3[m][a] = d U_m / d x_a */
        for (int m = 0; m < 3; ++m) {
            G3[m][0] = (U[3 * (c + si) + m] - U[3 * (c - si) + m]) / hx;
            G3[m][1] = (U[3 * (c + sj) + m] - U[3 * (c - sj) + m]) / hy;
            G3[m][2] = (U[3 * (c + 1) + m] - U[3 * (c - 1) + m]) / hz;
        }
        double diag = G3[0][0] * G3[0][0] + G3[1][1] * G3[1][1] + G3[2][2] * G3[2][2];
        double cross = G3[0][1] * G3[1][0] + G3[0][2] * G3[2][0] + G3[1][2] * G3[2][1];
        q[c] = -0.5 * (diag + 2 * cross);
    }
}

/* Time average of U over every padded cell, src/solver.h:1194-1204 with
 * calc_ax_plus_by<3> (src/mv.h:321-335): Utavg = (1 - w) Utavg + w U. */
void or_time_average(const or_grid *g, double *Utavg, const double *U, double weight) {
    const double a = 1. - weight, b = weight;
    const i64 n = 3 * g->n[0] * g->n[1] * g->n[2];
    for (i64 e = 0; e < n; ++e) Utavg[e] = a * Utavg[e] + b * U[e];
}

/* src/cfd.h:1109-1141 */
double or_max_cfl(const or_grid *g, const double *U, double dt) {
    double mx = 0;
    PAR_INTERIOR_MAX(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double v = fabs(U[3 * c + 0]) * dt / g->d[0][i] + fabs(U[3 * c + 1]) * dt / g->d[1][j] +
                   fabs(U[3 * c + 2]) * dt / g->d[2][k];
        if (v > mx) mx = v;
    }
    return mx;
}

/* ========================================================================== */
/* ALM */


/* AP properties: src/alm.h:35-103 (stations) + 105-142 (flattening).
 * af_cd/af_cl are [af_count][atk_count]; cd_tbl/cl_tbl are [ap_count][atk_count]. */
void or_build_ap_props(const double *af_r, const double *af_chord, const double *af_twist,
                       const double *af_cd, const double *af_cl, i64 af_count, i64 atk_count,
                       double R, double hub_R, i64 blades, i64 wt_count, i64 ap_per_blade,
                       double *ap, double *cd_tbl, double *cl_tbl) {
    i64 n = wt_count * blades * ap_per_blade;
    double dr = (R - hub_R) / ap_per_blade;
    for (i64 a = 0; a < n; ++a) {
        double *P = ap + OR_AP_STRIDE * a;
        memset(P, 0, sizeof(double) * OR_AP_STRIDE);
        double r = hub_R + (a % ap_per_blade + 0.5) * dr;
        P[3] = r;
        P[7] = dr;
        i64 s = floor_index(af_r, r, af_count);
        double *cd = cd_tbl + atk_count * a, *cl = cl_tbl + atk_count * a;
        if (s < 0 || s >= af_count - 1) {
            i64 e = s < 0 ? 0 : af_count - 1;
            P[4] = af_chord[e];
            P[5] = af_twist[e];
            for (i64 m = 0; m < atk_count; ++m) {
                cd[m] = af_cd[e * atk_count + m];
                cl[m] = af_cl[e * atk_count + m];
            }
        } else {
            double w1 = (r - af_r[s]) / (af_r[s + 1] - af_r[s]);
            double w0 = 1 - w1;
            P[4] = w0 * af_chord[s] + w1 * af_chord[s + 1];
            P[5] = w0 * af_twist[s] + w1 * af_twist[s + 1];
            for (i64 m = 0; m < atk_count; ++m) {
                cd[m] = w0 * af_cd[s * atk_count + m] + w1 * af_cd[(s + 1) * atk_count + m];
                cl[m] = w0 * af_cl[s * atk_count + m] + w1 * af_cl[(s + 1) * atk_count + m];
            }
        }
    }
}



/* src/alm.h:233-267 */
void or_update_ap_position(double *wt, i64 wt_count, double *ap, i64 ap_count,
                           i64 blade_per_wt, i64 ap_per_blade, double t) {
    for (i64 w = 0; w < wt_count; ++w) {
        double *T = wt + OR_WT_STRIDE * w;
        double A = T[9], B = T[10], C = T[11], D = T[12];
        T[7] = A * sin(B * t + C) + D;
        T[8] = A * B * cos(B * t + C);
    }
    i64 per_wt = ap_per_blade * blade_per_wt;
    for (i64 a = 0; a < ap_count; ++a) {
        double *P = ap + OR_AP_STRIDE * a;
        const double *T = wt + OR_WT_STRIDE * (a / per_wt);
        i64 blade = (a % per_wt) / ap_per_blade;
        double th0 = blade * (2 * PI_ / blade_per_wt);
        P[6] = th0 + t * T[13];
        double loc[3] = {T[3], T[4] + P[3] * cos(P[6]), T[5] + P[3] * sin(P[6])};
        rotate(loc, P, -T[7], (int)T[6]);
        P[0] += T[0];
        P[1] += T[1];
        P[2] += T[2];
    }
}

/* src/alm.h:269-414 (single rank). An optional collective blade pitch in P[13]
 * (degrees) is subtracted from the angle of attack; with P[13] = 0 the result is
 * the reference's. */
void or_calc_ap_force(const or_grid *g, const double *U, double *wt, i64 wt_count,
                      double *ap, i64 ap_count, i64 blade_per_wt, i64 ap_per_blade,
                      const double *cd_tbl, const double *cl_tbl, const double *atk,
                      i64 atk_count) {
    const i64 G = g->gc;
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    i64 per_wt = ap_per_blade * blade_per_wt;
    for (i64 a = 0; a < ap_count; ++a) {
        double *P = ap + OR_AP_STRIDE * a;
        const double *T = wt + OR_WT_STRIDE * (a / per_wt);
        i64 i = floor_index(X, P[0], g->n[0]);
        i64 j = floor_index(Y, P[1], g->n[1]);
        i64 k = floor_index(Z, P[2], g->n[2]);
        if (i >= G && i < g->n[0] - G && j >= G && j < g->n[1] - G && k >= G && k < g->n[2] - G) {
            double fx = (P[0] - X[i]) / (X[i + 1] - X[i]);
            double fy = (P[1] - Y[j]) / (Y[j + 1] - Y[j]);
            double fz = (P[2] - Z[k]) / (Z[k + 1] - Z[k]);
            double Uap[3];
            for (int m = 0; m < 3; ++m) {
                double v[8];
                for (int o = 0; o < 8; ++o)
                    v[o] = U[3 * lin(g, i + (o & 1), j + ((o >> 1) & 1), k + ((o >> 2) & 1)) + m];
                double v01 = v[0] * (1 - fx) + v[1] * fx, v23 = v[2] * (1 - fx) + v[3] * fx;
                double v45 = v[4] * (1 - fx) + v[5] * fx, v67 = v[6] * (1 - fx) + v[7] * fx;
                double lo = v01 * (1 - fy) + v23 * fy, hi = v45 * (1 - fy) + v67 * fy;
                Uap[m] = lo * (1 - fz) + hi * fz;
            }
            double rel[3] = {P[0] - T[0], P[1] - T[1], P[2] - T[2]};
            double Ur[3];
            rotate_U(rel, Uap, Ur, T[7], T[8], (int)T[6]);
            double ux = Ur[0];
            double ut = P[3] * T[13] + Ur[1] * sin(P[6]) - Ur[2] * cos(P[6]);
            double V2 = ux * ux + ut * ut;
            double phi = atan(ux / ut);
            double aoa = phi * 180 / PI_ - P[5] - P[13];
            double cd, cl;
            const double *CD = cd_tbl + atk_count * a, *CL = cl_tbl + atk_count * a;
            i64 s = floor_index(atk, aoa, atk_count);
            if (s < 0) { cd = CD[0]; cl = CL[0]; }
            else if (s >= atk_count - 1) { cd = CD[atk_count - 1]; cl = CL[atk_count - 1]; }
            else {
                double w = (aoa - atk[s]) / (atk[s + 1] - atk[s]);
                cd = CD[s] * (1 - w) + CD[s + 1] * w;
                cl = CL[s] * (1 - w) + CL[s + 1] * w;
            }
            double Fd = 0.5 * cd * V2 * P[4] * P[7];
            double Fl = 0.5 * cl * V2 * P[4] * P[7];
            double Fx = Fl * cos(phi) + Fd * sin(phi);
            double Ft = Fl * sin(phi) - Fd * cos(phi);
            Ft *= sgn(T[13]);
            double loc[3] = {-Fx, Ft * sin(P[6]), -Ft * cos(P[6])};
            rotate(loc, P + 8, -T[7], (int)T[6]);
            P[11] = fabs(Ft) * P[3];
            P[12] = Fx;
        } else {
            P[8] = P[9] = P[10] = 0;
            P[11] = P[12] = 0;
        }
    }
    for (i64 w = 0; w < wt_count; ++w) {
        double tq = 0, th = 0;
        for (i64 a = w * per_wt; a < (w + 1) * per_wt; ++a) {
            tq += ap[OR_AP_STRIDE * a + 11];
            th += ap[OR_AP_STRIDE * a + 12];
        }
        wt[OR_WT_STRIDE * w + 14] = tq;
        wt[OR_WT_STRIDE * w + 15] = th;
    }
}

/* dense Gaussian projection: src/alm.h:416-448 */
void or_distribute_ap_force(const or_grid *g, double *f, const double *ap, i64 ap_count,
                            double width) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    PAR_INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double sx = 0, sy = 0, sz = 0;
        double norm = 1 / ((width * sqrt(PI_)) * (width * sqrt(PI_)) * (width * sqrt(PI_)));
        for (i64 a = 0; a < ap_count; ++a) {
            const double *P = ap + OR_AP_STRIDE * a;
            double r2 = sq(X[i] - P[0]) + sq(Y[j] - P[1]) + sq(Z[k] - P[2]);
            double wgt = norm * exp(-r2 / sq(width));
            sx += wgt * P[8];
            sy += wgt * P[9];
            sz += wgt * P[10];
        }
        f[3 * c + 0] = sx;
        f[3 * c + 1] = sy;
        f[3 * c + 2] = sz;
    }
}

static void run_alm(const or_grid *g, or_state *s, double t) {
    if (s->ap_count <= 0) return;
    or_update_ap_position(s->wt, s->wt_count, s->ap, s->ap_count, s->blade_per_wt,
                          s->ap_per_blade, t);
    or_calc_ap_force(g, s->U, s->wt, s->wt_count, s->ap, s->ap_count, s->blade_per_wt,
                     s->ap_per_blade, s->cd_tbl, s->cl_tbl, s->atk, s->atk_count);
    or_distribute_ap_force(g, s->f, s->ap, s->ap_count, s->width);
}

/* the turbulent inflow of the step's time level (nothing without modes) */
static void inflow_at(const or_grid *g, or_state *s, double t) {
    if (s->inflow_n > 0)
        or_inflow_eval(g, s->inflow_modes, s->inflow_n, s->Uin, s->inflow_speed * t, s->inflow_cells,
                       s->inflow_face);
}

static void apply_JUbc_state(const or_grid *g, or_state *s) {
    or_apply_JUbc(g, s->JU, s->Uin);
    if (s->inflow_n > 0) or_apply_inflow_JU(g, s->JU, s->inflow_face);
}

/* ========================================================================== */
/* start-up pass after the fields are filled: src/solver.h:727-833 */
void or_state_initialize(const or_grid *g, or_state *s) {
    s->max_diag = or_build_A(g, s->A);
    or_set_solid(g, s->U, s->solid);
    inflow_at(g, s, s->dt * s->step);
    or_apply_Ubc(g, s->U, s->Uold, s->Uin, s->dt, s->retracted_floor, s->floor_start_x);
    if (s->inflow_n > 0) or_apply_inflow_U(g, s->U, s->inflow_cells);
    or_interpolate_JU(g, s->U, s->JU);
    apply_JUbc_state(g, s);
    or_nut(g, s->U, s->nut, s->Cs);
    or_divergence(g, s->JU, s->div);
    run_alm(g, s, s->dt * s->step);
    s->avg_div = or_l2_norm(g, s->div) / sqrt((double)n_interior(g));
    s->max_cfl = or_max_cfl(g, s->U, s->dt);
}

/* one time step: src/solver.h:1034-1190 (step counter advanced first, so the
 * first call integrates step 1 at t = dt) */
void or_state_step(const or_grid *g, or_state *s) {
    const i64 L = g->n[0] * g->n[1] * g->n[2];
    s->step += 1;
    memcpy(s->Uold, s->U, sizeof(double) * 3 * (size_t)L);
    inflow_at(g, s, s->dt * s->step);
    or_predictor(g, s->U, s->Uold, s->JU, s->nut, s->f, s->Re, s->dt);
    or_interpolate_JU(g, s->U, s->JU);
    apply_JUbc_state(g, s);
    or_poisson_rhs(g, s->JU, s->b, s->dt, s->max_diag);
    if (s->method == 1) {
        or_run_bicgstab(g, s->A, s->p, s->b, s->work, s->pc_kind, s->pc_omega, s->pc_it, &s->it,
                        s->max_it, &s->err, s->tol);
    } else {
        or_run_sor(g, s->A, s->p, s->b, s->work, s->omega, &s->it, s->max_it, &s->err, s->tol);
    }
    or_apply_pbc(g, s->p);
    or_project(g, s->U, s->JU, s->p, s->dt);
    or_set_solid(g, s->U, s->solid);
    or_apply_Ubc(g, s->U, s->Uold, s->Uin, s->dt, s->retracted_floor, s->floor_start_x);
    if (s->inflow_n > 0) or_apply_inflow_U(g, s->U, s->inflow_cells);
    apply_JUbc_state(g, s);
    or_nut(g, s->U, s->nut, s->Cs);
    or_divergence(g, s->JU, s->div);
    run_alm(g, s, s->dt * s->step);
    s->avg_div = or_l2_norm(g, s->div) / sqrt((double)n_interior(g));
    s->max_cfl = or_max_cfl(g, s->U, s->dt);
}
