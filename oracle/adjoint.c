/*
 * adjoint.c -- discrete adjoint (reverse mode) of the oracle's time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP adjoint kernels, used by
 * tests/ alone.
 *
 * The reference has no adjoint.  Each routine here is the transpose of the
 * Jacobian of one forward routine of oracle.c, written loop for loop in scatter
 * form: every forward statement y = f(x) becomes x_b += (df/dx)^T y_b, in the
 * reverse order.  Non-differentiable points (the MUSCL limiter's min/max, |J| in
 * the upwind flux, |F_t| in the torque, the table lookup's nodes) take the branch
 * the forward took.  The Poisson solve is differentiated as the exact solve
 * p = A^-1 b, so the adjoint solve is A^T y = p_b; since V*A is symmetric (V the
 * cell volumes), A^T = V A V^-1 and the forward solver is reused on V^-1 p_b.
 *
 * Pinned by tests/test_oracle_adjoint.py: every kernel by the dot-product identity
 * <y_b, J x_d> = <J^T y_b, x_d> with J x_d from central differences of the forward
 * routine, and the whole run by central differences of the time-averaged turbine
 * power with respect to yaw, collective pitch, chord and twist.
 */
#define _DEFAULT_SOURCE
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_util.h"

#define FACE_BOX(g, i, j, k)                                   \
    for (i64 i = (g)->gc - 1; i < (g)->n[0] - (g)->gc; ++i)    \
        for (i64 j = (g)->gc - 1; j < (g)->n[1] - (g)->gc; ++j) \
            for (i64 k = (g)->gc - 1; k < (g)->n[2] - (g)->gc; ++k)

static inline int is_interior(const or_grid *g, i64 i, i64 j, i64 k) {
    return i >= g->gc && i < g->n[0] - g->gc && j >= g->gc && j < g->n[1] - g->gc &&
           k >= g->gc && k < g->n[2] - g->gc;
}

/* ========================================================================== */
/* MUSCL limiter h = t*max(0, min(|a|, t*bb*b)), t = sgn(a): partials in a, b */
static void lim_d(double a, double b, double bb, double *da, double *db) {
    double t = sgn(a);
    *da = 0;
    *db = 0;
    if (t == 0) return;
    double inner = fmin(fabs(a), t * bb * b);
    if (!(inner > 0)) return;
    if (fabs(a) <= t * bb * b) *da = 1; /* h = t|a| = a */
    else *db = bb;                      /* h = t*t*bb*b = bb*b */
}

/* reverse of the convection of one component at one cell (oracle.c muscl_cell):
 * conv_b -> bars of the 5-point lines vb[d][0..4] and of the 6 face fluxes Jb */
static void muscl_cell_adj(const double vx[5], const double vy[5], const double vz[5],
                           const double J[6], double dxc, double dyc, double dzc, double conv_b,
                           double vb[3][5], double Jb[6]) {
    const double kap = 1. / 3.;
    const double bb = (3. - kap) / (1. - kap);
    const double q = 1. / 4.;
    const double *vs[3] = {vx, vy, vz};
    double vol = dxc * dyc * dzc;
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
        double Fpb = conv_b / vol, Fmb = -conv_b / vol;
        Jb[2 * d] += Fpb * 0.5 * ((pr + pl) - sgn(Jp) * (pr - pl));
        Jb[2 * d + 1] += Fmb * 0.5 * ((mr + ml) - sgn(Jm) * (mr - ml));
        double prb = Fpb * 0.5 * (Jp - fabs(Jp)), plb = Fpb * 0.5 * (Jp + fabs(Jp));
        double mrb = Fmb * 0.5 * (Jm - fabs(Jm)), mlb = Fmb * 0.5 * (Jm + fabs(Jm));
        double h1b = 0, h2b = 0, h3b = 0, h4b = 0, h5b = 0, h6b = 0;
        vb[d][3] += prb;
        h6b -= q * (1. - kap) * prb;
        h5b -= q * (1. + kap) * prb;
        vb[d][2] += plb;
        h3b += q * (1. - kap) * plb;
        h4b += q * (1. + kap) * plb;
        vb[d][2] += mrb;
        h4b -= q * (1. - kap) * mrb;
        h3b -= q * (1. + kap) * mrb;
        vb[d][1] += mlb;
        h1b += q * (1. - kap) * mlb;
        h2b += q * (1. + kap) * mlb;
        double a1b = 0, a2b = 0, a3b = 0, a4b = 0, da, db;
        lim_d(a4, a3, bb, &da, &db); a4b += h6b * da; a3b += h6b * db;
        lim_d(a3, a4, bb, &da, &db); a3b += h5b * da; a4b += h5b * db;
        lim_d(a3, a2, bb, &da, &db); a3b += h4b * da; a2b += h4b * db;
        lim_d(a2, a3, bb, &da, &db); a2b += h3b * da; a3b += h3b * db;
        lim_d(a2, a1, bb, &da, &db); a2b += h2b * da; a1b += h2b * db;
        lim_d(a1, a2, bb, &da, &db); a1b += h1b * da; a2b += h1b * db;
        vb[d][1] += a1b; vb[d][0] -= a1b;
        vb[d][2] += a2b; vb[d][1] -= a2b;
        vb[d][3] += a3b; vb[d][2] -= a3b;
        vb[d][4] += a4b; vb[d][3] -= a4b;
    }
}

/* reverse of or_predictor */
void or_predictor_adj(const or_grid *g, const double *Ub, const double *Uold, const double *JU,
                      const double *nut, double Re, double dt, double *Uold_b, double *JU_b,
                      double *nut_b, double *f_b) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    const i64 st[3] = {si, sj, 1};
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double J[6] = {JU[3 * c + 0], JU[3 * (c - si) + 0], JU[3 * c + 1],
                       JU[3 * (c - sj) + 1], JU[3 * c + 2], JU[3 * (c - 1) + 2]};
        double nu = 1 / Re + nut[c];
        double Jb[6] = {0, 0, 0, 0, 0, 0};
        for (int m = 0; m < 3; ++m) {
            double ub = Ub[3 * c + m];
            double vx[5], vy[5], vz[5];
            for (int o = -2; o <= 2; ++o) {
                vx[o + 2] = Uold[3 * (c + o * si) + m];
                vy[o + 2] = Uold[3 * (c + o * sj) + m];
                vz[o + 2] = Uold[3 * (c + o) + m];
            }
            Uold_b[3 * c + m] += ub;
            f_b[3 * c + m] += dt * ub;
            double vb[3][5] = {{0}};
            muscl_cell_adj(vx, vy, vz, J, DX[i], DY[j], DZ[k], -dt * ub, vb, Jb);
            double difb = dt * ub;
            double uc = vx[2];
            double ex = X[i + 1] - X[i], wx = X[i] - X[i - 1];
            double ey = Y[j + 1] - Y[j], wy = Y[j] - Y[j - 1];
            double ez = Z[k + 1] - Z[k], wz = Z[k] - Z[k - 1];
            double lap = ((vx[3] - uc) / ex - (uc - vx[1]) / wx) / DX[i] +
                         ((vy[3] - uc) / ey - (uc - vy[1]) / wy) / DY[j] +
                         ((vz[3] - uc) / ez - (uc - vz[1]) / wz) / DZ[k];
            nut_b[c] += difb * lap;
            double s = difb * nu;
            vb[0][3] += s / (ex * DX[i]); vb[0][1] += s / (wx * DX[i]); vb[0][2] -= s * (1 / ex + 1 / wx) / DX[i];
            vb[1][3] += s / (ey * DY[j]); vb[1][1] += s / (wy * DY[j]); vb[1][2] -= s * (1 / ey + 1 / wy) / DY[j];
            vb[2][3] += s / (ez * DZ[k]); vb[2][1] += s / (wz * DZ[k]); vb[2][2] -= s * (1 / ez + 1 / wz) / DZ[k];
            for (int d = 0; d < 3; ++d)
                for (int o = -2; o <= 2; ++o) Uold_b[3 * (c + o * st[d]) + m] += vb[d][o + 2];
        }
        JU_b[3 * c + 0] += Jb[0];
        JU_b[3 * (c - si) + 0] += Jb[1];
        JU_b[3 * c + 1] += Jb[2];
        JU_b[3 * (c - sj) + 1] += Jb[3];
        JU_b[3 * c + 2] += Jb[4];
        JU_b[3 * (c - 1) + 2] += Jb[5];
    }
}

/* reverse of or_interpolate_JU: U_b += I^T JU_b over the face box */
void or_interpolate_JU_adj(const or_grid *g, const double *JU_b, double *U_b) {
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    FACE_BOX(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double bx = 0.5 * DY[j] * DZ[k] * JU_b[3 * c + 0];
        U_b[3 * c + 0] += bx;
        U_b[3 * (c + si) + 0] += bx;
        double by = 0.5 * DX[i] * DZ[k] * JU_b[3 * c + 1];
        U_b[3 * c + 1] += by;
        U_b[3 * (c + sj) + 1] += by;
        double bz = 0.5 * DX[i] * DY[j] * JU_b[3 * c + 2];
        U_b[3 * c + 2] += bz;
        U_b[3 * (c + 1) + 2] += bz;
    }
}

/* reverse of or_apply_JUbc: the faces it overwrites carry no sensitivity */
void or_apply_JUbc_adj(const or_grid *g, double *JU_b) {
    const i64 G = g->gc, NX = g->n[0], NY = g->n[1], NZ = g->n[2];
    for (i64 j = G; j < NY - G; ++j)
        for (i64 k = G; k < NZ - G; ++k) JU_b[3 * lin(g, G - 1, j, k) + 0] = 0;
    for (i64 i = G; i < NX - G; ++i)
        for (i64 k = G; k < NZ - G; ++k) {
            JU_b[3 * lin(g, i, G - 1, k) + 1] = 0;
            JU_b[3 * lin(g, i, NY - G - 1, k) + 1] = 0;
        }
    for (i64 i = G; i < NX - G; ++i)
        for (i64 j = G; j < NY - G; ++j) {
            JU_b[3 * lin(g, i, j, G - 1) + 2] = 0;
            JU_b[3 * lin(g, i, j, NZ - G - 1) + 2] = 0;
        }
}

/* reverse of or_poisson_rhs */
void or_poisson_rhs_adj(const or_grid *g, const double *b_b, double *JU_b, double dt, double scale) {
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double s = b_b[c] / (dt * scale) / (g->d[0][i] * g->d[1][j] * g->d[2][k]);
        JU_b[3 * c + 0] += s;
        JU_b[3 * (c - si) + 0] -= s;
        JU_b[3 * c + 1] += s;
        JU_b[3 * (c - sj) + 1] -= s;
        JU_b[3 * c + 2] += s;
        JU_b[3 * (c - 1) + 2] -= s;
    }
}

/* reverse of or_apply_pbc: ghost bars fold onto the interior cells they copy */
void or_apply_pbc_adj(const or_grid *g, double *p_b) {
    const i64 G = g->gc, NX = g->n[0], NY = g->n[1], NZ = g->n[2];
    const double *DX = g->d[0];
    for (i64 i = G; i < NX - G; ++i)
        for (i64 j = G; j < NY - G; ++j) {
            p_b[lin(g, i, j, G)] += p_b[lin(g, i, j, G - 1)];
            p_b[lin(g, i, j, G - 1)] = 0;
            p_b[lin(g, i, j, NZ - G - 1)] += p_b[lin(g, i, j, NZ - G)];
            p_b[lin(g, i, j, NZ - G)] = 0;
        }
    for (i64 i = G; i < NX - G; ++i)
        for (i64 k = G; k < NZ - G; ++k) {
            p_b[lin(g, i, G, k)] += p_b[lin(g, i, G - 1, k)];
            p_b[lin(g, i, G - 1, k)] = 0;
            p_b[lin(g, i, NY - G - 1, k)] += p_b[lin(g, i, NY - G, k)];
            p_b[lin(g, i, NY - G, k)] = 0;
        }
    for (i64 j = G; j < NY - G; ++j)
        for (i64 k = G; k < NZ - G; ++k) {
            p_b[lin(g, G, j, k)] += p_b[lin(g, G - 1, j, k)];
            p_b[lin(g, G - 1, j, k)] = 0;
            i64 ii = NX - G - 1, io = NX - G;
            double hi = 0.5 * DX[ii], ho = 0.5 * DX[io];
            p_b[lin(g, ii, j, k)] -= p_b[lin(g, io, j, k)] * (ho / hi);
            p_b[lin(g, io, j, k)] = 0;
        }
}

/* reverse of or_project: U and JU pass through; p_b += -dt (Gc^T U_b + Gf^T JU_b) */
void or_project_adj(const or_grid *g, const double *U_b, const double *JU_b, double *p_b, double dt) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double bx = dt * U_b[3 * c + 0] / (X[i + 1] - X[i - 1]);
        p_b[c + si] -= bx;
        p_b[c - si] += bx;
        double by = dt * U_b[3 * c + 1] / (Y[j + 1] - Y[j - 1]);
        p_b[c + sj] -= by;
        p_b[c - sj] += by;
        double bz = dt * U_b[3 * c + 2] / (Z[k + 1] - Z[k - 1]);
        p_b[c + 1] -= bz;
        p_b[c - 1] += bz;
    }
    FACE_BOX(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double gx = dt * (DY[j] * DZ[k]) * JU_b[3 * c + 0] / (X[i + 1] - X[i]);
        p_b[c + si] -= gx;
        p_b[c] += gx;
        double gy = dt * (DX[i] * DZ[k]) * JU_b[3 * c + 1] / (Y[j + 1] - Y[j]);
        p_b[c + sj] -= gy;
        p_b[c] += gy;
        double gz = dt * (DX[i] * DY[j]) * JU_b[3 * c + 2] / (Z[k + 1] - Z[k]);
        p_b[c + 1] -= gz;
        p_b[c] += gz;
    }
}

/* reverse of or_set_solid */
void or_set_solid_adj(const or_grid *g, double *U_b, const double *solid) {
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        for (int m = 0; m < 3; ++m) U_b[3 * c + m] *= (1. - solid[c]);
    }
}

static void quad3_coef(double x0, double x1, double x2, double x3, double l[3]) {
    l[0] = ((x3 - x1) * (x3 - x2)) / ((x0 - x1) * (x0 - x2));
    l[1] = ((x3 - x0) * (x3 - x2)) / ((x1 - x0) * (x1 - x2));
    l[2] = ((x3 - x0) * (x3 - x1)) / ((x2 - x0) * (x2 - x1));
}

/* reverse of or_apply_Ubc, in place on U_b.  On entry U_b is the bar of U after
 * the boundary pass; on exit it is the bar of U before it (U_pre): wall ghosts are
 * folded onto the interior cells they extrapolate, the x+ ghost keeps only the bar
 * of its own advecting velocity U_pre.x (the stale ghost value), the x- ghosts are
 * cleared.  The convective outflow's dependence on Uold goes into Uold_b. */
void or_apply_Ubc_adj(const or_grid *g, double *U_b, const double *U_pre, const double *Uold,
                      double *Uold_b, double dt, i64 retracted_floor, double floor_start_x) {
    const i64 G = g->gc, NX = g->n[0], NY = g->n[1], NZ = g->n[2];
    const i64 si = NY * NZ;
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    /* z walls */
    for (i64 i = G; i < NX - G; ++i)
        for (i64 j = G; j < NY - G; ++j)
            for (int side = 0; side < 2; ++side) {
                i64 kin2 = side ? NZ - G - 2 : G + 1, kin1 = side ? NZ - G - 1 : G;
                i64 kg1 = side ? NZ - G : G - 1, kg2 = side ? NZ - G + 1 : G - 2;
                double zw = side ? Z[kin1] + 0.5 * DZ[kin1] : Z[kin1] - 0.5 * DZ[kin1];
                i64 c1 = lin(g, i, j, kin1), c2 = lin(g, i, j, kin2);
                int wall_is_u[3] = {0, 0, 0};
                if (side == 1 || (retracted_floor && X[i] < floor_start_x)) wall_is_u[0] = wall_is_u[1] = 1;
                i64 gh[2] = {lin(g, i, j, kg1), lin(g, i, j, kg2)};
                double xg[2] = {Z[kg1], Z[kg2]};
                for (int o = 0; o < 2; ++o) {
                    double l[3];
                    quad3_coef(Z[kin2], Z[kin1], zw, xg[o], l);
                    for (int m = 0; m < 3; ++m) {
                        double b = U_b[3 * gh[o] + m];
                        U_b[3 * c2 + m] += b * l[0];
                        U_b[3 * c1 + m] += b * l[1];
                        if (wall_is_u[m]) U_b[3 * c1 + m] += b * l[2];
                        U_b[3 * gh[o] + m] = 0;
                    }
                }
            }
    /* y walls */
    for (i64 i = G; i < NX - G; ++i)
        for (i64 k = G; k < NZ - G; ++k)
            for (int side = 0; side < 2; ++side) {
                i64 jin2 = side ? NY - G - 2 : G + 1, jin1 = side ? NY - G - 1 : G;
                i64 jg1 = side ? NY - G : G - 1, jg2 = side ? NY - G + 1 : G - 2;
                double yw = side ? Y[jin1] + 0.5 * DY[jin1] : Y[jin1] - 0.5 * DY[jin1];
                i64 c1 = lin(g, i, jin1, k), c2 = lin(g, i, jin2, k);
                const int wall_is_u[3] = {1, 0, 1};
                i64 gh[2] = {lin(g, i, jg1, k), lin(g, i, jg2, k)};
                double xg[2] = {Y[jg1], Y[jg2]};
                for (int o = 0; o < 2; ++o) {
                    double l[3];
                    quad3_coef(Y[jin2], Y[jin1], yw, xg[o], l);
                    for (int m = 0; m < 3; ++m) {
                        double b = U_b[3 * gh[o] + m];
                        U_b[3 * c2 + m] += b * l[0];
                        U_b[3 * c1 + m] += b * l[1];
                        if (wall_is_u[m]) U_b[3 * c1 + m] += b * l[2];
                        U_b[3 * gh[o] + m] = 0;
                    }
                }
            }
    /* x+: convective outflow */
    for (i64 i = NX - G; i < NX; ++i)
        for (i64 j = G; j < NY - G; ++j)
            for (i64 k = G; k < NZ - G; ++k) {
                i64 c = lin(g, i, j, k);
                double uo = U_pre[3 * c + 0];
                double uob = 0;
                for (int m = 0; m < 3; ++m) {
                    double a0 = Uold[3 * c + m], a1 = Uold[3 * (c - si) + m], a2 = Uold[3 * (c - 2 * si) + m];
                    double gr = (3 * a0 - 4 * a1 + a2) / 2;
                    double b = U_b[3 * c + m];
                    Uold_b[3 * c + m] += b;
                    uob -= b * dt * gr / DX[i];
                    double grb = -b * uo * dt / DX[i];
                    Uold_b[3 * c + m] += 1.5 * grb;
                    Uold_b[3 * (c - si) + m] -= 2.0 * grb;
                    Uold_b[3 * (c - 2 * si) + m] += 0.5 * grb;
                }
                U_b[3 * c + 0] = uob;
                U_b[3 * c + 1] = 0;
                U_b[3 * c + 2] = 0;
            }
    /* x-: fixed inflow */
    for (i64 i = 0; i < G; ++i)
        for (i64 j = G; j < NY - G; ++j)
            for (i64 k = G; k < NZ - G; ++k)
                for (int m = 0; m < 3; ++m) U_b[3 * lin(g, i, j, k) + m] = 0;
}

/* reverse of or_nut */
void or_nut_adj(const or_grid *g, const double *U, const double *nut_b, double *U_b, double Cs) {
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    const double *DX = g->d[0], *DY = g->d[1], *DZ = g->d[2];
    const i64 si = g->n[1] * g->n[2], sj = g->n[2];
    const i64 st[3] = {si, sj, 1};
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        double h[3] = {X[i + 1] - X[i - 1], Y[j + 1] - Y[j - 1], Z[k + 1] - Z[k - 1]};
        double vol = DX[i] * DY[j] * DZ[k];
        double G3[3][3];
        for (int m = 0; m < 3; ++m)
            for (int d = 0; d < 3; ++d) G3[m][d] = (U[3 * (c + st[d]) + m] - U[3 * (c - st[d]) + m]) / h[d];
        double S = sqrt(2 * sq(G3[0][0]) + 2 * sq(G3[1][1]) + 2 * sq(G3[2][2]) +
                        sq(G3[0][1] + G3[1][0]) + sq(G3[0][2] + G3[2][0]) + sq(G3[1][2] + G3[2][1]));
        if (!(S > 0)) continue;
        double nb = nut_b[c] * sq(Cs * cbrt(vol)) / S;
        for (int m = 0; m < 3; ++m)
            for (int d = 0; d < 3; ++d) {
                double dS = (m == d) ? 2 * G3[m][m] : (G3[m][d] + G3[d][m]);
                double gb = nb * dS / h[d];
                U_b[3 * (c + st[d]) + m] += gb;
                U_b[3 * (c - st[d]) + m] -= gb;
            }
    }
}

/* ========================================================================== */
/* ALM.  Rotation matrices of oracle_util.h's rotate() and their derivatives. */
static void rot_mat(double ang, int type, double M[3][3], double dM[3][3]) {
    double c = cos(ang), s = sin(ang);
    memset(M, 0, 9 * sizeof(double));
    memset(dM, 0, 9 * sizeof(double));
    switch (type) {
    case 0:
        M[0][0] = 1; M[1][1] = c; M[1][2] = s; M[2][1] = -s; M[2][2] = c;
        dM[1][1] = -s; dM[1][2] = c; dM[2][1] = -c; dM[2][2] = -s;
        break;
    case 1:
        M[0][0] = c; M[0][2] = -s; M[1][1] = 1; M[2][0] = -s; M[2][2] = c;
        dM[0][0] = -s; dM[0][2] = -c; dM[2][0] = -c; dM[2][2] = -s;
        break;
    case 2:
        M[0][0] = c; M[0][1] = s; M[1][0] = -s; M[1][1] = c; M[2][2] = 1;
        dM[0][0] = -s; dM[0][1] = c; dM[1][0] = -c; dM[1][1] = -s;
        break;
    default:
        M[0][0] = M[1][1] = M[2][2] = 1;
        break;
    }
}

/* Jacobian of oracle_util.h's rotate_U(xyz, U, ang, ow, type) in xyz, U and ang */
static void rotate_U_jac(const double xyz[3], const double U[3], double ang, double ow, int type,
                         double dX[3][3], double dU[3][3], double dA[3]) {
    double c = cos(ang), s = sin(ang);
    double x = xyz[0], y = xyz[1], z = xyz[2];
    double u = U[0], v = U[1], w = U[2];
    memset(dX, 0, 9 * sizeof(double));
    memset(dU, 0, 9 * sizeof(double));
    dA[0] = dA[1] = dA[2] = 0;
    switch (type) {
    case 0:
        dU[0][0] = 1;
        dX[1][1] = -ow * s; dX[1][2] = ow * c; dU[1][1] = c; dU[1][2] = s;
        dA[1] = ow * (-c * y - s * z) + (-s * v + c * w);
        dX[2][1] = -ow * c; dX[2][2] = -ow * s; dU[2][1] = -s + w; dU[2][2] = v;
        dA[2] = ow * (s * y - c * z) + (-c * v);
        break;
    case 1:
        dX[0][0] = -ow * s; dX[0][2] = -ow * c; dU[0][0] = c; dU[0][2] = -s;
        dA[0] = ow * (-c * x + s * z) + (-s * u - c * w);
        dU[1][1] = 1;
        dX[2][0] = ow * c; dX[2][2] = -ow * s; dU[2][0] = s; dU[2][2] = c;
        dA[2] = ow * (-s * x - c * z) + (c * u - s * w);
        break;
    case 2:
        dX[0][0] = -ow * s; dX[0][1] = ow * c; dU[0][0] = c; dU[0][1] = s;
        dA[0] = ow * (-c * x - s * y) + (-s * u + c * v);
        dX[1][1] = -ow * s; dU[1][0] = -s; dU[1][1] = c;
        dA[1] = ow * (s * s - c * c - c * y) + (-c * u - s * v);
        dU[2][2] = 1;
        break;
    default:
        dU[0][0] = dU[1][1] = dU[2][2] = 1;
        break;
    }
}

/* Reverse of one run_alm call at time t (positions, forces, projection) with the
 * objective term sum_w tq_w[w] * torque_w.  Recomputes the forward AP state from
 * U (mutating wt and ap exactly as the forward does), then accumulates
 * U_b += (dF/dU)^T (...) and, per AP a, ap_grad[4a + 0..3] += dJ/d(angle of its
 * turbine), dJ/d(pitch P[13]), dJ/d(chord P[4]), dJ/d(twist P[5]). */
void or_alm_adj(const or_grid *g, const double *U, double *wt, i64 wt_count, double *ap,
                i64 ap_count, i64 blade_per_wt, i64 ap_per_blade, const double *cd_tbl,
                const double *cl_tbl, const double *atk, i64 atk_count, double width, double t,
                const double *f_b, const double *tq_w, double *U_b, double *ap_grad) {
    if (ap_count <= 0) return;
    or_update_ap_position(wt, wt_count, ap, ap_count, blade_per_wt, ap_per_blade, t);
    or_calc_ap_force(g, U, wt, wt_count, ap, ap_count, blade_per_wt, ap_per_blade, cd_tbl, cl_tbl,
                     atk, atk_count);
    const i64 G = g->gc;
    const double *X = g->c[0], *Y = g->c[1], *Z = g->c[2];
    i64 per_wt = ap_per_blade * blade_per_wt;
    double *Fb = calloc((size_t)(3 * ap_count), sizeof(double));
    double *Xb = calloc((size_t)(3 * ap_count), sizeof(double));
    /* projection: f(c) = sum_a w(c, a) F_a */
    if (f_b) {
        double norm = 1 / ((width * sqrt(PI_)) * (width * sqrt(PI_)) * (width * sqrt(PI_)));
        INTERIOR(g, i, j, k) {
            i64 c = lin(g, i, j, k);
            const double *fb = f_b + 3 * c;
            if (fb[0] == 0 && fb[1] == 0 && fb[2] == 0) continue;
            for (i64 a = 0; a < ap_count; ++a) {
                const double *P = ap + OR_AP_STRIDE * a;
                double d0 = X[i] - P[0], d1 = Y[j] - P[1], d2 = Z[k] - P[2];
                double wgt = norm * exp(-(d0 * d0 + d1 * d1 + d2 * d2) / sq(width));
                Fb[3 * a + 0] += wgt * fb[0];
                Fb[3 * a + 1] += wgt * fb[1];
                Fb[3 * a + 2] += wgt * fb[2];
                double s = wgt * (fb[0] * P[8] + fb[1] * P[9] + fb[2] * P[10]) * 2 / sq(width);
                Xb[3 * a + 0] += s * d0;
                Xb[3 * a + 1] += s * d1;
                Xb[3 * a + 2] += s * d2;
            }
        }
    }
    for (i64 a = 0; a < ap_count; ++a) {
        double *P = ap + OR_AP_STRIDE * a;
        const double *T = wt + OR_WT_STRIDE * (a / per_wt);
        int type = (int)T[6];
        i64 i = floor_index(X, P[0], g->n[0]);
        i64 j = floor_index(Y, P[1], g->n[1]);
        i64 k = floor_index(Z, P[2], g->n[2]);
        double Pb[3] = {Xb[3 * a + 0], Xb[3 * a + 1], Xb[3 * a + 2]};
        double angb = 0;
        if (i >= G && i < g->n[0] - G && j >= G && j < g->n[1] - G && k >= G && k < g->n[2] - G) {
            /* forward recomputation */
            double hx = X[i + 1] - X[i], hy = Y[j + 1] - Y[j], hz = Z[k + 1] - Z[k];
            double fx = (P[0] - X[i]) / hx, fy = (P[1] - Y[j]) / hy, fz = (P[2] - Z[k]) / hz;
            double v[3][8], Uap[3];
            i64 cell[8];
            for (int o = 0; o < 8; ++o) cell[o] = lin(g, i + (o & 1), j + ((o >> 1) & 1), k + ((o >> 2) & 1));
            for (int m = 0; m < 3; ++m) {
                for (int o = 0; o < 8; ++o) v[m][o] = U[3 * cell[o] + m];
                double v01 = v[m][0] * (1 - fx) + v[m][1] * fx, v23 = v[m][2] * (1 - fx) + v[m][3] * fx;
                double v45 = v[m][4] * (1 - fx) + v[m][5] * fx, v67 = v[m][6] * (1 - fx) + v[m][7] * fx;
                double lo = v01 * (1 - fy) + v23 * fy, hi = v45 * (1 - fy) + v67 * fy;
                Uap[m] = lo * (1 - fz) + hi * fz;
            }
            double rel[3] = {P[0] - T[0], P[1] - T[1], P[2] - T[2]};
            double Ur[3];
            rotate_U(rel, Uap, Ur, T[7], T[8], type);
            double sth = sin(P[6]), cth = cos(P[6]);
            double ux = Ur[0];
            double ut = P[3] * T[13] + Ur[1] * sth - Ur[2] * cth;
            double V2 = ux * ux + ut * ut;
            double phi = atan(ux / ut);
            double aoa = phi * 180 / PI_ - P[5] - P[13];
            const double *CD = cd_tbl + atk_count * a, *CL = cl_tbl + atk_count * a;
            i64 s = floor_index(atk, aoa, atk_count);
            double cd, cl, scd = 0, scl = 0;
            if (s < 0) { cd = CD[0]; cl = CL[0]; }
            else if (s >= atk_count - 1) { cd = CD[atk_count - 1]; cl = CL[atk_count - 1]; }
            else {
                double h = atk[s + 1] - atk[s];
                double wv = (aoa - atk[s]) / h;
                cd = CD[s] * (1 - wv) + CD[s + 1] * wv;
                cl = CL[s] * (1 - wv) + CL[s + 1] * wv;
                scd = (CD[s + 1] - CD[s]) / h;
                scl = (CL[s + 1] - CL[s]) / h;
            }
            double K = 0.5 * V2 * P[4] * P[7];
            double Fd = cd * K, Fl = cl * K;
            double cphi = cos(phi), sphi = sin(phi);
            double Fx = Fl * cphi + Fd * sphi;
            double sg = sgn(T[13]);
            double Ft = (Fl * sphi - Fd * cphi) * sg;
            double M[3][3], dM[3][3];
            rot_mat(-T[7], type, M, dM);
            double loc[3] = {-Fx, Ft * sth, -Ft * cth};
            /* reverse */
            const double *Fbar = Fb + 3 * a;
            double locb[3];
            for (int r = 0; r < 3; ++r) {
                locb[r] = M[0][r] * Fbar[0] + M[1][r] * Fbar[1] + M[2][r] * Fbar[2];
                for (int q = 0; q < 3; ++q) angb -= Fbar[q] * dM[q][r] * loc[r];
            }
            double Fxb = -locb[0];
            double Ftb = locb[1] * sth - locb[2] * cth + tq_w[a / per_wt] * sgn(Ft) * P[3];
            double Flb = 0, Fdb = 0, phib = 0;
            Flb += Ftb * sg * sphi;
            Fdb -= Ftb * sg * cphi;
            phib += Ftb * sg * (Fl * cphi + Fd * sphi);
            Flb += Fxb * cphi;
            Fdb += Fxb * sphi;
            phib += Fxb * (-Fl * sphi + Fd * cphi);
            double cdb = Fdb * K, clb = Flb * K, Kb = Fdb * cd + Flb * cl;
            double V2b = Kb * 0.5 * P[4] * P[7];
            ap_grad[4 * a + 2] += Kb * 0.5 * V2 * P[7];
            double aoab = cdb * scd + clb * scl;
            phib += aoab * 180 / PI_;
            ap_grad[4 * a + 3] -= aoab;
            ap_grad[4 * a + 1] -= aoab;
            double uxb = 2 * ux * V2b + phib * ut / V2;
            double utb = 2 * ut * V2b - phib * ux / V2;
            double Urb[3] = {uxb, utb * sth, -utb * cth};
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
            for (int m = 0; m < 3; ++m) {
                const double *vv = v[m];
                double ub = Uapb[m];
                for (int o = 0; o < 8; ++o) {
                    double wx = (o & 1) ? fx : 1 - fx, wy = ((o >> 1) & 1) ? fy : 1 - fy, wz = ((o >> 2) & 1) ? fz : 1 - fz;
                    U_b[3 * cell[o] + m] += ub * wx * wy * wz;
                }
                fxb += ub * (((vv[1] - vv[0]) * (1 - fy) + (vv[3] - vv[2]) * fy) * (1 - fz) +
                             ((vv[5] - vv[4]) * (1 - fy) + (vv[7] - vv[6]) * fy) * fz);
                double lo_x0 = vv[0] * (1 - fx) + vv[1] * fx, lo_x1 = vv[2] * (1 - fx) + vv[3] * fx;
                double hi_x0 = vv[4] * (1 - fx) + vv[5] * fx, hi_x1 = vv[6] * (1 - fx) + vv[7] * fx;
                fyb += ub * ((lo_x1 - lo_x0) * (1 - fz) + (hi_x1 - hi_x0) * fz);
                fzb += ub * ((hi_x0 * (1 - fy) + hi_x1 * fy) - (lo_x0 * (1 - fy) + lo_x1 * fy));
            }
            Pb[0] += fxb / hx;
            Pb[1] += fyb / hy;
            Pb[2] += fzb / hz;
        }
        /* position: P = base + rotate(loc_p, -angle) */
        double lp[3] = {T[3], T[4] + P[3] * cos(P[6]), T[5] + P[3] * sin(P[6])};
        double M[3][3], dM[3][3];
        rot_mat(-T[7], type, M, dM);
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) angb -= Pb[r] * dM[r][q] * lp[q];
        ap_grad[4 * a + 0] += angb;
    }
    free(Fb);
    free(Xb);
}

/* Adjoint Poisson solve: b_b = A^-T p_b = V A^-1 (V^-1 p_b).  y is the solver's
 * vector (warm start in, solution out), tol_rel relative to |V^-1 p_b|.
 * work: 10 vectors of length L. */
void or_poisson_adj(const or_grid *g, const double *A, const double *p_b, double *y, double *b_b,
                    double *work, double tol_rel, i64 max_it, i64 *it, double *err) {
    const i64 L = g->n[0] * g->n[1] * g->n[2];
    double *rhs = work + 9 * (size_t)L;
    memset(rhs, 0, sizeof(double) * (size_t)L);
    memset(work, 0, sizeof(double) * 9 * (size_t)L);
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        rhs[c] = p_b[c] / (g->d[0][i] * g->d[1][j] * g->d[2][k]);
    }
    double nr = or_l2_norm(g, rhs) / sqrt((double)n_interior(g));
    memset(b_b, 0, sizeof(double) * (size_t)L);
    *it = 0;
    *err = 0;
    if (nr == 0) {
        memset(y, 0, sizeof(double) * (size_t)L);
        return;
    }
    or_run_bicgstab(g, A, y, rhs, work, 1, 1.2, 2, it, max_it, err, tol_rel * nr);
    INTERIOR(g, i, j, k) {
        i64 c = lin(g, i, j, k);
        b_b[c] = y[c] * (g->d[0][i] * g->d[1][j] * g->d[2][k]);
    }
}

/* ========================================================================== */
/* Whole run: from an initialized state, `steps` forward steps (all states kept),
 * J = sum_n sum_w tq_w[w] torque_w(n) over steps 1..steps, then the reverse sweep.
 * ap_grad [ap_count][4] receives dJ/d(turbine angle via AP), d pitch, d chord,
 * d twist.  state_b (optional, 10*L) receives the bars of the initial state taken
 * as independent inputs: U (3L, with ghosts), JU (3L), nut (L), f (3L).  On
 * return s holds the final forward state except its AP records. */
void or_adjoint_run(const or_grid *g, or_state *s, i64 steps, const double *tq_w, double adj_tol,
                    double *J, double *ap_grad, double *state_b) {
    const i64 L = g->n[0] * g->n[1] * g->n[2];
    const size_t L3 = 3 * (size_t)L;
    double *tU = malloc(sizeof(double) * L3 * (size_t)(steps + 1));
    double *tJU = malloc(sizeof(double) * L3 * (size_t)(steps + 1));
    double *tN = malloc(sizeof(double) * (size_t)L * (size_t)(steps + 1));
    memcpy(tU, s->U, sizeof(double) * L3);
    memcpy(tJU, s->JU, sizeof(double) * L3);
    memcpy(tN, s->nut, sizeof(double) * (size_t)L);
    *J = 0;
    for (i64 n = 1; n <= steps; ++n) {
        or_state_step(g, s);
        for (i64 w = 0; w < s->wt_count; ++w) *J += tq_w[w] * s->wt[OR_WT_STRIDE * w + 14];
        memcpy(tU + L3 * n, s->U, sizeof(double) * L3);
        memcpy(tJU + L3 * n, s->JU, sizeof(double) * L3);
        memcpy(tN + (size_t)L * n, s->nut, sizeof(double) * (size_t)L);
    }
    memset(ap_grad, 0, sizeof(double) * 4 * (size_t)s->ap_count);
    double *Ub = calloc(L3, sizeof(double)), *Ubn = calloc(L3, sizeof(double));
    double *JUb = calloc(L3, sizeof(double)), *JUbn = calloc(L3, sizeof(double));
    double *fb = calloc(L3, sizeof(double)), *fbn = calloc(L3, sizeof(double));
    double *nb = calloc((size_t)L, sizeof(double)), *nbn = calloc((size_t)L, sizeof(double));
    double *pb = calloc((size_t)L, sizeof(double)), *bb = calloc((size_t)L, sizeof(double));
    double *y = calloc((size_t)L, sizeof(double)), *work = calloc(10 * (size_t)L, sizeof(double));
    for (i64 n = steps; n >= 1; --n) {
        const double *Un = tU + L3 * n, *Um = tU + L3 * (n - 1);
        const double *JUm = tJU + L3 * (n - 1), *Nm = tN + (size_t)L * (n - 1);
        memset(Ubn, 0, sizeof(double) * L3);
        memset(JUbn, 0, sizeof(double) * L3);
        memset(fbn, 0, sizeof(double) * L3);
        memset(nbn, 0, sizeof(double) * (size_t)L);
        or_alm_adj(g, Un, s->wt, s->wt_count, s->ap, s->ap_count, s->blade_per_wt, s->ap_per_blade,
                   s->cd_tbl, s->cl_tbl, s->atk, s->atk_count, s->width, s->dt * n, fb, tq_w, Ub, ap_grad);
        or_nut_adj(g, Un, nb, Ub, s->Cs);
        or_apply_JUbc_adj(g, JUb);
        or_apply_Ubc_adj(g, Ub, Um, Um, Ubn, s->dt, s->retracted_floor, s->floor_start_x);
        or_set_solid_adj(g, Ub, s->solid);
        memset(pb, 0, sizeof(double) * (size_t)L);
        or_project_adj(g, Ub, JUb, pb, s->dt);
        or_apply_pbc_adj(g, pb);
        i64 it;
        double err;
        or_poisson_adj(g, s->A, pb, y, bb, work, adj_tol, 20000, &it, &err);
        or_poisson_rhs_adj(g, bb, JUb, s->dt, s->max_diag);
        or_apply_JUbc_adj(g, JUb);
        or_interpolate_JU_adj(g, JUb, Ub);
        for (i64 i = 0; i < g->n[0]; ++i)
            for (i64 j = 0; j < g->n[1]; ++j)
                for (i64 k = 0; k < g->n[2]; ++k) {
                    if (is_interior(g, i, j, k)) continue;
                    i64 c = lin(g, i, j, k);
                    for (int m = 0; m < 3; ++m) {
                        Ubn[3 * c + m] += Ub[3 * c + m];
                        Ub[3 * c + m] = 0;
                    }
                }
        or_predictor_adj(g, Ub, Um, JUm, Nm, s->Re, s->dt, Ubn, JUbn, nbn, fbn);
        double *t;
        t = Ub; Ub = Ubn; Ubn = t;
        t = JUb; JUb = JUbn; JUbn = t;
        t = fb; fb = fbn; fbn = t;
        t = nb; nb = nbn; nbn = t;
    }
    if (state_b) {
        memcpy(state_b, Ub, sizeof(double) * L3);
        memcpy(state_b + L3, JUb, sizeof(double) * L3);
        memcpy(state_b + 2 * L3, nb, sizeof(double) * (size_t)L);
        memcpy(state_b + 2 * L3 + L, fb, sizeof(double) * L3);
    }
    /* f_0 from the initialisation pass at t = 0 (its torque is not in J) */
    double *zero_w = calloc((size_t)(s->wt_count > 0 ? s->wt_count : 1), sizeof(double));
    or_alm_adj(g, tU, s->wt, s->wt_count, s->ap, s->ap_count, s->blade_per_wt, s->ap_per_blade,
               s->cd_tbl, s->cl_tbl, s->atk, s->atk_count, s->width, 0.0, fb, zero_w, Ub, ap_grad);
    free(zero_w);
    free(Ub); free(Ubn); free(JUb); free(JUbn); free(fb); free(fbn); free(nb); free(nbn);
    free(pb); free(bb); free(y); free(work);
    free(tU); free(tJU); free(tN);
}
