/*
 * oracle_util.h -- small helpers shared by the oracle's forward restatement
 * (oracle.c) and its discrete adjoint (adjoint.c).  TEST INFRASTRUCTURE ONLY.
 */
#ifndef DALM_ORACLE_UTIL_H
#define DALM_ORACLE_UTIL_H

#include <math.h>

#include "oracle.h"

#define PI_ M_PI

/* linear index, z fastest (src/util.h:10-13) */
static inline i64 lin(const or_grid *g, i64 i, i64 j, i64 k) {
    return (i * g->n[1] + j) * g->n[2] + k;
}
/* integer sign, used as a double multiplier (src/util.h:70-74) */
static inline double sgn(double v) { return (double)((v > 0.0) - (v < 0.0)); }
static inline double sq(double v) { return v * v; }

#define INTERIOR(g, i, j, k)                               \
    for (i64 i = (g)->gc; i < (g)->n[0] - (g)->gc; ++i)    \
        for (i64 j = (g)->gc; j < (g)->n[1] - (g)->gc; ++j) \
            for (i64 k = (g)->gc; k < (g)->n[2] - (g)->gc; ++k)

/* The OpenMP build of the oracle (_build/liboracle_omp.so), timed as the CPU
 * baseline on every host core, shares the outer i loop of the cell loops among
 * threads; the serial build, the checker, ignores these pragmas and keeps the
 * reference's order.  Maxima use OpenMP's reduction clause on the variable `mx`
 * (exact in any order); sums use fixed-order per-thread partials (oracle.c,
 * interior_dot), so the OpenMP build sums in another order than the reference but
 * always in the same one for a given thread count. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#define OMP_FOR_MAX _Pragma("omp parallel for schedule(static) reduction(max : mx)")
#else
#define OMP_FOR
#define OMP_FOR_MAX
#endif
#define PAR_INTERIOR(g, i, j, k) OMP_FOR INTERIOR(g, i, j, k)
#define PAR_INTERIOR_MAX(g, i, j, k) OMP_FOR_MAX INTERIOR(g, i, j, k)

static inline i64 n_interior(const or_grid *g) {
    return (g->n[0] - 2 * g->gc) * (g->n[1] - 2 * g->gc) * (g->n[2] - 2 * g->gc);
}

/* Lagrange quadratic through (x0,f0),(x1,f1),(x2,f2) evaluated at x3: src/util.h:190-201 */
static inline double quad3(double x0, double x1, double x2, double f0, double f1, double f2,
                           double x3) {
    double l0 = ((x3 - x1) * (x3 - x2)) / ((x0 - x1) * (x0 - x2));
    double l1 = ((x3 - x0) * (x3 - x2)) / ((x1 - x0) * (x1 - x2));
    double l2 = ((x3 - x0) * (x3 - x1)) / ((x2 - x0) * (x2 - x1));
    return f0 * l0 + f1 * l1 + f2 * l2;
}

/* binary search for the last index with a[idx] <= v: src/util.h:108-135 */
static inline i64 floor_index(const double *a, double v, i64 n) {
    if (v < a[0]) return -1;
    if (v >= a[n - 1]) return n - 1;
    i64 lo = 0, hi = n - 1, hit = -1;
    while (lo <= hi) {
        i64 mid = lo + (hi - lo) / 2;
        if (a[mid] <= v) { hit = mid; lo = mid + 1; }
        else hi = mid - 1;
    }
    return hit;
}

/* rotation by `ang` about one axis: src/euler_angle.h:34-62 */
static inline void rotate(const double in[3], double out[3], double ang, int type) {
    double c = cos(ang), s = sin(ang);
    double x = in[0], y = in[1], z = in[2];
    switch (type) {
    case 0: out[0] = x; out[1] = c * y + s * z; out[2] = -s * y + c * z; break;
    case 1: out[0] = c * x - s * z; out[1] = y; out[2] = -s * x + c * z; break;
    case 2: out[0] = c * x + s * y; out[1] = -s * x + c * y; out[2] = z; break;
    default: out[0] = x; out[1] = y; out[2] = z; break;
    }
}

/* velocity seen in the rotor frame, with the reference's algebra kept as is
 * (src/euler_angle.h:64-127, including the Roll z term v*w, the Yaw y term
 * -c*s - s*y and the Pitch z sign) */
static inline void rotate_U(const double xyz[3], const double U[3], double out[3], double ang,
                     double ow, int type) {
    double c = cos(ang), s = sin(ang);
    double x = xyz[0], y = xyz[1], z = xyz[2];
    double u = U[0], v = U[1], w = U[2];
    switch (type) {
    case 0:
        out[0] = 0. + u;
        out[1] = ow * (-s * y + c * z) + (c * v + s * w);
        out[2] = ow * (-c * y - s * z) + (-s * v + v * w);
        break;
    case 1:
        out[0] = ow * (-s * x - c * z) + (c * u - s * w);
        out[1] = 0. + v;
        out[2] = ow * (c * x - s * z) + (s * u + c * w);
        break;
    case 2:
        out[0] = ow * (-s * x + c * y) + (c * u + s * v);
        out[1] = ow * (-c * s - s * y) + (-s * u + c * v);
        out[2] = 0. + w;
        break;
    default: out[0] = u; out[1] = v; out[2] = w; break;
    }
}

#endif
