// solid.hip -- the solid fraction of a turbulence-generating grid
// (set_turbulence_generating_grid, src/solver.h:52-159), computed once at start-up.
//
// The grid is two layers of bars across the inflow: "vertical" bars of width `bar`
// at y = py + n * mesh over x in [px - thick, px], and "horizontal" bars at
// z = pz + n * mesh over x in [px - 2 thick, px - thick].  A cell's solid fraction
// is the volume of its box [c - d/2, c + d/2] (per direction) covered by the nearest
// bar of each layer, over the box volume.  set_solid_U (U *= 1 - solid) then acts
// on it every step (k_solid / km_solid).
//
// FP64 in the reference's operation order, compiled with -ffp-contract=off: the
// field matches the reference bit for bit.  One pass over the interior, 8 B written
// per cell.
#include "ctx.h"

namespace {

// length of [h1, t1] intersected with [h2, t2] (src/util.h:84-90)
__device__ __forceinline__ double overlap(double h1, double t1, double h2, double t2) {
    const double l1 = t1 - h1, l2 = t2 - h2, x1 = t1 - h2, x2 = t2 - h1;
    return fmax(0., fmin(l1, fmin(l2, fmin(x1, x2))));
}

// nearest bar centre: the reference rounds to an integer bar index first
__device__ __forceinline__ double nearest_bar(double c, double centre, double mesh) {
    const long long n = (long long)round((c - centre) / mesh);
    return (double)n * mesh + centre;
}

__global__ __launch_bounds__(256) void k_tgg_solid(Dims d, const double* __restrict__ x,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   const double* __restrict__ dx, const double* __restrict__ dy,
                                                   const double* __restrict__ dz, double thick, double mesh,
                                                   double bar, double px, double py, double pz,
                                                   double* __restrict__ solid) {
    const long long n = (long long)d.nx * d.ny * d.nz;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int k = d.gc + (int)(e % d.nz);
        const int j = d.gc + (int)((e / d.nz) % d.ny);
        const int i = d.gc + (int)(e / ((long long)d.nz * d.ny));
        const double xc = x[i], yc = y[j], zc = z[k];
        const double dxc = dx[i], dyc = dy[j], dzc = dz[k];
        const double xv = overlap(xc - 0.5 * dxc, xc + 0.5 * dxc, px - thick, px);
        const double by = nearest_bar(yc, py, mesh);
        const double yi = overlap(yc - 0.5 * dyc, yc + 0.5 * dyc, by - 0.5 * bar, by + 0.5 * bar);
        const double xh = overlap(xc - 0.5 * dxc, xc + 0.5 * dxc, px - 2 * thick, px - thick);
        const double bz = nearest_bar(zc, pz, mesh);
        const double zi = overlap(zc - 0.5 * dzc, zc + 0.5 * dzc, bz - 0.5 * bar, bz + 0.5 * bar);
        const double occupied = xv * yi * dzc + xh * zi * dyc;
        solid[d.at(i, j, k)] = occupied / (dxc * dyc * dzc);
    }
}

}  // namespace

int dalm_set_turbulence_grid(dalm_ctx* ctx, double thick, double mesh, double bar, const double placement[3]) {
    if (!ctx) return DALM_BAD_ARG;
    if (!placement || !(mesh > 0) || !(thick >= 0) || !(bar >= 0))
        return dalm_fail_msg(ctx, DALM_BAD_ARG, "dalm_set_turbulence_grid: need mesh > 0, thick >= 0, bar >= 0");
    long long blocks = (ctx->N + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_tgg_solid, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, ctx->d, ctx->x, ctx->y,
                       ctx->z, ctx->dx, ctx->dy, ctx->dz, thick, mesh, bar, placement[0], placement[1],
                       placement[2], ctx->solid);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->has_solid = true;  // set_solid_U now runs every step
    return DALM_OK;
}
