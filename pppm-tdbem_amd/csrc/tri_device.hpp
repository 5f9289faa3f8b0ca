// Per-triangle device functions shared by the mesh-update kernels and the FP64 quadrature:
// triangle geometry and binning (update_triangles_kernel, particle_grid.cu:40-56;
// getGridCoord / getGridBaseCoord, particle_grid.h:169-176), the grid-cache recompute test
// (GridCache::need_recompute, pppm_cache.h:64-69) and the trilinear far-field weights
// (update_interpolation_weight_kernel, pppm_cache.cu:139-181).
#pragma once
#include "internal.hpp"

namespace pppm_amd
{

constexpr double TRI1_X = 0.3333333333333330;  // gauss/triangle.h:6-14
constexpr double TRI1_W = 1.0;

struct TriD
{
    d3 v[3];
    d3 n;
    double jac;
};

__device__ inline f3 ld3(const float *v, int i) { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }

__device__ inline TriD tri_d(const float *verts, int i0, int i1, int i2)
{
    TriD t;
    t.v[0] = D(ld3(verts, i0));
    t.v[1] = D(ld3(verts, i1));
    t.v[2] = D(ld3(verts, i2));
    d3 cr = crossd(t.v[1] - t.v[0], t.v[2] - t.v[0]);
    t.jac = sqrt(dotd(cr, cr));
    t.n = {cr.x / t.jac, cr.y / t.jac, cr.z / t.jac};
    return t;
}
__device__ inline d3 l2g(double x1, double x2, const d3 *v) { return (1 - x1 - x2) * v[0] + x1 * v[1] + x2 * v[2]; }
__device__ inline d3 l2g2(double x1, double x2, const d3 *v) { return (1 - x1) * v[0] + (x1 - x2) * v[1] + x2 * v[2]; }

__device__ inline d3 cell_center_d(int x, int y, int z, const WeightArgs &a)
{
    return {((double)x + 0.5) * a.h + a.minx, ((double)y + 0.5) * a.h + a.miny, ((double)z + 0.5) * a.h + a.minz};
}
__device__ inline f3 cell_center_f(int x, int y, int z, const WeightArgs &a)
{
    return f3{((float)x + 0.5f) * a.h, ((float)y + 0.5f) * a.h, ((float)z + 0.5f) * a.h} + f3{a.minx, a.miny, a.minz};
}

// geometry and cells of triangle i; returns error bits (the cells fall back to (2, 2, 2))
__device__ inline int tri_geometry(const float *__restrict__ verts, const int *__restrict__ tris, int i,
                                   const BuildArgs &a, TriGeom &g)
{
    const int i0 = tris[3 * i], i1 = tris[3 * i + 1], i2 = tris[3 * i + 2];
    int3_ c = {2, 2, 2}, b = {2, 2, 2};
    int e = 0;
    if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= a.nv || i1 >= a.nv || i2 >= a.nv)
    {
        e = ERR_BAD_INDEX;
        g.center = g.normal = {0, 0, 0};
        g.area = 0;
    }
    else
    {
        const f3 v0 = ld3(verts, i0), v1 = ld3(verts, i1), v2 = ld3(verts, i2);
        g.center = (v0 + v1 + v2) / 3.0f;
        const f3 cr = crossf(v1 - v0, v2 - v0);
        const float inv = 1.0f / sqrtf(dotf(cr, cr));  // normalize() = v * rsqrt(dot(v, v)) (helper_math.h:1459)
        g.normal = cr * inv;
        g.area = lengthf(cr) / 2.0f;
        const f3 mn = {a.minx, a.miny, a.minz};
        const f3 q = (g.center - mn) / a.h;
        c = {(int)q.x, (int)q.y, (int)q.z};
        const f3 cc = f3{((float)c.x + 0.5f) * a.h, ((float)c.y + 0.5f) * a.h, ((float)c.z + 0.5f) * a.h} + mn;
        const f3 d = g.center - cc;
        b = {c.x - (d.x < 0), c.y - (d.y < 0), c.z - (d.z < 0)};
        const int N = a.N;
        if (!(q.x >= 0 && q.y >= 0 && q.z >= 0) || c.x < 2 || c.y < 2 || c.z < 2 || c.x > N - 3 || c.y > N - 3 ||
            c.z > N - 3)
        {
            e = ERR_OUT_OF_GRID;
            c = {2, 2, 2};
            b = {2, 2, 2};
        }
    }
    g.cell3 = c;
    g.base3 = b;
    g.cell = lin3(c.x, c.y, c.z, a.N);
    g.base = lin3(b.x, b.y, b.z, a.N);
    return e;
}

// GridCache::need_recompute; meta = centre(3) normal(3) area of the last computation
__device__ inline bool grid_needs_recompute(const TriGeom &g, const float *__restrict__ m, float h)
{
    const f3 oc = {m[0], m[1], m[2]}, on = {m[3], m[4], m[5]};
    const float oa = m[6];
    return (double)lengthf(oc - g.center) > RECOMPUTE_THRESHOLD * (double)h ||
           (double)lengthf(on - g.normal) > RECOMPUTE_THRESHOLD ||
           (double)fabsf(oa - g.area) > RECOMPUTE_THRESHOLD * (double)g.area;
}

// the 8 far-field interpolation weights of triangle i about its base cell (order 000..111),
// scaled by the 1-point quadrature weight of the triangle
__device__ inline void interp_weights(const float *__restrict__ verts, const int *__restrict__ tris, int i,
                                      int3_ b, const WeightArgs &a, float *__restrict__ out8)
{
    const TriD T = tri_d(verts, tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]);
    const d3 p = l2g(TRI1_X, TRI1_X, T.v);
    const d3 bc = cell_center_d(b.x, b.y, b.z, a);
    const double h = a.h;
    const double dx = (p.x - bc.x) / h, dy = (p.y - bc.y) / h, dz = (p.z - bc.z) / h;
    const double w[8] = {(1 - dx) * (1 - dy) * (1 - dz), (1 - dx) * (1 - dy) * dz, (1 - dx) * dy * (1 - dz),
                         (1 - dx) * dy * dz,             dx * (1 - dy) * (1 - dz), dx * (1 - dy) * dz,
                         dx * dy * (1 - dz),             dx * dy * dz};
    const double sc = 0.5 * TRI1_W * T.jac;
    for (int q = 0; q < 8; q++)
        out8[q] = (float)(sc * w[q]);
}

}  // namespace pppm_amd
