// Cell classification and fresh cells for moving meshes (gfx950): integer/label work over
// the grid, bandwidth- and latency-bound, no MFMA.
//
// Reference behaviour: fill_in_nearest_kernel, cell_classfication_kernel,
// post_cell_classification_kernel, apply_cell_type_kernel, fill_cell_data
// (src/algorithm/ghost_cell_util.cu:7-230); get_nearest_triangle_point_coeff
// (src/core/helper_math.h:1603-1800); get_fresh_cell_list, solve_fresh_history
// (src/algorithm/ghost_cell.cu:51-196); set_solid_cell_zero_kernel (ghost_cell.cu:547-555).
//
// Labels are reproduced exactly: the nearest-triangle search keeps the reference's window
// order and strict comparison over cell lists sorted by triangle id, the vertex normals sum
// their faces in ascending order, the column fill is a wave-per-column ballot scan with the
// same semantics as the serial z loop (including its relabelling of the leading unknown run),
// and ghost indices come from a prefix scan of the pre-fill ghost flags in linear order.
#include "internal.hpp"

namespace pppm_amd
{

namespace
{
constexpr int B = 256;
constexpr int WPB = B / WAVE;

__device__ inline f3 ld3(const float *p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(float *p, size_t i, f3 v)
{
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

__device__ inline f3 cell_center(int x, int y, int z, const CellArgs &a)
{
    return f3{((float)x + 0.5f) * a.h, ((float)y + 0.5f) * a.h, ((float)z + 0.5f) * a.h} + f3{a.minx, a.miny, a.minz};
}

// ---------------------------------------------------------------- closest point on a triangle
// Eberly, "Distance from a Point to a Triangle" (Geometric Tools, 2020): minimise
// |v0 + s e0 + t e1 - p|^2 over the triangle along the zero set of the gradient's
// directional derivative; the float operations are those of the reference's routine.
struct P2
{
    float s, t;
};
__device__ inline P2 min_on_edge02(float a11, float b1)
{
    const int neg = b1 < 0, past = a11 + b1 <= 0;
    return {0.0f, (float)neg * ((float)past - (float)(1 - past) * (b1 / a11))};
}
__device__ inline P2 min_on_edge12(float a01, float a11, float b1, float f10, float f01)
{
    const float g0 = a01 + b1 - f10, g1 = a11 + b1 - f01;
    const int neg = g0 < 0, past = g1 <= 0;
    const float t = (float)neg * ((float)past + (float)(1 - past) * g0 / (g0 - g1));
    return {1 - t, t};
}
__device__ inline P2 min_inside(P2 q0, float g0, P2 q1, float g1)
{
    const float z = g0 / (g0 - g1), w = 1 - z;
    return {w * q0.s + z * q1.s, w * q0.t + z * q1.t};
}
__device__ f3 closest_coeff(f3 p, f3 v0, f3 v1, f3 v2)
{
    const f3 d = p - v0, e0 = v1 - v0, e1 = v2 - v0;
    const float a00 = dotf(e0, e0), a01 = dotf(e0, e1), a11 = dotf(e1, e1);
    const float b0 = -dotf(d, e0), b1 = -dotf(d, e1);
    const float f00 = b0, f10 = b0 + a00, f01 = b0 + a01;
    P2 q0{0, 0}, q1{0, 0}, r{0, 0};
    float g0, g1;
    if (f00 >= 0)
    {
        if (f01 >= 0)
            r = min_on_edge02(a11, b1);
        else
        {
            q0 = {0, f00 / (f00 - f01)};
            q1.s = f01 / (f01 - f10);
            q1.t = 1 - q1.s;
            const float dt1 = q1.t - q0.t;
            g0 = dt1 * (a11 * q0.t + b1);
            if (g0 >= 0)
                r = min_on_edge02(a11, b1);
            else
            {
                g1 = dt1 * (a01 * q1.s + a11 * q1.t + b1);
                r = g1 <= 0 ? min_on_edge12(a01, a11, b1, f10, f01) : min_inside(q0, g0, q1, g1);
            }
        }
    }
    else if (f01 <= 0)
    {
        if (f10 <= 0)
            r = min_on_edge12(a01, a11, b1, f10, f01);
        else
        {
            q0 = {f00 / (f00 - f10), 0};
            q1.s = f01 / (f01 - f10);
            q1.t = 1 - q1.s;
            g0 = q1.t * (a01 * q0.s + b1);
            if (g0 >= 0)
                r = q0;
            else
            {
                g1 = q1.t * (a01 * q1.s + a11 * q1.t + b1);
                r = g1 <= 0 ? min_on_edge12(a01, a11, b1, f10, f01) : min_inside(q0, g0, q1, g1);
            }
        }
    }
    else if (f10 <= 0)
    {
        q0 = {0, f00 / (f00 - f01)};
        q1.s = f01 / (f01 - f10);
        q1.t = 1 - q1.s;
        const float dt1 = q1.t - q0.t;
        g0 = dt1 * (a11 * q0.t + b1);
        if (g0 >= 0)
            r = min_on_edge02(a11, b1);
        else
        {
            g1 = dt1 * (a01 * q1.s + a11 * q1.t + b1);
            r = g1 <= 0 ? min_on_edge12(a01, a11, b1, f10, f01) : min_inside(q0, g0, q1, g1);
        }
    }
    else
    {
        q0 = {f00 / (f00 - f10), 0};
        q1 = {0, f00 / (f00 - f01)};
        g0 = q1.t * (a01 * q0.s + b1);
        if (g0 >= 0)
            r = q0;
        else
        {
            g1 = q1.t * (a11 * q1.t + b1);
            r = g1 <= 0 ? min_on_edge02(a11, b1) : min_inside(q0, g0, q1, g1);
        }
    }
    return {1 - r.s - r.t, r.s, r.t};
}

// ---------------------------------------------------------------- vertex -> faces, normals
__global__ __launch_bounds__(B) void k_vf_count(const int *__restrict__ tris, int nt, int *cnt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt)
        return;
    for (int k = 0; k < 3; k++)
        atomicAdd(&cnt[tris[3 * i + k]], 1);
}
__global__ __launch_bounds__(B) void k_vf_fill(const int *__restrict__ tris, int nt, const int *__restrict__ start,
                                               int *fill, int *items)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt)
        return;
    for (int k = 0; k < 3; k++)
    {
        const int v = tris[3 * i + k];
        items[start[v] + atomicAdd(&fill[v], 1)] = i;
    }
}
// sorts each vertex's face list (ascending face id) and sums the angle-weighted normal in that
// order (ghost_cell_util.cu:56-80)
__global__ __launch_bounds__(B) void k_vertex_normals(const float *__restrict__ verts, const int *__restrict__ tris,
                                                      const TriGeom *__restrict__ geom, int nv,
                                                      const int *__restrict__ start, int *items, float *vnormal)
{
    const int v = blockIdx.x * B + threadIdx.x;
    if (v >= nv)
        return;
    const int s = start[v], e = start[v + 1];
    for (int i = s + 1; i < e; i++)
    {
        const int key = items[i];
        int j = i - 1;
        while (j >= s && items[j] > key)
        {
            items[j + 1] = items[j];
            j--;
        }
        items[j + 1] = key;
    }
    const f3 pv = ld3(verts, v);
    f3 n = {0, 0, 0};
    for (int i = s; i < e; i++)
    {
        const int f = items[i];
        int o[2], m = 0;
        for (int k = 0; k < 3; k++)
        {
            const int w = tris[3 * f + k];
            if (w != v && m < 2)
                o[m++] = w;
        }
        if (m < 2)
            continue;  // degenerate face that repeats this vertex
        const f3 e1 = ld3(verts, o[0]) - pv, e2 = ld3(verts, o[1]) - pv;
        const float ang = (float)acos((double)(dotf(e1, e2) / (lengthf(e1) * lengthf(e2))));
        n = n + geom[f].normal * ang;
    }
    st3(vnormal, v, n * (1.0f / sqrtf(dotf(n, n))));
}

// ---------------------------------------------------------------- nearest triangle, side
// fill_in_nearest_kernel: thread per cell, 7^3 window in (dx, dy, dz) order, strict "<".
__global__ __launch_bounds__(B) void k_cell_nearest(CellArgs a, const float *__restrict__ verts,
                                                    const int *__restrict__ tris, const int *__restrict__ cell_start,
                                                    const int *__restrict__ cell_items,
                                                    const float *__restrict__ vnormal, int *__restrict__ ctype,
                                                    int *__restrict__ cnearest, float *__restrict__ cpoint,
                                                    float *__restrict__ creflect, float *__restrict__ cdist)
{
    const int N = a.N;
    const long long c = (long long)blockIdx.x * B + threadIdx.x;
    if (c >= (long long)N * N * N)
        return;
    const int x = (int)(c / ((long long)N * N)), y = (int)(c / N % N), z = (int)(c % N);
    const f3 ctr = cell_center(x, y, z, a);
    float best = 3.402823466e+38F;
    int bi = -1;
    f3 bp = {0, 0, 0}, bc = {0, 0, 0};
    for (int dx = -3; dx <= 3; dx++)
    {
        const int cx = x + dx;
        if (cx < 0 || cx >= N)
            continue;
        for (int dy = -3; dy <= 3; dy++)
        {
            const int cy = y + dy;
            if (cy < 0 || cy >= N)
                continue;
            for (int dz = -3; dz <= 3; dz++)
            {
                const int cz = z + dz;
                if (cz < 0 || cz >= N)
                    continue;
                const int cl = lin3(cx, cy, cz, N);
                const int q1 = cell_start[cl + 1];
                for (int q = cell_start[cl]; q < q1; q++)
                {
                    const int f = cell_items[q];
                    const f3 v0 = ld3(verts, tris[3 * f]), v1 = ld3(verts, tris[3 * f + 1]),
                             v2 = ld3(verts, tris[3 * f + 2]);
                    const f3 w = closest_coeff(ctr, v0, v1, v2);
                    const f3 np = w.x * v0 + w.y * v1 + w.z * v2;
                    const float len = lengthf(ctr - np);
                    if (len < best)
                    {
                        best = len;
                        bi = f;
                        bp = np;
                        bc = w;
                    }
                }
            }
        }
    }
    cdist[c] = best;
    cnearest[c] = bi;
    if (bi < 0)
    {
        ctype[c] = CELL_UNKNOWN;
        st3(cpoint, c, f3{0, 0, 0});
        st3(creflect, c, f3{0, 0, 0});
        return;
    }
    st3(cpoint, c, bp);
    st3(creflect, c, 2.0f * bp - ctr);
    const f3 nrm = bc.x * ld3(vnormal, tris[3 * bi]) + bc.y * ld3(vnormal, tris[3 * bi + 1]) +
                   bc.z * ld3(vnormal, tris[3 * bi + 2]);
    const float direction = dotf(nrm, ctr - bp);
    const f3 dv = bp - ctr;
    const float hh = a.h / 2;
    const bool shell = a.thin_shell && fabsf(dv.x) < hh && fabsf(dv.y) < hh && fabsf(dv.z) < hh;
    ctype[c] = (direction <= 0 || shell) ? CELL_SOLID : CELL_AIR;
}

// cell_classfication_kernel: SOLID with an AIR 6-neighbour -> GHOST; flag for the index scan
__global__ __launch_bounds__(B) void k_ghost_mark(int N, const int *__restrict__ ctype, int *__restrict__ typ2,
                                                  int *__restrict__ flag)
{
    const long long c = (long long)blockIdx.x * B + threadIdx.x;
    if (c >= (long long)N * N * N)
        return;
    const int x = (int)(c / ((long long)N * N)), y = (int)(c / N % N), z = (int)(c % N);
    int t = ctype[c];
    if (t == CELL_SOLID)
    {
        const bool air = (x + 1 < N && ctype[lin3(x + 1, y, z, N)] == CELL_AIR) ||
                         (x > 0 && ctype[lin3(x - 1, y, z, N)] == CELL_AIR) ||
                         (y + 1 < N && ctype[lin3(x, y + 1, z, N)] == CELL_AIR) ||
                         (y > 0 && ctype[lin3(x, y - 1, z, N)] == CELL_AIR) ||
                         (z + 1 < N && ctype[c + 1] == CELL_AIR) || (z > 0 && ctype[c - 1] == CELL_AIR);
        if (air)
            t = CELL_GHOST;
    }
    typ2[c] = t;
    flag[c] = t == CELL_GHOST;
}

// post_cell_classification_kernel as a wave-per-column ballot scan. An unknown cell takes the
// value handed on by the last known cell before it in z: the first known cell of the column
// hands on its own type, later ones hand on SOLID (for SOLID or GHOST) or AIR. Unknown cells
// before the first known one take that cell's type; a column with no known cell is AIR.
__global__ __launch_bounds__(B) void k_column_fill(int N, int *__restrict__ typ)
{
    const int col = blockIdx.x * WPB + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (col >= N * N)
        return;
    int *c = typ + (size_t)col * N;
    int first = -1, fraw = CELL_AIR;
    for (int z0 = 0; z0 < N && first < 0; z0 += WAVE)
    {
        const int z = z0 + lane;
        const int v = z < N ? c[z] : CELL_UNKNOWN;
        const unsigned long long m = __ballot(v != CELL_UNKNOWN);
        if (m)
        {
            const int l = __ffsll((long long)m) - 1;
            first = z0 + l;
            fraw = __shfl(v, l, WAVE);
        }
    }
    bool have = false;
    int carry = CELL_AIR;
    for (int z0 = 0; z0 < N; z0 += WAVE)
    {
        const int z = z0 + lane;
        const int v = z < N ? c[z] : CELL_UNKNOWN;
        const bool known = v != CELL_UNKNOWN;
        const unsigned long long m = __ballot(known);
        const int hand = z == first ? v : ((v == CELL_SOLID || v == CELL_GHOST) ? CELL_SOLID : CELL_AIR);
        const unsigned long long before = m & ((1ull << lane) - 1ull);
        const int src = before ? 63 - __clzll((long long)before) : 0;
        const int from = __shfl(hand, src, WAVE);
        if (z < N && !known)
            c[z] = before ? from : (have ? carry : fraw);
        if (m)
        {
            carry = __shfl(hand, 63 - __clzll((long long)m), WAVE);
            have = true;
        }
    }
}

// apply_cell_type_kernel: final types, ghost index = inclusive scan of the pre-fill flags - 1
__global__ __launch_bounds__(B) void k_cell_apply(long long NN, const int *__restrict__ typ2,
                                                  const int *__restrict__ flag, const int *__restrict__ scan,
                                                  int *__restrict__ ctype, int *__restrict__ cghost)
{
    const long long c = (long long)blockIdx.x * B + threadIdx.x;
    if (c >= NN)
        return;
    const int t = typ2[c];
    ctype[c] = t;
    cghost[c] = t == CELL_GHOST ? scan[c] + flag[c] - 1 : -1;
}

// ---------------------------------------------------------------- fresh cells
__global__ __launch_bounds__(B) void k_fresh_flag(long long NN, int nt, const int *__restrict__ told,
                                                  const int *__restrict__ tnew, const int *__restrict__ cnearest,
                                                  int *__restrict__ flag, int *err)
{
    const long long c = (long long)blockIdx.x * B + threadIdx.x;
    if (c >= NN)
        return;
    const bool f = told[c] != CELL_AIR && tnew[c] == CELL_AIR;
    flag[c] = f;
    if (f && (cnearest[c] < 0 || cnearest[c] >= nt))
        atomicOr(err, 1);
}

__device__ inline bool air2(int x, int y, int z, int N, const int *told, const int *tnew)
{
    if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N)
        return false;
    const int l = lin3(x, y, z, N);
    return told[l] == CELL_AIR && tnew[l] == CELL_AIR;
}

// solve_fresh_history (ghost_cell.cu:75-196): thread per fresh cell. Fresh cells read only
// cells that are AIR before and after the update, which no thread writes.
__global__ __launch_bounds__(B) void k_fresh_fill(CellArgs a, const StepState *__restrict__ st,
                                                  const int *__restrict__ list, const int *__restrict__ n_ptr,
                                                  const int *__restrict__ err, const int *__restrict__ told,
                                                  const int *__restrict__ tnew, const int *__restrict__ cnearest,
                                                  const float *__restrict__ cpoint, const float *__restrict__ acc_now,
                                                  const float *__restrict__ acc_old, float *__restrict__ grid3)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= *n_ptr)
        return;
    const int N = a.N;
    const size_t nn = (size_t)N * N * N;
    const int t = st->t;
    float *g0 = grid3 + (size_t)(((t % 3) + 3) % 3) * nn;
    float *g1 = grid3 + (size_t)((((t - 1) % 3) + 3) % 3) * nn;
    const int c = list[i];
    if (*err)
    {
        g0[c] = 1e10f;
        return;
    }
    const int x = c / (N * N), y = c / N % N, z = c % N;
    const int tri = cnearest[c];
    const float accs[2] = {acc_now[tri], acc_old[tri]};
    const f3 xb = ld3(cpoint, c), xf = cell_center(x, y, z, a);
    const f3 xr = xb + 2.0f * (xf - xb);
    // getGridBaseCoord (particle_grid.h:171-176)
    const f3 mn = {a.minx, a.miny, a.minz};
    const f3 q = (xr - mn) / a.h;
    const int3_ cc = {(int)q.x, (int)q.y, (int)q.z};
    const f3 d = xr - cell_center(cc.x, cc.y, cc.z, a);
    const int3_ base = {cc.x - (d.x < 0), cc.y - (d.y < 0), cc.z - (d.z < 0)};
    float w[8];
    bool any = false;
    for (int k = 0; k < 8; k++)
    {
        const int nx = base.x + (k >> 2), ny = base.y + ((k >> 1) & 1), nz = base.z + (k & 1);
        if (air2(nx, ny, nz, N, told, tnew))
        {
            const float dist = lengthf(cell_center(nx, ny, nz, a) - xr);
            w[k] = (float)(1.0 / (double)(dist * dist + EPS_F));
            any = true;
        }
        else
            w[k] = 0;
    }
    if (!any)
    {
        float best = 1e10f, p1 = 0, p2 = 0;
        for (int dx = -1; dx < 3; dx++)
            for (int dy = -1; dy < 3; dy++)
                for (int dz = -1; dz < 3; dz++)
                {
                    const int nx = base.x + dx, ny = base.y + dy, nz = base.z + dz;
                    if (!air2(nx, ny, nz, N, told, tnew))
                        continue;
                    const float dist = lengthf(cell_center(nx, ny, nz, a) - xr);
                    if (dist < best)
                    {
                        best = dist;
                        p1 = g0[lin3(nx, ny, nz, N)];
                        p2 = g1[lin3(nx, ny, nz, N)];
                    }
                }
        g0[c] = p1;
        g1[c] = p2;
        return;
    }
    float sum = 0;
    for (int k = 0; k < 8; k++)
        sum += w[k];
    for (int k = 0; k < 8; k++)
        w[k] /= sum;
    const float len = lengthf(xf - xr);
    float *lv[2] = {g0, g1};
    for (int l = 0; l < 2; l++)
    {
        float pr = 0;
        for (int k = 0; k < 8; k++)
        {
            const int nx = base.x + (k >> 2), ny = base.y + ((k >> 1) & 1), nz = base.z + (k & 1);
            const bool in = nx >= 0 && ny >= 0 && nz >= 0 && nx < N && ny < N && nz < N;
            pr += w[k] * (in ? lv[l][lin3(nx, ny, nz, N)] : 0.0f);
        }
        lv[l][c] = pr - 1.225f * accs[l] * len;  // AIR_DENSITY (macro.h:84)
    }
}

__global__ __launch_bounds__(B) void k_solid_zero(long long NN, const StepState *__restrict__ st,
                                                  const int *__restrict__ ctype, float *__restrict__ grid3)
{
    const long long c = (long long)blockIdx.x * B + threadIdx.x;
    if (c >= NN)
        return;
    const int t = st->t;
    if (ctype[c] == CELL_SOLID)
        grid3[(size_t)(((t % 3) + 3) % 3) * NN + c] = 0;
}

inline int nblk(long long n) { return n > 0 ? (int)((n + B - 1) / B) : 1; }
}  // namespace

hipError_t launch_vertex_faces(const int *tris, int nt, int nv, int *cnt, int *start, int *fill, int *items,
                               int *scan_tmp, hipStream_t s)
{
    hipError_t e = launch_fill_int(cnt, nv, 0, s);
    if (e != hipSuccess)
        return e;
    e = launch_fill_int(fill, nv, 0, s);
    if (e != hipSuccess)
        return e;
    k_vf_count<<<nblk(nt), B, 0, s>>>(tris, nt, cnt);
    e = scan_exclusive(cnt, start, nv, scan_tmp, s);
    if (e != hipSuccess)
        return e;
    k_vf_fill<<<nblk(nt), B, 0, s>>>(tris, nt, start, fill, items);
    return hipGetLastError();
}
hipError_t launch_vertex_normals(const float *verts, const int *tris, const TriGeom *geom, int nv, const int *start,
                                 int *items, float *vnormal, hipStream_t s)
{
    k_vertex_normals<<<nblk(nv), B, 0, s>>>(verts, tris, geom, nv, start, items, vnormal);
    return hipGetLastError();
}
hipError_t launch_classify(const CellArgs &a, const float *verts, const int *tris, const int *cell_start,
                           const int *cell_items, const float *vnormal, int *ctype, int *cnearest, float *cpoint,
                           float *creflect, float *cdist, int *typ2, int *flag, int *scan, int *scan_tmp, int *cghost,
                           hipStream_t s)
{
    const long long NN = (long long)a.N * a.N * a.N;
    k_cell_nearest<<<nblk(NN), B, 0, s>>>(a, verts, tris, cell_start, cell_items, vnormal, ctype, cnearest, cpoint,
                                          creflect, cdist);
    k_ghost_mark<<<nblk(NN), B, 0, s>>>(a.N, ctype, typ2, flag);
    k_column_fill<<<nblk((long long)a.N * a.N * WAVE), B, 0, s>>>(a.N, typ2);
    hipError_t e = scan_exclusive(flag, scan, (int)NN, scan_tmp, s);
    if (e != hipSuccess)
        return e;
    k_cell_apply<<<nblk(NN), B, 0, s>>>(NN, typ2, flag, scan, ctype, cghost);
    return hipGetLastError();
}
hipError_t launch_fresh_flags(long long NN, int nt, const int *told, const int *tnew, const int *cnearest, int *flag,
                              int *err, hipStream_t s)
{
    k_fresh_flag<<<nblk(NN), B, 0, s>>>(NN, nt, told, tnew, cnearest, flag, err);
    return hipGetLastError();
}
hipError_t launch_fresh_fill(const CellArgs &a, const StepState *st, const int *list, const int *n_ptr, int cap,
                             const int *err, const int *told, const int *tnew, const int *cnearest,
                             const float *cpoint, const float *acc_now, const float *acc_old, float *grid3,
                             hipStream_t s)
{
    k_fresh_fill<<<nblk(cap), B, 0, s>>>(a, st, list, n_ptr, err, told, tnew, cnearest, cpoint, acc_now, acc_old,
                                          grid3);
    return hipGetLastError();
}
hipError_t launch_solid_zero(long long NN, const StepState *st, const int *ctype, float *grid3, hipStream_t s)
{
    k_solid_zero<<<nblk(NN), B, 0, s>>>(NN, st, ctype, grid3);
    return hipGetLastError();
}

}  // namespace pppm_amd
