// Experiment outputs read from the pressure grid (gfx950): far-field (FFAT) maps accumulated on
// the six faces of a cube around the grid centre, and the colour image of one grid plane with
// the mesh's cut drawn over it.
//
// Reference behaviour: collect_ffat_map (experiments/static/main.cu:60-80; the same map through
// get_coord in experiments/acoustic/main.cu:46-74): for face b = 2 i + (sign > 0) of axis i,
// map(b, x, y) += |p(centre + sign w e_i + (x - w) e_{i+1} + (y - w) e_{i+2})| with the axes taken
// cyclically and centre = (N/2, N/2, N/2). One thread per (x, y) adds its six faces, so every
// entry has a single writer and the sums are the reference's.
#include "internal.hpp"

#include "detmath.hpp"

namespace pppm_amd
{

namespace
{
// thread (x, y) of a 2w x 2w face; y along threadIdx.x, the map's contiguous index
__global__ __launch_bounds__(256) void k_ffat(const StepState *__restrict__ st, const float *__restrict__ grid3,
                                              int N, int w, float *__restrict__ map)
{
    const int W = 2 * w;
    const int y = blockIdx.x * 64 + (threadIdx.x & 63);
    const int x = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= W)
        return;
    const int t = st->t;
    const size_t nn = (size_t)N * N * N;
    const float *g = grid3 + (size_t)(((t % 3) + 3) % 3) * nn;
    const int c = N / 2;
    const int j = x - w, k = y - w;
#pragma unroll
    for (int b = 0; b < 6; b++)
    {
        const int i = b >> 1;
        const int sgn = (b & 1) ? 1 : -1;
        int p[3] = {c, c, c};
        p[i] += sgn * w;
        p[(i + 1) % 3] += j;
        p[(i + 2) % 3] += k;
        float *m = map + ((size_t)b * W + x) * W + y;
        *m = *m + fabsf(g[((size_t)p[0] * N + p[1]) * N + p[2]]);
    }
}
// ------------------------------------------------------------------ slice images
// The reference's image of one grid plane (save_grid, src/utils/visualize.h:120-132 ->
// RenderElement::assign / write_image, src/utils/visualize.cu:5-52 -> CudaRender::setData,
// add_mesh_to_images, src/ui/window.cu:28-120, 122-330 -> write_to_png,
// src/utils/array_writer.h:68-105), written as a W x W RGBA image, W = N * (1080 / N), in the
// PNG's row order (top row first). Pixel (x, y) of the plane image maps to PNG row W - 1 - y,
// column x. FP32 and FP64 operations follow the reference's; float-to-byte conversions
// truncate and saturate; e^x is the deterministic det_exp.
struct SliceArgs
{
    int N, axis, index, f, W;
    float h, vmax, cell_line, mesh_line;
    f3 mn, mx, plane_pos;
};

__device__ inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ inline unsigned char to_u8(float v)
{
    const float t = truncf(v);
    return (unsigned char)(t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t));
}

// JetColor (window.cu:28-57)
__device__ inline uchar4 jet(float v, float vmin, float vmax)
{
    float cx = 1.0f, cy = 1.0f, cz = 1.0f;
    if (v < vmin)
        v = vmin;
    if (v > vmax)
        v = vmax;
    const double dv = vmax - vmin;
    if (v < (vmin + 0.25 * dv))
    {
        cx = 0;
        cy = (float)(4 * (v - vmin) / dv);
    }
    else if (v < (vmin + 0.5 * dv))
    {
        cx = 0;
        cz = (float)(1 + 4 * (vmin + 0.25 * dv - v) / dv);
    }
    else if (v < (vmin + 0.75 * dv))
    {
        cx = (float)(4 * (v - vmin - 0.5 * dv) / dv);
        cz = 0;
    }
    else
    {
        cy = (float)(1 + 4 * (vmin + 0.75 * dv - v) / dv);
        cz = 0;
    }
    return make_uchar4(to_u8(cx * 255), to_u8(cy * 255), to_u8(cz * 255), 255);
}

// linePlaneIntersection (window.cu:124-153); `contact` holds the slot's previous content
__device__ inline bool line_plane(f3 &contact, f3 start, f3 end, f3 pn, f3 pp)
{
    const f3 dir = end - start;
    const f3 normal = pn * (1.0f / sqrtf(dot3(pn, pn)));
    const f3 to_start = start - pp;
    const float l = len3(dir);
    const float denom = dot3(normal, dir);
    if (fabsf(denom) < 1e-6)
        return false;
    const float numer = dot3(normal, to_start);
    const float t = -(numer / denom);
    const float eps = 1e-4 * l;
    if (t <= 0 - eps || t >= 1.0f + eps)
        return false;
    const f3 hit = start + t * dir;
    if (len3(contact - hit) < eps)
        return false;
    contact = hit;
    return true;
}

// get_pixel_coord (window.cu:155-173); axis 0: YZ plane, 1: XZ, 2: XY
__device__ inline void pixel_coord(f3 p, const SliceArgs &a, float &u, float &v)
{
    const f3 s{(p.x - a.mn.x) / (a.mx.x - a.mn.x), (p.y - a.mn.y) / (a.mx.y - a.mn.y), (p.z - a.mn.z) / (a.mx.z - a.mn.z)};
    const float W = (float)a.W;
    if (a.axis == 2)
    {
        u = s.x * W;
        v = s.y * W;
    }
    else if (a.axis == 1)
    {
        u = s.x * W;
        v = s.z * W;
    }
    else
    {
        u = s.y * W;
        v = s.z * W;
    }
}

// point_line_distance (window.cu:175-197)
__device__ inline float point_line(float px, float py, float sx, float sy, float ex, float ey)
{
    if (sqrtf((ex - sx) * (ex - sx) + (ey - sy) * (ey - sy)) < 1e-6)
        return sqrtf((px - sx) * (px - sx) + (py - sy) * (py - sy));
    const float dx = ex - sx, dy = ey - sy;
    const float qx = px - sx, qy = py - sy;
    float t = (qx * dx + qy * dy) / (dx * dx + dy * dy);
    if (t < 0)
        t = 0;
    else if (t > 1)
        t = 1;
    const float cx = sx + t * dx, cy = sy + t * dy;
    return sqrtf((px - cx) * (px - cx) + (py - cy) * (py - cy));
}

// add_mesh_kernel (window.cu:199-283): the triangle's cut with the plane as an anti-aliased
// segment; per pixel the largest coverage (0..255) and the number of segments near it
__global__ void k_mesh_lines(SliceArgs a, const float *__restrict__ verts, const int *__restrict__ tris, int nt,
                             int *__restrict__ cover, int *__restrict__ count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nt)
        return;
    const f3 v0{verts[3 * tris[3 * i]], verts[3 * tris[3 * i] + 1], verts[3 * tris[3 * i] + 2]};
    const f3 v1{verts[3 * tris[3 * i + 1]], verts[3 * tris[3 * i + 1] + 1], verts[3 * tris[3 * i + 1] + 2]};
    const f3 v2{verts[3 * tris[3 * i + 2]], verts[3 * tris[3 * i + 2] + 1], verts[3 * tris[3 * i + 2] + 2]};
    const f3 pn{a.axis == 0 ? 1.0f : 0.0f, a.axis == 1 ? 1.0f : 0.0f, a.axis == 2 ? 1.0f : 0.0f};
    f3 c[5];
    for (int k = 0; k < 5; k++)
        c[k] = f3{0, 0, 0};
    int n = 0;
    if (line_plane(c[n], v0, v1, pn, a.plane_pos))
        n++;
    if (line_plane(c[n], v1, v2, pn, a.plane_pos))
        n++;
    if (line_plane(c[n], v2, v0, pn, a.plane_pos))
        n++;
    if (n < 2)
        return;
    if (len3(c[0] - c[1]) < 1e-4 && n > 2)
        c[1] = c[2];
    float u1, w1, u2, w2;
    pixel_coord(c[0], a, u1, w1);
    pixel_coord(c[1], a, u2, w2);
    const float ext = a.mesh_line * 4;
    const int x0 = max(0, (int)(fminf(u1, u2) - ext)), x1 = min(a.W - 1, (int)(fmaxf(u1, u2) + ext));
    const int y0 = max(0, (int)(fminf(w1, w2) - ext)), y1 = min(a.W - 1, (int)(fmaxf(w1, w2) + ext));
    for (int px = x0; px <= x1; px++)
        for (int py = y0; py <= y1; py++)
        {
            const float d = point_line(px + 0.5f, py + 0.5f, u1, w1, u2, w2);
            float alpha = 1.0f;
            if (d > a.mesh_line)
            {
                const double q = (double)((d - a.mesh_line) * 3);
                alpha = (float)det_exp(-0.5 * (q * q));
            }
            atomicMax(&cover[(size_t)px * a.W + py], (int)(alpha * 255));
            atomicAdd(&count[(size_t)px * a.W + py], 1);
        }
}

// preprocess_image_data (window.cu:60-120) then add_mesh_post_kernel (window.cu:285-309), one
// thread per pixel, written in PNG order
__global__ void k_slice_pixels(SliceArgs a, const StepState *__restrict__ st, const float *__restrict__ grid3,
                               const int *__restrict__ cover, const int *__restrict__ count,
                               uchar4 *__restrict__ out)
{
    const int y = blockIdx.x * 256 + threadIdx.x, x = blockIdx.y;
    if (y >= a.W || x >= a.W)
        return;
    const int N = a.N, f = a.f;
    const size_t nn = (size_t)N * N * N;
    const int t = st->t;
    const float *g = grid3 + (size_t)(((t % 3) + 3) % 3) * nn;
    auto at = [&](int i, int j) {
        const int p = a.axis == 0 ? lin3(a.index, i, j, N) : (a.axis == 1 ? lin3(i, a.index, j, N) : lin3(i, j, a.index, N));
        return g[p];
    };
    const int xd = x / f, yd = y / f;
    float value = at(xd, yd);
    if (xd > 0 && xd < N - 1 && yd > 0 && yd < N - 1)
    {
        const float xfi = (float)((float)x / f - 0.5), yfi = (float)((float)y / f - 0.5);
        const int xi = (int)xfi, yi = (int)yfi;
        const float xr = xfi - xi, yr = yfi - yi;
        const float v1 = at(xi, yi), v2 = at(xi + 1, yi), v3 = at(xi, yi + 1), v4 = at(xi + 1, yi + 1);
        value = (1 - xr) * (1 - yr) * v1 + xr * (1 - yr) * v2 + (1 - xr) * yr * v3 + xr * yr * v4;
    }
    uchar4 pc = jet(value, -a.vmax, a.vmax);
    const float k = 2.0f;
    pc.x = to_u8(pc.x / k);
    pc.y = to_u8(pc.y / k);
    pc.z = to_u8(pc.z / k);
    const float cxp = (float)(xd * f) + f / 2.0f, cyp = (float)(yd * f) + f / 2.0f;
    const float dist = f / 2.0f - fmaxf(fabsf((x + 0.5f) - cxp), fabsf((y + 0.5f) - cyp));
    const float gs = dist <= a.cell_line ? 0.4f : 0.0f;
    uchar4 o = make_uchar4(to_u8(pc.x * (1 - gs) + 255 * gs), to_u8(pc.y * (1 - gs) + 255 * gs),
                           to_u8(pc.z * (1 - gs) + 255 * gs), 255);
    const size_t pix = (size_t)x * a.W + y;
    if (count[pix] > 0)
    {
        const float al = cover[pix] / 255.0f;
        o = make_uchar4(to_u8(o.x * (1.0f - al) + 255 * al), to_u8(o.y * (1.0f - al) + 255 * al),
                        to_u8(o.z * (1.0f - al) + 255 * al), to_u8(o.w * (1.0f - al) + 255 * al));
    }
    out[(size_t)(a.W - 1 - y) * a.W + x] = o;
}
}  // namespace

hipError_t launch_slice_image(int N, float h, const float *mn3, int axis, int index, float vmax,
                              const StepState *st, const float *grid3, const float *verts, const int *tris, int nt,
                              int *cover, int *count, unsigned char *rgba, hipStream_t s)
{
    SliceArgs a{};
    a.N = N;
    a.axis = axis;
    a.index = index;
    a.f = 1080 / N;
    a.W = N * a.f;
    a.h = h;
    a.vmax = vmax;
    a.cell_line = 1.5f;  // CudaRender::setData's line width (visualize.h:73)
    a.mesh_line = 3.0f;  // add_mesh_to_images' line width (visualize.h:79)
    a.mn = f3{mn3[0], mn3[1], mn3[2]};
    // max_pos = min_pos + (N, N, N) * h (particle_grid.h:114)
    a.mx = f3{mn3[0] + (float)N * h, mn3[1] + (float)N * h, mn3[2] + (float)N * h};
    // plane position: the grid centre moved along the normal to the slice (visualize.h:76-78)
    const f3 centre{(a.mx.x + a.mn.x) / 2, (a.mx.y + a.mn.y) / 2, (a.mx.z + a.mn.z) / 2};
    const float off = (float)index - N / 2.0f + 0.5f;
    const f3 nrm{axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
    const f3 step = (off * nrm) * h;
    a.plane_pos = centre + step;
    const long long WW = (long long)a.W * a.W;
    hipError_t e = launch_fill_int(cover, WW, 0, s);
    if (e == hipSuccess)
        e = launch_fill_int(count, WW, 0, s);
    if (e != hipSuccess)
        return e;
    if (nt > 0)
        k_mesh_lines<<<(nt + 255) / 256, 256, 0, s>>>(a, verts, tris, nt, cover, count);
    k_slice_pixels<<<dim3((unsigned)((a.W + 255) / 256), (unsigned)a.W), 256, 0, s>>>(a, st, grid3, cover, count,
                                                                                    (uchar4 *)rgba);
    return hipGetLastError();
}

hipError_t launch_ffat(const StepState *st, const float *grid3, int N, int w, float *map, hipStream_t s)
{
    const int W = 2 * w;
    const dim3 grid((unsigned)((W + 63) / 64), (unsigned)((W + 3) / 4));
    k_ffat<<<grid, 256, 0, s>>>(st, grid3, N, w, map);
    return hipGetLastError();
}

}  // namespace pppm_amd
