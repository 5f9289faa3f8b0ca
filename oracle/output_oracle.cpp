// CPU restatement of the reference's experiment outputs — TEST INFRASTRUCTURE ONLY.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
//
//   orc_ffat_collect   collect_ffat_map   experiments/static/main.cu:60-80,
//                                         experiments/acoustic/main.cu:46-74 (get_coord)
//   orc_slice_image    save_grid          src/utils/visualize.h:120-132, visualize.cu:5-52;
//                      JetColor, preprocess_image_data, linePlaneIntersection,
//                      get_pixel_coord, point_line_distance, add_mesh_kernel,
//                      add_mesh_post_kernel   src/ui/window.cu:28-330;
//                      write_to_png's orientation   src/utils/array_writer.h:94-105
// Built with -ffp-contract=off; float and double operations in the reference's order;
// float-to-byte conversions truncate and saturate; e^x is det_exp (detmath.h).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include "detmath.h"

namespace
{
struct V3
{
    float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len3(V3 a) { return std::sqrt(dot3(a, a)); }
inline unsigned char to_u8(float v)
{
    const float t = std::trunc(v);
    return (unsigned char)(t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t));
}
struct RGBA
{
    unsigned char x, y, z, w;
};

RGBA jet(float v, float vmin, float vmax)
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
    return {to_u8(cx * 255), to_u8(cy * 255), to_u8(cz * 255), 255};
}

bool line_plane(V3 &contact, V3 start, V3 end, V3 pn, V3 pp)
{
    const V3 dir = end - start;
    const V3 normal = pn * (1.0f / std::sqrt(dot3(pn, pn)));
    const V3 to_start = start - pp;
    const float l = len3(dir);
    const float denom = dot3(normal, dir);
    if (std::fabs(denom) < 1e-6)
        return false;
    const float numer = dot3(normal, to_start);
    const float t = -(numer / denom);
    const float eps = 1e-4 * l;
    if (t <= 0 - eps || t >= 1.0f + eps)
        return false;
    const V3 hit = start + t * dir;
    if (len3(contact - hit) < eps)
        return false;
    contact = hit;
    return true;
}

float point_line(float px, float py, float sx, float sy, float ex, float ey)
{
    if (std::sqrt((ex - sx) * (ex - sx) + (ey - sy) * (ey - sy)) < 1e-6)
        return std::sqrt((px - sx) * (px - sx) + (py - sy) * (py - sy));
    const float dx = ex - sx, dy = ey - sy;
    const float qx = px - sx, qy = py - sy;
    float t = (qx * dx + qy * dy) / (dx * dx + dy * dy);
    if (t < 0)
        t = 0;
    else if (t > 1)
        t = 1;
    const float cx = sx + t * dx, cy = sy + t * dy;
    return std::sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy));
}
}  // namespace

extern "C" {

// grid: N^3 floats, index (x N + y) N + z; map: 6 x 2w x 2w, accumulated in place.
// Face b loops axis i = 0..2 and sign = -1, +1 in that order (b = 2 i + (sign > 0)).
void orc_ffat_collect(const float *grid, int N, int w, float *map)
{
    const int W = 2 * w;
    const int c = N / 2;
    int b = -1;
    for (int i = 0; i < 3; i++)
        for (int sign = -1; sign <= 1; sign += 2)
        {
            b++;
            for (int x = 0; x < W; x++)
                for (int y = 0; y < W; y++)
                {
                    int p[3] = {c, c, c};
                    p[i] += sign * w;
                    p[(i + 1) % 3] += x - w;
                    p[(i + 2) % 3] += y - w;
                    float &m = map[((size_t)b * W + x) * W + y];
                    m += std::fabs(grid[((size_t)p[0] * N + p[1]) * N + p[2]]);
                }
        }
}

// grid: N^3 (x-major), the level to draw; verts/tris: the mesh (nt may be 0); rgba: W x W x 4
// bytes with W = N * (1080 / N), written in PNG row order (top row first)
void orc_slice_image(const float *grid, int N, float h, const float *mn3, int axis, int index, float vmax,
                     const float *verts, const int *tris, int nt, unsigned char *rgba)
{
    const int f = 1080 / N, W = N * f;
    const V3 mn{mn3[0], mn3[1], mn3[2]};
    const V3 mx{mn3[0] + (float)N * h, mn3[1] + (float)N * h, mn3[2] + (float)N * h};
    const V3 centre{(mx.x + mn.x) / 2, (mx.y + mn.y) / 2, (mx.z + mn.z) / 2};
    const float off = (float)index - N / 2.0f + 0.5f;
    const V3 nrm{axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
    const V3 pp = centre + (off * nrm) * h;
    const float mesh_line = 3.0f, cell_line = 1.5f;
    std::vector<int> cover((size_t)W * W, 0), count((size_t)W * W, 0);
    for (int i = 0; i < nt; i++)
    {
        auto vx = [&](int k) { return V3{verts[3 * tris[3 * i + k]], verts[3 * tris[3 * i + k] + 1], verts[3 * tris[3 * i + k] + 2]}; };
        const V3 v0 = vx(0), v1 = vx(1), v2 = vx(2);
        V3 c[5] = {};
        int n = 0;
        if (line_plane(c[n], v0, v1, nrm, pp))
            n++;
        if (line_plane(c[n], v1, v2, nrm, pp))
            n++;
        if (line_plane(c[n], v2, v0, nrm, pp))
            n++;
        if (n < 2)
            continue;
        if (len3(c[0] - c[1]) < 1e-4 && n > 2)
            c[1] = c[2];
        float uv[2][2];
        for (int k = 0; k < 2; k++)
        {
            const V3 p = c[k];
            const V3 sc{(p.x - mn.x) / (mx.x - mn.x), (p.y - mn.y) / (mx.y - mn.y), (p.z - mn.z) / (mx.z - mn.z)};
            const float Wf = (float)W;
            if (axis == 2)
            {
                uv[k][0] = sc.x * Wf;
                uv[k][1] = sc.y * Wf;
            }
            else if (axis == 1)
            {
                uv[k][0] = sc.x * Wf;
                uv[k][1] = sc.z * Wf;
            }
            else
            {
                uv[k][0] = sc.y * Wf;
                uv[k][1] = sc.z * Wf;
            }
        }
        const float ext = mesh_line * 4;
        const int x0 = std::max(0, (int)(std::fmin(uv[0][0], uv[1][0]) - ext));
        const int x1 = std::min(W - 1, (int)(std::fmax(uv[0][0], uv[1][0]) + ext));
        const int y0 = std::max(0, (int)(std::fmin(uv[0][1], uv[1][1]) - ext));
        const int y1 = std::min(W - 1, (int)(std::fmax(uv[0][1], uv[1][1]) + ext));
        for (int px = x0; px <= x1; px++)
            for (int py = y0; py <= y1; py++)
            {
                const float d = point_line(px + 0.5f, py + 0.5f, uv[0][0], uv[0][1], uv[1][0], uv[1][1]);
                float alpha = 1.0f;
                if (d > mesh_line)
                {
                    const double q = (double)((d - mesh_line) * 3);
                    alpha = (float)orc_detmath::det_exp(-0.5 * (q * q));
                }
                int &cv = cover[(size_t)px * W + py];
                cv = std::max(cv, (int)(alpha * 255));
                count[(size_t)px * W + py]++;
            }
    }
    auto at = [&](int i, int j) {
        const size_t p = axis == 0 ? ((size_t)index * N + i) * N + j
                                   : (axis == 1 ? ((size_t)i * N + index) * N + j : ((size_t)i * N + j) * N + index);
        return grid[p];
    };
    for (int x = 0; x < W; x++)
        for (int y = 0; y < W; y++)
        {
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
            RGBA pc = jet(value, -vmax, vmax);
            const float k = 2.0f;
            pc.x = to_u8(pc.x / k);
            pc.y = to_u8(pc.y / k);
            pc.z = to_u8(pc.z / k);
            const float cxp = (float)(xd * f) + f / 2.0f, cyp = (float)(yd * f) + f / 2.0f;
            const float dist = f / 2.0f - std::fmax(std::fabs((x + 0.5f) - cxp), std::fabs((y + 0.5f) - cyp));
            const float gs = dist <= cell_line ? 0.4f : 0.0f;
            RGBA o{to_u8(pc.x * (1 - gs) + 255 * gs), to_u8(pc.y * (1 - gs) + 255 * gs),
                   to_u8(pc.z * (1 - gs) + 255 * gs), 255};
            const size_t pix = (size_t)x * W + y;
            if (count[pix] > 0)
            {
                const float al = cover[pix] / 255.0f;
                o = RGBA{to_u8(o.x * (1.0f - al) + 255 * al), to_u8(o.y * (1.0f - al) + 255 * al),
                         to_u8(o.z * (1.0f - al) + 255 * al), to_u8(o.w * (1.0f - al) + 255 * al)};
            }
            unsigned char *q = rgba + ((size_t)(W - 1 - y) * W + x) * 4;
            q[0] = o.x;
            q[1] = o.y;
            q[2] = o.z;
            q[3] = o.w;
        }
}

}  // extern "C"
