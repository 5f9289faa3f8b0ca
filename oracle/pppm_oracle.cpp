/*
 * CPU oracle for the PPPM time step — TEST INFRASTRUCTURE ONLY (see pppm_oracle.h).
 *
 * Restates, from the reference's behaviour, in plain C++:
 *   FDTD inner stencil        src/algorithm/fdtd.cu:7-33
 *   ABC boundary (4x4 solve)  src/algorithm/fdtd.cu:36-150
 *   split-field PML           src/algorithm/fdtd.cu:162-237
 *   PPPM step phases          src/algorithm/pppm.cu:8-225, pppm.h:137-145
 *   weight caches             src/algorithm/pppm_cache.cu:7-289, pppm_cache.h:10-155
 *   TDBEM (CQ/BDF2 + FFT)     src/algorithm/bem.h:12-385
 *   quadrature                src/helmholtz/integrand.h:10-267, potential.h:18-101
 *   binning + neighbour lists src/algorithm/particle_grid.cu:40-241, particle_grid.h:165-185
 *   direct (brute) near field src/algorithm/pppm_direct.h:15-58, pppm_direct.cu:6-48
 * Orders that the reference leaves to atomics (atomic_append, atomicAdd_block) are
 * fixed here: triangles within a cell in ascending index, sums in list order.
 * Build with -ffp-contract=off: every float/double operation is a separate IEEE op.
 */
#include "pppm_oracle.h"

#include "detmath.h"
#include "linear_model.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace
{

constexpr double PI_R = 3.14159265359;  // macro.h:80 (not M_PI)
constexpr int STEPS = 16;               // CMakeLists.txt:27 STEP_NUM
constexpr int RING = 32;                // bem.h:55 History = CircularArray<float, 2*STEP_NUM>
constexpr float EPS_F = 1e-8f;          // macro.h:99
constexpr int GRID_CACHE = 27;          // pppm_cache.h:9
constexpr double RECOMPUTE_THRESHOLD = 0.01;  // pppm_cache.h:8

// LINE_GAUSS_NUM = 3 table (gauss/line.h:26-37), TRI_GAUSS_NUM = 1 (gauss/triangle.h:6-14)
constexpr double LINE3_X[3] = {0.774596669241483, 0.000000000000000, -0.774596669241483};
constexpr double LINE3_W[3] = {0.555555555555554, 0.888888888888889, 0.555555555555554};
constexpr double TRI1_X = 0.3333333333333330;
constexpr double TRI1_W = 1.0;

// ------------------------------------------------------------------ vectors
struct f3
{
    float x, y, z;
};
static inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float length(f3 a) { return sqrtf(dot(a, a)); }

struct d3
{
    double x, y, z;
};
static inline d3 D(f3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
static inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 cross(d3 a, d3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct i3
{
    int x, y, z;
};
static inline bool operator==(i3 a, i3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline bool operator!=(i3 a, i3 b) { return !(a == b); }
static inline i3 operator+(i3 a, i3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }

struct cd
{
    double re, im;
};
struct cf
{
    float re, im;
};

// ------------------------------------------------------------------ FP16 (RNE), bem.h:171-191
static uint16_t f2h(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u)
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u)
        return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u)
    {
        if (ax < 0x33000000u)
            return (uint16_t)sign;
        uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        int e = (int)(ax >> 23);
        int shift = 126 - e;
        uint32_t m = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (m & 1u)))
            m++;
        return (uint16_t)(sign | m);
    }
    uint32_t e = (ax >> 23) - 127 + 15;
    uint32_t mant = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u)))
        h++;
    return (uint16_t)(sign | h);
}

static float h2f(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    uint32_t x;
    if (e == 0)
    {
        if (m == 0)
            x = sign;
        else
        {
            int sh = 0;
            while (!(m & 0x400u))
            {
                m <<= 1;
                sh++;
            }
            m &= 0x3ffu;
            x = sign | ((uint32_t)(127 - 15 + 1 - sh) << 23) | (m << 13);
        }
    }
    else if (e == 31)
        x = sign | 0x7f800000u | (m << 13);
    else
        x = sign | ((e - 15 + 127) << 23) | (m << 13);
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// LayerWeightHalf, bem.h:163-169 (72 bytes)
struct LWH
{
    uint16_t s[STEPS];
    float ss;
    uint16_t d[STEPS];
    float sd;
};
static_assert(sizeof(LWH) == 72, "LayerWeightHalf is 72 B");

// LayerWeightHalf::set, bem.h:171-191
static LWH quantize(const float *s, const float *d)
{
    LWH w;
    float ms = 0, md = 0;
    for (int i = 0; i < STEPS; i++)
    {
        ms = std::max(ms, std::fabs(s[i]));
        md = std::max(md, std::fabs(d[i]));
    }
    if (ms == 0)
        ms = 1;
    if (md == 0)
        md = 1;
    for (int i = 0; i < STEPS; i++)
    {
        w.s[i] = f2h(s[i] / ms);
        w.d[i] = f2h(d[i] / md);
    }
    w.ss = ms;
    w.sd = md;
    return w;
}

// LayerWeightHalf::convolution, bem.h:203-213
static inline float conv(const LWH &w, const float *nring, const float *dring, int t)
{
    float r = 0;
    for (int k = 0; k < STEPS; k++)
    {
        int slot = ((t - k) % RING + RING) % RING;
        float a = -h2f(w.s[k]) * w.ss * nring[slot];
        float b = h2f(w.d[k]) * w.sd * dring[slot];
        r += a + b;
    }
    return r;
}

// ------------------------------------------------------------------ quadrature, FP64 mode
static inline void cexp_d(double re, double im, double &o_re, double &o_im)
{
    // deterministic exp / sincos (detmath.h), the same operations as the device build
    const double e = orc_detmath::det_exp(re);
    double sn, cs;
    orc_detmath::det_sincos(im, &sn, &cs);
    o_re = e * cs;
    o_im = e * sn;
}

// single_layer_potential / double_layer_potential, potential.h:22-47
static inline cd layer64(d3 src, d3 trg, d3 n, cd k, int type)
{
    if (type == ORC_SINGLE_LAYER)
    {
        d3 s = trg - src;
        double r2 = dot(s, s);
        if (r2 == 0)
            return {0, 0};
        double r = std::sqrt(r2);
        double er, ei;
        cexp_d(-r * k.im, r * k.re, er, ei);
        double q = 4 * PI_R * r;
        return {er / q, ei / q};
    }
    d3 s = src - trg;
    double r2 = dot(s, s);
    if (r2 == 0)
        return {0, 0};
    double r = std::sqrt(r2);
    double zr = -r * k.im, zi = r * k.re;  // i*r*k
    double er, ei;
    cexp_d(zr, zi, er, ei);
    double q = 4 * PI_R * r2 * r;
    double ar = -er / q, ai = -ei / q;
    double br = 1 - zr, bi = -zi;
    double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
    double dn = dot(s, n);
    return {pr * dn, pi * dn};
}

struct TriD
{
    d3 v[3];
    d3 n;  // unit normal, triangle_norm (integrand.h:17-23)
    double jac;
};
static TriD tri_d(f3 a, f3 b, f3 c)
{
    TriD t;
    t.v[0] = D(a);
    t.v[1] = D(b);
    t.v[2] = D(c);
    d3 cr = cross(t.v[1] - t.v[0], t.v[2] - t.v[0]);
    t.jac = std::sqrt(dot(cr, cr));
    t.n = {cr.x / t.jac, cr.y / t.jac, cr.z / t.jac};
    return t;
}
// local_to_global (integrand.h:40-44), local_to_global2 (integrand.h:46-50)
static inline d3 l2g(double x1, double x2, const d3 *v) { return (1 - x1 - x2) * v[0] + x1 * v[1] + x2 * v[2]; }
static inline d3 l2g2(double x1, double x2, const d3 *v) { return (1 - x1) * v[0] + (x1 - x2) * v[1] + x2 * v[2]; }

// potential_integrand with a batch of scaled targets, integrand.h:209-222, potential.h:55-92
static cd f2p64(const TriD &t, const d3 *pts, const double *scale, int npts, cd k, int type)
{
    d3 p = l2g(TRI1_X, TRI1_X, t.v);
    double sr = 0, si = 0;
    for (int i = 0; i < npts; i++)
    {
        cd g = layer64(p, pts[i], t.n, k, type);
        sr += g.re * scale[i];
        si += g.im * scale[i];
    }
    double f = 0.5 * TRI1_W * t.jac;
    return {f * sr, f * si};
}

// singular_point, integrand.h:52-147
static cd sing_point64(double xsi, double e1, double e2, double e3, double w, const d3 *tv, const d3 *sv, d3 tn,
                       int nn, cd k, int type)
{
    double rr = 0, ri = 0;
    xsi = 0.5 * (xsi + 1);
    e1 = 0.5 * (e1 + 1);
    e2 = 0.5 * (e2 + 1);
    e3 = 0.5 * (e3 + 1);
    auto add = [&](d3 a, d3 b, double f) {
        cd g = layer64(a, b, tn, k, type);
        rr += g.re * f;
        ri += g.im * f;
    };
    double jw;
    switch (nn)
    {
        case 3:
        {
            jw = xsi * xsi * xsi * e1 * e1 * e2;
            double e12 = e1 * e2, e123 = e1 * e2 * e3;
            add(l2g2(xsi, xsi * (1.0 - e1 + e12), tv), l2g2(xsi * (1.0 - e123), xsi * (1.0 - e1), sv), 1);
            add(l2g2(xsi * (1.0 - e123), xsi * (1.0 - e1), tv), l2g2(xsi, xsi * (1.0 - e1 + e12), sv), 1);
            add(l2g2(xsi, xsi * (e1 - e12 + e123), tv), l2g2(xsi * (1.0 - e12), xsi * (e1 - e12), sv), 1);
            add(l2g2(xsi * (1.0 - e12), xsi * (e1 - e12), tv), l2g2(xsi, xsi * (e1 - e12 + e123), sv), 1);
            add(l2g2(xsi * (1.0 - e123), xsi * (e1 - e123), tv), l2g2(xsi, xsi * (e1 - e12), sv), 1);
            add(l2g2(xsi, xsi * (e1 - e12), tv), l2g2(xsi * (1.0 - e123), xsi * (e1 - e123), sv), 1);
            break;
        }
        case 2:
        {
            jw = xsi * xsi * xsi * e1 * e1;
            double e12 = e1 * e2, e123 = e1 * e2 * e3;
            add(l2g2(xsi, xsi * e1 * e3, tv), l2g2(xsi * (1.0 - e12), xsi * e1 * (1.0 - e2), sv), 1);
            add(l2g2(xsi, xsi * e1, tv), l2g2(xsi * (1.0 - e123), xsi * e1 * e2 * (1 - e3), sv), e2);
            add(l2g2(xsi * (1.0 - e12), xsi * e1 * (1.0 - e2), tv), l2g2(xsi, xsi * e123, sv), e2);
            add(l2g2(xsi * (1.0 - e123), xsi * e12 * (1.0 - e3), tv), l2g2(xsi, xsi * e1, sv), e2);
            add(l2g2(xsi * (1.0 - e123), xsi * e1 * (1.0 - e2 * e3), tv), l2g2(xsi, xsi * e12, sv), e2);
            break;
        }
        case 1:
        {
            jw = xsi * xsi * xsi * e2;
            add(l2g2(xsi, xsi * e1, tv), l2g2(xsi * e2, xsi * e2 * e3, sv), 1);
            add(l2g2(xsi * e2, xsi * e2 * e3, tv), l2g2(xsi, xsi * e1, sv), 1);
            break;
        }
        default:
            return {0, 0};  // no case 0 (integrand.h:68-147)
    }
    double f = jw * w;
    return {rr * f, ri * f};
}

static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; i++)
    {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; it++)
        {
            double p0 = 1, p1 = z;
            for (int j = 2; j <= n; j++)
            {
                double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            double dp = n * (z * p1 - p0) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16)
                break;
        }
        double p0 = 1, p1 = z;
        for (int j = 2; j <= n; j++)
        {
            double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        double dp = n * (z * p1 - p0) / (z * z - 1);
        x[i] = z;
        w[i] = 2 / ((1 - z * z) * dp * dp);
    }
}

// singular_integrand, integrand.h:149-172 (trial = src, test = trg)
static cd sing64(const TriD &tr, const TriD &ts, cd k, int nn, int type, const double *gx, const double *gw, int ng)
{
    double rr = 0, ri = 0;
    for (int a = 0; a < ng; a++)
        for (int b = 0; b < ng; b++)
            for (int c = 0; c < ng; c++)
                for (int d = 0; d < ng; d++)
                {
                    cd p = sing_point64(gx[a], gx[b], gx[c], gx[d], gw[a] * gw[b] * gw[c] * gw[d], tr.v, ts.v, tr.n,
                                        nn, k, type);
                    rr += p.re;
                    ri += p.im;
                }
    double f = tr.jac * ts.jac / 16;
    return {rr * f, ri * f};
}

// regular_integrand, integrand.h:174-190
static cd reg64(const TriD &tr, const TriD &ts, cd k, int type)
{
    d3 a = l2g(TRI1_X, TRI1_X, tr.v);
    d3 b = l2g(TRI1_X, TRI1_X, ts.v);
    cd g = layer64(a, b, tr.n, k, type);
    double f = 0.25 * TRI1_W * TRI1_W * tr.jac * ts.jac;
    return {f * g.re, f * g.im};
}

static inline int common_vertices(const int *a, const int *b)
{
    // triangle_common_vertex_num, integrand.h:10-14
    return (a[0] == b[0]) + (a[0] == b[1]) + (a[0] == b[2]) + (a[1] == b[0]) + (a[1] == b[1]) + (a[1] == b[2]) +
           (a[2] == b[0]) + (a[2] == b[1]) + (a[2] == b[2]);
}

static inline f3 centroid_f(f3 a, f3 b, f3 c) { return (a + b + c) / 3.0f; }

// face2FaceIntegrand, integrand.h:228-259 (decisions in FP32 as the reference does)
static cd f2f64(const f3 *V, const int *src, const int *trg, cd k, int type, float threshold, int ng = 3)
{
    f3 sv[3] = {V[src[0]], V[src[1]], V[src[2]]};
    f3 tv[3] = {V[trg[0]], V[trg[1]], V[trg[2]]};
    int nn = common_vertices(src, trg);
    float dist = length(centroid_f(sv[0], sv[1], sv[2]) - centroid_f(tv[0], tv[1], tv[2]));
    if (dist < threshold && nn == 0)
        return {0, 0};  // both self-integrals take the singular branch with neighbor_num 0 -> 0
    TriD S = tri_d(sv[0], sv[1], sv[2]);
    TriD T = tri_d(tv[0], tv[1], tv[2]);
    if (nn == 0)
        return reg64(S, T, k, type);
    std::vector<double> gx, gw;
    if (ng == 3)
    {
        gx.assign(LINE3_X, LINE3_X + 3);
        gw.assign(LINE3_W, LINE3_W + 3);
    }
    else
        gauss_legendre(ng, gx, gw);
    return sing64(S, T, k, nn, type, gx.data(), gw.data(), ng);
}

// ------------------------------------------------------------------ quadrature, REF32 mode
// Follows the reference's float arithmetic and its thrust::complex promotions to double.
static inline cf cexp_f(float re, float im)
{
    float e = expf(re);
    return {e * cosf(im), e * sinf(im)};
}
static inline cf ikr_f(float r, cf k)
{
    // cpx(0, 1) * r * k  ->  (0, r) * k
    float zr = 0.0f * k.re - r * k.im;
    float zi = 0.0f * k.im + r * k.re;
    return {zr, zi};
}
// returns the complex<double> value that the reference adds into a complex<float>
static inline cd layer32(f3 src, f3 trg, f3 n, cf k, int type)
{
    if (type == ORC_SINGLE_LAYER)
    {
        f3 s = trg - src;
        float r2 = s.x * s.x + s.y * s.y + s.z * s.z;
        if (r2 == 0)
            return {0, 0};
        float r = sqrtf(r2);
        cf e = cexp_f(ikr_f(r, k).re, ikr_f(r, k).im);
        double q = 4 * PI_R * (double)r;
        return {(double)e.re / q, (double)e.im / q};
    }
    f3 s = src - trg;
    float r2 = s.x * s.x + s.y * s.y + s.z * s.z;
    if (r2 == 0)
        return {0, 0};
    float r = sqrtf(r2);
    cf z = ikr_f(r, k);
    cf e = cexp_f(z.re, z.im);
    double q = 4 * PI_R * (double)r2 * (double)r;
    double ar = -(double)e.re / q, ai = -(double)e.im / q;
    cf b = {1.0f - z.re, -z.im};  // (1 - ikr) as complex<float>
    double pr = ar * b.re - ai * b.im, pi = ar * b.im + ai * b.re;
    float dn = dot(s, n);
    return {pr * dn, pi * dn};
}
static inline f3 l2g_f(float x1, float x2, const f3 *v) { return (1 - x1 - x2) * v[0] + x1 * v[1] + x2 * v[2]; }
static inline f3 l2g2_f(float x1, float x2, const f3 *v) { return (1 - x1) * v[0] + (x1 - x2) * v[1] + x2 * v[2]; }
static inline f3 tri_norm_f(const f3 *v)
{
    f3 n = cross(v[1] - v[0], v[2] - v[0]);
    return n / length(n);
}
static inline float jac_f(const f3 *v) { return length(cross(v[1] - v[0], v[2] - v[0])); }

static cf f2p32(const f3 *v, const f3 *pts, const float *scale, int npts, cf k, int type)
{
    f3 n = tri_norm_f(v);
    float jac = jac_f(v);
    float gx = (float)TRI1_X;
    f3 p = l2g_f(gx, gx, v);
    cf pot = {0, 0};
    for (int i = 0; i < npts; i++)
    {
        cd g = layer32(p, pts[i], n, k, type);
        pot.re = (float)((double)pot.re + g.re * (double)scale[i]);
        pot.im = (float)((double)pot.im + g.im * (double)scale[i]);
    }
    double f = 0.5 * (double)(float)TRI1_W * (double)jac;
    return {(float)(f * pot.re), (float)(f * pot.im)};
}

static cf sing_point32(float xsi, float e1, float e2, float e3, float w, const f3 *tv, const f3 *sv, f3 tn, int nn,
                       cf k, int type)
{
    cf res = {0, 0};
    xsi = (float)(0.5 * (xsi + 1));
    e1 = (float)(0.5 * (e1 + 1));
    e2 = (float)(0.5 * (e2 + 1));
    e3 = (float)(0.5 * (e3 + 1));
    auto add = [&](f3 a, f3 b, bool by_e2) {
        cd g = layer32(a, b, tn, k, type);
        cf gf = {(float)g.re, (float)g.im};
        if (by_e2)
        {
            gf.re = gf.re * e2;
            gf.im = gf.im * e2;
        }
        res.re = res.re + gf.re;
        res.im = res.im + gf.im;
    };
    float jw;
    switch (nn)
    {
        case 3:
        {
            jw = xsi * xsi * xsi * e1 * e1 * e2;
            float e12 = e1 * e2, e123 = e1 * e2 * e3;
            add(l2g2_f(xsi, (float)(xsi * (1.0 - e1 + e12)), tv), l2g2_f((float)(xsi * (1.0 - e123)), (float)(xsi * (1.0 - e1)), sv), false);
            add(l2g2_f((float)(xsi * (1.0 - e123)), (float)(xsi * (1.0 - e1)), tv), l2g2_f(xsi, (float)(xsi * (1.0 - e1 + e12)), sv), false);
            add(l2g2_f(xsi, xsi * (e1 - e12 + e123), tv), l2g2_f((float)(xsi * (1.0 - e12)), xsi * (e1 - e12), sv), false);
            add(l2g2_f((float)(xsi * (1.0 - e12)), xsi * (e1 - e12), tv), l2g2_f(xsi, xsi * (e1 - e12 + e123), sv), false);
            add(l2g2_f((float)(xsi * (1.0 - e123)), xsi * (e1 - e123), tv), l2g2_f(xsi, xsi * (e1 - e12), sv), false);
            add(l2g2_f(xsi, xsi * (e1 - e12), tv), l2g2_f((float)(xsi * (1.0 - e123)), xsi * (e1 - e123), sv), false);
            break;
        }
        case 2:
        {
            jw = xsi * xsi * xsi * e1 * e1;
            float e12 = e1 * e2, e123 = e1 * e2 * e3;
            add(l2g2_f(xsi, xsi * e1 * e3, tv), l2g2_f((float)(xsi * (1.0 - e12)), (float)(xsi * e1 * (1.0 - e2)), sv), false);
            add(l2g2_f(xsi, xsi * e1, tv), l2g2_f((float)(xsi * (1.0 - e123)), xsi * e1 * e2 * (1 - e3), sv), true);
            add(l2g2_f((float)(xsi * (1.0 - e12)), (float)(xsi * e1 * (1.0 - e2)), tv), l2g2_f(xsi, xsi * e123, sv), true);
            add(l2g2_f((float)(xsi * (1.0 - e123)), (float)(xsi * e12 * (1.0 - e3)), tv), l2g2_f(xsi, xsi * e1, sv), true);
            add(l2g2_f((float)(xsi * (1.0 - e123)), (float)(xsi * e1 * (1.0 - e2 * e3)), tv), l2g2_f(xsi, xsi * e12, sv), true);
            break;
        }
        case 1:
        {
            jw = xsi * xsi * xsi * e2;
            add(l2g2_f(xsi, xsi * e1, tv), l2g2_f(xsi * e2, xsi * e2 * e3, sv), false);
            add(l2g2_f(xsi * e2, xsi * e2 * e3, tv), l2g2_f(xsi, xsi * e1, sv), false);
            break;
        }
        default:
            return {0, 0};
    }
    res.re = res.re * jw * w;
    res.im = res.im * jw * w;
    return res;
}

static cf f2f32(const f3 *V, const int *src, const int *trg, cf k, int type, float threshold)
{
    f3 sv[3] = {V[src[0]], V[src[1]], V[src[2]]};
    f3 tv[3] = {V[trg[0]], V[trg[1]], V[trg[2]]};
    int nn = common_vertices(src, trg);
    float dist = length(centroid_f(sv[0], sv[1], sv[2]) - centroid_f(tv[0], tv[1], tv[2]));
    if (dist < threshold && nn == 0)
        return {0, 0};
    float sj = jac_f(sv), tj = jac_f(tv);
    f3 sn = tri_norm_f(sv);
    if (nn == 0)
    {
        float gx = (float)TRI1_X;
        f3 a = l2g_f(gx, gx, sv), b = l2g_f(gx, gx, tv);
        cd g = layer32(a, b, sn, k, type);
        cf gf = {(float)g.re, (float)g.im};
        double f = 0.25 * (double)(float)TRI1_W * (double)(float)TRI1_W * (double)sj * (double)tj;
        return {(float)(f * gf.re), (float)(f * gf.im)};
    }
    float gx[3] = {(float)LINE3_X[0], (float)LINE3_X[1], (float)LINE3_X[2]};
    float gw[3] = {(float)LINE3_W[0], (float)LINE3_W[1], (float)LINE3_W[2]};
    cf res = {0, 0};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++)
                for (int d = 0; d < 3; d++)
                {
                    cf p = sing_point32(gx[a], gx[b], gx[c], gx[d], gw[a] * gw[b] * gw[c] * gw[d], sv, tv, sn, nn, k,
                                        type);
                    res.re = res.re + p.re;
                    res.im = res.im + p.im;
                }
    return {res.re * sj * tj / 16, res.im * sj * tj / 16};
}

// ------------------------------------------------------------------ TDBEM, bem.h:253-385
struct Tdbem
{
    int mode;
    double lambda;  // FP32-rounded in REF32 mode
    cd wave[STEPS];
    float threshold;
};

static Tdbem tdbem_init(int mode, float dt, float c, float grid_size)
{
    Tdbem b;
    b.mode = mode;
    if (mode == ORC_REF32)
    {
        double lam = std::max(std::pow((double)(dt * c), 3.0 / STEPS), std::pow((double)EPS_F, 1.0 / (2 * STEPS)));
        float lf = (float)lam;
        b.lambda = lf;
        for (int k = 0; k < STEPS; k++)
        {
            // s_k = lambda * exp(-2*PI*i/16*k) (complex<double>, stored to complex<float>)
            double ang = (-2 * PI_R / STEPS) * k;
            cf s = {(float)((double)lf * std::cos(ang)), (float)((double)lf * std::sin(ang))};
            // BDF2(s) = 0.5*((s*s) - 4*s + 3)  (macro.h:90)
            cf ss = {s.re * s.re - s.im * s.im, s.re * s.im + s.im * s.re};
            cf t = {ss.re - 4 * s.re, ss.im - 4 * s.im};
            t.re = t.re + 3;
            double br = 0.5 * t.re, bi = 0.5 * t.im;
            float dc = dt * c;
            br /= dc;
            bi /= dc;
            b.wave[k] = {(double)(float)(-bi), (double)(float)br};  // * i
        }
    }
    else
    {
        double lam = std::max(std::pow((double)dt * (double)c, 3.0 / STEPS), std::pow((double)EPS_F, 1.0 / (2 * STEPS)));
        b.lambda = lam;
        for (int k = 0; k < STEPS; k++)
        {
            double ang = (-2 * PI_R / STEPS) * k;
            double sr = lam * std::cos(ang), si = lam * std::sin(ang);
            double tr = sr * sr - si * si - 4 * sr + 3, ti = 2 * sr * si - 4 * si;
            double dc = (double)dt * (double)c;
            double br = 0.5 * tr / dc, bi = 0.5 * ti / dc;
            b.wave[k] = {-bi, br};
        }
    }
    b.threshold = (float)(grid_size * 0.1);
    return b;
}

// FFT<step> recursion, bem.h:12-38 (positive-exponent radix-2 DIT), any precision
template <class R>
struct cR
{
    R re, im;
};
template <class R, bool ROUND_F>
static void fft_rec(const cR<R> *in, cR<R> *out, cR<R> *buf, int step)
{
    if (step == STEPS)
    {
        out[0] = in[0];
        return;
    }
    int ns = step * 2;
    fft_rec<R, ROUND_F>(in, out, buf, ns);
    fft_rec<R, ROUND_F>(in + step, out + step, buf + step, ns);
    for (int i = 0; i < STEPS; i += ns)
    {
        buf[i / 2] = out[i];
        buf[(STEPS + i) / 2] = out[i + step];
    }
    for (int i = 0; i < STEPS / 2; i += step)
    {
        double ang = 2 * PI_R * i / STEPS;
        double cr = std::cos(ang), ci = std::sin(ang);
        cR<R> b = buf[i + STEPS / 2];
        double tr = cr * (double)b.re - ci * (double)b.im;
        double ti = cr * (double)b.im + ci * (double)b.re;
        cR<R> a = buf[i];
        out[i] = {(R)((double)a.re + tr), (R)((double)a.im + ti)};
        out[i + STEPS / 2] = {(R)((double)a.re - tr), (R)((double)a.im - ti)};
    }
}

// scaledFFT, bem.h:292-304. REF32: float buffers, float scaling; FP64: double throughout.
static void scaled_fft(int mode, const cd *in, double lambda, double *out)
{
    if (mode == ORC_REF32)
    {
        cR<float> vin[STEPS], o[STEPS], bf[STEPS];
        for (int i = 0; i < STEPS; i++)
            vin[i] = {(float)in[i].re, (float)in[i].im};
        fft_rec<float, true>(vin, o, bf, 1);
        float temp = 1.0f, lf = (float)lambda;
        for (int i = 0; i < STEPS; i++)
        {
            out[i] = (double)(o[i].re / STEPS * temp);
            temp /= lf;
        }
    }
    else
    {
        cR<double> vin[STEPS], o[STEPS], bf[STEPS];
        for (int i = 0; i < STEPS; i++)
            vin[i] = {in[i].re, in[i].im};
        fft_rec<double, false>(vin, o, bf, 1);
        double temp = 1.0;
        for (int i = 0; i < STEPS; i++)
        {
            out[i] = o[i].re / STEPS * temp;
            temp /= lambda;
        }
    }
}

// mirror k > 8 as conj(v[16-k]) (bem.h:313-316)
static inline void mirror(cd *v)
{
    for (int k = STEPS / 2 + 1; k < STEPS; k++)
        v[k] = {v[STEPS - k].re, -v[STEPS - k].im};
}

// ------------------------------------------------------------------ scene
struct Tri
{
    int idx[3];
    f3 center, normal;
    float area;
    i3 cell, base;
};

struct GridEntry
{
    LWH fdtd[GRID_CACHE];
    LWH bem[GRID_CACHE];
    f3 center, normal;
    float area;
    bool valid = false;
};

struct FaceEntry
{
    LWH w;
    float normal_angle, distance;
};

}  // namespace

struct orc_scene
{
    int N;
    float h, dt, c, rc;
    f3 minp;
    int mode;
    int pml;
    float damp;
    Tdbem bem;
    std::vector<float> grid[3];
    std::vector<float> pmlv[3];  // 6 floats per cell (v1[3], v2[3])
    std::vector<float> far;
    int t = -1;

    std::vector<f3> verts;
    std::vector<Tri> tris;
    int nt = 0;
    bool mesh_set = false;

    // binning
    std::vector<int> cell_start, cell_items;  // CSR over N^3 by triangle cell
    std::vector<int> base_start, base_items;  // CSR over N^3 by base cell
    std::vector<int> c3_cells, c3_start, c3_pairs;
    std::vector<int> b4_cells, b4_start, b4_tris;
    std::vector<int> b4_of_tri;  // index into b4_cells for each triangle's base cell

    // caches
    std::vector<GridEntry> gcache;
    std::unordered_map<long long, FaceEntry> fcache;
    std::vector<float> interp;    // T x 8
    std::vector<double> denom;    // T (0.5*area - factor_sum, double)
    long long grid_recomputes = 0, face_recomputes = 0;

    // histories (T x 32)
    std::vector<float> hn, hd;
    std::vector<float> face_far;

    // cell classification (GhostCellSolver::cell_data / cell_data_old, ghost_cell.h:125-127)
    std::vector<int> ctype, ctype_old, cghost, cnearest;
    std::vector<f3> cpoint, creflect;
    std::vector<float> cdist;
    int ghost_count = 0;
    bool classified = false;

    inline int lin(int x, int y, int z) const { return (x * N + y) * N + z; }
    inline int lin(i3 c) const { return lin(c.x, c.y, c.z); }
    inline std::vector<float> &G(int rel) { return grid[((t + rel) % 3 + 3) % 3]; }
    inline const std::vector<float> &G(int rel) const { return grid[((t + rel) % 3 + 3) % 3]; }
    inline std::vector<float> &P(int rel) { return pmlv[((t + rel) % 3 + 3) % 3]; }
    // ParticleGrid::getCenter (particle_grid.h:163-167)
    inline f3 center_f(i3 c) const
    {
        return f3{((float)c.x + 0.5f) * h, ((float)c.y + 0.5f) * h, ((float)c.z + 0.5f) * h} + minp;
    }
    inline d3 center_d(i3 c) const
    {
        return {((double)c.x + 0.5) * h + minp.x, ((double)c.y + 0.5) * h + minp.y, ((double)c.z + 0.5) * h + minp.z};
    }
    inline int ring(int tt) const { return ((tt % RING) + RING) % RING; }
};

namespace
{

// ------------------------------------------------------------------ FDTD
static void fdtd_inner(orc_scene &s)
{
    // fdtd_inner_kernel + laplacian, fdtd.cu:7-33 (FP64 arithmetic on FP32 storage)
    int N = s.N;
    const std::vector<float> &g = s.G(0), &gm = s.G(-1);
    std::vector<float> &gp = s.G(1);
    double h = s.h, dt = s.dt, c = s.c;
    double coef = c * c * dt * dt;
    double hh = h * h;
#pragma omp parallel for schedule(static)
    for (int i = 2; i < N - 2; i++)
        for (int j = 2; j < N - 2; j++)
            for (int k = 2; k < N - 2; k++)
            {
                double sum = 0;
                sum += g[s.lin(i - 1, j, k)];
                sum += g[s.lin(i + 1, j, k)];
                sum += g[s.lin(i, j - 1, k)];
                sum += g[s.lin(i, j + 1, k)];
                sum += g[s.lin(i, j, k - 1)];
                sum += g[s.lin(i, j, k + 1)];
                sum -= 6 * g[s.lin(i, j, k)];
                double lap = sum / hh;
                double p = 2 * g[s.lin(i, j, k)];
                p += coef * lap;
                p += -gm[s.lin(i, j, k)];
                gp[s.lin(i, j, k)] = (float)p;
            }
}

// Solve4x4Linear, fdtd.cu:36-86
static void solve4(const float a[4][4], const float b[4], float res[4])
{
    float s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    float s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    float s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    float s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    float s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    float s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    float c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    float c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    float c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    float c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    float c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    float c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    float invdet = (float)(1.0 / (double)(s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0));
    float inv[4][4];
    inv[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * invdet;
    inv[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * invdet;
    inv[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * invdet;
    inv[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * invdet;
    inv[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * invdet;
    inv[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * invdet;
    inv[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * invdet;
    inv[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * invdet;
    inv[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * invdet;
    inv[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * invdet;
    inv[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * invdet;
    inv[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * invdet;
    inv[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * invdet;
    inv[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * invdet;
    inv[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * invdet;
    inv[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * invdet;
    for (int i = 0; i < 4; i++)
    {
        res[i] = 0;
        for (int j = 0; j < 4; j++)
            res[i] += inv[i][j] * b[j];
    }
}

// ab_condition_solve, fdtd.cu:88-137
static void abc_cell(orc_scene &s, int i, int j, int k)
{
    int N = s.N;
    std::vector<float> &g = s.G(0);
    const std::vector<float> &gm = s.G(-1);
    std::vector<float> &gp = s.G(1);
    float dt = s.dt, h = s.h, c = s.c;
    int co[3] = {i, j, k};
    int judge[3];
    for (int d = 0; d < 3; d++)
        judge[d] = -((int)(co[d] == 1)) + ((int)(co[d] == N - 2));
    auto at = [&](const std::vector<float> &a, int dd, int sgn) {
        int q[3] = {i, j, k};
        q[dd] += sgn;
        return a[s.lin(q[0], q[1], q[2])];
    };
    float a[4][4], b[4];
    float coef = c * c * dt * dt / (h * h);
    a[3][3] = 1;
    for (int d = 0; d < 3; d++)
        a[3][d] = -coef * (float)(judge[d] != 0);
    float gc = g[s.lin(i, j, k)], gmc = gm[s.lin(i, j, k)];
    b[3] = -gmc + 2 * gc - coef * (6 * gc);
    for (int d = 0; d < 3; d++)
        b[3] += coef * at(g, d, +1) * (float)(judge[d] <= 0) + coef * at(g, d, -1) * (float)(judge[d] >= 0);
    float coef1 = 1 / (dt * dt), coef2 = c / (2 * h * dt), coef3 = -(c * c) / (2 * h * h);
    for (int idx = 0; idx < 3; idx++)
    {
        int x = idx, y = (idx + 1) % 3, z = (idx + 2) % 3;
        a[idx][3] = coef1;
        a[idx][x] = coef2;
        a[idx][y] = coef3 * (float)(judge[y] != 0);
        a[idx][z] = coef3 * (float)(judge[z] != 0);
        b[idx] = -coef1 * (gmc - 2 * gc) +
                 -coef2 * (at(g, x, +1) * (float)(-(int)(judge[x] < 0)) + at(g, x, -1) * (float)(-(int)(judge[x] > 0)) +
                           at(gm, x, +1) * (float)(-judge[x]) + at(gm, x, -1) * (float)judge[x]) +
                 -coef3 * (at(g, y, +1) * (float)(judge[y] <= 0) + at(g, y, -1) * (float)(judge[y] >= 0) +
                           at(g, z, +1) * (float)(judge[z] <= 0) + at(g, z, -1) * (float)(judge[z] >= 0) - 4 * gc);
    }
    float r[4];
    solve4(a, b, r);
    gp[s.lin(i, j, k)] = r[3];
    for (int d = 0; d < 3; d++)
        if (judge[d] != 0)
        {
            int q[3] = {i, j, k};
            q[d] += judge[d];
            g[s.lin(q[0], q[1], q[2])] = r[d];
        }
}

static void fdtd_boundary(orc_scene &s)
{
    // fdtd_boundary_kernel, fdtd.cu:139-150: every cell of the index-1 / N-2 shell once
    int N = s.N;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++)
            for (int k = 1; k < N - 1; k++)
                if (i == 1 || i == N - 2 || j == 1 || j == N - 2 || k == 1 || k == N - 2)
                    abc_cell(s, i, j, k);
}

// get_damp / get_sign / pml_step_kernel, fdtd.cu:162-232
static inline float pml_damp(const orc_scene &s, int x, int y, int z)
{
    int N = s.N, w = s.pml;
    if (x < w || x >= N - w || y < w || y >= N - w || z < w || z >= N - w)
        return s.damp;
    return 0;
}

static void pml_step(orc_scene &s)
{
    int N = s.N;
    const std::vector<float> &g = s.G(0), &gm = s.G(-1);
    std::vector<float> &gp = s.G(1);
    const std::vector<float> &P = s.P(0);
    std::vector<float> &Pn = s.P(1);
    float c = s.c, h = s.h, dt = s.dt;
    for (int x = 1; x < N - 1; x++)
        for (int y = 1; y < N - 1; y++)
            for (int z = 1; z < N - 1; z++)
            {
                int co[3] = {x, y, z};
                int sg[3];
                for (int i = 0; i < 3; i++)
                    sg[i] = co[i] > N / 2 ? 1 : -1;
                float cd_ = pml_damp(s, x, y, z);
                auto off = [&](int i, int d) {
                    int q[3] = {x, y, z};
                    q[i] += d;
                    return s.lin(q[0], q[1], q[2]);
                };
                auto dmp = [&](int i, int d) {
                    int q[3] = {x, y, z};
                    q[i] += d;
                    return pml_damp(s, q[0], q[1], q[2]);
                };
                int me = s.lin(x, y, z);
                float rhs = 0;
                for (int i = 0; i < 3; i++)
                {
                    rhs += 1 / (h * h) * (g[off(i, 1)] - 2 * g[me] + g[off(i, -1)]);
                    rhs += 1 / h *
                           (cd_ * P[off(i, sg[i]) * 6 + 3 + i] - dmp(i, -sg[i]) * P[off(i, -sg[i]) * 6 + i]);
                }
                gp[me] = g[me] * 2 - gm[me] + c * c * dt * dt * rhs;
                if (cd_ == 0)
                    continue;
                for (int i = 0; i < 3; i++)
                {
                    float r = 0;
                    r += -1.0f / 2.0f * (dmp(i, -sg[i]) * P[off(i, -sg[i]) * 6 + i] + cd_ * P[me * 6 + i]);
                    r += -1.0f / (h * 2.0f) * (g[off(i, sg[i])] - g[off(i, -sg[i])]);
                    Pn[me * 6 + i] = P[me * 6 + i] + c * dt * r;
                }
                for (int i = 0; i < 3; i++)
                {
                    float r = 0;
                    r += -1.0f / 2.0f * (dmp(i, -sg[i]) * P[me * 6 + 3 + i] + cd_ * P[off(i, sg[i]) * 6 + 3 + i]);
                    r += -1.0f / (h * 2.0f) * (g[off(i, sg[i])] - g[off(i, -sg[i])]);
                    Pn[me * 6 + 3 + i] = P[me * 6 + 3 + i] + c * dt * r;
                }
            }
}

// ------------------------------------------------------------------ mesh / binning
static inline i3 grid_coord(const orc_scene &s, f3 p)
{
    // getGridCoord: make_int3((pos - min_pos) / grid_size), particle_grid.h:169
    f3 q = (p - s.minp) / s.h;
    return {(int)q.x, (int)q.y, (int)q.z};
}
static inline i3 base_coord(const orc_scene &s, f3 p)
{
    // getGridBaseCoord, particle_grid.h:171-176
    i3 c = grid_coord(s, p);
    f3 d = p - s.center_f(c);
    return {c.x - (d.x < 0), c.y - (d.y < 0), c.z - (d.z < 0)};
}

static int update_triangles(orc_scene &s)
{
    // update_triangles_kernel geometry, particle_grid.cu:40-56
    for (int i = 0; i < s.nt; i++)
    {
        Tri &t = s.tris[i];
        f3 v0 = s.verts[t.idx[0]], v1 = s.verts[t.idx[1]], v2 = s.verts[t.idx[2]];
        t.center = (v0 + v1 + v2) / 3.0f;
        f3 cr = cross(v1 - v0, v2 - v0);
        float inv = 1.0f / sqrtf(dot(cr, cr));  // normalize() = v * rsqrtf(dot(v, v)), helper_math.h:1459
        t.normal = cr * inv;
        t.area = length(cr) / 2.0f;
        t.cell = grid_coord(s, t.center);
        t.base = base_coord(s, t.center);
        int N = s.N;
        if (t.cell.x < 2 || t.cell.y < 2 || t.cell.z < 2 || t.cell.x > N - 3 || t.cell.y > N - 3 || t.cell.z > N - 3)
            return -3;
    }
    return 0;
}

static void csr_by(const orc_scene &s, bool base, std::vector<int> &start, std::vector<int> &items)
{
    int NN = s.N * s.N * s.N;
    start.assign(NN + 1, 0);
    for (int i = 0; i < s.nt; i++)
        start[s.lin(base ? s.tris[i].base : s.tris[i].cell) + 1]++;
    for (int c = 0; c < NN; c++)
        start[c + 1] += start[c];
    items.assign(s.nt, 0);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < s.nt; i++)
        items[fill[s.lin(base ? s.tris[i].base : s.tris[i].cell)]++] = i;
}

static void build_lists(orc_scene &s)
{
    int N = s.N;
    csr_by(s, false, s.cell_start, s.cell_items);
    csr_by(s, true, s.base_start, s.base_items);
    // init_neigbor_list_size_kernel + fill_neighbor_list_3_kernel, particle_grid.cu:139-196
    s.c3_cells.clear();
    s.c3_start.assign(1, 0);
    s.c3_pairs.clear();
    for (int x = 1; x < N - 1; x++)
        for (int y = 1; y < N - 1; y++)
            for (int z = 1; z < N - 1; z++)
            {
                size_t before = s.c3_pairs.size();
                for (int idx = 0; idx < 27; idx++)
                {
                    int dx = idx % 3 - 1, dy = idx / 3 % 3 - 1, dz = idx / 9 - 1;
                    int nb = s.lin(x + dx, y + dy, z + dz);
                    for (int q = s.cell_start[nb]; q < s.cell_start[nb + 1]; q++)
                    {
                        int tri = s.cell_items[q];
                        // weight_idx(grid, tri_coord) = grid - tri + 1 (pppm_cache.h:36-40)
                        int o = (1 - dx) + 3 * (1 - dy) + 9 * (1 - dz);
                        s.c3_pairs.push_back(tri * 27 + o);
                    }
                }
                if (s.c3_pairs.size() > before)
                {
                    s.c3_cells.push_back(s.lin(x, y, z));
                    s.c3_start.push_back((int)s.c3_pairs.size());
                }
            }
    // base_coord_nonempty + fill_neighbor_list_4_kernel, particle_grid.cu:198-228
    s.b4_cells.clear();
    s.b4_start.assign(1, 0);
    s.b4_tris.clear();
    std::vector<int> b4_index(N * N * N, -1);
    for (int x = 1; x < N - 1; x++)
        for (int y = 1; y < N - 1; y++)
            for (int z = 1; z < N - 1; z++)
            {
                int b = s.lin(x, y, z);
                if (s.base_start[b + 1] == s.base_start[b])
                    continue;
                for (int idx = 0; idx < 64; idx++)
                {
                    int dx = idx % 4 - 1, dy = idx / 4 % 4 - 1, dz = idx / 16 - 1;
                    int nb = s.lin(x + dx, y + dy, z + dz);
                    for (int q = s.cell_start[nb]; q < s.cell_start[nb + 1]; q++)
                        s.b4_tris.push_back(s.cell_items[q]);
                }
                b4_index[b] = (int)s.b4_cells.size();
                s.b4_cells.push_back(b);
                s.b4_start.push_back((int)s.b4_tris.size());
            }
    s.b4_of_tri.resize(s.nt);
    for (int i = 0; i < s.nt; i++)
        s.b4_of_tri[i] = b4_index[s.lin(s.tris[i].base)];
}

// ------------------------------------------------------------------ weight construction
static inline i3 weight_offset(int o) { return {o % 3 - 1, o / 3 % 3 - 1, o / 9 - 1}; }  // pppm_cache.h:42-52

// one complex "laplace weight" pre-FFT for a batch of scaled targets (batch_grid_near_field, pppm_cache.cu:32-51)
static void batch_f2p(const orc_scene &s, const Tri &t, const i3 *cells, const float *scales, int n, cd *vs,
                      cd *vd)
{
    f3 tv[3] = {s.verts[t.idx[0]], s.verts[t.idx[1]], s.verts[t.idx[2]]};
    if (s.mode == ORC_REF32)
    {
        f3 pts[16];
        for (int i = 0; i < n; i++)
            pts[i] = s.center_f(cells[i]);
        for (int k = 0; k <= STEPS / 2; k++)
        {
            cf kk = {(float)s.bem.wave[k].re, (float)s.bem.wave[k].im};
            cf a = f2p32(tv, pts, scales, n, kk, ORC_SINGLE_LAYER);
            cf b = f2p32(tv, pts, scales, n, kk, ORC_DOUBLE_LAYER);
            vs[k] = {a.re, a.im};
            vd[k] = {b.re, b.im};
        }
    }
    else
    {
        TriD T = tri_d(tv[0], tv[1], tv[2]);
        d3 pts[16];
        double sc[16];
        for (int i = 0; i < n; i++)
        {
            pts[i] = s.center_d(cells[i]);
            sc[i] = scales[i];
        }
        for (int k = 0; k <= STEPS / 2; k++)
        {
            vs[k] = f2p64(T, pts, sc, n, s.bem.wave[k], ORC_SINGLE_LAYER);
            vd[k] = f2p64(T, pts, sc, n, s.bem.wave[k], ORC_DOUBLE_LAYER);
        }
    }
    mirror(vs);
    mirror(vd);
}

// update_grid_cache_kernel, pppm_cache.cu:64-123
static void grid_entry(orc_scene &s, int ti)
{
    const Tri &t = s.tris[ti];
    GridEntry &e = s.gcache[ti];
    float c = s.c, dt = s.dt, h = s.h;
    float scale = c * c * dt * dt / (h * h);
    for (int o = 0; o < GRID_CACHE; o++)
    {
        i3 nb = t.cell + weight_offset(o);
        i3 dst[8] = {nb,
                     nb + i3{-1, 0, 0},
                     nb + i3{1, 0, 0},
                     nb + i3{0, -1, 0},
                     nb + i3{0, 1, 0},
                     nb + i3{0, 0, -1},
                     nb + i3{0, 0, 1},
                     nb};
        float dsc[8] = {2, scale, scale, scale, scale, scale, scale, -6 * scale};
        i3 cells[8];
        float scs[8];
        int n = 0;
        for (int i = 0; i < 8; i++)
            if (dst[i] != t.cell)
            {
                cells[n] = dst[i];
                scs[n] = dsc[i];
                n++;
            }
        cd vs[STEPS], vd[STEPS];
        batch_f2p(s, t, cells, scs, n, vs, vd);
        double ws[STEPS], wd[STEPS];
        scaled_fft(s.mode, vs, s.bem.lambda, ws);
        scaled_fft(s.mode, vd, s.bem.lambda, wd);
        // w.move(-1), bem.h:128-140
        for (int i = STEPS - 1; i >= 1; i--)
        {
            ws[i] = ws[i - 1];
            wd[i] = wd[i - 1];
        }
        ws[0] = wd[0] = 0;
        double gs[STEPS] = {0}, gd[STEPS] = {0};
        bool self = (nb == t.cell);
        if (!self)
        {
            // add_grid_near_field(w, nb, -1, -2) and (w2, nb, 1, 0): laplace_weight to the cell centre
            i3 one[1] = {nb};
            float sc1[1] = {1.0f};
            cd us[STEPS], ud[STEPS];
            batch_f2p(s, t, one, sc1, 1, us, ud);
            scaled_fft(s.mode, us, s.bem.lambda, gs);
            scaled_fft(s.mode, ud, s.bem.lambda, gd);
        }
        float fs[STEPS], fd[STEPS], bs[STEPS], bd[STEPS];
        if (s.mode == ORC_REF32)
        {
            for (int i = 0; i < STEPS; i++)
            {
                fs[i] = (float)ws[i];
                fd[i] = (float)wd[i];
            }
            if (!self)
                for (int i = 2; i < STEPS; i++)
                {
                    fs[i] += (float)gs[i - 2] * -1.0f;
                    fd[i] += (float)gd[i - 2] * -1.0f;
                }
        }
        else
        {
            for (int i = 0; i < STEPS; i++)
            {
                double a = ws[i], b = wd[i];
                if (!self && i >= 2)
                {
                    a += gs[i - 2] * -1.0;
                    b += gd[i - 2] * -1.0;
                }
                fs[i] = (float)a;
                fd[i] = (float)b;
            }
        }
        for (int i = 0; i < STEPS; i++)
        {
            bs[i] = (float)gs[i];
            bd[i] = (float)gd[i];
        }
        e.fdtd[o] = quantize(fs, fd);
        e.bem[o] = quantize(bs, bd);
    }
    e.center = t.center;
    e.normal = t.normal;
    e.area = t.area;
    e.valid = true;
}

// GridCache::need_recompute, pppm_cache.h:64-69
static bool grid_dirty(const orc_scene &s, const GridEntry &e, const Tri &t)
{
    return ((double)length(e.center - t.center) > RECOMPUTE_THRESHOLD * (double)s.h ||
            (double)length(e.normal - t.normal) > RECOMPUTE_THRESHOLD ||
            (double)std::fabs(e.area - t.area) > RECOMPUTE_THRESHOLD * (double)t.area);
}

// update_interpolation_weight_kernel, pppm_cache.cu:157-181
static void interp_weights(orc_scene &s)
{
    s.interp.assign((size_t)s.nt * 8, 0.0f);
    for (int i = 0; i < s.nt; i++)
    {
        const Tri &t = s.tris[i];
        f3 v[3] = {s.verts[t.idx[0]], s.verts[t.idx[1]], s.verts[t.idx[2]]};
        if (s.mode == ORC_REF32)
        {
            float jac = jac_f(v);
            float gx = (float)TRI1_X;
            f3 p = l2g_f(gx, gx, v);
            f3 d = (p - s.center_f(t.base)) / s.h;
            float w[8] = {(1 - d.x) * (1 - d.y) * (1 - d.z), (1 - d.x) * (1 - d.y) * d.z, (1 - d.x) * d.y * (1 - d.z),
                          (1 - d.x) * d.y * d.z,             d.x * (1 - d.y) * (1 - d.z), d.x * (1 - d.y) * d.z,
                          d.x * d.y * (1 - d.z),             d.x * d.y * d.z};
            float sc = (float)(0.5 * (double)(float)TRI1_W * (double)jac);
            for (int q = 0; q < 8; q++)
                s.interp[i * 8 + q] = 0.0f + sc * w[q];
        }
        else
        {
            TriD T = tri_d(v[0], v[1], v[2]);
            d3 p = l2g(TRI1_X, TRI1_X, T.v);
            d3 bc = s.center_d(t.base);
            double hh = s.h;
            double dx = (p.x - bc.x) / hh, dy = (p.y - bc.y) / hh, dz = (p.z - bc.z) / hh;
            double w[8] = {(1 - dx) * (1 - dy) * (1 - dz), (1 - dx) * (1 - dy) * dz, (1 - dx) * dy * (1 - dz),
                           (1 - dx) * dy * dz,             dx * (1 - dy) * (1 - dz), dx * (1 - dy) * dz,
                           dx * dy * (1 - dz),             dx * dy * dz};
            double sc = 0.5 * TRI1_W * T.jac;
            for (int q = 0; q < 8; q++)
                s.interp[i * 8 + q] = (float)(sc * w[q]);
        }
    }
}

// update_particle_cache_kernel, pppm_cache.cu:234-271  (src = neighbour j, dst = centre c)
static LWH face_weight(const orc_scene &s, int ci, int ji)
{
    const Tri &c = s.tris[ci], &j = s.tris[ji];
    cd vs[STEPS], vd[STEPS];
    for (int k = 0; k <= STEPS / 2; k++)
    {
        if (s.mode == ORC_REF32)
        {
            cf kk = {(float)s.bem.wave[k].re, (float)s.bem.wave[k].im};
            cf a = f2f32(s.verts.data(), j.idx, c.idx, kk, ORC_SINGLE_LAYER, s.bem.threshold);
            cf b = f2f32(s.verts.data(), j.idx, c.idx, kk, ORC_DOUBLE_LAYER, s.bem.threshold);
            vs[k] = {a.re, a.im};
            vd[k] = {b.re, b.im};
        }
        else
        {
            vs[k] = f2f64(s.verts.data(), j.idx, c.idx, s.bem.wave[k], ORC_SINGLE_LAYER, s.bem.threshold);
            vd[k] = f2f64(s.verts.data(), j.idx, c.idx, s.bem.wave[k], ORC_DOUBLE_LAYER, s.bem.threshold);
        }
    }
    mirror(vs);
    mirror(vd);
    i3 cells[8];
    float scs[8];
    int n = 0;
    for (int dx = 0; dx < 2; dx++)
        for (int dy = 0; dy < 2; dy++)
            for (int dz = 0; dz < 2; dz++)
            {
                int widx = dx * 4 + dy * 2 + dz;
                i3 cc = c.base + i3{dx, dy, dz};
                f3 ctr = s.center_f(cc);
                double lim = (double)s.h * 1.5;
                if ((double)std::fabs(ctr.x - j.center.x) > lim || (double)std::fabs(ctr.y - j.center.y) > lim ||
                    (double)std::fabs(ctr.z - j.center.z) > lim)
                {
                    cells[n] = cc;
                    scs[n] = -s.interp[ci * 8 + widx];
                    n++;
                }
            }
    cd us[STEPS], ud[STEPS];
    batch_f2p(s, j, cells, scs, n, us, ud);
    for (int k = 0; k < STEPS; k++)
    {
        if (s.mode == ORC_REF32)
        {
            vs[k] = {(double)((float)vs[k].re + (float)us[k].re), (double)((float)vs[k].im + (float)us[k].im)};
            vd[k] = {(double)((float)vd[k].re + (float)ud[k].re), (double)((float)vd[k].im + (float)ud[k].im)};
        }
        else
        {
            vs[k] = {vs[k].re + us[k].re, vs[k].im + us[k].im};
            vd[k] = {vd[k].re + ud[k].re, vd[k].im + ud[k].im};
        }
    }
    double ws[STEPS], wd[STEPS];
    scaled_fft(s.mode, vs, s.bem.lambda, ws);
    scaled_fft(s.mode, vd, s.bem.lambda, wd);
    float fs[STEPS], fd[STEPS];
    for (int i = 0; i < STEPS; i++)
    {
        fs[i] = (float)ws[i];
        fd[i] = (float)wd[i];
    }
    return quantize(fs, fd);
}

static void update_caches(orc_scene &s, bool first)
{
    // GridCache::update_cache, pppm_cache.cu:125-134
    std::vector<int> dirty;
    for (int i = 0; i < s.nt; i++)
    {
        const GridEntry &e = s.gcache[i];
        if (first || !e.valid || grid_dirty(s, e, s.tris[i]))
            dirty.push_back(i);
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t q = 0; q < dirty.size(); q++)
        grid_entry(s, dirty[q]);
    s.grid_recomputes = (long long)dirty.size();
    // FaceCache::update_cache, pppm_cache.cu:273-289
    interp_weights(s);
    // compute_face_compute_list_kernel, pppm_cache.cu:183-232: new or moved pairs
    struct Job
    {
        long long key;
        int ci, ji;
        float normal_angle, distance;
    };
    std::vector<Job> jobs;
    for (int ci = 0; ci < s.nt; ci++)
    {
        int b = s.b4_of_tri[ci];
        const Tri &c = s.tris[ci];
        for (int q = s.b4_start[b]; q < s.b4_start[b + 1]; q++)
        {
            int ji = s.b4_tris[q];
            const Tri &j = s.tris[ji];
            float normal_angle = dot(j.normal, c.normal);
            float distance = length(j.center - c.center);
            long long key = (long long)ci * s.nt + ji;
            auto it = s.fcache.find(key);
            bool recompute = false;
            if (it == s.fcache.end())
                recompute = true;
            else
            {
                const FaceEntry &e = it->second;
                recompute = (double)std::fabs(e.normal_angle - normal_angle) > RECOMPUTE_THRESHOLD ||
                            (double)std::fabs(e.distance - distance) > RECOMPUTE_THRESHOLD * (double)distance;
            }
            if (recompute)
                jobs.push_back({key, ci, ji, normal_angle, distance});
        }
    }
    std::vector<LWH> done(jobs.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t q = 0; q < jobs.size(); q++)
        done[q] = face_weight(s, jobs[q].ci, jobs[q].ji);
    for (size_t q = 0; q < jobs.size(); q++)
    {
        FaceEntry fe;
        fe.normal_angle = jobs[q].normal_angle;
        fe.distance = jobs[q].distance;
        fe.w = done[q];
        s.fcache[jobs[q].key] = fe;
    }
    s.face_recomputes = (long long)jobs.size();
    s.denom.assign(s.nt, 0.0);
    for (int ci = 0; ci < s.nt; ci++)
    {
        int b = s.b4_of_tri[ci];
        float factor_sum = 0;
        for (int q = s.b4_start[b]; q < s.b4_start[b + 1]; q++)
        {
            const LWH &w = s.fcache[(long long)ci * s.nt + s.b4_tris[q]].w;
            factor_sum += h2f(w.d[0]) * w.sd;  // face_factor, pppm.cu:167; summed as in pppm.cu:181-185
        }
        s.denom[ci] = 0.5 * (double)s.tris[ci].area - (double)factor_sum;
    }
}

// ------------------------------------------------------------------ step phases
static void phase_far(orc_scene &s)
{
    // solve_fdtd_far, pppm.cu:8-51
    int t = s.t;
    const std::vector<float> &g = s.G(0);
    s.far = g;
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t q = 0; q < s.c3_cells.size(); q++)
    {
        float acc = 0;
        for (int p = s.c3_start[q]; p < s.c3_start[q + 1]; p++)
        {
            int tri = s.c3_pairs[p] / 27, o = s.c3_pairs[p] % 27;
            acc += conv(s.gcache[tri].fdtd[o], &s.hn[tri * RING], &s.hd[tri * RING], t);
        }
        int cell = s.c3_cells[q];
        s.far[cell] = g[cell] - acc;
    }
}

static void phase_neumann(orc_scene &s, const float *nm)
{
    // set_neumann_condition_kernel, pppm.cu:206-213
    int slot = s.ring(s.t);
    for (int i = 0; i < s.nt; i++)
        s.hn[i * RING + slot] = nm[i];
}

static void phase_dirichlet(orc_scene &s)
{
    int t = s.t, slot = s.ring(t), N = s.N;
    // solve_face_far_kernel, pppm.cu:116-144
    s.face_far.assign(s.nt, 0.0f);
    for (int i = 0; i < s.nt; i++)
    {
        const Tri &tr = s.tris[i];
        float acc = 0;
        for (int dx = 0; dx < 2; dx++)
            for (int dy = 0; dy < 2; dy++)
                for (int dz = 0; dz < 2; dz++)
                {
                    int w = dx * 4 + dy * 2 + dz;
                    acc += s.interp[i * 8 + w] * s.far[s.lin(tr.base.x + dx, tr.base.y + dy, tr.base.z + dz)];
                }
        s.face_far[i] = acc;
        s.hd[i * RING + slot] = 0;
    }
    (void)N;
    // solve_face_near_kernel + update_dirichlet_kernel, pppm.cu:146-193
    std::vector<float> newd(s.nt);
#pragma omp parallel for schedule(dynamic, 64)
    for (int ci = 0; ci < s.nt; ci++)
    {
        int b = s.b4_of_tri[ci];
        float near = 0;
        for (int q = s.b4_start[b]; q < s.b4_start[b + 1]; q++)
        {
            int ji = s.b4_tris[q];
            const LWH &w = s.fcache[(long long)ci * s.nt + ji].w;
            near += conv(w, &s.hn[ji * RING], &s.hd[ji * RING], t);
        }
        float num = s.face_far[ci] * s.rc + near;
        newd[ci] = (float)((double)num / s.denom[ci]);
    }
    for (int ci = 0; ci < s.nt; ci++)
        s.hd[ci * RING + slot] = newd[ci];
}

static void phase_near(orc_scene &s)
{
    // solve_fdtd_near, pppm.cu:53-114
    int t = s.t;
    std::vector<float> &g = s.G(0);
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t q = 0; q < s.c3_cells.size(); q++)
    {
        float acc = 0;
        for (int p = s.c3_start[q]; p < s.c3_start[q + 1]; p++)
        {
            int tri = s.c3_pairs[p] / 27, o = s.c3_pairs[p] % 27;
            acc += conv(s.gcache[tri].bem[o], &s.hn[tri * RING], &s.hd[tri * RING], t);
        }
        int cell = s.c3_cells[q];
        g[cell] = (float)((double)s.far[cell] + (double)acc * 0.98);
    }
}

static void fdtd_step(orc_scene &s)
{
    // FDTD::step, fdtd.h:73-87
    if (s.pml > 0)
        pml_step(s);
    else
    {
        fdtd_inner(s);
        fdtd_boundary(s);
    }
    s.t++;
}

// ------------------------------------------------------------------ direct path (pppm_direct.h:15-58)
static float laplace_point(const orc_scene &s, int tri, f3 pt, int t)
{
    // TDBEM::laplace (bem.h:356-363) with a float LayerWeight
    const Tri &tr = s.tris[tri];
    f3 tv[3] = {s.verts[tr.idx[0]], s.verts[tr.idx[1]], s.verts[tr.idx[2]]};
    cd vs[STEPS], vd[STEPS];
    for (int k = 0; k <= STEPS / 2; k++)
    {
        if (s.mode == ORC_REF32)
        {
            cf kk = {(float)s.bem.wave[k].re, (float)s.bem.wave[k].im};
            float one = 1.0f;
            cf a = f2p32(tv, &pt, &one, 1, kk, ORC_SINGLE_LAYER);
            cf b = f2p32(tv, &pt, &one, 1, kk, ORC_DOUBLE_LAYER);
            vs[k] = {a.re, a.im};
            vd[k] = {b.re, b.im};
        }
        else
        {
            TriD T = tri_d(tv[0], tv[1], tv[2]);
            d3 p = D(pt);
            double one = 1.0;
            vs[k] = f2p64(T, &p, &one, 1, s.bem.wave[k], ORC_SINGLE_LAYER);
            vd[k] = f2p64(T, &p, &one, 1, s.bem.wave[k], ORC_DOUBLE_LAYER);
        }
    }
    mirror(vs);
    mirror(vd);
    double ws[STEPS], wd[STEPS];
    scaled_fft(s.mode, vs, s.bem.lambda, ws);
    scaled_fft(s.mode, vd, s.bem.lambda, wd);
    float r = 0;
    for (int k = 0; k < STEPS; k++)
    {
        int slot = s.ring(t - k);
        r += -(float)ws[k] * s.hn[tri * RING + slot] + (float)wd[k] * s.hd[tri * RING + slot];
    }
    return r;
}

static float near_field_direct(const orc_scene &s, i3 src_center, i3 dst, int t)
{
    float v = 0;
    f3 pt = s.center_f(dst);
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++)
            {
                i3 src = src_center + i3{dx, dy, dz};
                if (src == dst)
                    continue;
                int cl = s.lin(src);
                for (int q = s.cell_start[cl]; q < s.cell_start[cl + 1]; q++)
                    v += laplace_point(s, s.cell_items[q], pt, t);
            }
    return v;
}

// ------------------------------------------------------------------ cell classification
// GhostCellSolver's cell data (ghost_cell_util.cu:7-230) and fresh cells (ghost_cell.cu:51-218,
// 547-560). Cell types follow the reference's enum (ghost_cell.h:11-17).
enum : int
{
    CELL_AIR = 0,
    CELL_SOLID = 1,
    CELL_GHOST = 2,
    CELL_UNKNOWN = 3
};

struct P2
{
    float s, t;
};

// Closest point of a triangle to a point, as barycentric weights of (v0, v1, v2): Eberly,
// "Distance from a Point to a Triangle" (Geometric Tools, 2020) — minimise
// |v0 + s e0 + t e1 - p|^2 over the triangle by following the zero set of the gradient's
// directional derivative. Same float operations as get_nearest_triangle_point_coeff
// (helper_math.h:1603-1800), on which the labels depend.
static inline P2 min_on_edge02(float a11, float b1)
{
    const int neg = b1 < 0, past = a11 + b1 <= 0;
    return {0.0f, (float)neg * ((float)past - (float)(1 - past) * (b1 / a11))};
}
static inline P2 min_on_edge12(float a01, float a11, float b1, float f10, float f01)
{
    const float g0 = a01 + b1 - f10, g1 = a11 + b1 - f01;
    const int neg = g0 < 0, past = g1 <= 0;
    const float t = (float)neg * ((float)past + (float)(1 - past) * g0 / (g0 - g1));
    return {1 - t, t};
}
static inline P2 min_inside(P2 q0, float g0, P2 q1, float g1)
{
    const float z = g0 / (g0 - g1), w = 1 - z;
    return {w * q0.s + z * q1.s, w * q0.t + z * q1.t};
}
static f3 closest_coeff(f3 p, f3 v0, f3 v1, f3 v2)
{
    const f3 d = p - v0, e0 = v1 - v0, e1 = v2 - v0;
    const float a00 = dot(e0, e0), a01 = dot(e0, e1), a11 = dot(e1, e1);
    const float b0 = -dot(d, e0), b1 = -dot(d, e1);
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

// angle-weighted vertex normals over the faces that use each vertex, in ascending face
// order (fill_in_nearest_kernel, ghost_cell_util.cu:56-80; the reference's list order is
// that of atomic appends)
static std::vector<f3> vertex_normals(const orc_scene &s)
{
    const int nv = (int)s.verts.size();
    std::vector<std::vector<int>> faces(nv);
    for (int i = 0; i < s.nt; i++)
        for (int k = 0; k < 3; k++)
            faces[s.tris[i].idx[k]].push_back(i);
    std::vector<f3> out(nv, f3{0, 0, 0});
    for (int v = 0; v < nv; v++)
    {
        f3 n{0, 0, 0};
        for (int f : faces[v])
        {
            const Tri &t = s.tris[f];
            int o[2], m = 0;
            for (int k = 0; k < 3; k++)
                if (t.idx[k] != v)
                    o[m++] = t.idx[k];
            const f3 e1 = s.verts[o[0]] - s.verts[v], e2 = s.verts[o[1]] - s.verts[v];
            const float ang = (float)std::acos((double)(dot(e1, e2) / (length(e1) * length(e2))));
            n = n + t.normal * ang;
        }
        out[v] = n * (1.0f / sqrtf(dot(n, n)));
    }
    return out;
}

// fill_cell_data (ghost_cell_util.cu:207-230): nearest triangle in the 7^3 window, side of
// the surface, ghost cells, column fill of unknown cells, ghost indices
static int classify(orc_scene &s, bool thin_shell)
{
    const int N = s.N, NN = N * N * N;
    if (s.classified)
        s.ctype_old = s.ctype;
    else
        s.ctype_old.assign(NN, CELL_AIR);
    s.ctype.assign(NN, CELL_UNKNOWN);
    s.cnearest.assign(NN, -1);
    s.cpoint.assign(NN, f3{0, 0, 0});
    s.creflect.assign(NN, f3{0, 0, 0});
    s.cdist.assign(NN, 3.402823466e+38F);
    const std::vector<f3> vn = vertex_normals(s);
#pragma omp parallel for schedule(dynamic, 1)
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++)
            for (int z = 0; z < N; z++)
            {
                const f3 ctr = s.center_f(i3{x, y, z});
                float best = 3.402823466e+38F;
                int bi = -1;
                f3 bp{0, 0, 0}, bc{0, 0, 0};
                for (int dx = -3; dx <= 3; dx++)
                    for (int dy = -3; dy <= 3; dy++)
                        for (int dz = -3; dz <= 3; dz++)
                        {
                            const int cx = x + dx, cy = y + dy, cz = z + dz;
                            if (cx < 0 || cy < 0 || cz < 0 || cx >= N || cy >= N || cz >= N)
                                continue;
                            const int cl = s.lin(cx, cy, cz);
                            for (int q = s.cell_start[cl]; q < s.cell_start[cl + 1]; q++)
                            {
                                const int f = s.cell_items[q];
                                const Tri &t = s.tris[f];
                                const f3 v0 = s.verts[t.idx[0]], v1 = s.verts[t.idx[1]], v2 = s.verts[t.idx[2]];
                                const f3 cf3 = closest_coeff(ctr, v0, v1, v2);
                                const f3 np = cf3.x * v0 + cf3.y * v1 + cf3.z * v2;
                                const float len = length(ctr - np);
                                if (len < best)
                                {
                                    best = len;
                                    bi = f;
                                    bp = np;
                                    bc = cf3;
                                }
                            }
                        }
                const int c = s.lin(x, y, z);
                if (bi < 0)
                    continue;  // UNKNOWN
                s.cdist[c] = best;
                s.cnearest[c] = bi;
                s.cpoint[c] = bp;
                s.creflect[c] = 2.0f * bp - ctr;
                const Tri &t = s.tris[bi];
                const f3 nrm = bc.x * vn[t.idx[0]] + bc.y * vn[t.idx[1]] + bc.z * vn[t.idx[2]];
                const float direction = dot(nrm, ctr - bp);
                const f3 dv = bp - ctr;
                const float hh = s.h / 2;
                const bool shell = thin_shell && std::fabs(dv.x) < hh && std::fabs(dv.y) < hh && std::fabs(dv.z) < hh;
                s.ctype[c] = (direction <= 0 || shell) ? CELL_SOLID : CELL_AIR;
            }
    // ghost cells: SOLID with an AIR 6-neighbour (cell_classfication_kernel)
    std::vector<int> typ(s.ctype), flag(NN, 0);
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++)
            for (int z = 0; z < N; z++)
            {
                const int c = s.lin(x, y, z);
                if (s.ctype[c] != CELL_SOLID)
                    continue;
                const int nb[6][3] = {{x + 1, y, z}, {x - 1, y, z}, {x, y + 1, z},
                                      {x, y - 1, z}, {x, y, z + 1}, {x, y, z - 1}};
                for (auto &q : nb)
                    if (q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < N && q[1] < N && q[2] < N &&
                        s.ctype[s.lin(q[0], q[1], q[2])] == CELL_AIR)
                        typ[c] = CELL_GHOST;
                flag[c] = typ[c] == CELL_GHOST;
            }
    // column fill along z (post_cell_classification_kernel)
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++)
        {
            int cur = CELL_AIR;
            bool found = false;
            for (int z = 0; z < N; z++)
            {
                const int c = s.lin(x, y, z);
                if (!found && typ[c] != CELL_UNKNOWN)
                {
                    found = true;
                    cur = typ[c];
                    for (int k = 0; k < z; k++)
                        typ[s.lin(x, y, k)] = cur;
                    continue;
                }
                if (typ[c] == CELL_SOLID || typ[c] == CELL_GHOST)
                    cur = CELL_SOLID;
                else if (typ[c] == CELL_AIR)
                    cur = CELL_AIR;
                else
                    typ[c] = cur;
            }
        }
    // ghost index = inclusive scan of the pre-fill ghost flags - 1 (apply_cell_type_kernel)
    s.cghost.assign(NN, -1);
    int run = 0;
    for (int c = 0; c < NN; c++)
    {
        run += flag[c];
        s.ctype[c] = typ[c];
        if (typ[c] == CELL_GHOST)
            s.cghost[c] = run - 1;
    }
    s.ghost_count = run;
    s.classified = true;
    return run;
}

// get_fresh_cell_list + solve_fresh_history (ghost_cell.cu:51-196): cells that turned AIR
// get values at t and t-1 from the AIR (old and new) cells around their mirror point.
// Returns the number of fresh cells, or -1 if one of them has no nearest triangle (the
// reference then writes the 1e10 sentinel into every fresh cell at t).
static int fill_fresh(orc_scene &s, const float *acc_now, const float *acc_old)
{
    const int N = s.N, NN = N * N * N;
    std::vector<int> fresh;
    for (int c = 0; c < NN; c++)
        if (s.ctype_old[c] != CELL_AIR && s.ctype[c] == CELL_AIR)
            fresh.push_back(c);
    bool err = false;
    for (int c : fresh)
        if (s.cnearest[c] < 0 || s.cnearest[c] >= s.nt)
            err = true;
    auto air2 = [&](i3 q) {
        if (q.x < 0 || q.y < 0 || q.z < 0 || q.x >= N || q.y >= N || q.z >= N)
            return false;
        const int l = s.lin(q);
        return s.ctype_old[l] == CELL_AIR && s.ctype[l] == CELL_AIR;
    };
    std::vector<float> &g0 = s.G(0), &g1 = s.G(-1);
    for (int c : fresh)
    {
        if (err)
        {
            g0[c] = 1e10f;
            continue;
        }
        const i3 co{c / (N * N), c / N % N, c % N};
        const int tri = s.cnearest[c];
        const float accs[2] = {acc_now[tri], acc_old[tri]};
        std::vector<float> *lv[2] = {&g0, &g1};
        const f3 xb = s.cpoint[c], xf = s.center_f(co);
        const f3 xr = xb + 2.0f * (xf - xb);
        const i3 base = base_coord(s, xr);
        i3 nb[8];
        float w[8];
        bool any = false;
        for (int dx = 0; dx < 2; dx++)
            for (int dy = 0; dy < 2; dy++)
                for (int dz = 0; dz < 2; dz++)
                {
                    const int k = dx * 4 + dy * 2 + dz;
                    nb[k] = base + i3{dx, dy, dz};
                    if (air2(nb[k]))
                    {
                        const float dist = length(s.center_f(nb[k]) - xr);
                        w[k] = (float)(1.0 / (double)(dist * dist + 1e-8f));
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
                        const i3 q = base + i3{dx, dy, dz};
                        if (!air2(q))
                            continue;
                        const float dist = length(s.center_f(q) - xr);
                        if (dist < best)
                        {
                            best = dist;
                            p1 = g0[s.lin(q)];
                            p2 = g1[s.lin(q)];
                        }
                    }
            g0[c] = p1;
            g1[c] = p2;
            continue;
        }
        float sum = 0;
        for (int k = 0; k < 8; k++)
            sum += w[k];
        for (int k = 0; k < 8; k++)
            w[k] /= sum;
        const float len = length(xf - xr);
        for (int i = 0; i < 2; i++)
        {
            float pr = 0;
            for (int k = 0; k < 8; k++)
            {
                // neighbours outside the grid carry weight 0 and are not read
                const i3 q = nb[k];
                const bool in = q.x >= 0 && q.y >= 0 && q.z >= 0 && q.x < N && q.y < N && q.z < N;
                pr += w[k] * (in ? (*lv[i])[s.lin(q)] : 0.0f);
            }
            (*lv[i])[c] = pr - 1.225f * accs[i] * len;
        }
    }
    return err ? -1 : (int)fresh.size();
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

orc_scene *orc_create(int res, float dl, float dt, const float *min_xyz, float c, float reflect_coeff,
                      int weight_mode, int pml_width)
{
    if (res < 6 || dl <= 0 || dt <= 0)
        return nullptr;
    orc_scene *s = new orc_scene();
    s->N = res;
    s->h = dl;
    s->dt = dt;
    s->c = c;
    s->rc = reflect_coeff;
    s->minp = min_xyz ? f3{min_xyz[0], min_xyz[1], min_xyz[2]} : f3{0, 0, 0};
    s->mode = weight_mode;
    s->pml = pml_width;
    s->damp = 2.0f / dl;
    size_t NN = (size_t)res * res * res;
    for (int i = 0; i < 3; i++)
    {
        s->grid[i].assign(NN, 0.0f);
        if (pml_width > 0)
            s->pmlv[i].assign(NN * 6, 0.0f);
    }
    s->far.assign(NN, 0.0f);
    s->bem = tdbem_init(weight_mode, dt, c, dl);
    return s;
}

void orc_destroy(orc_scene *s) { delete s; }

int orc_set_mesh(orc_scene *s, const float *verts, int nv, const int *tris, int nt)
{
    if (!s || nv <= 0 || nt <= 0)
        return -1;
    s->verts.resize(nv);
    for (int i = 0; i < nv; i++)
        s->verts[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    s->nt = nt;
    s->tris.assign(nt, Tri{});
    for (int i = 0; i < nt; i++)
        for (int q = 0; q < 3; q++)
        {
            int v = tris[3 * i + q];
            if (v < 0 || v >= nv)
                return -2;
            s->tris[i].idx[q] = v;
        }
    int rc = update_triangles(*s);
    if (rc)
        return rc;
    build_lists(*s);
    s->gcache.assign(nt, GridEntry{});
    s->fcache.clear();
    update_caches(*s, true);
    s->hn.assign((size_t)nt * RING, 0.0f);
    s->hd.assign((size_t)nt * RING, 0.0f);
    s->mesh_set = true;
    return 0;
}

int orc_update_mesh(orc_scene *s, const float *verts)
{
    if (!s || !s->mesh_set)
        return -1;
    for (size_t i = 0; i < s->verts.size(); i++)
        s->verts[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    int rc = update_triangles(*s);
    if (rc)
        return rc;
    build_lists(*s);
    update_caches(*s, false);
    return 0;
}

int orc_fdtd_step(orc_scene *s)
{
    fdtd_step(*s);
    return 0;
}
int orc_solve_far(orc_scene *s)
{
    if (!s->mesh_set)
        return -1;
    phase_far(*s);
    return 0;
}
int orc_set_neumann(orc_scene *s, const float *nm)
{
    if (!s->mesh_set)
        return -1;
    phase_neumann(*s, nm);
    return 0;
}
int orc_update_dirichlet(orc_scene *s)
{
    if (!s->mesh_set)
        return -1;
    phase_dirichlet(*s);
    return 0;
}
int orc_solve_near(orc_scene *s)
{
    if (!s->mesh_set)
        return -1;
    phase_near(*s);
    return 0;
}

int orc_step(orc_scene *s, const float *neumann)
{
    if (!s || !s->mesh_set)
        return -1;
    fdtd_step(*s);
    phase_far(*s);
    phase_neumann(*s, neumann);
    phase_dirichlet(*s);
    phase_near(*s);
    return 0;
}

int orc_solve_far_direct(orc_scene *s)
{
    // solve_fdtd_far_simple = fdtd.step + direct_fdtd_far (pppm.cu:227-233, pppm_direct.cu:27-48)
    if (!s->mesh_set)
        return -1;
    fdtd_step(*s);
    int N = s->N, t = s->t;
    float c = s->c, dt = s->dt;
    const std::vector<float> &g = s->G(0);
    s->far = g;
    for (int x = 1; x < N - 1; x++)
        for (int y = 1; y < N - 1; y++)
            for (int z = 1; z < N - 1; z++)
            {
                i3 co = {x, y, z};
                float f = 0;
                f += 2 * near_field_direct(*s, co, co, t - 1);
                float lap = 0;
                lap += near_field_direct(*s, co, co + i3{-1, 0, 0}, t - 1);
                lap += near_field_direct(*s, co, co + i3{1, 0, 0}, t - 1);
                lap += near_field_direct(*s, co, co + i3{0, -1, 0}, t - 1);
                lap += near_field_direct(*s, co, co + i3{0, 1, 0}, t - 1);
                lap += near_field_direct(*s, co, co + i3{0, 0, -1}, t - 1);
                lap += near_field_direct(*s, co, co + i3{0, 0, 1}, t - 1);
                lap -= 6 * near_field_direct(*s, co, co, t - 1);
                lap = lap / (s->h * s->h);
                f += (c * c * dt * dt) * lap;
                f -= near_field_direct(*s, co, co, t - 2);
                s->far[s->lin(co)] = g[s->lin(co)] - f;
            }
    return 0;
}

int orc_solve_near_direct(orc_scene *s)
{
    // correction_fdtd_near_field_kernel, pppm_direct.cu:6-18 (no 0.98 factor)
    if (!s->mesh_set)
        return -1;
    int N = s->N, t = s->t;
    std::vector<float> &g = s->G(0);
    for (int x = 1; x < N - 1; x++)
        for (int y = 1; y < N - 1; y++)
            for (int z = 1; z < N - 1; z++)
            {
                i3 co = {x, y, z};
                g[s->lin(co)] = s->far[s->lin(co)] + near_field_direct(*s, co, co, t);
            }
    return 0;
}

int orc_time_idx(const orc_scene *s) { return s->t; }
void orc_set_time_idx(orc_scene *s, int t) { s->t = t; }

int orc_get_grid(const orc_scene *s, int rel, float *out)
{
    const std::vector<float> &g = s->G(rel);
    std::memcpy(out, g.data(), g.size() * sizeof(float));
    return 0;
}
int orc_set_grid(orc_scene *s, int rel, const float *in)
{
    std::vector<float> &g = s->G(rel);
    std::memcpy(g.data(), in, g.size() * sizeof(float));
    return 0;
}
int orc_get_far(const orc_scene *s, float *out)
{
    std::memcpy(out, s->far.data(), s->far.size() * sizeof(float));
    return 0;
}
float orc_read_cell(const orc_scene *s, int x, int y, int z) { return s->G(0)[s->lin(x, y, z)]; }

int orc_get_history(const orc_scene *s, float *neumann, float *dirichlet)
{
    if (neumann)
        std::memcpy(neumann, s->hn.data(), s->hn.size() * sizeof(float));
    if (dirichlet)
        std::memcpy(dirichlet, s->hd.data(), s->hd.size() * sizeof(float));
    return 0;
}
int orc_set_history(orc_scene *s, const float *neumann, const float *dirichlet)
{
    if (neumann)
        std::memcpy(s->hn.data(), neumann, s->hn.size() * sizeof(float));
    if (dirichlet)
        std::memcpy(s->hd.data(), dirichlet, s->hd.size() * sizeof(float));
    return 0;
}
void orc_reset(orc_scene *s)
{
    for (int i = 0; i < 3; i++)
    {
        std::fill(s->grid[i].begin(), s->grid[i].end(), 0.0f);
        std::fill(s->pmlv[i].begin(), s->pmlv[i].end(), 0.0f);
    }
    s->t = -1;
    std::fill(s->hn.begin(), s->hn.end(), 0.0f);
    std::fill(s->hd.begin(), s->hd.end(), 0.0f);
}

int orc_stats(const orc_scene *s, long long *out)
{
    out[0] = s->nt;
    out[1] = (long long)s->c3_cells.size();
    out[2] = (long long)s->c3_pairs.size();
    out[3] = (long long)s->b4_cells.size();
    long long p4 = 0;
    for (int i = 0; i < s->nt; i++)
    {
        int b = s->b4_of_tri[i];
        p4 += s->b4_start[b + 1] - s->b4_start[b];
    }
    out[4] = p4;
    out[5] = s->grid_recomputes;
    out[6] = s->face_recomputes;
    return 0;
}

int orc_get_triangles(const orc_scene *s, float *o)
{
    for (int i = 0; i < s->nt; i++)
    {
        const Tri &t = s->tris[i];
        float v[13] = {t.center.x, t.center.y, t.center.z, t.normal.x, t.normal.y, t.normal.z, t.area,
                       (float)t.cell.x, (float)t.cell.y, (float)t.cell.z, (float)t.base.x, (float)t.base.y,
                       (float)t.base.z};
        std::memcpy(o + 13 * i, v, sizeof(v));
    }
    return 0;
}

int orc_get_c3(const orc_scene *s, int *cells, int *start, int *pairs)
{
    if (cells)
        std::memcpy(cells, s->c3_cells.data(), s->c3_cells.size() * sizeof(int));
    if (start)
        std::memcpy(start, s->c3_start.data(), s->c3_start.size() * sizeof(int));
    if (pairs)
        std::memcpy(pairs, s->c3_pairs.data(), s->c3_pairs.size() * sizeof(int));
    return 0;
}

int orc_get_b4(const orc_scene *s, int *cells, int *start, int *tris)
{
    if (cells)
        std::memcpy(cells, s->b4_cells.data(), s->b4_cells.size() * sizeof(int));
    if (start)
        std::memcpy(start, s->b4_start.data(), s->b4_start.size() * sizeof(int));
    if (tris)
        std::memcpy(tris, s->b4_tris.data(), s->b4_tris.size() * sizeof(int));
    return 0;
}

int orc_export_grid_weights(const orc_scene *s, void *recs)
{
    LWH *o = (LWH *)recs;
    for (int i = 0; i < s->nt; i++)
        for (int q = 0; q < GRID_CACHE; q++)
        {
            o[(i * GRID_CACHE + q) * 2 + 0] = s->gcache[i].fdtd[q];
            o[(i * GRID_CACHE + q) * 2 + 1] = s->gcache[i].bem[q];
        }
    return 0;
}

int orc_import_grid_weights(orc_scene *s, const void *recs)
{
    const LWH *o = (const LWH *)recs;
    for (int i = 0; i < s->nt; i++)
        for (int q = 0; q < GRID_CACHE; q++)
        {
            s->gcache[i].fdtd[q] = o[(i * GRID_CACHE + q) * 2 + 0];
            s->gcache[i].bem[q] = o[(i * GRID_CACHE + q) * 2 + 1];
        }
    return 0;
}

int orc_export_face_weights(const orc_scene *s, int *start, int *nbr, void *recs)
{
    LWH *o = (LWH *)recs;
    long long p = 0;
    for (int ci = 0; ci < s->nt; ci++)
    {
        if (start)
            start[ci] = (int)p;
        int b = s->b4_of_tri[ci];
        for (int q = s->b4_start[b]; q < s->b4_start[b + 1]; q++, p++)
        {
            int ji = s->b4_tris[q];
            if (nbr)
                nbr[p] = ji;
            if (o)
                o[p] = s->fcache.at((long long)ci * s->nt + ji).w;
        }
    }
    if (start)
        start[s->nt] = (int)p;
    return 0;
}

int orc_import_face_weights(orc_scene *s, const void *recs)
{
    const LWH *o = (const LWH *)recs;
    long long p = 0;
    for (int ci = 0; ci < s->nt; ci++)
    {
        int b = s->b4_of_tri[ci];
        float factor_sum = 0;
        for (int q = s->b4_start[b]; q < s->b4_start[b + 1]; q++, p++)
        {
            int ji = s->b4_tris[q];
            FaceEntry &e = s->fcache[(long long)ci * s->nt + ji];
            e.w = o[p];
            factor_sum += h2f(e.w.d[0]) * e.w.sd;
        }
        s->denom[ci] = 0.5 * (double)s->tris[ci].area - (double)factor_sum;
    }
    return 0;
}

int orc_get_interp_weights(const orc_scene *s, float *out8)
{
    std::memcpy(out8, s->interp.data(), s->interp.size() * sizeof(float));
    return 0;
}

int orc_classify(orc_scene *s, int thin_shell)
{
    if (!s->mesh_set)
        return -1;
    return classify(*s, thin_shell != 0);
}

int orc_get_cells(const orc_scene *s, int *type, int *ghost_idx, int *nearest, float *point3, float *reflect3,
                  float *dist)
{
    if (!s->classified)
        return -1;
    const size_t NN = s->ctype.size();
    if (type)
        std::memcpy(type, s->ctype.data(), NN * sizeof(int));
    if (ghost_idx)
        std::memcpy(ghost_idx, s->cghost.data(), NN * sizeof(int));
    if (nearest)
        std::memcpy(nearest, s->cnearest.data(), NN * sizeof(int));
    if (point3)
        std::memcpy(point3, s->cpoint.data(), NN * sizeof(f3));
    if (reflect3)
        std::memcpy(reflect3, s->creflect.data(), NN * sizeof(f3));
    if (dist)
        std::memcpy(dist, s->cdist.data(), NN * sizeof(float));
    return 0;
}

// The scene's step as flat arrays (weights decoded as the step reads them: (float)h * scale)
bool orc_linear_model(const orc_scene *s, LinearModel &m)
{
    if (!s->mesh_set || s->pml > 0)
        return false;
    const int N = s->N, nt = s->nt;
    m.N = N;
    m.nt = nt;
    m.rc = s->rc;
    const double c = s->c, dt = s->dt, h = s->h;
    m.coef = c * c * dt * dt;
    m.hh = h * h;
    const float fc = s->c, fdt = s->dt, fh = s->h;
    m.acoef = fc * fc * fdt * fdt / (fh * fh);
    m.coef1 = 1 / (fdt * fdt);
    m.coef2 = fc / (2 * fh * fdt);
    m.coef3 = -(fc * fc) / (2 * fh * fh);
    m.c3_cells = s->c3_cells;
    m.c3_start = s->c3_start;
    m.c3_pairs = s->c3_pairs;
    const size_t R = (size_t)nt * 27;
    m.wfS.resize(R * STEPS);
    m.wfD.resize(R * STEPS);
    m.wbS.resize(R * STEPS);
    m.wbD.resize(R * STEPS);
    for (int i = 0; i < nt; i++)
        for (int o = 0; o < 27; o++)
        {
            const LWH &f = s->gcache[i].fdtd[o], &b = s->gcache[i].bem[o];
            const size_t r = ((size_t)i * 27 + o) * STEPS;
            for (int k = 0; k < STEPS; k++)
            {
                m.wfS[r + k] = h2f(f.s[k]) * f.ss;
                m.wfD[r + k] = h2f(f.d[k]) * f.sd;
                m.wbS[r + k] = h2f(b.s[k]) * b.ss;
                m.wbD[r + k] = h2f(b.d[k]) * b.sd;
            }
        }
    std::vector<int> c3_of(s->grid[0].size(), -1);
    for (size_t q = 0; q < s->c3_cells.size(); q++)
        c3_of[s->c3_cells[q]] = (int)q;
    m.f_start.assign(nt + 1, 0);
    m.f_nbr.clear();
    m.fS.clear();
    m.fD.clear();
    m.den.assign(nt, 0.0);
    m.corner.assign((size_t)nt * 8, -1);
    m.interp.assign(s->interp.begin(), s->interp.end());
    for (int ci = 0; ci < nt; ci++)
    {
        const int b = s->b4_of_tri[ci];
        double fac = 0;
        for (int q = s->b4_start[b]; q < s->b4_start[b + 1]; q++)
        {
            const int ji = s->b4_tris[q];
            const LWH &w = s->fcache.at((long long)ci * nt + ji).w;
            m.f_nbr.push_back(ji);
            for (int k = 0; k < STEPS; k++)
            {
                m.fS.push_back(h2f(w.s[k]) * w.ss);
                m.fD.push_back(h2f(w.d[k]) * w.sd);
            }
            fac += (double)(h2f(w.d[0]) * w.sd);
        }
        m.f_start[ci + 1] = (int)m.f_nbr.size();
        m.den[ci] = 0.5 * (double)s->tris[ci].area - fac;
        const i3 bc = s->tris[ci].base;
        for (int w = 0; w < 8; w++)
            m.corner[ci * 8 + w] = c3_of[s->lin(bc.x + (w >> 2), bc.y + ((w >> 1) & 1), bc.z + (w & 1))];
    }
    return true;
}

int orc_set_cells(orc_scene *s, const int *type, const int *type_old, const int *nearest, const float *point3)
{
    // test hook: cell data as classify / the previous classify would leave it, so that the
    // fresh-cell fill can be checked on a hand-built stencil
    if (!s->mesh_set)
        return -1;
    const size_t NN = (size_t)s->N * s->N * s->N;
    s->ctype.assign(type, type + NN);
    s->ctype_old.assign(type_old, type_old + NN);
    s->cnearest.assign(nearest, nearest + NN);
    s->cpoint.resize(NN);
    std::memcpy(s->cpoint.data(), point3, NN * sizeof(f3));
    s->cghost.assign(NN, -1);
    s->creflect.assign(NN, f3{0, 0, 0});
    s->cdist.assign(NN, 0.0f);
    s->classified = true;
    return 0;
}

int orc_fill_fresh(orc_scene *s, const float *acc_now, const float *acc_old)
{
    if (!s->classified)
        return -2;
    return fill_fresh(*s, acc_now, acc_old);
}

int orc_set_solid_zero(orc_scene *s)
{
    // set_solid_cell_zero_kernel, ghost_cell.cu:547-555
    if (!s->classified)
        return -1;
    std::vector<float> &g = s->G(0);
    for (size_t c = 0; c < g.size(); c++)
        if (s->ctype[c] == CELL_SOLID)
            g[c] = 0;
    return 0;
}

int orc_vertex_normals(const orc_scene *s, float *out3)
{
    if (!s->mesh_set)
        return -1;
    const std::vector<f3> vn = vertex_normals(*s);
    std::memcpy(out3, vn.data(), vn.size() * sizeof(f3));
    return 0;
}

double orc_det_exp(double x) { return orc_detmath::det_exp(x); }
void orc_det_sincos(double x, double *s, double *c) { orc_detmath::det_sincos(x, s, c); }

void orc_closest_coeff(const float *p3, const float *tri9, float *out3)
{
    const f3 c = closest_coeff(f3{p3[0], p3[1], p3[2]}, f3{tri9[0], tri9[1], tri9[2]}, f3{tri9[3], tri9[4], tri9[5]},
                               f3{tri9[6], tri9[7], tri9[8]});
    out3[0] = c.x;
    out3[1] = c.y;
    out3[2] = c.z;
}

unsigned short orc_f2h(float f) { return f2h(f); }
float orc_h2f(unsigned short h) { return h2f(h); }

void orc_tdbem_init(int mode, float dt, float c, double *lambda, double *wave)
{
    Tdbem b = tdbem_init(mode, dt, c, 1.0f);
    *lambda = b.lambda;
    for (int k = 0; k < STEPS; k++)
    {
        wave[2 * k] = b.wave[k].re;
        wave[2 * k + 1] = b.wave[k].im;
    }
}

void orc_scaled_fft(int mode, const double *in, double lambda, double *out)
{
    cd v[STEPS];
    for (int k = 0; k < STEPS; k++)
        v[k] = {in[2 * k], in[2 * k + 1]};
    scaled_fft(mode, v, lambda, out);
}

void orc_scaled_idft(const double *in, double lambda, double *out)
{
    // TDBEM::scaledIDFT, bem.h:279-290, evaluated in FP64
    for (int k = 0; k < STEPS; k++)
    {
        double rr = 0;
        for (int i = 0; i < STEPS; i++)
        {
            double a = 2 * PI_R / STEPS * k * i;
            rr += in[2 * i] * std::cos(a) - in[2 * i + 1] * std::sin(a);
        }
        out[k] = rr / STEPS * std::pow(lambda, -k);
    }
}

void orc_face2point(int mode, const float *tri9, const float *pt3, double kre, double kim, int type, double *out)
{
    f3 v[3] = {{tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]}, {tri9[6], tri9[7], tri9[8]}};
    f3 p = {pt3[0], pt3[1], pt3[2]};
    if (mode == ORC_REF32)
    {
        float one = 1.0f;
        cf r = f2p32(v, &p, &one, 1, cf{(float)kre, (float)kim}, type);
        out[0] = r.re;
        out[1] = r.im;
    }
    else
    {
        TriD T = tri_d(v[0], v[1], v[2]);
        d3 pd = D(p);
        double one = 1.0;
        cd r = f2p64(T, &pd, &one, 1, cd{kre, kim}, type);
        out[0] = r.re;
        out[1] = r.im;
    }
}

void orc_face2face(int mode, const float *verts, const int *src3, const int *trg3, double kre, double kim, int type,
                   float threshold, double *out)
{
    int mx = std::max({src3[0], src3[1], src3[2], trg3[0], trg3[1], trg3[2]});
    std::vector<f3> V(mx + 1);
    for (int i = 0; i <= mx; i++)
        V[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    if (mode == ORC_REF32)
    {
        cf r = f2f32(V.data(), src3, trg3, cf{(float)kre, (float)kim}, type, threshold);
        out[0] = r.re;
        out[1] = r.im;
    }
    else
    {
        cd r = f2f64(V.data(), src3, trg3, cd{kre, kim}, type, threshold);
        out[0] = r.re;
        out[1] = r.im;
    }
}

void orc_laplace_weight_point(int mode, float dt, float c, float grid_size, const float *tri9, const float *pt3,
                              int type, double *out16)
{
    Tdbem b = tdbem_init(mode, dt, c, grid_size);
    cd v[STEPS];
    for (int k = 0; k <= STEPS / 2; k++)
    {
        double o[2];
        orc_face2point(mode, tri9, pt3, b.wave[k].re, b.wave[k].im, type, o);
        v[k] = {o[0], o[1]};
    }
    mirror(v);
    scaled_fft(mode, v, b.lambda, out16);
}

void orc_singular_fp64(const float *verts, const int *src3, const int *trg3, double kre, double kim, int type,
                       int gauss_n, double *out)
{
    int mx = std::max({src3[0], src3[1], src3[2], trg3[0], trg3[1], trg3[2]});
    std::vector<f3> V(mx + 1);
    for (int i = 0; i <= mx; i++)
        V[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    f3 sv[3] = {V[src3[0]], V[src3[1]], V[src3[2]]};
    f3 tv[3] = {V[trg3[0]], V[trg3[1]], V[trg3[2]]};
    int nn = common_vertices(src3, trg3);
    TriD S = tri_d(sv[0], sv[1], sv[2]);
    TriD T = tri_d(tv[0], tv[1], tv[2]);
    std::vector<double> gx, gw;
    gauss_legendre(gauss_n, gx, gw);
    cd r = sing64(S, T, cd{kre, kim}, nn, type, gx.data(), gw.data(), gauss_n);
    out[0] = r.re;
    out[1] = r.im;
}

}  // extern "C"
