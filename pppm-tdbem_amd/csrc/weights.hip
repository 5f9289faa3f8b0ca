// FP64 construction of the near-field TDBEM weights on the GPU.
//
// Reference behaviour (computed there in FP32 complex arithmetic):
//   TDBEM::init / scaledFFT / laplace_weight     src/algorithm/bem.h:264-351
//   Helmholtz layer potentials                    src/helmholtz/potential.h:22-101
//   face2Point / face2Face integrands (1-point,   src/helmholtz/integrand.h:52-267
//     regular 1x1, Duffy-singular 3^4 Gauss)
//   GridCache  (27 offsets x {fdtd, bem})         src/algorithm/pppm_cache.cu:53-134
//   trilinear interpolation weights               src/algorithm/pppm_cache.cu:139-181
//   FaceCache  (face pair weights + dirty test)   src/algorithm/pppm_cache.cu:183-289
//   LayerWeightHalf::set (FP16 + FP32 scale)      src/algorithm/bem.h:171-191
//
// Here every integral, the wave numbers, lambda and the inverse DFT are FP64; the
// weight series is rounded to FP32 and quantised to FP16 with a max-abs FP32 scale,
// exactly as the reference stores it. Storage: per record 16 x half2 {S_k, D_k}
// (64 B, one half cache line) plus a float2 {scale_S, scale_D}.
#include "detmath.hpp"
#include "internal.hpp"
#include "tri_device.hpp"

#include <algorithm>
#include <cmath>

namespace pppm_amd
{

namespace
{
constexpr int B = 128;

__constant__ double c_gx[3] = {0.774596669241483, 0.000000000000000, -0.774596669241483};  // gauss/line.h:26-37
__constant__ double c_gw[3] = {0.555555555555554, 0.888888888888889, 0.555555555555554};

struct cdbl
{
    double re, im;
};

__device__ inline cdbl cexp_d(double re, double im)
{
    const double e = det_exp(re);
    double sn, cs;
    det_sincos(im, &sn, &cs);
    return {e * cs, e * sn};
}

// single / double layer kernels, potential.h:22-47
__device__ inline cdbl layer64(d3 src, d3 trg, d3 n, double kr, double ki, int type)
{
    if (type == 0)
    {
        d3 s = trg - src;
        double r2 = dotd(s, s);
        if (r2 == 0)
            return {0, 0};
        double r = sqrt(r2);
        cdbl e = cexp_d(-r * ki, r * kr);
        double q = 4 * PI_R * r;
        return {e.re / q, e.im / q};
    }
    d3 s = src - trg;
    double r2 = dotd(s, s);
    if (r2 == 0)
        return {0, 0};
    double r = sqrt(r2);
    double zr = -r * ki, zi = r * kr;
    cdbl e = cexp_d(zr, zi);
    double q = 4 * PI_R * r2 * r;
    double ar = -e.re / q, ai = -e.im / q;
    double br = 1 - zr, bi = -zi;
    double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
    double dn = dotd(s, n);
    return {pr * dn, pi * dn};
}


// singular_point (integrand.h:52-147); trial = tv, test = sv
__device__ inline cdbl sing_point64(double xsi, double e1, double e2, double e3, double w, const d3 *tv,
                                    const d3 *sv, d3 tn, int nn, double kr, double ki, int type)
{
    double rr = 0, ri = 0;
    xsi = 0.5 * (xsi + 1);
    e1 = 0.5 * (e1 + 1);
    e2 = 0.5 * (e2 + 1);
    e3 = 0.5 * (e3 + 1);
    double jw = 0;
#define ADD(A, Bp, F)                                      \
    {                                                      \
        cdbl g_ = layer64((A), (Bp), tn, kr, ki, type);    \
        rr += g_.re * (F);                                 \
        ri += g_.im * (F);                                 \
    }
    if (nn == 3)
    {
        jw = xsi * xsi * xsi * e1 * e1 * e2;
        double e12 = e1 * e2, e123 = e1 * e2 * e3;
        ADD(l2g2(xsi, xsi * (1.0 - e1 + e12), tv), l2g2(xsi * (1.0 - e123), xsi * (1.0 - e1), sv), 1.0);
        ADD(l2g2(xsi * (1.0 - e123), xsi * (1.0 - e1), tv), l2g2(xsi, xsi * (1.0 - e1 + e12), sv), 1.0);
        ADD(l2g2(xsi, xsi * (e1 - e12 + e123), tv), l2g2(xsi * (1.0 - e12), xsi * (e1 - e12), sv), 1.0);
        ADD(l2g2(xsi * (1.0 - e12), xsi * (e1 - e12), tv), l2g2(xsi, xsi * (e1 - e12 + e123), sv), 1.0);
        ADD(l2g2(xsi * (1.0 - e123), xsi * (e1 - e123), tv), l2g2(xsi, xsi * (e1 - e12), sv), 1.0);
        ADD(l2g2(xsi, xsi * (e1 - e12), tv), l2g2(xsi * (1.0 - e123), xsi * (e1 - e123), sv), 1.0);
    }
    else if (nn == 2)
    {
        jw = xsi * xsi * xsi * e1 * e1;
        double e12 = e1 * e2, e123 = e1 * e2 * e3;
        ADD(l2g2(xsi, xsi * e1 * e3, tv), l2g2(xsi * (1.0 - e12), xsi * e1 * (1.0 - e2), sv), 1.0);
        ADD(l2g2(xsi, xsi * e1, tv), l2g2(xsi * (1.0 - e123), xsi * e1 * e2 * (1 - e3), sv), e2);
        ADD(l2g2(xsi * (1.0 - e12), xsi * e1 * (1.0 - e2), tv), l2g2(xsi, xsi * e123, sv), e2);
        ADD(l2g2(xsi * (1.0 - e123), xsi * e12 * (1.0 - e3), tv), l2g2(xsi, xsi * e1, sv), e2);
        ADD(l2g2(xsi * (1.0 - e123), xsi * e1 * (1.0 - e2 * e3), tv), l2g2(xsi, xsi * e12, sv), e2);
    }
    else if (nn == 1)
    {
        jw = xsi * xsi * xsi * e2;
        ADD(l2g2(xsi, xsi * e1, tv), l2g2(xsi * e2, xsi * e2 * e3, sv), 1.0);
        ADD(l2g2(xsi * e2, xsi * e2 * e3, tv), l2g2(xsi, xsi * e1, sv), 1.0);
    }
    else
        return {0, 0};
#undef ADD
    double f = jw * w;
    return {rr * f, ri * f};
}

__device__ inline int common_vertices(int a0, int a1, int a2, int b0, int b1, int b2)
{
    return (a0 == b0) + (a0 == b1) + (a0 == b2) + (a1 == b0) + (a1 == b1) + (a1 == b2) + (a2 == b0) + (a2 == b1) +
           (a2 == b2);
}

// face2FaceIntegrand (integrand.h:228-259); the near-coincidence test stays in FP32. nn =
// common vertices; SING selects the code path compiled: the singular rules (nn > 0) or the
// regular 1-point rule (nn == 0), so the regular kernel carries none of the singular code's
// registers.
template <bool SING>
__device__ inline cdbl f2f64(const float *verts, const int *src, const int *trg, int nn, double kr, double ki,
                             int type, float threshold)
{
    const f3 s0 = ld3(verts, src[0]), s1 = ld3(verts, src[1]), s2 = ld3(verts, src[2]);
    const f3 t0 = ld3(verts, trg[0]), t1 = ld3(verts, trg[1]), t2 = ld3(verts, trg[2]);
    const float dist = lengthf((s0 + s1 + s2) / 3.0f - (t0 + t1 + t2) / 3.0f);
    if (dist < threshold && nn == 0)
        return {0, 0};
    const TriD S = tri_d(verts, src[0], src[1], src[2]);
    const TriD T = tri_d(verts, trg[0], trg[1], trg[2]);
    if constexpr (!SING)
    {
        d3 a = l2g(TRI1_X, TRI1_X, S.v), b = l2g(TRI1_X, TRI1_X, T.v);
        cdbl g = layer64(a, b, S.n, kr, ki, type);
        double f = 0.25 * TRI1_W * TRI1_W * S.jac * T.jac;
        return {f * g.re, f * g.im};
    }
    else
    {
        double rr = 0, ri = 0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++)
                    {
                        cdbl p = sing_point64(c_gx[a], c_gx[b], c_gx[c], c_gx[d],
                                              c_gw[a] * c_gw[b] * c_gw[c] * c_gw[d], S.v, T.v, S.n, nn, kr, ki, type);
                        rr += p.re;
                        ri += p.im;
                    }
        double f = S.jac * T.jac / 16;
        return {rr * f, ri * f};
    }
}

// Inverse transform with the reference's radix-2 decimation-in-time butterflies
// (FFT<step>, bem.h:12-38), written in place on bit-reversed input. Each butterfly
// a +- w*b uses the same operands, twiddle and operation order as the recursive form,
// so results are bit-identical to it. The order matters: with PI = 3.14159265359 the
// twiddle angles carry ~1e-13 rad errors and lambda^-k (up to ~5.6e3 at k = 15)
// amplifies them to FP16-rounding scale, so another DFT factorisation would flip stored
// FP16 codes. Twiddles come from a host-computed table.
__device__ inline int bitrev4(int j) { return ((j & 1) << 3) | ((j & 2) << 1) | ((j & 4) >> 1) | ((j & 8) >> 3); }

// scaledFFT (bem.h:292-304) of a conjugate-symmetric series given by its first 9 entries
__device__ inline void scaled_idft(const cdbl *v9, const Tdbem &b, double *out)
{
    cdbl x[STEPS];
#pragma unroll
    for (int j = 0; j < STEPS; j++)
    {
        const int k = bitrev4(j);
        x[j] = k <= 8 ? v9[k] : cdbl{v9[16 - k].re, -v9[16 - k].im};
    }
#pragma unroll
    for (int L = 2; L <= STEPS; L <<= 1)
    {
        const int half = L >> 1, tws = STEPS / L;
#pragma unroll
        for (int g = 0; g < STEPS; g += L)
#pragma unroll
            for (int m = 0; m < half; m++)
            {
                const double cr = b.tw_re[m * tws], ci = b.tw_im[m * tws];
                const cdbl y = x[g + m + half];
                const double tr = cr * y.re - ci * y.im;
                const double ti = cr * y.im + ci * y.re;
                const cdbl a = x[g + m];
                x[g + m] = {a.re + tr, a.im + ti};
                x[g + m + half] = {a.re - tr, a.im - ti};
            }
    }
#pragma unroll
    for (int n = 0; n < STEPS; n++)
        out[n] = x[n].re / STEPS * b.inv_pow[n];
}

// LayerWeightHalf::set (bem.h:171-191) into the half2 {S, D} record layout
__device__ inline void quantize_store(const float *s, const float *d, uint32_t *w16, float2 *sc)
{
    float ms = 0, md = 0;
    for (int i = 0; i < STEPS; i++)
    {
        ms = fmaxf(ms, fabsf(s[i]));
        md = fmaxf(md, fabsf(d[i]));
    }
    if (ms == 0)
        ms = 1;
    if (md == 0)
        md = 1;
    for (int i = 0; i < STEPS; i++)
    {
        uint32_t hs = f32_to_f16_bits(s[i] / ms);
        uint32_t hd = f32_to_f16_bits(d[i] / md);
        w16[i] = hs | (hd << 16);
    }
    *sc = make_float2(ms, md);
}



// A device-side count, or 0 when the mesh update has flagged an error (an invalid vertex
// index, or a list that ran out of capacity and was left incomplete): downstream kernels
// then do no work instead of reading unwritten entries.
__device__ inline int live_count(const int *n_ptr, const int *err) { return *err != 0 ? 0 : *n_ptr; }

// ---------------------------------------------------------------- grid cache
// Both layer kernels of one (source point, target) pair: they share r and e^{ikr}, so the
// values are those of two layer64 calls. Split in two so that the target's part is computed once
// for the 9 wave numbers.
// orders one wave's LDS writes before its lanes' reads of them
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The wave-number-free part, per (source point, target): r = |trg - src| (0: the target is the
// source point, both kernels 0), the denominators 4 pi r and 4 pi r^2 r, and (src - trg) . n.
struct LayerTarget
{
    double r, q0, q1, dn;
};
__device__ inline LayerTarget layer_target(d3 src, d3 trg, d3 n)
{
    const d3 s0 = trg - src;
    const double r2 = dotd(s0, s0);
    if (r2 == 0)
        return {0, 0, 0, 0};
    const double r = sqrt(r2);
    const d3 s1 = src - trg;  // = -s0 exactly, so dotd(s1, s1) == r2
    return {r, 4 * PI_R * r, 4 * PI_R * r2 * r, dotd(s1, n)};
}
__device__ inline void layer64_pair(const LayerTarget &t, double kr, double ki, cdbl &g0, cdbl &g1)
{
    if (t.r == 0)
    {
        g0 = g1 = {0, 0};
        return;
    }
    const double zr = -t.r * ki, zi = t.r * kr;
    const cdbl e = cexp_d(zr, zi);
    g0 = {e.re / t.q0, e.im / t.q0};
    const double ar = -e.re / t.q1, ai = -e.im / t.q1;
    const double br = 1 - zr, bi = -zi;
    const double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
    g1 = {pr * t.dn, pi * t.dn};
}

// The target cells of one triangle's 27 stencil batches, as offsets from its cell: the
// 3x3x3 block plus one more step along a single axis (81 cells); idx maps an offset in
// [-2, 2]^3 to its slot.
struct Targets81
{
    signed char off[81][3];
    signed char idx[125];
};
constexpr Targets81 make_targets81()
{
    Targets81 t{};
    for (int i = 0; i < 125; i++)
        t.idx[i] = -1;
    int n = 0;
    for (int x = -2; x <= 2; x++)
        for (int y = -2; y <= 2; y++)
            for (int z = -2; z <= 2; z++)
            {
                const int ax = x < 0 ? -x : x, ay = y < 0 ? -y : y, az = z < 0 ? -z : z;
                const int twos = (ax == 2) + (ay == 2) + (az == 2);
                if (twos > 1)
                    continue;
                t.off[n][0] = (signed char)x;
                t.off[n][1] = (signed char)y;
                t.off[n][2] = (signed char)z;
                t.idx[(x + 2) * 25 + (y + 2) * 5 + (z + 2)] = (signed char)n;
                n++;
            }
    return t;
}
__constant__ Targets81 c_t81 = make_targets81();
// Threads per block of the grid-cache weights: phase 1 in two rounds over the 729 evaluations,
// phases 2-4 on the first 27 x 16 (16-lane groups, one offset each) while the last wave loads the
// next triangle's inputs. Two blocks are resident per CU (up to 128 VGPRs), so one block's
// latency-bound phases 2-4 overlap the other's FP64-bound phase 1.
constexpr int GW = 512;
static_assert(2 * GW >= GC * 9 && GW % WAVE == 0 && WAVE % STEPS == 0 && GC * STEPS <= GW - WAVE,
              "phase 1 in two rounds; phases 2-4 in one, leaving the last wave free");

// Phases 2-4 of one stencil offset o in the registers of a 16-lane group: lane jj holds entry jj
// of the bit-reversed input of the group's four series (o, value), i.e. wave number
// kr = bitrev(jj), whose batch sums it forms itself (kr > 8: the conjugate of 16 - kr's); after
// the butterflies lane jj holds tap jj of the four series, and tap q of the W^fdtd record takes
// taps q - 1 and q - 2 from its neighbour lanes. Every value gets the operations and order of
// the per-(o, k) loop.
__device__ inline void gw_post(int o, int jj, const cdbl (*__restrict__ gl)[81][2], const GridTriIn &ti,
                               const WeightArgs &a, const signed char *s_idx, const double *s_twr,
                               const double *s_twi, const double *s_inv, uint32_t *__restrict__ wf,
                               float2 *__restrict__ sf, uint32_t *__restrict__ wb, float2 *__restrict__ sb,
                               float *__restrict__ meta)
{
    const int kr = bitrev4(jj), k = kr <= 8 ? kr : STEPS - kr;
    const int ox = o % 3 - 1, oy = o / 3 % 3 - 1, oz = o / 9 - 1;
    // phase 2: the 8-target FDTD-stencil batch in the reference's target order, and the
    // cell-centre target
    cdbl x[4];
    {
        const float s = a.c * a.c * a.dt * a.dt / (a.h * a.h);
        const int dx[8] = {0, -1, 1, 0, 0, 0, 0, 0}, dy[8] = {0, 0, 0, -1, 1, 0, 0, 0},
                  dz[8] = {0, 0, 0, 0, 0, -1, 1, 0};
        const float dsc[8] = {2, s, s, s, s, s, s, -6 * s};
        double sr0 = 0, si0 = 0, sr1 = 0, si1 = 0;
#pragma unroll
        for (int q = 0; q < 8; q++)
        {
            const int ex = ox + dx[q], ey = oy + dy[q], ez = oz + dz[q];
            if (ex == 0 && ey == 0 && ez == 0)
                continue;  // triangle_in_grid_cache: no target in the triangle's own cell (pppm_cache.cu:7-19)
            const int slot = s_idx[(ex + 2) * 25 + (ey + 2) * 5 + (ez + 2)];
            const double sc = dsc[q];
            sr0 += gl[k][slot][0].re * sc;
            si0 += gl[k][slot][0].im * sc;
            sr1 += gl[k][slot][1].re * sc;
            si1 += gl[k][slot][1].im * sc;
        }
        const double f = 0.5 * TRI1_W * ti.jac;
        x[0] = {f * sr0, f * si0};
        x[1] = {f * sr1, f * si1};
        if (!(ox == 0 && oy == 0 && oz == 0))
        {
            const int slot = s_idx[(ox + 2) * 25 + (oy + 2) * 5 + (oz + 2)];
            double cr0 = 0, ci0 = 0, cr1 = 0, ci1 = 0;
            cr0 += gl[k][slot][0].re * 1.0;
            ci0 += gl[k][slot][0].im * 1.0;
            cr1 += gl[k][slot][1].re * 1.0;
            ci1 += gl[k][slot][1].im * 1.0;
            x[2] = {f * cr0, f * ci0};
            x[3] = {f * cr1, f * ci1};
        }
        else
            x[2] = x[3] = {0, 0};
    }
    if (kr > 8)
    {
#pragma unroll
        for (int v = 0; v < 4; v++)
            x[v].im = -x[v].im;
    }
    // phase 3: scaled_idft of the four series; each radix-2 stage exchanges butterfly partners
    // with a lane shuffle, and both partners form the twiddled term from the same operands, so
    // every entry gets the operations of the sequential butterflies
#pragma unroll
    for (int L = 2; L <= STEPS; L <<= 1)
    {
        const int half = L >> 1, tws = STEPS / L;
        const bool lower = (jj & half) == 0;
        const int m = jj & (half - 1);
        const double cr = s_twr[m * tws], ci = s_twi[m * tws];
#pragma unroll
        for (int v = 0; v < 4; v++)
        {
            const cdbl other = {__shfl_xor(x[v].re, half, STEPS), __shfl_xor(x[v].im, half, STEPS)};
            const cdbl y = lower ? other : x[v], a0 = lower ? x[v] : other;
            const double tr_ = cr * y.re - ci * y.im;
            const double ti_ = cr * y.im + ci * y.re;
            x[v] = lower ? cdbl{a0.re + tr_, a0.im + ti_} : cdbl{a0.re - tr_, a0.im - ti_};
        }
    }
    const bool self = (o == 13);
    double X[4];
#pragma unroll
    for (int v = 0; v < 4; v++)  // offset (0, 0, 0) has no centre target: its centre taps are exactly 0
        X[v] = (v >= 2 && self) ? 0.0 : x[v].re / STEPS * s_inv[jj];
    // phase 4: W^fdtd = taps shifted by one minus the centre taps shifted by two, W^bem = the
    // centre taps (GridCache, pppm_cache.cu:99-116); the max-abs scales of LayerWeightHalf::set
    // are reduced over the 16 lanes of a record (exact, order-free)
    const int q = jj;
    const double p0 = __shfl(X[0], q >= 1 ? q - 1 : 0, STEPS), p1 = __shfl(X[1], q >= 1 ? q - 1 : 0, STEPS);
    const double p2 = __shfl(X[2], q >= 2 ? q - 2 : 0, STEPS), p3 = __shfl(X[3], q >= 2 ? q - 2 : 0, STEPS);
    double xf = q >= 1 ? p0 : 0.0;
    double yf = q >= 1 ? p1 : 0.0;
    if (!self && q >= 2)
    {
        xf += p2 * -1.0;
        yf += p3 * -1.0;
    }
    const float fs[2] = {(float)xf, (float)X[2]}, fd[2] = {(float)yf, (float)X[3]};
    float ms[2], md[2];
#pragma unroll
    for (int r = 0; r < 2; r++)
    {
        ms[r] = fabsf(fs[r]);
        md[r] = fabsf(fd[r]);
    }
#pragma unroll
    for (int off = STEPS / 2; off >= 1; off >>= 1)
    {
#pragma unroll
        for (int r = 0; r < 2; r++)
        {
            ms[r] = fmaxf(ms[r], __shfl_xor(ms[r], off, STEPS));
            md[r] = fmaxf(md[r], __shfl_xor(md[r], off, STEPS));
        }
    }
    const size_t rec = (size_t)ti.tri * GC + o;
#pragma unroll
    for (int r = 0; r < 2; r++)
    {
        if (ms[r] == 0)
            ms[r] = 1;
        if (md[r] == 0)
            md[r] = 1;
        (r == 0 ? wf : wb)[rec * STEPS + q] = f32_to_f16_bits(fs[r] / ms[r]) | (f32_to_f16_bits(fd[r] / md[r]) << 16);
        if (q == 0)
            (r == 0 ? sf : sb)[rec] = make_float2(ms[r], md[r]);
    }
    if (o == 0 && q == 0)
    {
        float *m = meta + 7 * ti.tri;
#pragma unroll
        for (int i = 0; i < 7; i++)
            m[i] = ti.meta[i];
    }
}

// update_grid_cache_kernel (pppm_cache.cu:64-123) for the dirty triangles; block b takes the
// dirty triangles b, b + gridDim.x, ... in four phases:
//  1. both layer kernels at the 81 distinct target cells of the 27 stencil batches, for each
//     wave number k <= 8 (one complex exponential per target, shared by the two layers), into
//     LDS;
//  2. per (offset o, k): the 8-target FDTD-stencil batch, summed in the reference's target
//     order, and the cell-centre target -> 4 complex values (single / double layer of each);
//  3. per (o, value): the scaled inverse DFT to 16 real taps;
//  4. per (o, record, tap): W^fdtd = taps shifted by one minus the centre taps shifted by two
//     (GridCache, pppm_cache.cu:99-116), W^bem = the centre taps; FP16 + scale each.
// The inputs come from the record the binning pass wrote for the triangle; during phases 2-4 of
// triangle j the last wave copies triangle j + 1's record (its number was loaded one iteration
// ago), computes the wave-number-free part of its 81 targets (distance, denominators, normal
// product) for phase 1, and loads triangle j + 2's number. Every value is computed with the operations and order
// of a per-(o, k) loop, so the records do not depend on how the work is spread over threads.
__global__ __launch_bounds__(GW) void k_grid_weights(const GridTriIn *__restrict__ gin, const int *__restrict__ list,
                                                     const int *__restrict__ n_ptr, WeightArgs a, Tdbem bem,
                                                     uint32_t *__restrict__ wf, float2 *__restrict__ sf,
                                                     uint32_t *__restrict__ wb, float2 *__restrict__ sb,
                                                     float *__restrict__ meta, const int *__restrict__ err)
{
    __shared__ cdbl gl[9][81][2];
    __shared__ GridTriIn tin[2];       // inputs of triangles j, j + 1 (slot = index mod 2)
    __shared__ LayerTarget tgt[2][81];  // target parts of triangles j, j + 1
    __shared__ int s_tri[2];           // numbers of triangles j + 1, j + 2 (slot = index mod 2)
    // per-block copies of the target table, the wave numbers, twiddles and lambda powers
    __shared__ signed char s_idx[125];
    __shared__ double s_wr[9], s_wi[9], s_twr[STEPS / 2], s_twi[STEPS / 2], s_inv[STEPS];
    // the count is read once per block: the error flag can be raised by the face-cache kernels
    // running beside this one, and every thread must take the same trip count through the
    // barriers below
    __shared__ int s_n;
    constexpr int RW = (int)(sizeof(GridTriIn) / 8);  // words of an input record, one per lane
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const bool loader = tid >= GW - WAVE;
    const int G = (int)gridDim.x;
    if (tid == 0)
        s_n = live_count(n_ptr, err);
    __syncthreads();
    const int n = s_n;
    if ((int)blockIdx.x >= n)
        return;  // the launch covers every CU; blocks without a triangle do nothing
    if (tid < 125)
        s_idx[tid] = c_t81.idx[tid];
    if (tid < 9)
    {
        s_wr[tid] = bem.wave_re[tid];
        s_wi[tid] = bem.wave_im[tid];
    }
    if (tid < STEPS / 2)
    {
        s_twr[tid] = bem.tw_re[tid];
        s_twi[tid] = bem.tw_im[tid];
    }
    if (tid < STEPS)
        s_inv[tid] = bem.inv_pow[tid];
    // the loader wave's target parts of the triangle in input slot b (written by this wave)
    auto targets = [&](int b) {
        wave_sync();
        const GridTriIn &ti = tin[b];
        const d3 p{ti.px, ti.py, ti.pz}, nn{ti.nx, ti.ny, ti.nz};
        for (int t = lane; t < 81; t += WAVE)
        {
            const d3 pt = cell_center_d(ti.cx + c_t81.off[t][0], ti.cy + c_t81.off[t][1], ti.cz + c_t81.off[t][2], a);
            tgt[b][t] = layer_target(p, pt, nn);
        }
    };
    if (loader)
    {
        // triangle 0's inputs and target parts, and triangle 1's number
        int t0 = lane == 0 ? list[blockIdx.x] : 0;
        if (lane == 1 && (int)blockIdx.x + G < n)
            s_tri[1] = list[blockIdx.x + G];
        t0 = __shfl(t0, 0);
        if (lane < RW)
            reinterpret_cast<unsigned long long *>(&tin[0])[lane] =
                reinterpret_cast<const unsigned long long *>(gin + t0)[lane];
        targets(0);
    }
    __syncthreads();
    // this thread's phase-1 items (at most two): wave number and target
    const int k0 = tid / 81, jt0 = tid % 81, k1 = (tid + GW) / 81, jt1 = (tid + GW) % 81;
    const bool second = tid + GW < 9 * 81;
    for (int j = 0;; j++)
    {
        const int li = (int)blockIdx.x + j * G;
        if (li >= n)
            break;
        const GridTriIn &ti = tin[j & 1];
        layer64_pair(tgt[j & 1][jt0], s_wr[k0], s_wi[k0], gl[k0][jt0][0], gl[k0][jt0][1]);
        if (second)
            layer64_pair(tgt[j & 1][jt1], s_wr[k1], s_wi[k1], gl[k1][jt1][0], gl[k1][jt1][1]);
        __syncthreads();
        if (loader)
        {
            if (li + G < n && lane < RW)
                reinterpret_cast<unsigned long long *>(&tin[(j + 1) & 1])[lane] =
                    reinterpret_cast<const unsigned long long *>(gin + s_tri[(j + 1) & 1])[lane];
            if (li + 2 * G < n && lane == RW)
                s_tri[j & 1] = list[li + 2 * G];
            if (li + G < n)
                targets((j + 1) & 1);
        }
        else if (tid < GC * STEPS)
            gw_post(tid / STEPS, tid % STEPS, gl, ti, a, s_idx, s_twr, s_twi, s_inv, wf, sf, wb, sb, meta);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- face cache
__device__ inline unsigned hash64(unsigned long long k, int cap)
{
    return (unsigned)((k * 0x9E3779B97F4A7C15ull) >> 32) & (unsigned)(cap - 1);
}
constexpr unsigned long long EMPTY_KEY = ~0ull;

// compute_face_compute_list_kernel + get_cache_element_with_check (pppm_cache.cu:183-232,
// pppm_cache.h:134-147): (c, j) -> persistent slot; dirty if new or moved past the thresholds.
// dirty[p] is 0 or 1; pairs past the live count write 0 (the flags are scanned over the
// capacity).
__device__ inline int lookup_pair(int p, const TriGeom *__restrict__ geom, const int *__restrict__ owner,
                                  const int *__restrict__ face_nbr, const int *__restrict__ n_ptr,
                                  unsigned long long *hkeys, int *hvals, int hcap, int *pool_count, int pool_cap,
                                  float2 *fmeta, int *face_slot, int *dirty, int *err_flag, int nt,
                                  const PrevFaces &prev)
{
    // lists past an overflow (or an invalid mesh) were not written: touch nothing
    const bool live = p < live_count(n_ptr, err_flag);
    int c = -1, j = -1;
    bool ok = false;
    if (live)
    {
        c = owner[p];
        j = face_nbr[p];
        ok = (unsigned)c < (unsigned)nt && (unsigned)j < (unsigned)nt;
        if (!ok)
            atomicOr(err_flag, ERR_CAPACITY);
    }
    int slot = -1, hpos = -1;  // hpos: hash entry of a key that needs a new slot
    bool fresh = false;
    if (ok && prev.start)
    {
        // the same (c, j) at the same place in c's list of the previous update: its slot,
        // without a hash probe (a slot belongs to its key for the handle's life)
        const int k = p - prev.cur_start[c];
        const int o0 = prev.start[c], o1 = prev.start[c + 1];
        if (k < o1 - o0 && prev.nbr[o0 + k] == j)
            slot = prev.slot[o0 + k];
        if (slot >= pool_cap)
            slot = -1;
    }
    const unsigned long long key = ((unsigned long long)c << 32) | (unsigned)j;
    unsigned h = hash64(key, hcap);
    for (int probe = 0; ok && slot < 0 && probe < hcap; probe++)
    {
        unsigned long long k = hkeys[h];
        if (k == EMPTY_KEY)
        {
            const unsigned long long was = atomicCAS(&hkeys[h], EMPTY_KEY, key);
            if (was == EMPTY_KEY)
                k = key, fresh = true;
            else
                k = was;
        }
        if (k == key)
        {
            if (!fresh)
                slot = hvals[h];
            if (slot < 0)  // new key, or a key left without a slot by an earlier overflow
            {
                hpos = (int)h;
                fresh = true;
            }
            break;
        }
        h = (h + 1) & (unsigned)(hcap - 1);
    }
    // new slots: one atomic per wave, handed out in lane order, so the 64 B records of the
    // wave's consecutive pairs are contiguous and the gathers that walk the lists read whole
    // cache lines (no key is looked up twice in one update, so its entry gets its slot here)
    const unsigned long long need = __ballot(hpos >= 0);
    if (need)
    {
        const int lane = (int)(threadIdx.x & 63);
        const int leader = __builtin_ctzll(need);
        int base = 0;
        if (lane == leader)
            base = atomicAdd(pool_count, (int)__popcll(need));
        base = __shfl(base, leader);
        if (hpos >= 0)
        {
            slot = base + (int)__popcll(need & ((1ull << lane) - 1ull));
            if (slot >= pool_cap)
            {
                atomicOr(err_flag, ERR_CAPACITY);
                slot = -1;
            }
            hvals[hpos] = slot;
        }
    }
    if (!live)
    {
        dirty[p] = 0;
        return 0;
    }
    if (!ok)
    {
        dirty[p] = 0;
        face_slot[p] = -1;
        return 0;
    }
    face_slot[p] = slot;
    if (slot < 0)
    {
        dirty[p] = 0;
        return 0;
    }
    const TriGeom gc = geom[c], gj = geom[j];
    const float normal_angle = dotf(gj.normal, gc.normal);
    const float distance = lengthf(gj.center - gc.center);
    bool d = fresh;
    if (!fresh)
    {
        const float2 m = fmeta[slot];
        d = (double)fabsf(m.x - normal_angle) > RECOMPUTE_THRESHOLD ||
            (double)fabsf(m.y - distance) > RECOMPUTE_THRESHOLD * (double)distance;
    }
    if (d)
    {
        // the slot's meta is written by k_face_finish once its weights exist; a new slot
        // holds +inf until then, so an update that stops before the weights are computed
        // leaves the pair dirty for the next one
        if (fresh)
            fmeta[slot] = make_float2(INFINITY, INFINITY);
        dirty[p] = 1;
        return 1;
    }
    dirty[p] = 0;
    return 0;
}

__global__ __launch_bounds__(256) void k_face_lookup(const TriGeom *__restrict__ geom, const int *__restrict__ owner,
                                                     const int *__restrict__ face_nbr, const int *__restrict__ n_ptr,
                                                     int cap, unsigned long long *hkeys, int *hvals, int hcap,
                                                     int *pool_count, int pool_cap, float2 *fmeta, int *face_slot,
                                                     int *dirty, int *err_flag, int nt, PrevFaces prev, int *tsum)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    int f = 0;
    if (p < cap)
        f = lookup_pair(p, geom, owner, face_nbr, n_ptr, hkeys, hvals, hcap, pool_count, pool_cap, fmeta, face_slot,
                        dirty, err_flag, nt, prev);
    // one atomic per wave for the scan's tile sums (64 pairs never straddle a tile)
    const unsigned long long bal = __ballot(f);
    if (tsum && (threadIdx.x & 63) == 0 && bal)
        atomicAdd(&tsum[(p & ~63) / SCAN_TILE], (int)__popcll(bal));
}

// Face-cache integrals (pppm_cache.cu:234-271) per (dirty pair, k <= 8, layer): the
// face-face integral plus the interpolated far-field representation subtracted for corners
// farther than 1.5h. Regular pairs (no common vertex) take 16 lanes per item: lane 0
// evaluates the face-face term, lane 1 + w the far-field term of corner w, and lane 0 adds
// the corner products in corner order, as potential_integrand does.
__global__ __launch_bounds__(B) void k_face_regular(const float *__restrict__ verts, const int *__restrict__ tris,
                                                    const TriGeom *__restrict__ geom,
                                                    const int *__restrict__ dirty_pairs, const int *__restrict__ n_ptr,
                                                    const int *__restrict__ owner, const int *__restrict__ face_nbr,
                                                    const float *__restrict__ interp, WeightArgs a, Tdbem bem,
                                                    cdbl *__restrict__ tmp, const int *__restrict__ err)
{
    constexpr int L = 16;
    const long long lanes = (long long)live_count(n_ptr, err) * 18 * L;
    // the stride is a multiple of 16, so the 16 lanes of an item stay together
    for (long long e = (long long)blockIdx.x * B + threadIdx.x; e < lanes; e += (long long)gridDim.x * B)
    {
        const long long i = e / L;
        const int sub = (int)(e % L);
        const int d = (int)(i / 18), r = (int)(i % 18), layer = r / 9, k = r % 9;
        const int p = dirty_pairs[d];
        const int c = owner[p], j = face_nbr[p];
        const int src[3] = {tris[3 * j], tris[3 * j + 1], tris[3 * j + 2]};
        const int trg[3] = {tris[3 * c], tris[3 * c + 1], tris[3 * c + 2]};
        const int nn = common_vertices(src[0], src[1], src[2], trg[0], trg[1], trg[2]);
        const bool regular = nn == 0;  // the same for the 16 lanes of the item
        const double kr = bem.wave_re[k], ki = bem.wave_im[k];
        double pr = 0, pi = 0;
        bool far_corner = false;
        if (regular && sub == 0)
        {
            const cdbl v = f2f64<false>(verts, src, trg, nn, kr, ki, layer, bem.threshold);
            pr = v.re;
            pi = v.im;
        }
        else if (regular && sub <= 8)
        {
            const int w = sub - 1;
            const TriGeom gc = geom[c], gj = geom[j];
            const double lim = (double)a.h * 1.5;
            const int x = gc.base3.x + (w >> 2), y = gc.base3.y + ((w >> 1) & 1), z = gc.base3.z + (w & 1);
            const f3 ctr = cell_center_f(x, y, z, a);
            if ((double)fabsf(ctr.x - gj.center.x) > lim || (double)fabsf(ctr.y - gj.center.y) > lim ||
                (double)fabsf(ctr.z - gj.center.z) > lim)
            {
                const TriD J = tri_d(verts, src[0], src[1], src[2]);
                const d3 pj = l2g(TRI1_X, TRI1_X, J.v);
                const cdbl g = layer64(pj, cell_center_d(x, y, z, a), J.n, kr, ki, layer);
                const double sc = (double)(-interp[8 * c + w]);
                pr = g.re * sc;
                pi = g.im * sc;
                far_corner = true;
            }
        }
        double sr = 0, si = 0;
        bool any = false;
#pragma unroll
        for (int w = 0; w < 8; w++)
        {
            const double xr = __shfl(pr, w + 1, L), xi = __shfl(pi, w + 1, L);
            const bool fw = __shfl(far_corner ? 1 : 0, w + 1, L) != 0;
            if (fw)
            {
                sr += xr;
                si += xi;
                any = true;
            }
        }
        if (regular && sub == 0)
        {
            cdbl v = {pr, pi};
            if (any)
            {
                const TriD J = tri_d(verts, src[0], src[1], src[2]);
                const double f = 0.5 * TRI1_W * J.jac;
                v.re += f * sr;
                v.im += f * si;
            }
            tmp[i] = v;
        }
    }
}

// Singular pairs (common vertices): one thread per (pair, k, layer), the 81-point rules in
// sequence, then the corner terms.
__global__ __launch_bounds__(B) void k_face_singular(const float *__restrict__ verts, const int *__restrict__ tris,
                                                      const TriGeom *__restrict__ geom,
                                                      const int *__restrict__ dirty_pairs,
                                                      const int *__restrict__ n_ptr, const int *__restrict__ owner,
                                                      const int *__restrict__ face_nbr,
                                                      const float *__restrict__ interp, WeightArgs a, Tdbem bem,
                                                      cdbl *__restrict__ tmp, const int *__restrict__ err)
{
    const long long items = (long long)live_count(n_ptr, err) * 18;
    for (long long i = (long long)blockIdx.x * B + threadIdx.x; i < items; i += (long long)gridDim.x * B)
    {
        const int d = (int)(i / 18), r = (int)(i % 18), layer = r / 9, k = r % 9;
        const int p = dirty_pairs[d];
        const int c = owner[p], j = face_nbr[p];
        const int src[3] = {tris[3 * j], tris[3 * j + 1], tris[3 * j + 2]};
        const int trg[3] = {tris[3 * c], tris[3 * c + 1], tris[3 * c + 2]};
        const int nn = common_vertices(src[0], src[1], src[2], trg[0], trg[1], trg[2]);
        if (nn == 0)
            continue;  // k_face_regular computes this pair
        const double kr = bem.wave_re[k], ki = bem.wave_im[k];
        cdbl v = f2f64<true>(verts, src, trg, nn, kr, ki, layer, bem.threshold);
        // the interpolated far field of the corners farther than 1.5h: potential_integrand
        // (integrand.h:209-222) over the accepted corners in corner order, accumulated on the fly
        const TriGeom gc = geom[c], gj = geom[j];
        const double lim = (double)a.h * 1.5;
        const TriD J = tri_d(verts, src[0], src[1], src[2]);
        const d3 pj = l2g(TRI1_X, TRI1_X, J.v);
        double sr = 0, si = 0;
        bool any = false;
        for (int w = 0; w < 8; w++)
        {
            const int x = gc.base3.x + (w >> 2), y = gc.base3.y + ((w >> 1) & 1), z = gc.base3.z + (w & 1);
            const f3 ctr = cell_center_f(x, y, z, a);
            if ((double)fabsf(ctr.x - gj.center.x) > lim || (double)fabsf(ctr.y - gj.center.y) > lim ||
                (double)fabsf(ctr.z - gj.center.z) > lim)
            {
                const cdbl g = layer64(pj, cell_center_d(x, y, z, a), J.n, kr, ki, layer);
                const double sc = (double)(-interp[8 * c + w]);
                sr += g.re * sc;
                si += g.im * sc;
                any = true;
            }
        }
        if (any)
        {
            const double f = 0.5 * TRI1_W * J.jac;
            v.re += f * sr;
            v.im += f * si;
        }
        tmp[i] = v;
    }
}

__global__ __launch_bounds__(B) void k_face_finish(const int *__restrict__ dirty_pairs,
                                                   const int *__restrict__ n_ptr, const int *__restrict__ face_slot,
                                                   const TriGeom *__restrict__ geom, const int *__restrict__ owner,
                                                   const int *__restrict__ face_nbr, Tdbem bem,
                                                   const cdbl *__restrict__ tmp, uint32_t *__restrict__ fw,
                                                   float2 *__restrict__ fsc, float2 *__restrict__ fmeta,
                                                   const int *__restrict__ err)
{
    const int n = live_count(n_ptr, err);
    for (int d = blockIdx.x * B + threadIdx.x; d < n; d += gridDim.x * B)
    {
        const int p = dirty_pairs[d];
        const int slot = face_slot[p];
        const TriGeom gc = geom[owner[p]], gj = geom[face_nbr[p]];
        fmeta[slot] = make_float2(dotf(gj.normal, gc.normal), lengthf(gj.center - gc.center));
        cdbl v[9];
        double ws[STEPS], wd[STEPS];
        for (int k = 0; k < 9; k++)
            v[k] = tmp[(size_t)d * 18 + k];
        scaled_idft(v, bem, ws);
        for (int k = 0; k < 9; k++)
            v[k] = tmp[(size_t)d * 18 + 9 + k];
        scaled_idft(v, bem, wd);
        float fs[STEPS], fd[STEPS];
        for (int q = 0; q < STEPS; q++)
        {
            fs[q] = (float)ws[q];
            fd[q] = (float)wd[q];
        }
        quantize_store(fs, fd, fw + (size_t)slot * STEPS, fsc + slot);
    }
}

__global__ __launch_bounds__(256) void k_rehash(const unsigned long long *__restrict__ okeys,
                                                const int *__restrict__ ovals, int ocap, unsigned long long *nkeys,
                                                int *nvals, int ncap)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ocap)
        return;
    const unsigned long long key = okeys[i];
    if (key == EMPTY_KEY)
        return;
    unsigned h = hash64(key, ncap);
    for (int probe = 0; probe < ncap; probe++)
    {
        if (atomicCAS(&nkeys[h], EMPTY_KEY, key) == EMPTY_KEY)
        {
            nvals[h] = ovals[i];
            return;
        }
        h = (h + 1) & (unsigned)(ncap - 1);
    }
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long *a, long long n, unsigned long long v)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        a[i] = v;
}

inline int nb(long long n, int b) { return n > 0 ? (int)((n + b - 1) / b) : 1; }
// grid-stride launches for work whose size lives in device memory: cap bounds the items
inline int nb_capped(long long cap_items, int b) { return (int)std::min<long long>(nb(cap_items, b), 8192); }
// grid-stride launch sized from a host estimate of the live count (the previous update's),
// at least 64 blocks, at most the capacity-based size: any live count is still covered
inline int nb_hint(long long hint_items, long long cap_items, int b)
{
    return std::max(std::min(nb(hint_items, b), nb_capped(cap_items, b)), std::min(64, nb_capped(cap_items, b)));
}
}  // namespace

void tdbem_init(Tdbem &b, float dt, float c, float grid_size)
{
    // TDBEM::init (bem.h:264-277) in FP64
    const double dc = (double)dt * (double)c;
    const double lam = std::max(std::pow(dc, 3.0 / STEPS), std::pow((double)EPS_F, 1.0 / (2 * STEPS)));
    b.lambda = lam;
    for (int k = 0; k < STEPS; k++)
    {
        const double ang = (-2 * PI_R / STEPS) * k;
        const double sr = lam * std::cos(ang), si = lam * std::sin(ang);
        const double tr = sr * sr - si * si - 4 * sr + 3, ti = 2 * sr * si - 4 * si;  // BDF2 (macro.h:90)
        const double br = 0.5 * tr / dc, bi = 0.5 * ti / dc;
        b.wave_re[k] = -bi;  // i * BDF2(s_k) / (c dt)
        b.wave_im[k] = br;
    }
    for (int j = 0; j < STEPS / 2; j++)
    {
        const double ang = 2 * PI_R * j / STEPS;
        b.tw_re[j] = std::cos(ang);
        b.tw_im[j] = std::sin(ang);
    }
    double t = 1.0;
    for (int n = 0; n < STEPS; n++)
    {
        b.inv_pow[n] = t;
        t /= lam;
    }
    b.threshold = (float)(grid_size * 0.1);
}

size_t face_tmp_bytes(int max_dirty) { return (size_t)max_dirty * 18 * sizeof(cdbl); }

hipError_t launch_grid_weights(const GridTriIn *gin, const int *list, const int *n_ptr, int cap, const WeightArgs &a,
                               const Tdbem &b, uint32_t *wf, float2 *sf, uint32_t *wb, float2 *sb, float *meta,
                               const int *err, hipStream_t s)
{
    // blocks stride over the dirty triangles (a device-side count), two per CU (two 512-thread
    // blocks are resident per CU), so a jump in the count (a whole moving object past the
    // recompute thresholds at once) still fills the device; blocks past the count exit at once
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            n = 0;
        return n > 0 ? n : 256;
    }();
    const long long blocks = std::max<long long>(1, std::min<long long>(cap, 2LL * cus));
    k_grid_weights<<<(unsigned)blocks, GW, 0, s>>>(gin, list, n_ptr, a, b, wf, sf, wb, sb, meta, err);
    return hipGetLastError();
}
hipError_t launch_face_lookup(const TriGeom *geom, const int *tris, const int *owner, const int *face_nbr,
                              const int *n_ptr, int cap, unsigned long long *hkeys, int *hvals, int hcap,
                              int *pool_count, int pool_cap, float2 *fmeta, int *face_slot, int *dirty, int *err_flag,
                              int nt, const PrevFaces &prev, int *tsum, hipStream_t s)
{
    k_face_lookup<<<nb(cap, 256), 256, 0, s>>>(geom, owner, face_nbr, n_ptr, cap, hkeys, hvals, hcap, pool_count,
                                               pool_cap, fmeta, face_slot, dirty, err_flag, nt, prev, tsum);
    return hipGetLastError();
}
hipError_t launch_face_weights(const float *verts, const int *tris, const TriGeom *geom, const int *dirty_pairs,
                               const int *n_ptr, int cap, const int *owner, const int *face_nbr, const int *face_slot,
                               const float *interp, const WeightArgs &a, const Tdbem &b, void *tmp, uint32_t *fw,
                               float2 *fsc, float2 *fmeta, const int *err, long long hint_dirty, hipStream_t s)
{
    const long long hint = hint_dirty < 0 ? cap : std::min<long long>(hint_dirty, cap);
    k_face_regular<<<nb_hint(hint * 18 * 16, (long long)cap * 18 * 16, B), B, 0, s>>>(
        verts, tris, geom, dirty_pairs, n_ptr, owner, face_nbr, interp, a, b, (cdbl *)tmp, err);
    k_face_singular<<<nb_hint(hint * 18, (long long)cap * 18, B), B, 0, s>>>(
        verts, tris, geom, dirty_pairs, n_ptr, owner, face_nbr, interp, a, b, (cdbl *)tmp, err);
    k_face_finish<<<nb_hint(hint, cap, B), B, 0, s>>>(dirty_pairs, n_ptr, face_slot, geom, owner, face_nbr, b,
                                                  (const cdbl *)tmp, fw, fsc, fmeta, err);
    return hipGetLastError();
}
hipError_t launch_rehash(const unsigned long long *okeys, const int *ovals, int ocap, unsigned long long *nkeys,
                         int *nvals, int ncap, hipStream_t s)
{
    k_rehash<<<nb(ocap, 256), 256, 0, s>>>(okeys, ovals, ocap, nkeys, nvals, ncap);
    return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long *a, long long n, unsigned long long v, hipStream_t s)
{
    k_fill_u64<<<nb(n, 256), 256, 0, s>>>(a, n, v);
    return hipGetLastError();
}

}  // namespace pppm_amd
