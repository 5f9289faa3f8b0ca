// Adjoint of the PPPM time step: the transposed step run backwards in time, for gradients of
// a listener loss with respect to the per-triangle Neumann inputs (surface accelerations).
//
// The forward step (update_grid_and_face, src/algorithm/pppm.h:137-145, kernels of
// src/algorithm/fdtd.cu:7-150 and pppm.cu:8-204) is linear in the grid levels, the history
// rings and the Neumann input, with coefficients fixed by the mesh (FP16 weights x FP32 scales,
// interpolation weights, the ABC's 4x4 systems). Its adjoint therefore needs no forward state:
// only the cotangents of the listener samples. Reversing step t, in this order:
//   1. listener cotangents into lambda_P(t); C3 cells: u_c = lambda_P(t)[c], lambda_P(t)[c] = 0
//      (the near field overwrote those cells, pppm.cu:99)
//   2. near field transposed: lambda_N/D(tri, t-k) += 0.98 (-S_k, D_k)(tri, o) u_c(tri, o)
//   3. Dirichlet update: v_c = lambda_D(c, t) / (A_c / 2 - sum_j D_0(c, j)), lambda_D(c, t) = 0
//   4. Dirichlet transposed: lambda_far(x) = u_x + sum rc w(c, x) v_c over the triangles whose
//      base-cell corner x is; lambda_N/D(j, t-k) += (-S_k, D_k [k>0])(c, j) v_c over the face
//      pairs (c, j)
//   5. far field transposed: lambda_P(t)[c] += lambda_far(c); lambda_N/D(tri, t-k) +=
//      (S_k, -D_k)(tri, o) lambda_far(c(tri, o)), k >= 1; the gradient of N(tri, t) is final
//      (no later reverse step touches it) and is written out
//   6. FDTD transposed: the ABC shell cells solve A^T b = (ghost cotangents, lambda_P(t)) and
//      every cell gathers the transposed leapfrog and ABC coefficients of the cells that read it:
//      lambda_P(t-1) += g terms, lambda_P(t-2) = gm terms (its ghost layer holds the cotangents
//      of the ghost values the ABC wrote, fdtd.cu:132-136)
// Every reverse gather is a gather over a transposed list built once per mesh (per triangle:
// its 27 C3 cells; per triangle j: the face pairs (c, j); per C3 cell: the triangles whose
// base cell it is a corner of), so each output has one writer and the sums have a fixed order.
// The reverse time index lives in device memory, so a captured graph replays any step.
#include "internal.hpp"

#include <algorithm>

namespace pppm_amd
{

namespace
{

constexpr int BLK = 256;

__device__ inline float *alevel(float *l3, int t, int rel, size_t nn)
{
    const int l = ((t + rel) % 3 + 3) % 3;
    return l3 + (size_t)l * nn;
}

__device__ inline float wS(uint32_t w, float2 sc) { return f16_bits_to_f32(w) * sc.x; }
__device__ inline float wD(uint32_t w, float2 sc) { return f16_bits_to_f32(w >> 16) * sc.y; }

// 1. listener cotangents and the C3 cells' cotangents u (the near field's input)
__device__ __forceinline__ void adj_read(const AdjScene &d, int blk)
{
    const int t = d.ast->t;
    const size_t nn = (size_t)d.N * d.N * d.N;
    float *lam = alevel(d.lam3, t, 0, nn);
    const int c = blk * BLK + threadIdx.x;
    const float *go = d.gout ? d.gout + (size_t)(t - d.t_first) * d.gout_stride : nullptr;
    if (c < d.n_c3)
    {
        const int q = d.c3_cells[c];
        float u = lam[q];
        if (go)
            for (int l = 0; l < d.n_listen; l++)
                if (d.lc3[l] == c)
                    u += go[l];
        d.unear[c] = u;
        lam[q] = 0.0f;
    }
    else if (c == d.n_c3 && go)
    {
        // listeners outside the corrected cells: one thread, in listener order
        for (int l = 0; l < d.n_listen; l++)
            if (d.lc3[l] < 0)
                lam[d.lcell[l]] += go[l];
    }
}

// The 16 tap threads of a triangle are one aligned 16-lane group (a block's first thread is at
// a multiple of 16 within its triangle range); lane k fetches the per-record values of records
// k and k + 16 of the triangle's 27, and every lane takes record o's from lane o % 16 by shuffle.
struct RecVals
{
    float v[2];
    float2 sc[2];
};
__device__ __forceinline__ RecVals rec_vals(const int *__restrict__ tc3, const float *__restrict__ val,
                                            const float2 *__restrict__ scales, int tri, int k)
{
    RecVals r;
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        const int o = k + STEPS * h;
        r.v[h] = 0.0f;
        r.sc[h] = make_float2(0.0f, 0.0f);
        if (o < GC)
        {
            const int rec = tri * GC + o;
            r.v[h] = val[tc3[rec]];
            r.sc[h] = scales[rec];
        }
    }
    return r;
}
__device__ __forceinline__ void rec_take(const RecVals &r, int o, float &v, float2 &sc)
{
    const int h = o / STEPS, src = o % STEPS;
    v = __shfl(r.v[h], src, STEPS);
    sc.x = __shfl(r.sc[h].x, src, STEPS);
    sc.y = __shfl(r.sc[h].y, src, STEPS);
}

// 2. near field transposed, one thread per (triangle, tap): the 27 records of a triangle are
// contiguous, so the 16 threads of a triangle read 64 B of each record
__device__ __forceinline__ void adj_near(const AdjScene &d, int blk)
{
    const int i = blk * BLK + threadIdx.x;
    if (i >= d.nt * STEPS)
        return;  // whole 16-lane groups
    const int t = d.ast->t;
    const int tri = i / STEPS, k = i % STEPS;
    const RecVals rv = rec_vals(d.tc3, d.unear, d.sb, tri, k);
    float an = 0.0f, ad = 0.0f;
#pragma unroll
    for (int o = 0; o < GC; o++)
    {
        const int rec = tri * GC + o;
        float u;
        float2 sc;
        rec_take(rv, o, u, sc);
        const uint32_t w = d.wb[(size_t)rec * STEPS + k];
        an += wS(w, sc) * u;
        ad += wD(w, sc) * u;
    }
    const int slot = tri * RING + ((t - k) & (RING - 1));
    d.ln[slot] += -0.98f * an;
    d.ld[slot] += 0.98f * ad;
}

// 3. the Dirichlet update's divisor
__device__ __forceinline__ void adj_dirv(const AdjScene &d, int blk)
{
    const int c = blk * BLK + threadIdx.x;
    if (c >= d.nt)
        return;
    const int slot = c * RING + (d.ast->t & (RING - 1));
    d.v[c] = (float)((double)d.ld[slot] / d.den[c]);
    d.ld[slot] = 0.0f;
}

// blocks of the C3 part of the Dirichlet and far-field kernels (the triangle part starts at a
// block boundary, so its 16-lane groups are aligned)
__device__ __forceinline__ int c3_part_blocks(const AdjScene &d) { return (d.n_c3 + BLK - 1) / BLK; }

// 4. Dirichlet transposed: C3 cells [0, n_c3) gather the face-far-field interpolation; then one
// thread per (triangle j, tap) gathers the face pairs (c, j) (records in that order: a
// triangle's 16 lanes read consecutive 64 B records; lane k fetches pair eb + k's cell value and
// scales, shared by shuffles)
__device__ __forceinline__ void adj_dir(const AdjScene &d, int blk)
{
    const int t = d.ast->t;
    const int nb3 = c3_part_blocks(d);
    if (blk < nb3)
    {
        const int i = blk * BLK + threadIdx.x;
        if (i < d.n_c3)
        {
            float a = d.unear[i];
            for (int e = d.ct_start[i]; e < d.ct_start[i + 1]; e++)
                a += d.ct_w[e] * d.v[d.ct_c[e]];
            d.lfar[i] = a;
        }
        return;
    }
    const int q = (blk - nb3) * BLK + threadIdx.x;
    if (q >= d.nt * STEPS)
        return;  // whole 16-lane groups
    const int j = q / STEPS, k = q % STEPS;
    const int e0 = d.ft_start[j], e1 = d.ft_start[j + 1];
    float an = 0.0f, ad = 0.0f;
    for (int eb = e0; eb < e1; eb += STEPS)
    {
        const int el = eb + k;
        float vl = 0.0f;
        float2 sl = make_float2(0.0f, 0.0f);
        if (el < e1)
        {
            vl = d.v[d.ft_c[el]];
            sl = d.tfsc[el];
        }
        const int n = min(STEPS, e1 - eb);
        for (int m = 0; m < n; m++)
        {
            const float vc = __shfl(vl, m, STEPS);
            const float2 sc = make_float2(__shfl(sl.x, m, STEPS), __shfl(sl.y, m, STEPS));
            const uint32_t w = d.tfw[(size_t)(eb + m) * STEPS + k];
            an += wS(w, sc) * vc;
            ad += wD(w, sc) * vc;
        }
    }
    const int rs = j * RING + ((t - k) & (RING - 1));
    d.ln[rs] += -an;
    if (k > 0)  // D_j[t] enters the face near field as 0 (pppm.cu:143)
        d.ld[rs] += ad;
}

// 5. far field transposed; the (triangle, tap 0) threads write the finished gradient of N(t)
__device__ __forceinline__ void adj_far(const AdjScene &d, int blk)
{
    const int t = d.ast->t;
    const size_t nn = (size_t)d.N * d.N * d.N;
    const int nb3 = c3_part_blocks(d);
    if (blk < nb3)
    {
        const int i = blk * BLK + threadIdx.x;
        if (i < d.n_c3)
        {
            float *lam = alevel(d.lam3, t, 0, nn);
            lam[d.c3_cells[i]] += d.lfar[i];
        }
        return;
    }
    const int q = (blk - nb3) * BLK + threadIdx.x;
    if (q >= d.nt * STEPS)
        return;  // whole 16-lane groups
    const int tri = q / STEPS, k = q % STEPS;
    // every lane of the group takes part in the shuffles; tap 0 then writes the gradient instead
    const RecVals rv = rec_vals(d.tc3, d.lfar, d.sf, tri, k);
    float an = 0.0f, ad = 0.0f;
#pragma unroll
    for (int o = 0; o < GC; o++)
    {
        const int rec = tri * GC + o;
        float lf;
        float2 sc;
        rec_take(rv, o, lf, sc);
        const uint32_t w = d.wf[(size_t)rec * STEPS + k];
        an += wS(w, sc) * lf;
        ad += wD(w, sc) * lf;
    }
    if (k == 0)
    {
        const int slot = tri * RING + (t & (RING - 1));
        d.grad[(size_t)(t - d.t_first) * d.grad_stride + tri] = d.ln[slot];
        d.ln[slot] = 0.0f;
        return;
    }
    const int slot = tri * RING + ((t - k) & (RING - 1));
    d.ln[slot] += an;
    d.ld[slot] += -ad;
}

// ABC shell geometry (fdtd.cu:88-137): judge_d = -1 at index 1, +1 at N-2, 0 otherwise
struct Judge
{
    int j[3];
};
__device__ inline bool shell_cell(int x, int y, int z, int N)
{
    const bool in = x >= 1 && y >= 1 && z >= 1 && x <= N - 2 && y <= N - 2 && z <= N - 2;
    return in && (x == 1 || y == 1 || z == 1 || x == N - 2 || y == N - 2 || z == N - 2);
}
__device__ inline Judge judge_of(int x, int y, int z, int N)
{
    const int co[3] = {x, y, z};
    Judge J;
    for (int d = 0; d < 3; d++)
        J.j[d] = -((int)(co[d] == 1)) + ((int)(co[d] == N - 2));
    return J;
}

struct AbcCoef
{
    float coef, coef1, coef2, coef3;
};
// the ABC's FP32 coefficients, computed as in ab_condition_solve (fdtd.cu:94-118)
__device__ inline AbcCoef abc_coef(const FdtdParams &p)
{
    const float dt = p.dt, h = p.h, c = p.c;
    AbcCoef k;
    k.coef = c * c * dt * dt / (h * h);
    k.coef1 = 1 / (dt * dt);
    k.coef2 = c / (2 * h * dt);
    k.coef3 = -(c * c) / (2 * h * h);
    return k;
}

// 6a. per shell cell: b = A^{-T} (ghost cotangents, lambda_P(t)) in FP64 (Gaussian elimination
// with partial pivoting on the 4x4 transposed system)
__device__ __forceinline__ void adj_abc(const AdjScene &d, int blk)
{
    const int N = d.fp.N;
    const long long S = N - 2;
    const long long fx = 2 * S * S, fy = 2 * (S - 2) * S, fz = 2 * (S - 2) * (S - 2);
    long long idx = (long long)blk * BLK + threadIdx.x;
    int i, j, k;
    if (idx < fx)
    {
        i = idx / (S * S) ? N - 2 : 1;
        const long long r = idx % (S * S);
        j = 1 + (int)(r / S);
        k = 1 + (int)(r % S);
    }
    else if ((idx -= fx) < fy)
    {
        j = idx / ((S - 2) * S) ? N - 2 : 1;
        const long long r = idx % ((S - 2) * S);
        i = 2 + (int)(r / S);
        k = 1 + (int)(r % S);
    }
    else if ((idx -= fy) < fz)
    {
        k = idx / ((S - 2) * (S - 2)) ? N - 2 : 1;
        const long long r = idx % ((S - 2) * (S - 2));
        i = 2 + (int)(r / (S - 2));
        j = 2 + (int)(r % (S - 2));
    }
    else
        return;
    const int t = d.ast->t;
    const size_t nn = (size_t)N * N * N;
    const float *lam = alevel(d.lam3, t, 0, nn);
    const float *lam1 = alevel(d.lam3, t, -1, nn);
    const Judge J = judge_of(i, j, k, N);
    const int strd[3] = {N * N, N, 1};
    const int me = lin3(i, j, k, N);
    const AbcCoef q = abc_coef(d.fp);
    // A (fdtd.cu:94-111), transposed into m, with the right-hand side in column 4
    float a[4][4];
    a[3][3] = 1;
    for (int e = 0; e < 3; e++)
        a[3][e] = -q.coef * (float)(J.j[e] != 0);
    for (int r = 0; r < 3; r++)
    {
        const int y = (r + 1) % 3, z = (r + 2) % 3;
        a[r][3] = q.coef1;
        a[r][r] = q.coef2;
        a[r][y] = q.coef3 * (float)(J.j[y] != 0);
        a[r][z] = q.coef3 * (float)(J.j[z] != 0);
    }
    double m[4][5];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            m[r][c] = a[c][r];
    for (int e = 0; e < 3; e++)
        m[e][4] = J.j[e] != 0 ? (double)lam1[me + J.j[e] * strd[e]] : 0.0;
    m[3][4] = lam[me];
    for (int c = 0; c < 4; c++)
    {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(m[r][c]) > fabs(m[piv][c]))
                piv = r;
        if (piv != c)
            for (int e = 0; e < 5; e++)
            {
                const double tmp = m[c][e];
                m[c][e] = m[piv][e];
                m[piv][e] = tmp;
            }
        for (int r = c + 1; r < 4; r++)
        {
            const double f = m[r][c] / m[c][c];
            for (int e = c; e < 5; e++)
                m[r][e] -= f * m[c][e];
        }
    }
    double b[4];
    for (int r = 3; r >= 0; r--)
    {
        double s = m[r][4];
        for (int e = r + 1; e < 4; e++)
            s -= m[r][e] * b[e];
        b[r] = s / m[r][r];
    }
    d.bb[me] = make_float4((float)b[0], (float)b[1], (float)b[2], (float)b[3]);
}

// 6b. per cell: the transposed leapfrog (interior cells that read it) and the transposed ABC
// (shell cells that read it, with the coefficients of fdtd.cu:94-118 for each neighbour slot).
// The blocks first cover the deep interior (every index in [3, N-4]: no shell cell reads the
// cell or its neighbours, the transposed leapfrog alone), then the frame of the 3 outer layers
// of every face, so no wave mixes the two paths.
__device__ __forceinline__ void adj_fdtd_frame_cell(const AdjScene &d, int y0, int y1, int y2)
{
    const int N = d.fp.N;
    const size_t nn = (size_t)N * N * N;
    const int id = lin3(y0, y1, y2, N);
    const int yc[3] = {y0, y1, y2};
    const int t = d.ast->t;
    const float *lam = alevel(d.lam3, t, 0, nn);
    float *lam1 = alevel(d.lam3, t, -1, nn);
    float *lam2 = alevel(d.lam3, t, -2, nn);
    int out = 0;
    for (int e = 0; e < 3; e++)
        out += yc[e] < 1 || yc[e] > N - 2;
    if (out >= 2)  // edge and corner ghosts: never read by the forward step
    {
        lam2[id] = 0.0f;
        return;
    }
    const int strd[3] = {N * N, N, 1};
    auto interior = [&](int a, int b, int c) { return a >= 2 && b >= 2 && c >= 2 && a <= N - 3 && b <= N - 3 && c <= N - 3; };
    const double cc = 2.0 - 6.0 * d.fp.coef / d.fp.hh, cn = d.fp.coef / d.fp.hh;
    double ag = 0.0, am = 0.0;
    if (out == 0)
    {
        // leapfrog: p(t)[x] = 2 g[x] + cn (sum of the 6 neighbours - 6 g[x]) - gm[x] (fdtd.cu:7-33)
        if (interior(y0, y1, y2))
        {
            ag += cc * (double)lam[id];
            am -= (double)lam[id];
        }
        for (int e = 0; e < 3; e++)
            for (int s = -1; s <= 1; s += 2)
            {
                int x[3] = {y0, y1, y2};
                x[e] += s;
                if (interior(x[0], x[1], x[2]))
                    ag += cn * (double)lam[id + s * strd[e]];
            }
    }
    const AbcCoef q = abc_coef(d.fp);
    // shell cells x in {y, y +- e_d} that read y as g or gm
    for (int slot = 0; slot < 7; slot++)
    {
        int x[3] = {y0, y1, y2};
        int e = -1, s = 0;  // y = x + s e_e (s = 0: y = x)
        if (slot > 0)
        {
            e = (slot - 1) >> 1;
            s = ((slot - 1) & 1) ? 1 : -1;
            x[e] -= s;
        }
        if (!shell_cell(x[0], x[1], x[2], N))
            continue;
        const Judge J = judge_of(x[0], x[1], x[2], N);
        const float4 b4 = d.bb[lin3(x[0], x[1], x[2], N)];
        const double b[4] = {b4.x, b4.y, b4.z, b4.w};
        if (s == 0)
        {
            // g[x]: row 3: 2 - 6 coef; rows 0-2: 2 coef1 + 4 coef3. gm[x]: row 3: -1; rows 0-2: -coef1
            ag += (double)(2 - q.coef * 6) * b[3] + (double)(2 * q.coef1 + 4 * q.coef3) * (b[0] + b[1] + b[2]);
            am += -b[3] - (double)q.coef1 * (b[0] + b[1] + b[2]);
        }
        else
        {
            const int jd = J.j[e];
            // y = x + e_d: judge <= 0 terms (and judge < 0 for row d); y = x - e_d: judge >= 0 / > 0
            const bool side = s > 0 ? jd <= 0 : jd >= 0;
            const bool strict = s > 0 ? jd < 0 : jd > 0;
            const int o1 = (e + 1) % 3, o2 = (e + 2) % 3;
            ag += (side ? (double)q.coef * b[3] : 0.0) + (strict ? (double)q.coef2 * b[e] : 0.0) -
                  (side ? (double)q.coef3 * (b[o1] + b[o2]) : 0.0);
            // gm[x + s e_d] enters row d with coef2 * judge (s = +1) or -coef2 * judge (s = -1)
            am += (double)q.coef2 * (double)(s * jd) * b[e];
        }
    }
    if (out == 0)
        lam1[id] += (float)ag;
    lam2[id] = (float)am;
}

__device__ __forceinline__ void adj_fdtd(const AdjScene &d, int blk)
{
    const int N = d.fp.N;
    const unsigned M = (unsigned)(N - 6), MM = M * M;
    const int nbI = (int)((MM * M + BLK - 1) / BLK);
    if (blk < nbI)
    {
        const unsigned u = (unsigned)blk * BLK + threadIdx.x;
        if (u >= MM * M)
            return;
        const unsigned x = u / MM, r = u - x * MM, y = r / M, z = r - y * M;
        const int id = lin3(3 + (int)x, 3 + (int)y, 3 + (int)z, N);
        const size_t nn = (size_t)N * N * N;
        const int t = d.ast->t;
        const float *lam = alevel(d.lam3, t, 0, nn);
        float *lam1 = alevel(d.lam3, t, -1, nn);
        float *lam2 = alevel(d.lam3, t, -2, nn);
        // the general path's operations, in its order
        const double cc = 2.0 - 6.0 * d.fp.coef / d.fp.hh, cn = d.fp.coef / d.fp.hh;
        const int sN2 = N * N;
        double ag = 0.0, am = 0.0;
        ag += cc * (double)lam[id];
        am -= (double)lam[id];
        ag += cn * (double)lam[id - sN2];
        ag += cn * (double)lam[id + sN2];
        ag += cn * (double)lam[id - N];
        ag += cn * (double)lam[id + N];
        ag += cn * (double)lam[id - 1];
        ag += cn * (double)lam[id + 1];
        lam1[id] += (float)ag;
        lam2[id] = (float)am;
        return;
    }
    // frame: the 6 outer x planes (6 N^2 cells), then in the other x planes the 6 outer y rows
    // (6 M N), then in the other (x, y) the 6 outer z cells (6 M^2); outer index k -> {0,1,2,N-3,N-2,N-1}
    unsigned f = (unsigned)(blk - nbI) * BLK + threadIdx.x;
    const unsigned uN = (unsigned)N, NN2 = uN * uN;
    auto outer = [&](unsigned k) { return k < 3 ? (int)k : N - 6 + (int)k; };
    int x, y, z;
    if (f < 6 * NN2)
    {
        const unsigned p = f / NN2, r = f - p * NN2;
        x = outer(p);
        y = (int)(r / uN);
        z = (int)(r - (unsigned)y * uN);
    }
    else if ((f -= 6 * NN2) < 6 * M * uN)
    {
        const unsigned xi = f / (6 * uN), r = f - xi * 6 * uN, q = r / uN;
        x = 3 + (int)xi;
        y = outer(q);
        z = (int)(r - q * uN);
    }
    else if ((f -= 6 * M * uN) < 6 * MM)
    {
        const unsigned xi = f / (6 * M), r = f - xi * 6 * M, yi = r / 6;
        x = 3 + (int)xi;
        y = 3 + (int)yi;
        z = outer(r - yi * 6);
    }
    else
        return;
    adj_fdtd_frame_cell(d, x, y, z);
}

__global__ void k_adj_back(AdjState *st) { st->t = st->t - 1; }

// the reverse step's kernels, in order (kind = index)
enum AdjKind
{
    kRead,
    kNear,
    kDirv,
    kDir,
    kFar,
    kAbc,
    kFdtd
};

template <int KIND>
__device__ __forceinline__ void adj_kernel(const AdjScene &d, int blk)
{
    if constexpr (KIND == kRead)
        adj_read(d, blk);
    else if constexpr (KIND == kNear)
        adj_near(d, blk);
    else if constexpr (KIND == kDirv)
        adj_dirv(d, blk);
    else if constexpr (KIND == kDir)
        adj_dir(d, blk);
    else if constexpr (KIND == kFar)
        adj_far(d, blk);
    else if constexpr (KIND == kAbc)
        adj_abc(d, blk);
    else
        adj_fdtd(d, blk);
}

template <int KIND>
__global__ __launch_bounds__(BLK) void k_adj(AdjScene d)
{
    adj_kernel<KIND>(d, (int)blockIdx.x);
}

// Batched: one launch covers the blocks of every scene; virtual block v belongs to scene
// m.scene_of[v] as its local block v - m.first[scene]. Blocks go to the 8 XCDs in turn and block
// b takes virtual block (b % 8) * m.per + b / 8, so each XCD works through one contiguous run of
// scenes.
template <int KIND>
__global__ __launch_bounds__(BLK) void k_batch_adj(const AdjScene *__restrict__ ds, BatchMap m)
{
    const int v = (int)(blockIdx.x & 7) * m.per + (int)(blockIdx.x >> 3);
    if (v >= m.total)
        return;
    const int sc = m.scene_of[v];
    adj_kernel<KIND>(ds[sc], v - m.first[sc]);
}

__global__ void k_batch_adj_back(const AdjScene *__restrict__ ds, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        ds[i].ast->t = ds[i].ast->t - 1;
}

}  // namespace

int adj_blocks(const AdjScene &d, int kind)
{
    const int N = d.fp.N;
    const long long S = N - 2;
    const long long shell = 2 * S * S + 2 * (S - 2) * S + 2 * (S - 2) * (S - 2);
    const long long nn = (long long)N * N * N;
    long long items = 0;
    switch (kind)
    {
    case kRead: items = d.n_c3 + 1; break;  // the extra thread: listeners outside the C3 cells
    case kNear: items = (long long)d.nt * STEPS; break;
    case kDirv: items = d.nt; break;
    case kDir:
    case kFar:  // the C3 part's blocks, then the triangle part's
        return (int)std::max<long long>(1, (d.n_c3 + BLK - 1) / BLK + ((long long)d.nt * STEPS + BLK - 1) / BLK);
    case kAbc: items = shell; break;
    default:  // the deep interior's blocks, then the frame's
    {
        const long long M = N - 6;
        return (int)((M * M * M + BLK - 1) / BLK + (nn - M * M * M + BLK - 1) / BLK);
    }
    }
    return (int)std::max<long long>(1, (items + BLK - 1) / BLK);
}

hipError_t enqueue_adjoint_step(const AdjScene &d, hipStream_t s)
{
    k_adj<kRead><<<adj_blocks(d, kRead), BLK, 0, s>>>(d);
    k_adj<kNear><<<adj_blocks(d, kNear), BLK, 0, s>>>(d);
    k_adj<kDirv><<<adj_blocks(d, kDirv), BLK, 0, s>>>(d);
    k_adj<kDir><<<adj_blocks(d, kDir), BLK, 0, s>>>(d);
    k_adj<kFar><<<adj_blocks(d, kFar), BLK, 0, s>>>(d);
    k_adj<kAbc><<<adj_blocks(d, kAbc), BLK, 0, s>>>(d);
    k_adj<kFdtd><<<adj_blocks(d, kFdtd), BLK, 0, s>>>(d);
    k_adj_back<<<1, 1, 0, s>>>(d.ast);
    return hipGetLastError();
}

hipError_t enqueue_batch_adjoint_step(const AdjScene *ds, const BatchMap *maps, int n, hipStream_t s)
{
    k_batch_adj<kRead><<<8 * maps[kRead].per, BLK, 0, s>>>(ds, maps[kRead]);
    k_batch_adj<kNear><<<8 * maps[kNear].per, BLK, 0, s>>>(ds, maps[kNear]);
    k_batch_adj<kDirv><<<8 * maps[kDirv].per, BLK, 0, s>>>(ds, maps[kDirv]);
    k_batch_adj<kDir><<<8 * maps[kDir].per, BLK, 0, s>>>(ds, maps[kDir]);
    k_batch_adj<kFar><<<8 * maps[kFar].per, BLK, 0, s>>>(ds, maps[kFar]);
    k_batch_adj<kAbc><<<8 * maps[kAbc].per, BLK, 0, s>>>(ds, maps[kAbc]);
    k_batch_adj<kFdtd><<<8 * maps[kFdtd].per, BLK, 0, s>>>(ds, maps[kFdtd]);
    k_batch_adj_back<<<(n + 255) / 256, 256, 0, s>>>(ds, n);
    return hipGetLastError();
}

}  // namespace pppm_amd
