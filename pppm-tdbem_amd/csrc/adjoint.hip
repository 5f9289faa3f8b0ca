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
__global__ __launch_bounds__(BLK) void k_adj_read(AdjScene d)
{
    const int t = d.ast->t;
    const size_t nn = (size_t)d.N * d.N * d.N;
    float *lam = alevel(d.lam3, t, 0, nn);
    const int c = blockIdx.x * BLK + threadIdx.x;
    const float *go = d.gout ? d.gout + (size_t)(t - d.t_first) * d.n_listen : nullptr;
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

// 2. near field transposed, one thread per (triangle, tap): the 27 records of a triangle are
// contiguous, so the 16 threads of a triangle read 64 B of each record
__global__ __launch_bounds__(BLK) void k_adj_near(AdjScene d)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= d.nt * STEPS)
        return;
    const int t = d.ast->t;
    const int tri = i / STEPS, k = i % STEPS;
    float an = 0.0f, ad = 0.0f;
    for (int o = 0; o < GC; o++)
    {
        const int rec = tri * GC + o;
        const float u = d.unear[d.tc3[rec]];
        const uint32_t w = d.wb[(size_t)rec * STEPS + k];
        const float2 sc = d.sb[rec];
        an += wS(w, sc) * u;
        ad += wD(w, sc) * u;
    }
    const int slot = tri * RING + ((t - k) & (RING - 1));
    d.ln[slot] += -0.98f * an;
    d.ld[slot] += 0.98f * ad;
}

// 3. the Dirichlet update's divisor
__global__ __launch_bounds__(BLK) void k_adj_dirv(AdjScene d)
{
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c >= d.nt)
        return;
    const int slot = c * RING + (d.ast->t & (RING - 1));
    d.v[c] = (float)((double)d.ld[slot] / d.den[c]);
    d.ld[slot] = 0.0f;
}

// 4. Dirichlet transposed: C3 cells [0, n_c3) gather the face-far-field interpolation; then one
// thread per (triangle j, tap) gathers the face pairs (c, j)
__global__ __launch_bounds__(BLK) void k_adj_dir(AdjScene d)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    const int t = d.ast->t;
    if (i < d.n_c3)
    {
        float a = d.unear[i];
        for (int e = d.ct_start[i]; e < d.ct_start[i + 1]; e++)
            a += d.ct_w[e] * d.v[d.ct_c[e]];
        d.lfar[i] = a;
        return;
    }
    const int q = i - d.n_c3;
    if (q >= d.nt * STEPS)
        return;
    const int j = q / STEPS, k = q % STEPS;
    float an = 0.0f, ad = 0.0f;
    for (int e = d.ft_start[j]; e < d.ft_start[j + 1]; e++)
    {
        const int slot = d.ft_slot[e];
        const float vc = d.v[d.ft_c[e]];
        const uint32_t w = d.fw[(size_t)slot * STEPS + k];
        const float2 sc = d.fsc[slot];
        an += wS(w, sc) * vc;
        ad += wD(w, sc) * vc;
    }
    const int rs = j * RING + ((t - k) & (RING - 1));
    d.ln[rs] += -an;
    if (k > 0)  // D_j[t] enters the face near field as 0 (pppm.cu:143)
        d.ld[rs] += ad;
}

// 5. far field transposed; the (triangle, tap 0) threads write the finished gradient of N(t)
__global__ __launch_bounds__(BLK) void k_adj_far(AdjScene d)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    const int t = d.ast->t;
    const size_t nn = (size_t)d.N * d.N * d.N;
    if (i < d.n_c3)
    {
        float *lam = alevel(d.lam3, t, 0, nn);
        lam[d.c3_cells[i]] += d.lfar[i];
        return;
    }
    const int q = i - d.n_c3;
    if (q >= d.nt * STEPS)
        return;
    const int tri = q / STEPS, k = q % STEPS;
    if (k == 0)
    {
        const int slot = tri * RING + (t & (RING - 1));
        d.grad[(size_t)(t - d.t_first) * d.grad_stride + tri] = d.ln[slot];
        d.ln[slot] = 0.0f;
        return;
    }
    float an = 0.0f, ad = 0.0f;
    for (int o = 0; o < GC; o++)
    {
        const int rec = tri * GC + o;
        const float lf = d.lfar[d.tc3[rec]];
        const uint32_t w = d.wf[(size_t)rec * STEPS + k];
        const float2 sc = d.sf[rec];
        an += wS(w, sc) * lf;
        ad += wD(w, sc) * lf;
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
__global__ __launch_bounds__(BLK) void k_adj_abc(AdjScene d)
{
    const int N = d.fp.N;
    const long long S = N - 2;
    const long long fx = 2 * S * S, fy = 2 * (S - 2) * S, fz = 2 * (S - 2) * (S - 2);
    long long idx = (long long)blockIdx.x * BLK + threadIdx.x;
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
// (shell cells that read it, with the coefficients of fdtd.cu:94-118 for each neighbour slot)
__global__ __launch_bounds__(BLK) void k_adj_fdtd(AdjScene d)
{
    const int N = d.fp.N;
    const size_t nn = (size_t)N * N * N;
    const long long id = (long long)blockIdx.x * BLK + threadIdx.x;
    if (id >= (long long)nn)
        return;
    const int y0 = (int)(id / ((long long)N * N)), y1 = (int)(id / N % N), y2 = (int)(id % N);
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

__global__ void k_adj_back(AdjState *st) { st->t = st->t - 1; }

}  // namespace

int adj_c3_blocks(const AdjScene &d) { return (d.n_c3 + 1 + BLK - 1) / BLK; }

hipError_t enqueue_adjoint_step(const AdjScene &d, hipStream_t s)
{
    const int N = d.fp.N;
    const long long S = N - 2;
    const long long shell = 2 * S * S + 2 * (S - 2) * S + 2 * (S - 2) * (S - 2);
    const long long nn = (long long)N * N * N;
    const int tri_blocks = (d.nt * STEPS + BLK - 1) / BLK;
    const int c3_blocks = (d.n_c3 + BLK - 1) / BLK;
    k_adj_read<<<adj_c3_blocks(d), BLK, 0, s>>>(d);
    k_adj_near<<<tri_blocks, BLK, 0, s>>>(d);
    k_adj_dirv<<<(d.nt + BLK - 1) / BLK, BLK, 0, s>>>(d);
    k_adj_dir<<<(int)((d.n_c3 + (long long)d.nt * STEPS + BLK - 1) / BLK), BLK, 0, s>>>(d);
    k_adj_far<<<(int)((d.n_c3 + (long long)d.nt * STEPS + BLK - 1) / BLK), BLK, 0, s>>>(d);
    k_adj_abc<<<(int)((shell + BLK - 1) / BLK), BLK, 0, s>>>(d);
    k_adj_fdtd<<<(int)((nn + BLK - 1) / BLK), BLK, 0, s>>>(d);
    k_adj_back<<<1, 1, 0, s>>>(d.ast);
    (void)c3_blocks;
    return hipGetLastError();
}

}  // namespace pppm_amd
