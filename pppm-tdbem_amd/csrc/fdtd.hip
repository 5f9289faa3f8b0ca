// FDTD pressure update for gfx950: 7-point leapfrog stencil (FP64 arithmetic on FP32
// storage) fused with the first-order absorbing boundary, or the split-field PML.
//
// Reference behaviour: FDTD::step (src/algorithm/fdtd.h:73-87), fdtd_inner_kernel and
// laplacian (src/algorithm/fdtd.cu:7-33), ab_condition_solve / Solve4x4Linear
// (fdtd.cu:36-137), fdtd_boundary_kernel (fdtd.cu:139-150), pml_step_kernel (fdtd.cu:162-232).
//
// One launch covers the whole step: the first blocks stream the interior (indices
// 2..N-3 in every dimension, z contiguous so a wave reads 256 B rows), the remaining
// blocks solve the 4x4 ABC system for the index-1 / N-2 shell. Interior and shell
// write disjoint cells of level t+1; the shell also writes the ghost layer of level t,
// which no thread of this launch reads except the writing thread itself (masked
// terms of its own system), so the fusion is race-free. From an interior edge of 128 cells
// (grids beyond the Infinity Cache) the interior is instead marched along x by its own
// launch before the shell launch (k_fdtd_march). t is read from device memory (so the
// launches can be replayed from a graph); a one-thread kernel advances it after.
#include "internal.hpp"

#include <algorithm>

namespace pppm_amd
{

namespace
{

constexpr int BLK = 256;
// interior edge (N - 4) from which the interior is marched along x instead of one thread per cell
constexpr int kMarchMin = 128;

__device__ inline float *level(float *g3, int t, int rel, size_t nn)
{
    int l = ((t + rel) % 3 + 3) % 3;
    return g3 + (size_t)l * nn;
}

// Solve4x4Linear (fdtd.cu:36-86): cofactor inverse in FP32, FP64 reciprocal of the determinant.
__device__ inline void solve4(const float a[4][4], const float b[4], float res[4])
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
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        float r = 0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            r += inv[i][j] * b[j];
        res[i] = r;
    }
}

// ab_condition_solve (fdtd.cu:88-137) for one shell cell
__device__ inline void abc_cell(float *g, const float *gm, float *gp, int N, int i, int j, int k, const FdtdParams &p)
{
    const int co[3] = {i, j, k};
    int judge[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
        judge[d] = -((int)(co[d] == 1)) + ((int)(co[d] == N - 2));
    const int strd[3] = {N * N, N, 1};
    const int me = lin3(i, j, k, N);
    const float dt = p.dt, h = p.h, c = p.c;
    float a[4][4], b[4];
    float coef = c * c * dt * dt / (h * h);
    a[3][3] = 1;
#pragma unroll
    for (int d = 0; d < 3; d++)
        a[3][d] = -coef * (float)(judge[d] != 0);
    const float gc = g[me], gmc = gm[me];
    b[3] = -gmc + 2 * gc - coef * (6 * gc);
#pragma unroll
    for (int d = 0; d < 3; d++)
        b[3] += coef * g[me + strd[d]] * (float)(judge[d] <= 0) + coef * g[me - strd[d]] * (float)(judge[d] >= 0);
    float coef1 = 1 / (dt * dt), coef2 = c / (2 * h * dt), coef3 = -(c * c) / (2 * h * h);
#pragma unroll
    for (int idx = 0; idx < 3; idx++)
    {
        const int x = idx, y = (idx + 1) % 3, z = (idx + 2) % 3;
        a[idx][3] = coef1;
        a[idx][x] = coef2;
        a[idx][y] = coef3 * (float)(judge[y] != 0);
        a[idx][z] = coef3 * (float)(judge[z] != 0);
        b[idx] = -coef1 * (gmc - 2 * gc) +
                 -coef2 * (g[me + strd[x]] * (float)(-(int)(judge[x] < 0)) +
                           g[me - strd[x]] * (float)(-(int)(judge[x] > 0)) + gm[me + strd[x]] * (float)(-judge[x]) +
                           gm[me - strd[x]] * (float)judge[x]) +
                 -coef3 * (g[me + strd[y]] * (float)(judge[y] <= 0) + g[me - strd[y]] * (float)(judge[y] >= 0) +
                           g[me + strd[z]] * (float)(judge[z] <= 0) + g[me - strd[z]] * (float)(judge[z] >= 0) -
                           4 * gc);
    }
    float r[4];
    solve4(a, b, r);
    gp[me] = r[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (judge[d] != 0)
            g[me + judge[d] * strd[d]] = r[d];
}

// Leapfrog update of one interior cell from its six neighbours, in the reference's FP64
// operation order: sum x-, x+, y-, y+, z-, z+, then -6 * centre; / h^2; 2 p + coef lap - p_prev
__device__ __forceinline__ float leapfrog(float xm, float xp, float ym, float yp, float zm, float zp, float gc,
                                          float gmv, const FdtdParams &p)
{
    double sum = 0;
    sum += xm;
    sum += xp;
    sum += ym;
    sum += yp;
    sum += zm;
    sum += zp;
    sum -= 6 * gc;
    const double lap = sum / p.hh;
    double v = 2 * gc;
    v += p.coef * lap;
    v += -gmv;
    return (float)v;
}

// Interior of a large grid: a block owns MTY x MTZ (y, z) columns of one chunk of x planes and
// marches them along x, so the x-neighbours come from registers and the y/z neighbours of the
// current plane from cache: each level-t value is fetched from HBM about once instead of once
// per plane that reads it. A thread holds MCPT columns 64 apart, so a wave covers MTZ z-values
// of one row from a 256-byte boundary (cells outside 2..N-3 idle, reading a live cell); the 16
// rows of a block are marched by its 16 waves together, so only the two rows past the tile and
// the two cache lines past its ends come from other blocks. Distinct noalias parameters let the compiler issue the loads of the next planes ahead
// of this plane's store.
constexpr int MTY = 16;
constexpr int MBLK = 64 * MTY;
constexpr int MCPT = 4;            // z columns per thread, 64 apart: a tile is 16 rows x 256 z
constexpr int MTZ = 64 * MCPT;
__device__ __forceinline__ void march_columns(const float *__restrict__ g, const float *__restrict__ gm,
                                              float *__restrict__ gp, int N, int x0, int x1, int y, int z0,
                                              const FdtdParams &p)
{
    const int NN2 = N * N;
    bool on[MCPT];
    float gxm[MCPT], gc[MCPT], gxp[MCPT];
    int me = lin3(x0, y, 0, N);
#pragma unroll
    for (int c = 0; c < MCPT; c++)
    {
        const int z = z0 + 64 * c;
        on[c] = z >= 2 && z <= N - 3;
        const int zz = on[c] ? z : 2;
        gxm[c] = g[me + zz - NN2];
        gc[c] = g[me + zz];
        gxp[c] = g[me + zz + NN2];
    }
#pragma unroll 2
    for (int x = x0; x < x1; x++)
    {
#pragma unroll
        for (int c = 0; c < MCPT; c++)
        {
            const int z = on[c] ? z0 + 64 * c : 2;
            const int q = me + z;
            const float gym = g[q - N], gyp = g[q + N], gzm = g[q - 1], gzp = g[q + 1];
            const float gnx = g[q + 2 * NN2];  // x + 2 <= N - 1
            const float gmv = gm[q];
            const float v = leapfrog(gxm[c], gxp[c], gym, gyp, gzm, gzp, gc[c], gmv, p);
            if (on[c])
                gp[q] = v;
            gxm[c] = gc[c];
            gc[c] = gxp[c];
            gxp[c] = gnx;
        }
        me += NN2;
    }
}

// Marched-interior layout: tiles of MTY x MTZ columns times chunks of xch planes, tile index z
// fastest, then y, then the x chunk; n_tiles of them, launched as n_blocks (a multiple of 8).
struct MarchShape
{
    int ntz, nty, xch, n_tiles, n_blocks;
};

__global__ __launch_bounds__(MBLK) void k_fdtd_march(float *__restrict__ grid3, FdtdParams p,
                                                     const StepState *__restrict__ st, MarchShape ms,
                                                     const int *__restrict__ skip)
{
    if (skip && *skip)
        return;
    const int N = p.N;
    const size_t nn = (size_t)N * N * N;
    const int t = st->t;
    // blocks go to the 8 XCDs in turn; give each XCD a contiguous run of tiles so that
    // neighbouring tiles (and their shared halo rows) meet in one L2
    const int b = blockIdx.x, per = ms.n_blocks >> 3;
    const int tile = (b & 7) * per + (b >> 3);
    if (tile >= ms.n_tiles)
        return;
    const int tz = tile % ms.ntz, r = tile / ms.ntz;
    const int ty = r % ms.nty, xc = r / ms.nty;
    const int z0 = tz * MTZ + (int)(threadIdx.x & 63);
    const int y = 2 + ty * MTY + (int)(threadIdx.x >> 6);
    if (y > N - 3 || z0 > N - 3)
        return;
    const int x0 = 2 + xc * ms.xch;
    const int x1 = min(x0 + ms.xch, N - 2);
    march_columns(level(grid3, t, 0, nn), level(grid3, t, -1, nn), level(grid3, t, 1, nn), N, x0, x1, y, z0, p);
}

// Interior cells per thread of the thread-per-column path: XC consecutive x planes
constexpr int XC = 4;

// Interior columns (blocks [0, nb_inner)) and the ABC shell (the remaining blocks); nb_inner = 0
// when the interior is marched by another kernel. blk: the block's index within this scene's
// launch (or its share of a batched launch). An interior block covers 256 consecutive (y, z)
// cells of one chunk of XC x planes (a chunk takes ceil((N-4)^2 / 256) blocks; every index is
// 32-bit); each thread loads the XC + 2 values of its column, the y / z neighbours and the
// previous level of its XC cells before it writes any of them, so the loads of XC cells are in
// flight together.
__device__ __forceinline__ void fdtd_block(float *__restrict__ grid3, const FdtdParams &p, int t, int blk,
                                           int nb_inner)
{
    const int N = p.N;
    const size_t nn = (size_t)N * N * N;
    float *g = level(grid3, t, 0, nn);
    const float *gm = level(grid3, t, -1, nn);
    float *gp = level(grid3, t, 1, nn);
    if (blk < nb_inner)
    {
        const unsigned M = N - 4, MM = M * M;
        const unsigned bpp = (MM + BLK - 1) / BLK;
        const unsigned xb = (unsigned)blk / bpp;
        const unsigned q = ((unsigned)blk - xb * bpp) * BLK + threadIdx.x;
        if (q < MM)
        {
            const unsigned yq = q / M;
            const int y = 2 + (int)yq, z = 2 + (int)(q - yq * M), x0 = 2 + (int)xb * XC;
            const int NN2 = N * N;
            const int c0 = lin3(x0, y, z, N);
            float col[XC + 2], ym[XC], yp[XC], zm[XC], zp[XC], pm[XC];
#pragma unroll
            for (int i = 0; i < XC + 2; i++)
                col[i] = g[c0 + (min(x0 - 1 + i, N - 2) - x0) * NN2];  // planes past N - 2 are not used
#pragma unroll
            for (int i = 0; i < XC; i++)
            {
                const int me = c0 + (min(x0 + i, N - 3) - x0) * NN2;
                ym[i] = g[me - N];
                yp[i] = g[me + N];
                zm[i] = g[me - 1];
                zp[i] = g[me + 1];
                pm[i] = gm[me];
            }
#pragma unroll
            for (int i = 0; i < XC; i++)
                if (x0 + i <= N - 3)
                    // laplacian + update in FP64, sum order x-, x+, y-, y+, z-, z+, -6*centre (fdtd.cu:7-33)
                    gp[c0 + i * NN2] = leapfrog(col[i], col[i + 2], ym[i], yp[i], zm[i], zp[i], col[i + 1], pm[i], p);
        }
    }
    else
    {
        const int S = N - 2;
        // z faces in 16 x 16 tiles of (x, y), padded: a block's cells share their neighbour rows
        const int zt = (S - 2 + 15) / 16;
        const int fx = 2 * S * S, fy = 2 * (S - 2) * S, fz = 2 * zt * zt * 256;
        // blocks [nb_inner, nb_inner + nb_fxy): the x and y faces (rows along z, coalesced);
        // the rest: the z faces, one cell per row, so their blocks are handed to the XCDs in
        // contiguous runs of x planes and a plane's rows meet its neighbours' in one L2
        const int b = blk - nb_inner;
        const int nb_fxy = (fx + fy + BLK - 1) / BLK;
        int idx;
        if (b < nb_fxy)
        {
            idx = b * BLK + (int)threadIdx.x;
            if (idx >= fx + fy)
                idx = -1;
        }
        else
        {
            const int bz = b - nb_fxy, per = (fz + 8 * BLK - 1) / (8 * BLK);
            idx = fx + fy + ((bz & 7) * per + (bz >> 3)) * BLK + (int)threadIdx.x;
        }
        int i = -1, j = 0, k = 0;
        if (idx >= fx + fy + fz)
            idx = -1;
        if (idx < 0)
            ;
        else if (idx < fx)
        {
            const int side = idx / (S * S);
            const int r = idx - side * (S * S);
            i = side ? N - 2 : 1;
            j = 1 + r / S;
            k = 1 + r % S;
        }
        else if ((idx -= fx) < fy)
        {
            const int side = idx / ((S - 2) * S);
            const int r = idx - side * ((S - 2) * S);
            j = side ? N - 2 : 1;
            i = 2 + r / S;
            k = 1 + r % S;
        }
        else if ((idx -= fy) < fz)
        {
            const int side = idx / (zt * zt * 256);
            const int r = idx - side * (zt * zt * 256);
            const int tile = r >> 8, w = r & 255;
            k = side ? N - 2 : 1;
            i = 2 + (tile / zt) * 16 + (w >> 4);
            j = 2 + (tile % zt) * 16 + (w & 15);
            if (i > N - 3 || j > N - 3)
                i = -1;
        }
        if (i >= 0)
            abc_cell(g, gm, gp, N, i, j, k, p);
    }
}

// The FDTD of one scene (interior blocks, then the ABC shell) and, with a fuse, the history
// windows of the step's gathers (the last blocks; they read only the rings, which no FDTD block
// writes).
__global__ __launch_bounds__(BLK) void k_fdtd_abc(float *__restrict__ grid3, FdtdParams p, const StepState *__restrict__ st,
                                                  int nb_inner, int nb_fdtd, const int *__restrict__ skip, StepFuse f)
{
    // the update's status block for the host, before the skip: an update that raised an error
    // is reported too
    if (f.status_host && blockIdx.x == gridDim.x - 1 && (int)threadIdx.x < f.status_n)
        f.status_host[threadIdx.x] = f.status[threadIdx.x];
    if (skip && *skip)
        return;
    const int t = st->t;
    const int blk = (int)blockIdx.x;
    if (blk < nb_fdtd)
        fdtd_block(grid3, p, t, blk, nb_inner);
    else
    {
        // windows for the step that follows (t + 1): row k < 15 = ring value at t - k, row 15 =
        // ring slot t + 1 (overwritten by the far-field and Dirichlet launches before use)
        const int i = (blk - nb_fdtd) * BLK + (int)threadIdx.x;
        if (i < f.nt * STEPS)
        {
            const int tri = i / STEPS, k = i % STEPS;
            const int slot = (k == STEPS - 1 ? t + 1 : t - k) & (RING - 1);
            f.wn[i] = f.hn[tri * RING + slot];
            f.wd[i] = f.hd[tri * RING + slot];
        }
    }
}

// Interior of a small grid as LDS-staged 2.5D tiles: a block of TY x TZ threads owns TY x TZ
// (y, z) columns of one chunk of x planes and marches them along x. Per plane the block stages
// the level-t values of its tile plus a one-cell halo in y and z in LDS (double buffered), so
// each level-t value comes from memory about once ((TY+2)(TZ+2) / (TY TZ) with the halo); the
// x-neighbours stay in registers. Threads outside 2..N-3 only help with the staging. The
// update is the same leapfrog on the same seven values, so the results equal the one thread
// per cell kernel's bit for bit.
template <int TY, int TZ>
__device__ __forceinline__ void tile_march(float *__restrict__ grid3, const FdtdParams &p, int t, int tile, int xch,
                                           float *__restrict__ lds)
{
    constexpr int LY = TY + 2, LZ = TZ + 2, LN = LY * LZ, NT = TY * TZ;
    const int N = p.N, M = N - 4;
    const size_t nn = (size_t)N * N * N;
    const float *__restrict__ g = level(grid3, t, 0, nn);
    const float *__restrict__ gm = level(grid3, t, -1, nn);
    float *__restrict__ gp = level(grid3, t, 1, nn);
    const int ntz = (M + TZ - 1) / TZ, nty = (M + TY - 1) / TY;
    const int tz = tile % ntz, r = tile / ntz;
    const int ty = r % nty, xc = r / nty;
    const int y0 = 2 + ty * TY, z0 = 2 + tz * TZ;
    const int x0 = 2 + xc * xch, x1 = min(x0 + xch, N - 2);
    const int tid = threadIdx.x;
    const int ly = tid / TZ, lz = tid % TZ;
    const int y = y0 + ly, z = z0 + lz;
    const bool on = y <= N - 3 && z <= N - 3;
    const int me_l = (ly + 1) * LZ + lz + 1;
    // stage plane x (tile plus halo, clamped into the grid) into buf
    auto stage = [&](int x, float *buf) {
        const float *pl = g + (size_t)x * N * N;
        for (int i = tid; i < LN; i += NT)
        {
            const int gy = min(max(y0 - 1 + i / LZ, 0), N - 1), gz = min(max(z0 - 1 + i % LZ, 0), N - 1);
            buf[i] = pl[gy * N + gz];
        }
    };
    float *cur = lds, *nxt = lds + LN;
    stage(x0, cur);
    const int yy = on ? y : 2, zz = on ? z : 2;
    float gxm = g[lin3(x0 - 1, yy, zz, N)];
    __syncthreads();
    float gc = cur[me_l];
    for (int x = x0; x < x1; x++)
    {
        stage(x + 1, nxt);  // x + 1 <= N - 2: inside the grid
        const int q = lin3(x, yy, zz, N);
        const float gmv = gm[q];
        __syncthreads();
        const float gxp = nxt[me_l];
        const float v = leapfrog(gxm, gxp, cur[me_l - LZ], cur[me_l + LZ], cur[me_l - 1], cur[me_l + 1], gc, gmv, p);
        if (on)
            gp[q] = v;
        gxm = gc;
        gc = gxp;
        __syncthreads();  // every thread is done with cur before it is staged again
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
}

template <int TY, int TZ>
__global__ __launch_bounds__(TY *TZ) void k_batch_tile(const FdtdScene *__restrict__ ds, BatchMap m)
{
    __shared__ float lds[2 * (TY + 2) * (TZ + 2)];
    const int v = (int)(blockIdx.x & 7) * m.per + (int)(blockIdx.x >> 3);
    if (v >= m.total)
        return;
    const int sc = m.scene_of[v];
    const FdtdScene &d = ds[sc];
    tile_march<TY, TZ>(d.grid3, d.p, d.st->t, v - m.first[sc], d.xch, lds);
}

// the ABC shell of every scene, plus the interior of the scenes stepped one thread per cell
__global__ __launch_bounds__(BLK) void k_batch_fdtd(const FdtdScene *__restrict__ ds, BatchMap m)
{
    const int v = (int)(blockIdx.x & 7) * m.per + (int)(blockIdx.x >> 3);
    if (v >= m.total)
        return;
    const int sc = m.scene_of[v];
    const FdtdScene &d = ds[sc];
    fdtd_block(d.grid3, d.p, d.st->t, v - m.first[sc], d.nb_inner);
}

// get_damp (fdtd.cu:162-169)
__device__ inline float pml_damp(const FdtdParams &p, int x, int y, int z)
{
    const int N = p.N, w = p.pml;
    if (x < w || x >= N - w || y < w || y >= N - w || z < w || z >= N - w)
        return p.damp;
    return 0.0f;
}

// pml_step_kernel (fdtd.cu:179-232); PML state: 6 floats per cell (v1[3], v2[3]) per level
__global__ __launch_bounds__(BLK) void k_fdtd_pml(float *__restrict__ grid3, float *__restrict__ pml3, FdtdParams p,
                                                  const StepState *__restrict__ st, const int *__restrict__ skip)
{
    if (skip && *skip)
        return;
    const int N = p.N;
    const size_t nn = (size_t)N * N * N;
    const int t = st->t;
    const float *g = level(grid3, t, 0, nn);
    const float *gm = level(grid3, t, -1, nn);
    float *gp = level(grid3, t, 1, nn);
    const int l0 = ((t % 3) + 3) % 3, l1 = (((t + 1) % 3) + 3) % 3;
    const float *P = pml3 + (size_t)l0 * nn * 6;
    float *Pn = pml3 + (size_t)l1 * nn * 6;
    const int S = N - 2;
    const long long idx = (long long)blockIdx.x * BLK + threadIdx.x;
    if (idx < (long long)S * S * S)
    {
        const int x = 1 + (int)(idx / (S * S));
        const int rem = (int)(idx % (S * S));
        const int y = 1 + rem / S, z = 1 + rem % S;
        const int co[3] = {x, y, z};
        const int strd[3] = {N * N, N, 1};
        int sg[3];
#pragma unroll
        for (int i = 0; i < 3; i++)
            sg[i] = co[i] > N / 2 ? 1 : -1;
        const float cd = pml_damp(p, x, y, z);
        const int me = lin3(x, y, z, N);
        const float c = p.c, h = p.h, dt = p.dt;
        auto dmp = [&](int i, int d) {
            int q[3] = {x, y, z};
            q[i] += d;
            return pml_damp(p, q[0], q[1], q[2]);
        };
        float rhs = 0;
#pragma unroll
        for (int i = 0; i < 3; i++)
        {
            rhs += 1 / (h * h) * (g[me + strd[i]] - 2 * g[me] + g[me - strd[i]]);
            rhs += 1 / h *
                   (cd * P[(size_t)(me + sg[i] * strd[i]) * 6 + 3 + i] -
                    dmp(i, -sg[i]) * P[(size_t)(me - sg[i] * strd[i]) * 6 + i]);
        }
        gp[me] = g[me] * 2 - gm[me] + c * c * dt * dt * rhs;
        if (cd != 0)
        {
#pragma unroll
            for (int i = 0; i < 3; i++)
            {
                float r = 0;
                r += -1.0f / 2.0f * (dmp(i, -sg[i]) * P[(size_t)(me - sg[i] * strd[i]) * 6 + i] + cd * P[(size_t)me * 6 + i]);
                r += -1.0f / (h * 2.0f) * (g[me + sg[i] * strd[i]] - g[me - sg[i] * strd[i]]);
                Pn[(size_t)me * 6 + i] = P[(size_t)me * 6 + i] + c * dt * r;
            }
#pragma unroll
            for (int i = 0; i < 3; i++)
            {
                float r = 0;
                r += -1.0f / 2.0f *
                     (dmp(i, -sg[i]) * P[(size_t)me * 6 + 3 + i] + cd * P[(size_t)(me + sg[i] * strd[i]) * 6 + 3 + i]);
                r += -1.0f / (h * 2.0f) * (g[me + sg[i] * strd[i]] - g[me - sg[i] * strd[i]]);
                Pn[(size_t)me * 6 + 3 + i] = P[(size_t)me * 6 + 3 + i] + c * dt * r;
            }
        }
    }
}

}  // namespace

// (TY, TZ) of the tile variants (FdtdVariant::Tile*)
static void tile_shape(int variant, int &ty, int &tz)
{
    switch (variant)
    {
    case 1: ty = 8, tz = 64; break;
    case 2: ty = 16, tz = 32; break;
    case 3: ty = 4, tz = 64; break;
    default: ty = 8, tz = 32; break;
    }
}

int fdtd_shell_blocks(int N)
{
    const long long S = N - 2;
    const long long zt = (S - 2 + 15) / 16;
    const long long fxy = 2 * S * S + 2 * (S - 2) * S, fz = 2 * zt * zt * 256;
    return (int)((fxy + BLK - 1) / BLK) + 8 * (int)((fz + 8LL * BLK - 1) / (8LL * BLK));
}

int fdtd_inner_blocks(int N)
{
    const long long M = N - 4;
    return (int)(((M + XC - 1) / XC) * ((M * M + BLK - 1) / BLK));
}

int fdtd_tile_blocks(int N, int variant, int &xch)
{
    int ty, tz;
    tile_shape(variant, ty, tz);
    const int M = N - 4;
    const long long cols = (long long)((M + ty - 1) / ty) * ((M + tz - 1) / tz);
    // chunks of at least 8 planes, enough tiles that a scene alone spreads over the CUs
    const long long nxc = std::max<long long>(1, std::min<long long>((M + 7) / 8, (512 + cols - 1) / cols));
    xch = (int)((M + nxc - 1) / nxc);
    return (int)(cols * ((M + xch - 1) / xch));
}

hipError_t launch_batch_fdtd(const FdtdScene *ds, const BatchMap &m, hipStream_t s)
{
    k_batch_fdtd<<<8 * m.per, BLK, 0, s>>>(ds, m);
    return hipGetLastError();
}

hipError_t launch_batch_tile(int variant, const FdtdScene *ds, const BatchMap &m, hipStream_t s)
{
    switch (variant)
    {
    case 1: k_batch_tile<8, 64><<<8 * m.per, 8 * 64, 0, s>>>(ds, m); break;
    case 2: k_batch_tile<16, 32><<<8 * m.per, 16 * 32, 0, s>>>(ds, m); break;
    case 3: k_batch_tile<4, 64><<<8 * m.per, 4 * 64, 0, s>>>(ds, m); break;
    default: k_batch_tile<8, 32><<<8 * m.per, 8 * 32, 0, s>>>(ds, m); break;
    }
    return hipGetLastError();
}

hipError_t launch_fdtd(float *grid3, float *pml3, const FdtdParams &p, const StepState *st, hipStream_t s,
                       const int *skip, const StepFuse *fuse)
{
    const int N = p.N;
    if (p.pml > 0)
    {
        long long cells = (long long)(N - 2) * (N - 2) * (N - 2);
        int nb = (int)((cells + BLK - 1) / BLK);
        k_fdtd_pml<<<nb, BLK, 0, s>>>(grid3, pml3, p, st, skip);
        return hipGetLastError();
    }
    const long long M = N - 4, S = N - 2;
    int nb_inner = fdtd_inner_blocks(N);
    if (M >= kMarchMin)
    {
        // about 2048 tiles or more; chunks of at least 16 planes (each chunk re-reads 2 planes)
        MarchShape ms{};
        ms.ntz = (N - 3) / MTZ + 1;  // z tiles [MTZ tz, MTZ tz + MTZ) up to z = N - 3
        ms.nty = (int)((M + MTY - 1) / MTY);
        const long long cols = (long long)ms.ntz * ms.nty;
        const long long nxc = std::max<long long>(1, (2048 + cols - 1) / cols);
        ms.xch = (int)std::max<long long>(16, (M + nxc - 1) / nxc);
        ms.n_tiles = (int)(cols * ((M + ms.xch - 1) / ms.xch));
        ms.n_blocks = (ms.n_tiles + 7) / 8 * 8;
        k_fdtd_march<<<ms.n_blocks, MBLK, 0, s>>>(grid3, p, st, ms, skip);
        nb_inner = 0;
    }
    const long long zt = (S - 2 + 15) / 16;
    const long long fxy = 2 * S * S + 2 * (S - 2) * S, fz = 2 * zt * zt * 256;
    const int nb_shell = (int)((fxy + BLK - 1) / BLK) + 8 * (int)((fz + 8LL * BLK - 1) / (8LL * BLK));
    StepFuse f{};
    int nb_win = 0;
    if (fuse && fuse->wn)
    {
        f = *fuse;
        nb_win = (int)(((long long)f.nt * STEPS + BLK - 1) / BLK);
    }
    k_fdtd_abc<<<nb_inner + nb_shell + nb_win, BLK, 0, s>>>(grid3, p, st, nb_inner, nb_inner + nb_shell, skip, f);
    return hipGetLastError();
}

}  // namespace pppm_amd
