// Deterministic exclusive prefix sums of int arrays, used for the CSR layouts of the
// triangle bins and the 3x3x3 / 4x4x4 lists, and for the compaction of selected cells,
// triangles and face pairs (the reference uses thrust::remove_if for the same compactions,
// src/core/array.h:77-115).
//
// Reduce-then-scan with no inter-block waiting: the first kernel writes one sum per 1024-item
// tile, the second has every tile add up the sums of the tiles before it (one wave), scan its
// own items and, in the last tile, write the total. Past 512 tiles in a job that per-tile sum
// grows with the square of the tile count, so the tile offsets come instead from a scan of the
// tile sums (the same routine, one level down). Several independent arrays ("jobs") share the
// launches; jobs of at most 8192 items take a single launch.
// A job in flag mode scans (in != 0) and can write the compacted list of selected indices
// and an index map (rank or -1) in the same pass. Integer sums are exact, so the result does
// not depend on the order of additions. Only kernels are enqueued (graph-capturable).
#include "internal.hpp"

#include <algorithm>

namespace pppm_amd
{

namespace
{
constexpr int SB = 256;        // threads per block
constexpr int SI = 4;          // items per thread (multi-tile scans)
constexpr int SPB = SB * SI;   // items per tile
static_assert(SPB == SCAN_TILE, "tile size shared with the producers of tile sums");
constexpr int SI1 = 32;        // items per thread when one tile holds everything
constexpr int kDirectTiles = 512;  // most tiles per job whose offsets the tiles sum themselves

__device__ inline int wave_incl(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off)
            v += u;
    }
    return v;
}

__device__ inline int wave_total(int v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// index of the job that owns global tile `tile` (jobs are laid out in tile order)
__device__ inline int job_index(const ScanBatch &b, int tile)
{
    int k = 0;
    for (int q = 1; q < b.njobs; q++)
        if (tile >= b.job[q].tile0)
            k = q;
    return k;
}

__device__ inline const ScanJob &job_of(const ScanBatch &b, int tile) { return b.job[job_index(b, tile)]; }

__device__ inline int item_of(const ScanJob &J, long long q)
{
    const int x = q < J.n ? J.in[q] : 0;
    return J.flag_mode ? (x != 0 ? 1 : 0) : x;
}

// LDS slot of tile item j: one pad word per 32, so the blocked reads (thread t reads items
// t*IT .. t*IT+IT-1) hit distinct banks
__device__ inline int pad(int j) { return j + (j >> 5); }

// block-wide exclusive offset of this thread's items and the block total
__device__ inline int block_scan(int tsum, int &btotal)
{
    __shared__ int wtot[SB / 64];
    const int incl = wave_incl(tsum);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63)
        wtot[wid] = incl;
    __syncthreads();
    int woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SB / 64; w++)
    {
        if (w < wid)
            woff += wtot[w];
        tot += wtot[w];
    }
    btotal = tot;
    return woff + incl - tsum;
}

// tile sums: striped (coalesced) loads, any order of integer addition
__global__ __launch_bounds__(SB) void k_tile_sums(ScanBatch b, int *__restrict__ sums)
{
    const ScanJob &J = job_of(b, blockIdx.x);
    const long long base = (long long)(blockIdx.x - J.tile0) * SPB;
    int tsum = 0;
#pragma unroll
    for (int i = 0; i < SI; i++)
        tsum += item_of(J, base + i * SB + threadIdx.x);
    int btotal;
    (void)block_scan(tsum, btotal);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = btotal;
}

// sums == null: every job is a single tile (one block per job). offs != null: the exclusive
// offset of tile t of job k within its job is offs[t + k]; otherwise it is the sum of sums[]
// over the job's earlier tiles. IT items per thread. Items are loaded striped (coalesced)
// into LDS, scanned in blocked order, and the positions go back through LDS for striped stores.
template <int IT>
__global__ __launch_bounds__(SB) void k_tile_scan(ScanBatch b, const int *__restrict__ sums,
                                                  const int *__restrict__ offs)
{
    constexpr int T = SB * IT;
    __shared__ int s_off;
    __shared__ int sv[T + T / 32];
    const ScanJob &J = sums ? job_of(b, blockIdx.x) : b.job[blockIdx.x];
    const int tile = sums ? (int)blockIdx.x - J.tile0 : 0;
    const long long base = (long long)tile * T;
    const int tid = threadIdx.x;
    // this thread's items are loaded first, so their loads overlap the tile offset's
    int it[IT];
#pragma unroll
    for (int i = 0; i < IT; i++)
        it[i] = item_of(J, base + i * SB + tid);
    if (offs)
    {
        if (tid == 0)
            s_off = offs[blockIdx.x + job_index(b, blockIdx.x)];
    }
    else if (tid < 64)
    {
        // the sums of the job's earlier tiles (at most kDirectTiles of them on this path): all
        // loads issued before the first is added
        int acc = 0;
        if (sums)
        {
            int part[kDirectTiles / 64];
#pragma unroll
            for (int i = 0; i < kDirectTiles / 64; i++)
            {
                const int t = J.tile0 + i * 64 + tid;
                part[i] = t < (int)blockIdx.x ? sums[t] : 0;
            }
#pragma unroll
            for (int i = 0; i < kDirectTiles / 64; i++)
                acc += part[i];
        }
        acc = wave_total(acc);
        if (tid == 0)
            s_off = acc;
    }
#pragma unroll
    for (int i = 0; i < IT; i++)
        sv[pad(i * SB + tid)] = it[i];
    __syncthreads();
    int v[IT], tsum = 0;
#pragma unroll
    for (int i = 0; i < IT; i++)
    {
        v[i] = sv[pad(tid * IT + i)];
        tsum += v[i];
    }
    int btotal;
    int run = block_scan(tsum, btotal);  // its barrier also publishes s_off
    const int off = s_off;
    run += off;
    // blocked pass: compaction lists, and (position << 1 | selected) back into LDS
#pragma unroll
    for (int i = 0; i < IT; i++)
    {
        const long long q = base + tid * IT + i;
        if (J.list && v[i] && q < J.n)
            J.list[run] = (int)q;
        sv[pad(tid * IT + i)] = (run << 1) | (v[i] != 0 ? 1 : 0);
        run += v[i];
    }
    __syncthreads();
    if (J.pos || J.index_map)
#pragma unroll
        for (int i = 0; i < IT; i++)
        {
            const long long q = base + i * SB + tid;
            if (q < J.n)
            {
                const int e = sv[pad(i * SB + tid)];
                if (J.pos)
                    J.pos[q] = e >> 1;
                if (J.index_map)
                    J.index_map[q] = (e & 1) ? (e >> 1) : -1;
            }
        }
    const long long tiles = ((long long)J.n + T - 1) / T;
    if (tile == (tiles > 0 ? tiles - 1 : 0) && tid == 0)
    {
        if (J.pos)
            J.pos[J.n] = off + btotal;
        if (J.total)
            *J.total = off + btotal;
    }
}

}  // namespace

namespace
{
long long tiles_of(long long n) { return std::max<long long>(1, (n + SPB - 1) / SPB); }

// ints of tmp a batch with these job lengths uses: the tile sums, and past kDirectTiles tiles in
// a job, the tile offsets and the tmp of their own scan
size_t tmp_ints(const long long *n, int njobs)
{
    long long t = 0, most = 0, inner[SCAN_MAX_JOBS];
    for (int k = 0; k < njobs; k++)
    {
        inner[k] = tiles_of(n[k]);
        t += inner[k];
        most = std::max(most, inner[k]);
    }
    size_t need = (size_t)t + 16;
    if (most > kDirectTiles)
        need += (size_t)t + njobs + 16 + tmp_ints(inner, njobs);
    return need;
}
}  // namespace

size_t scan_tmp_ints(int n)
{
    const long long nn = n;
    return tmp_ints(&nn, 1);
}

size_t scan_batch_tmp_ints(const int *n, int njobs)
{
    long long nn[SCAN_MAX_JOBS];
    for (int k = 0; k < njobs && k < SCAN_MAX_JOBS; k++)
        nn[k] = n[k];
    return tmp_ints(nn, std::min(njobs, SCAN_MAX_JOBS));
}

// Jobs with at most SB * SI1 items each run in one launch (a block per job); otherwise a
// tile-sums launch and a tile-scan launch cover every job's tiles.
int scan_batch_layout(ScanBatch &b)
{
    int tiles = 0;
    for (int k = 0; k < b.njobs; k++)
    {
        b.job[k].tile0 = tiles;
        tiles += (int)std::max<long long>(1, ((long long)b.job[k].n + SPB - 1) / SPB);
    }
    return tiles;
}

bool scan_batch_single(const ScanBatch &b)
{
    for (int k = 0; k < b.njobs; k++)
        if (b.job[k].n > SB * SI1)
            return false;
    return true;
}

hipError_t scan_batch(ScanBatch b, int *tmp, hipStream_t s, const int *presums)
{
    if (b.njobs < 1 || b.njobs > SCAN_MAX_JOBS)
        return hipErrorInvalidValue;
    for (int k = 0; k < b.njobs; k++)
        if (b.job[k].n < 0 || b.job[k].n >= (1 << 30))  // positions are kept as (pos << 1 | bit)
            return hipErrorInvalidValue;
    const int tiles = scan_batch_layout(b);
    if (scan_batch_single(b))
    {
        k_tile_scan<SI1><<<b.njobs, SB, 0, s>>>(b, nullptr, nullptr);
        return hipGetLastError();
    }
    if (!presums)
    {
        k_tile_sums<<<tiles, SB, 0, s>>>(b, tmp);
        presums = tmp;
    }
    long long most = 0;
    for (int k = 0; k < b.njobs; k++)
        most = std::max(most, tiles_of(b.job[k].n));
    int *offs = nullptr;
    if (most > kDirectTiles)
    {
        // job k's tile offsets (and its total) land at offs + tile0 + k .. + tiles
        offs = tmp + tiles + 16;
        ScanBatch inner{};
        inner.njobs = b.njobs;
        for (int k = 0; k < b.njobs; k++)
            inner.job[k] = ScanJob{presums + b.job[k].tile0, (int)tiles_of(b.job[k].n), 0, offs + b.job[k].tile0 + k,
                                   nullptr, nullptr, nullptr, 0};
        const hipError_t e = scan_batch(inner, offs + tiles + b.njobs + 16, s, nullptr);
        if (e != hipSuccess)
            return e;
    }
    k_tile_scan<SI><<<tiles, SB, 0, s>>>(b, presums, offs);
    return hipGetLastError();
}

// out must hold n+1 ints; tmp must hold scan_tmp_ints(n) ints.
hipError_t scan_exclusive(const int *in, int *out, int n, int *tmp, hipStream_t s)
{
    if (n < 0)
        return hipErrorInvalidValue;
    ScanBatch b{};
    b.njobs = 1;
    b.job[0].in = in;
    b.job[0].n = n;
    b.job[0].pos = out;
    return scan_batch(b, tmp, s, nullptr);
}

}  // namespace pppm_amd
