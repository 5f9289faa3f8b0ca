// Deterministic exclusive prefix sum of int arrays, used for the CSR layouts of the
// triangle bins, the 3x3x3 / 4x4x4 lists and the compaction of non-empty cells (the
// reference uses thrust::remove_if for the same compactions, src/core/array.h:77-115).
//
// Reduce-then-scan in two launches and no inter-block waiting: the first kernel writes one
// sum per 4096-item tile, the second has every tile add up the sums of the tiles before it
// (at most a few hundred ints, one wave), scan its own items and, in the last tile, write
// the total to out[n]. Arrays up to 8192 items take a single launch. Integer sums are exact,
// so the result does not depend on the order of additions. Only kernels are enqueued, so
// the scan can be captured in a HIP graph.
#include "internal.hpp"

namespace pppm_amd
{

namespace
{
constexpr int SB = 256;        // threads per block
constexpr int SI = 16;         // items per thread (multi-tile scans)
constexpr int SPB = SB * SI;   // items per tile
constexpr int SI1 = 32;        // items per thread when one tile holds everything

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

template <int IT>
__device__ inline void load_tile(const int *__restrict__ in, long long n, int v[IT], int &tsum)
{
    const long long base = (long long)blockIdx.x * SB * IT + (long long)threadIdx.x * IT;
    tsum = 0;
#pragma unroll
    for (int i = 0; i < IT; i++)
    {
        const long long q = base + i;
        v[i] = q < n ? in[q] : 0;
        tsum += v[i];
    }
}

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

__global__ __launch_bounds__(SB) void k_tile_sums(const int *__restrict__ in, long long n, int *__restrict__ sums)
{
    int v[SI], tsum, btotal;
    load_tile<SI>(in, n, v, tsum);
    (void)block_scan(tsum, btotal);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = btotal;
}

// sums == null: one tile, no offset. IT items per thread.
template <int IT>
__global__ __launch_bounds__(SB) void k_tile_scan(const int *__restrict__ in, int *__restrict__ out, long long n,
                                                  const int *__restrict__ sums)
{
    __shared__ int s_off;
    if (threadIdx.x < 64)
    {
        int acc = 0;
        if (sums)
            for (int b0 = 0; b0 < (int)blockIdx.x; b0 += 64)
            {
                const int b = b0 + (int)threadIdx.x;
                acc += b < (int)blockIdx.x ? sums[b] : 0;
            }
        acc = wave_total(acc);
        if (threadIdx.x == 0)
            s_off = acc;
    }
    int v[IT], tsum, btotal;
    load_tile<IT>(in, n, v, tsum);
    int run = block_scan(tsum, btotal);  // its barrier also publishes s_off
    const int off = s_off;
    run += off;
    const long long base = (long long)blockIdx.x * SB * IT + (long long)threadIdx.x * IT;
#pragma unroll
    for (int i = 0; i < IT; i++)
    {
        const long long q = base + i;
        if (q < n)
            out[q] = run;
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        out[n] = off + btotal;
}

}  // namespace

size_t scan_tmp_ints(int n)
{
    const long long nb = ((long long)n + SPB - 1) / SPB;
    return (size_t)(nb > 1 ? nb : 1) + 16;
}

// out must hold n+1 ints; tmp must hold scan_tmp_ints(n) ints.
hipError_t scan_exclusive(const int *in, int *out, int n, int *tmp, hipStream_t s)
{
    if (n < 0)
        return hipErrorInvalidValue;
    if (n <= SB * SI1)
    {
        k_tile_scan<SI1><<<1, SB, 0, s>>>(in, out, n, nullptr);
        return hipGetLastError();
    }
    const long long nb = ((long long)n + SPB - 1) / SPB;
    k_tile_sums<<<(unsigned)nb, SB, 0, s>>>(in, n, tmp);
    k_tile_scan<SI><<<(unsigned)nb, SB, 0, s>>>(in, out, n, tmp);
    return hipGetLastError();
}

}  // namespace pppm_amd
