// Deterministic exclusive prefix sum of int arrays (multi-level, 1024 items per block).
// Used for the CSR layouts of the triangle bins, the 3x3x3 / 4x4x4 lists and the
// compaction of non-empty cells (the reference uses thrust::remove_if for the same
// compactions, src/core/array.h:77-115).
#include "internal.hpp"

namespace pppm_amd
{

namespace
{
constexpr int SB = 256;        // threads per block
constexpr int SI = 4;          // items per thread
constexpr int SPB = SB * SI;   // items per block

// wave64 inclusive scan of one value per lane
__device__ inline int wave_incl(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        int u = __shfl_up(v, off, 64);
        if (lane >= off)
            v += u;
    }
    return v;
}

// block-level exclusive scan; writes block total to sums[blockIdx.x]
__global__ __launch_bounds__(SB) void k_scan_block(const int *__restrict__ in, int *__restrict__ out, int n,
                                                    int *__restrict__ sums)
{
    __shared__ int wtot[SB / 64];
    const long long base = (long long)blockIdx.x * SPB + (long long)threadIdx.x * SI;
    int v[SI];
    int tsum = 0;
#pragma unroll
    for (int i = 0; i < SI; i++)
    {
        long long q = base + i;
        v[i] = q < n ? in[q] : 0;
        tsum += v[i];
    }
    int incl = wave_incl(tsum);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63)
        wtot[wid] = incl;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wid; w++)
        woff += wtot[w];
    int run = woff + incl - tsum;
#pragma unroll
    for (int i = 0; i < SI; i++)
    {
        long long q = base + i;
        if (q < n)
            out[q] = run;
        run += v[i];
    }
    if (threadIdx.x == SB - 1)
        sums[blockIdx.x] = woff + incl;
}

__global__ __launch_bounds__(SB) void k_scan_add(int *__restrict__ out, int n, const int *__restrict__ offs)
{
    const long long base = (long long)blockIdx.x * SPB + (long long)threadIdx.x * SI;
    const int o = offs[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SI; i++)
    {
        long long q = base + i;
        if (q < n)
            out[q] += o;
    }
}

// out[n] = total, read from src[nb] (the level below stores its total there); src == null -> 0
__global__ void k_write_total(int *out, int n, const int *src, int nb) { out[n] = src ? src[nb] : 0; }

}  // namespace

size_t scan_tmp_ints(int n)
{
    size_t tot = 0;
    long long m = n;
    do
    {
        long long nb = (m + SPB - 1) / SPB;
        if (nb < 1)
            nb = 1;
        tot += 2 * (size_t)nb + 2;  // block sums + their scan (nb+1)
        m = nb;
    } while (m > 1);
    return tot + 16;
}

// out must hold n+1 ints. Only kernels are enqueued (no memset/memcpy nodes), so the
// scan can be captured in a HIP graph.
hipError_t scan_exclusive(const int *in, int *out, int n, int *tmp, hipStream_t s)
{
    if (n <= 0)
    {
        k_write_total<<<1, 1, 0, s>>>(out, 0, nullptr, 0);
        return hipGetLastError();
    }
    int nb = (n + SPB - 1) / SPB;
    int *sums = tmp;            // nb
    int *sums_ex = tmp + nb;    // nb + 1
    k_scan_block<<<nb, SB, 0, s>>>(in, out, n, sums);
    if (nb == 1)
    {
        // total = sums[0]
        k_write_total<<<1, 1, 0, s>>>(out, n, sums, 0);
        return hipGetLastError();
    }
    hipError_t e = scan_exclusive(sums, sums_ex, nb, tmp + 2 * nb + 1, s);
    if (e != hipSuccess)
        return e;
    k_scan_add<<<nb, SB, 0, s>>>(out, n, sums_ex);
    k_write_total<<<1, 1, 0, s>>>(out, n, sums_ex, nb);
    return hipGetLastError();
}

}  // namespace pppm_amd
