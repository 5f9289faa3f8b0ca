// Experiment outputs read from the pressure grid (gfx950): far-field (FFAT) maps accumulated on
// the six faces of a cube around the grid centre.
//
// Reference behaviour: collect_ffat_map (experiments/static/main.cu:60-80; the same map through
// get_coord in experiments/acoustic/main.cu:46-74): for face b = 2 i + (sign > 0) of axis i,
// map(b, x, y) += |p(centre + sign w e_i + (x - w) e_{i+1} + (y - w) e_{i+2})| with the axes taken
// cyclically and centre = (N/2, N/2, N/2). One thread per (x, y) adds its six faces, so every
// entry has a single writer and the sums are the reference's.
#include "internal.hpp"

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
}  // namespace

hipError_t launch_ffat(const StepState *st, const float *grid3, int N, int w, float *map, hipStream_t s)
{
    const int W = 2 * w;
    const dim3 grid((unsigned)((W + 63) / 64), (unsigned)((W + 3) / 4));
    k_ffat<<<grid, 256, 0, s>>>(st, grid3, N, w, map);
    return hipGetLastError();
}

}  // namespace pppm_amd
