// CPU restatement of the reference's experiment outputs — TEST INFRASTRUCTURE ONLY.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
//
//   orc_ffat_collect   collect_ffat_map   experiments/static/main.cu:60-80,
//                                         experiments/acoustic/main.cu:46-74 (get_coord)
#include <cmath>
#include <cstddef>

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

}  // extern "C"
