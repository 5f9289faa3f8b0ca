// Triangle geometry, binning and neighbour lists on the GPU.
//
// Reference behaviour: update_triangles_kernel (src/algorithm/particle_grid.cu:40-88),
// getGridCoord / getGridBaseCoord (particle_grid.h:169-176), the per-cell lists and
// their compaction (particle_grid.cu:90-137), init_neigbor_list_size_kernel and
// fill_neighbor_list_{3,4}_kernel (particle_grid.cu:139-241).
//
// Layout: dense CSR over the N^3 cells (count -> exclusive scan -> fill) replaces the
// reference's fixed-capacity per-cell arrays (4112 B per cell). Triangles inside a cell
// are kept in ascending index order, which fixes the order the reference leaves to
// atomic_append. Non-empty cells are compacted in linear (x-major) order, as
// thrust::remove_if does in the reference.
#include "internal.hpp"

namespace pppm_amd
{

namespace
{
constexpr int B = 256;

__device__ inline f3 load3(const float *v, int i) { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }

__global__ __launch_bounds__(B) void k_tri_geometry(const float *__restrict__ verts, const int *__restrict__ tris,
                                                    TriGeom *__restrict__ geom, BuildArgs a, int *err_flag)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= a.nt)
        return;
    const int i0 = tris[3 * i], i1 = tris[3 * i + 1], i2 = tris[3 * i + 2];
    if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= a.nv || i1 >= a.nv || i2 >= a.nv)
    {
        atomicOr(err_flag, 2);
        return;
    }
    const f3 v0 = load3(verts, i0), v1 = load3(verts, i1), v2 = load3(verts, i2);
    TriGeom g;
    g.center = (v0 + v1 + v2) / 3.0f;
    const f3 cr = crossf(v1 - v0, v2 - v0);
    const float inv = 1.0f / sqrtf(dotf(cr, cr));  // normalize() = v * rsqrt(dot(v, v)) (helper_math.h:1459)
    g.normal = cr * inv;
    g.area = lengthf(cr) / 2.0f;
    const f3 mn = {a.minx, a.miny, a.minz};
    const f3 q = (g.center - mn) / a.h;
    int3_ c = {(int)q.x, (int)q.y, (int)q.z};
    const f3 cc = f3{((float)c.x + 0.5f) * a.h, ((float)c.y + 0.5f) * a.h, ((float)c.z + 0.5f) * a.h} + mn;
    const f3 d = g.center - cc;
    int3_ b = {c.x - (d.x < 0), c.y - (d.y < 0), c.z - (d.z < 0)};
    const int N = a.N;
    if (!(q.x >= 0 && q.y >= 0 && q.z >= 0) || c.x < 2 || c.y < 2 || c.z < 2 || c.x > N - 3 || c.y > N - 3 ||
        c.z > N - 3)
    {
        atomicOr(err_flag, 1);
        c = {2, 2, 2};
        b = {2, 2, 2};
    }
    g.cell3 = c;
    g.base3 = b;
    g.cell = lin3(c.x, c.y, c.z, N);
    g.base = lin3(b.x, b.y, b.z, N);
    geom[i] = g;
}

__global__ __launch_bounds__(B) void k_count_cells(const TriGeom *__restrict__ geom, int nt, int *cell_cnt,
                                                   int *base_cnt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt)
        return;
    atomicAdd(&cell_cnt[geom[i].cell], 1);
    atomicAdd(&base_cnt[geom[i].base], 1);
}

__global__ __launch_bounds__(B) void k_fill_cells(const TriGeom *__restrict__ geom, int nt,
                                                  const int *__restrict__ cell_start,
                                                  const int *__restrict__ base_start, int *cell_fill, int *base_fill,
                                                  int *cell_items, int *base_items)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt)
        return;
    const int c = geom[i].cell, b = geom[i].base;
    cell_items[cell_start[c] + atomicAdd(&cell_fill[c], 1)] = i;
    base_items[base_start[b] + atomicAdd(&base_fill[b], 1)] = i;
}

// insertion sort of each cell's (short) item list -> ascending triangle index
__global__ __launch_bounds__(B) void k_sort_cells(const int *__restrict__ start, int *items, int ncell)
{
    const int c = blockIdx.x * B + threadIdx.x;
    if (c >= ncell)
        return;
    const int s = start[c], e = start[c + 1];
    for (int i = s + 1; i < e; i++)
    {
        int v = items[i];
        int j = i - 1;
        while (j >= s && items[j] > v)
        {
            items[j + 1] = items[j];
            j--;
        }
        items[j + 1] = v;
    }
}

__global__ __launch_bounds__(B) void k_mark27(const TriGeom *__restrict__ geom, int nt, int N, int *flag)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt * 27)
        return;
    const int tri = i / 27, o = i % 27;
    const int3_ c = geom[tri].cell3;
    flag[lin3(c.x + o % 3 - 1, c.y + o / 3 % 3 - 1, c.z + o / 9 - 1, N)] = 1;
}

__global__ __launch_bounds__(B) void k_flag_base(const int *__restrict__ base_cnt, int N, int *flag)
{
    const int c = blockIdx.x * B + threadIdx.x;
    if (c >= N * N * N)
        return;
    const int x = c / (N * N), y = c / N % N, z = c % N;
    const bool inside = x >= 1 && y >= 1 && z >= 1 && x <= N - 2 && y <= N - 2 && z <= N - 2;
    flag[c] = (inside && base_cnt[c] > 0) ? 1 : 0;
}

__global__ __launch_bounds__(B) void k_scatter_selected(const int *__restrict__ flag, const int *__restrict__ pos,
                                                        int n, int *out, int *index_map)
{
    const int c = blockIdx.x * B + threadIdx.x;
    if (c >= n)
        return;
    if (flag[c])
    {
        out[pos[c]] = c;
        if (index_map)
            index_map[c] = pos[c];
    }
    else if (index_map)
        index_map[c] = -1;
}

// number of triangles in the width^3 block of cells starting at cell - 1
__global__ __launch_bounds__(B) void k_count_nbr(const int *__restrict__ cells, int ncells,
                                                 const int *__restrict__ cell_start, int N, int width, int *cnt)
{
    const int q = blockIdx.x * B + threadIdx.x;
    if (q >= ncells)
        return;
    const int c = cells[q];
    const int x = c / (N * N), y = c / N % N, z = c % N;
    int sum = 0;
    const int n3 = width * width * width;
    for (int idx = 0; idx < n3; idx++)
    {
        const int dx = idx % width - 1, dy = idx / width % width - 1, dz = idx / (width * width) - 1;
        const int nb = lin3(x + dx, y + dy, z + dz, N);
        sum += cell_start[nb + 1] - cell_start[nb];
    }
    cnt[q] = sum;
}

// 3x3x3 pair lists: entry = tri*27 + weight offset (grid - tri_cell + 1), pppm_cache.h:36-40
__global__ __launch_bounds__(B) void k_fill_c3(const int *__restrict__ cells, int ncells,
                                               const int *__restrict__ cell_start, const int *__restrict__ cell_items,
                                               const int *__restrict__ c3_start, int N, int *pairs)
{
    const int q = blockIdx.x * B + threadIdx.x;
    if (q >= ncells)
        return;
    const int c = cells[q];
    const int x = c / (N * N), y = c / N % N, z = c % N;
    int w = c3_start[q];
    for (int idx = 0; idx < 27; idx++)
    {
        const int dx = idx % 3 - 1, dy = idx / 3 % 3 - 1, dz = idx / 9 - 1;
        const int nb = lin3(x + dx, y + dy, z + dz, N);
        const int o = (1 - dx) + 3 * (1 - dy) + 9 * (1 - dz);
        for (int k = cell_start[nb]; k < cell_start[nb + 1]; k++)
            pairs[w++] = cell_items[k] * 27 + o;
    }
}

__global__ __launch_bounds__(B) void k_fill_b4(const int *__restrict__ cells, int ncells,
                                               const int *__restrict__ cell_start, const int *__restrict__ cell_items,
                                               const int *__restrict__ b4_start, int N, int *tris)
{
    const int q = blockIdx.x * B + threadIdx.x;
    if (q >= ncells)
        return;
    const int c = cells[q];
    const int x = c / (N * N), y = c / N % N, z = c % N;
    int w = b4_start[q];
    for (int idx = 0; idx < 64; idx++)
    {
        const int dx = idx % 4 - 1, dy = idx / 4 % 4 - 1, dz = idx / 16 - 1;
        const int nb = lin3(x + dx, y + dy, z + dz, N);
        for (int k = cell_start[nb]; k < cell_start[nb + 1]; k++)
            tris[w++] = cell_items[k];
    }
}

__global__ __launch_bounds__(B) void k_face_counts(const TriGeom *__restrict__ geom, int nt,
                                                   const int *__restrict__ b4_index, const int *__restrict__ b4_start,
                                                   int *cnt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt)
        return;
    const int b = b4_index[geom[i].base];
    cnt[i] = b4_start[b + 1] - b4_start[b];
}

__global__ __launch_bounds__(B) void k_face_fill(const TriGeom *__restrict__ geom, int nt,
                                                 const int *__restrict__ b4_index, const int *__restrict__ b4_start,
                                                 const int *__restrict__ b4_tris, const int *__restrict__ face_start,
                                                 int *face_nbr)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt)
        return;
    const int b = b4_index[geom[i].base];
    const int s = b4_start[b], e = b4_start[b + 1];
    int w = face_start[i];
    for (int k = s; k < e; k++)
        face_nbr[w++] = b4_tris[k];
}

// the 8 far-field interpolation corners of each triangle as C3 indices (order 000..111)
__global__ __launch_bounds__(B) void k_corner_map(const TriGeom *__restrict__ geom, int nt, int N,
                                                  const int *__restrict__ c3_index, int *corner)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt * 8)
        return;
    const int tri = i >> 3, w = i & 7;
    const int3_ b = geom[tri].base3;
    corner[i] = c3_index[lin3(b.x + (w >> 2), b.y + ((w >> 1) & 1), b.z + (w & 1), N)];
}

__global__ __launch_bounds__(B) void k_flag_nonzero(const int *__restrict__ in, int n, int *flag01)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n)
        flag01[i] = in[i] != 0 ? 1 : 0;
}

inline int nblk(long long n) { return (int)((n + B - 1) / B) > 0 ? (int)((n + B - 1) / B) : 1; }
}  // namespace

hipError_t launch_tri_geometry(const float *verts, const int *tris, TriGeom *geom, const BuildArgs &a, int *err_flag,
                               hipStream_t s)
{
    k_tri_geometry<<<nblk(a.nt), B, 0, s>>>(verts, tris, geom, a, err_flag);
    return hipGetLastError();
}
hipError_t launch_count_cells(const TriGeom *geom, int nt, int *cell_cnt, int *base_cnt, hipStream_t s)
{
    k_count_cells<<<nblk(nt), B, 0, s>>>(geom, nt, cell_cnt, base_cnt);
    return hipGetLastError();
}
hipError_t launch_fill_cells(const TriGeom *geom, int nt, const int *cell_start, const int *base_start,
                             int *cell_fill, int *base_fill, int *cell_items, int *base_items, hipStream_t s)
{
    k_fill_cells<<<nblk(nt), B, 0, s>>>(geom, nt, cell_start, base_start, cell_fill, base_fill, cell_items,
                                         base_items);
    return hipGetLastError();
}
hipError_t launch_sort_cells(const int *start, int *items, int ncell, hipStream_t s)
{
    k_sort_cells<<<nblk(ncell), B, 0, s>>>(start, items, ncell);
    return hipGetLastError();
}
hipError_t launch_mark27(const TriGeom *geom, int nt, int N, int *flag, hipStream_t s)
{
    k_mark27<<<nblk((long long)nt * 27), B, 0, s>>>(geom, nt, N, flag);
    return hipGetLastError();
}
hipError_t launch_flag_base(const int *base_cnt, int N, int *flag, hipStream_t s)
{
    k_flag_base<<<nblk((long long)N * N * N), B, 0, s>>>(base_cnt, N, flag);
    return hipGetLastError();
}
hipError_t launch_scatter_selected(const int *flag, const int *pos, int n, int *out, int *index_map, hipStream_t s)
{
    k_scatter_selected<<<nblk(n), B, 0, s>>>(flag, pos, n, out, index_map);
    return hipGetLastError();
}
hipError_t launch_count_nbr(const int *cells, int ncells, const int *cell_start, int N, int width, int *cnt,
                            hipStream_t s)
{
    k_count_nbr<<<nblk(ncells), B, 0, s>>>(cells, ncells, cell_start, N, width, cnt);
    return hipGetLastError();
}
hipError_t launch_fill_c3(const int *cells, int ncells, const int *cell_start, const int *cell_items,
                          const int *c3_start, int N, int *pairs, hipStream_t s)
{
    k_fill_c3<<<nblk(ncells), B, 0, s>>>(cells, ncells, cell_start, cell_items, c3_start, N, pairs);
    return hipGetLastError();
}
hipError_t launch_fill_b4(const int *cells, int ncells, const int *cell_start, const int *cell_items,
                          const int *b4_start, int N, int *tris, hipStream_t s)
{
    k_fill_b4<<<nblk(ncells), B, 0, s>>>(cells, ncells, cell_start, cell_items, b4_start, N, tris);
    return hipGetLastError();
}
hipError_t launch_face_counts(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start, int *cnt,
                              hipStream_t s)
{
    k_face_counts<<<nblk(nt), B, 0, s>>>(geom, nt, b4_index, b4_start, cnt);
    return hipGetLastError();
}
hipError_t launch_face_fill(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start,
                            const int *b4_tris, const int *face_start, int *face_nbr, hipStream_t s)
{
    k_face_fill<<<nblk(nt), B, 0, s>>>(geom, nt, b4_index, b4_start, b4_tris, face_start, face_nbr);
    return hipGetLastError();
}
hipError_t launch_flag_nonzero(const int *in, int n, int *flag01, hipStream_t s)
{
    k_flag_nonzero<<<nblk(n), B, 0, s>>>(in, n, flag01);
    return hipGetLastError();
}
hipError_t launch_corner_map(const TriGeom *geom, int nt, int N, const int *c3_index, int *corner, hipStream_t s)
{
    k_corner_map<<<nblk((long long)nt * 8), B, 0, s>>>(geom, nt, N, c3_index, corner);
    return hipGetLastError();
}

}  // namespace pppm_amd
