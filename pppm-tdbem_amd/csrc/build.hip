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
//
// Every launch has a host-known size (a capacity); the live counts stay in device memory
// (the scan totals) and kernels exit or write zeros past them, so a mesh update needs no
// device-to-host round trip and can be captured in a HIP graph.
#include "internal.hpp"

namespace pppm_amd
{

namespace
{
constexpr int B = 256;
constexpr int WPB = B / WAVE;

__device__ inline f3 load3(const float *v, int i) { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }

__device__ inline int wave_incl_scan(int v)
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

__global__ __launch_bounds__(B) void k_fill_int(int *__restrict__ a, long long n, int value)
{
    const long long i = (long long)blockIdx.x * B + threadIdx.x;
    if (i < n)
        a[i] = value;
}

// copies vertex row (t - t0) of a device series into the working vertex buffer
__global__ __launch_bounds__(B) void k_copy_verts(const float *__restrict__ series, long long stride, int t0,
                                                  const StepState *__restrict__ st, float *__restrict__ verts, int n)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n)
        verts[i] = series[(size_t)(st->t + 1 - t0) * stride + i];
}

__global__ __launch_bounds__(B) void k_tri_geometry(const float *__restrict__ verts, const int *__restrict__ tris,
                                                    TriGeom *__restrict__ geom, BuildArgs a, int *err_flag)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= a.nt)
        return;
    const int i0 = tris[3 * i], i1 = tris[3 * i + 1], i2 = tris[3 * i + 2];
    TriGeom g;
    int3_ c = {2, 2, 2}, b = {2, 2, 2};
    if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= a.nv || i1 >= a.nv || i2 >= a.nv)
    {
        atomicOr(err_flag, ERR_BAD_INDEX);
        g.center = g.normal = {0, 0, 0};
        g.area = 0;
    }
    else
    {
        const f3 v0 = load3(verts, i0), v1 = load3(verts, i1), v2 = load3(verts, i2);
        g.center = (v0 + v1 + v2) / 3.0f;
        const f3 cr = crossf(v1 - v0, v2 - v0);
        const float inv = 1.0f / sqrtf(dotf(cr, cr));  // normalize() = v * rsqrt(dot(v, v)) (helper_math.h:1459)
        g.normal = cr * inv;
        g.area = lengthf(cr) / 2.0f;
        const f3 mn = {a.minx, a.miny, a.minz};
        const f3 q = (g.center - mn) / a.h;
        c = {(int)q.x, (int)q.y, (int)q.z};
        const f3 cc = f3{((float)c.x + 0.5f) * a.h, ((float)c.y + 0.5f) * a.h, ((float)c.z + 0.5f) * a.h} + mn;
        const f3 d = g.center - cc;
        b = {c.x - (d.x < 0), c.y - (d.y < 0), c.z - (d.z < 0)};
        const int N = a.N;
        if (!(q.x >= 0 && q.y >= 0 && q.z >= 0) || c.x < 2 || c.y < 2 || c.z < 2 || c.x > N - 3 || c.y > N - 3 ||
            c.z > N - 3)
        {
            atomicOr(err_flag, ERR_OUT_OF_GRID);
            c = {2, 2, 2};
            b = {2, 2, 2};
        }
    }
    g.cell3 = c;
    g.base3 = b;
    g.cell = lin3(c.x, c.y, c.z, a.N);
    g.base = lin3(b.x, b.y, b.z, a.N);
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
                                                  int *cell_items, int *base_items, int *cell_rank, int *base_rank)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt)
        return;
    const int c = geom[i].cell, b = geom[i].base;
    const int rc = atomicAdd(&cell_fill[c], 1), rb = atomicAdd(&base_fill[b], 1);
    cell_items[cell_start[c] + rc] = i;
    base_items[base_start[b] + rb] = i;
    cell_rank[i] = rc;
    base_rank[i] = rb;
}

// insertion sort of each non-empty cell's (short) list -> ascending triangle index;
// the triangle that drew rank 0 in the fill owns the sort of its cell
__global__ __launch_bounds__(B) void k_sort_cells(const TriGeom *__restrict__ geom, int nt,
                                                  const int *__restrict__ start, int *items,
                                                  const int *__restrict__ rank, int use_base)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= nt || rank[i] != 0)
        return;
    const int c = use_base ? geom[i].base : geom[i].cell;
    const int s = start[c], e = start[c + 1];
    for (int k = s + 1; k < e; k++)
    {
        int v = items[k];
        int j = k - 1;
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

// one wave per segment: lane l = neighbour cell l of the width^3 block around the segment
// cell (offset -1 .. width-2 per axis, x fastest); counts or fills in that order. The grid
// is bounded (the live count is only known on the device): waves stride over the live
// segments, and the count variant also zeroes the counts of the capacity tail.
template <int WIDTH, bool FILL>
__global__ __launch_bounds__(B) void k_nbr_wave(const int *__restrict__ cells, const int *__restrict__ n_ptr, int cap,
                                                const int *__restrict__ cell_start,
                                                const int *__restrict__ cell_items, int N,
                                                const int *__restrict__ out_start, int out_cap, int *out, int *cnt,
                                                int *err_flag)
{
    constexpr int NB = WIDTH * WIDTH * WIDTH;
    static_assert(NB <= 64, "one lane per neighbour cell");
    const int lane = threadIdx.x & 63;
    const int n = min(*n_ptr, cap);
    if (!FILL)
        for (long long q = n + (long long)blockIdx.x * B + threadIdx.x; q < cap; q += (long long)gridDim.x * B)
            cnt[q] = 0;
    const int stride = gridDim.x * WPB;
    for (int q = blockIdx.x * WPB + (threadIdx.x >> 6); q < n; q += stride)
    {
        const int c = cells[q];
        const int x = c / (N * N), y = c / N % N, z = c % N;
        int s = 0, e = 0, dx = 0, dy = 0, dz = 0;
        if (lane < NB)
        {
            dx = lane % WIDTH - 1;
            dy = lane / WIDTH % WIDTH - 1;
            dz = lane / (WIDTH * WIDTH) - 1;
            const int nb = lin3(x + dx, y + dy, z + dz, N);
            s = cell_start[nb];
            e = cell_start[nb + 1];
        }
        const int len = e - s;
        const int incl = wave_incl_scan(len);
        if (!FILL)
        {
            if (lane == 63)
                cnt[q] = incl;
            continue;
        }
        const int base = out_start[q] + incl - len;
        const bool fits = out_start[q] + incl <= out_cap;
        if (lane == 63 && !fits)
            atomicOr(err_flag, ERR_CAPACITY);
        if (!fits)
            continue;
        // 3x3x3 lists store tri*27 + weight offset (grid - tri_cell + 1, pppm_cache.h:36-40)
        const int o = (1 - dx) + 3 * (1 - dy) + 9 * (1 - dz);
        for (int k = 0; k < len; k++)
        {
            const int tri = cell_items[s + k];
            out[base + k] = WIDTH == 3 ? tri * 27 + o : tri;
        }
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

// one wave per triangle: copies the 4x4x4 list of its base cell and the owner index
__global__ __launch_bounds__(B) void k_face_fill(const TriGeom *__restrict__ geom, int nt,
                                                 const int *__restrict__ b4_index, const int *__restrict__ b4_start,
                                                 const int *__restrict__ b4_tris, const int *__restrict__ face_start,
                                                 int cap, int *face_nbr, int *face_owner, int *err_flag)
{
    const int i = blockIdx.x * WPB + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // an earlier overflow (4x4x4 lists) leaves b4_tris incomplete: copy nothing
    if (i >= nt || *err_flag != 0)
        return;
    const int b = b4_index[geom[i].base];
    const int s = b4_start[b], e = b4_start[b + 1];
    const int w = face_start[i];
    if (w + (e - s) > cap)
    {
        if (lane == 0)
            atomicOr(err_flag, ERR_CAPACITY);
        return;
    }
    for (int k = lane; k < e - s; k += WAVE)
    {
        face_nbr[w + k] = b4_tris[s + k];
        face_owner[w + k] = i;
    }
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

// wave w of a segmented gather owns the segments whose first pair is in [64w, 64w+64):
// map[w] = smallest s in [0, n] with start[s] >= 64 w, n = the live segment count (read from
// device memory when n_ptr is given, else nseg). Searching only the live prefix matters:
// capacity-sized CSR arrays carry a tail of empty segments whose cell entries are not
// written, and a wave must never own them.
__global__ __launch_bounds__(B) void k_wave_map(const int *__restrict__ start, int nseg, const int *__restrict__ n_ptr,
                                                int waves, int *map)
{
    const int w = blockIdx.x * B + threadIdx.x;
    if (w > waves)
        return;
    const int p = w * WAVE;
    const int n = n_ptr ? min(*n_ptr, nseg) : nseg;
    int lo = 0, hi = n;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (start[mid] >= p)
            hi = mid;
        else
            lo = mid + 1;
    }
    map[w] = lo;
}

__global__ void k_advance(StepState *st) { st->t = st->t + 1; }

// gathers scalar counters that live in device memory into one status block
__global__ void k_status(StatusSrc src, int *out)
{
    const int i = threadIdx.x;
    if (i < src.n)
        out[i] = src.p[i] ? *src.p[i] : 0;
}

inline int nblk(long long n) { return n > 0 ? (int)((n + B - 1) / B) : 1; }
inline int nblk_waves(long long waves) { return waves > 0 ? (int)((waves + WPB - 1) / WPB) : 1; }
// list kernels whose live count is on the device: a bounded grid that strides
inline int nblk_strided(long long waves) { return std::min(nblk_waves(waves), 1024); }
}  // namespace

hipError_t launch_fill_int(int *a, long long n, int value, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    k_fill_int<<<nblk(n), B, 0, s>>>(a, n, value);
    return hipGetLastError();
}
hipError_t launch_copy_verts(const float *series, long long stride, int t0, const StepState *st, float *verts, int n,
                             hipStream_t s)
{
    k_copy_verts<<<nblk(n), B, 0, s>>>(series, stride, t0, st, verts, n);
    return hipGetLastError();
}
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
                             int *cell_fill, int *base_fill, int *cell_items, int *base_items, int *cell_rank,
                             int *base_rank, hipStream_t s)
{
    k_fill_cells<<<nblk(nt), B, 0, s>>>(geom, nt, cell_start, base_start, cell_fill, base_fill, cell_items,
                                         base_items, cell_rank, base_rank);
    return hipGetLastError();
}
hipError_t launch_sort_cells(const TriGeom *geom, int nt, const int *start, int *items, const int *rank,
                             int use_base, hipStream_t s)
{
    k_sort_cells<<<nblk(nt), B, 0, s>>>(geom, nt, start, items, rank, use_base);
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
hipError_t launch_nbr_count(const int *cells, const int *n_ptr, int cap, const int *cell_start, int N, int width,
                            int *cnt, hipStream_t s)
{
    if (width == 3)
        k_nbr_wave<3, false><<<nblk_strided(cap), B, 0, s>>>(cells, n_ptr, cap, cell_start, nullptr, N, nullptr, 0,
                                                            nullptr, cnt, nullptr);
    else
        k_nbr_wave<4, false><<<nblk_strided(cap), B, 0, s>>>(cells, n_ptr, cap, cell_start, nullptr, N, nullptr, 0,
                                                            nullptr, cnt, nullptr);
    return hipGetLastError();
}
hipError_t launch_nbr_fill(const int *cells, const int *n_ptr, int cap, const int *cell_start, const int *cell_items,
                           int N, int width, const int *out_start, int out_cap, int *out, int *err_flag, hipStream_t s)
{
    if (width == 3)
        k_nbr_wave<3, true><<<nblk_strided(cap), B, 0, s>>>(cells, n_ptr, cap, cell_start, cell_items, N, out_start,
                                                           out_cap, out, nullptr, err_flag);
    else
        k_nbr_wave<4, true><<<nblk_strided(cap), B, 0, s>>>(cells, n_ptr, cap, cell_start, cell_items, N, out_start,
                                                           out_cap, out, nullptr, err_flag);
    return hipGetLastError();
}
hipError_t launch_face_counts(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start, int *cnt,
                              hipStream_t s)
{
    k_face_counts<<<nblk(nt), B, 0, s>>>(geom, nt, b4_index, b4_start, cnt);
    return hipGetLastError();
}
hipError_t launch_face_fill(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start,
                            const int *b4_tris, const int *face_start, int cap, int *face_nbr, int *face_owner,
                            int *err_flag, hipStream_t s)
{
    k_face_fill<<<nblk_waves(nt), B, 0, s>>>(geom, nt, b4_index, b4_start, b4_tris, face_start, cap, face_nbr,
                                              face_owner, err_flag);
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
hipError_t launch_wave_map(const int *start, int nseg, const int *n_ptr, int waves, int *map, hipStream_t s)
{
    k_wave_map<<<nblk(waves + 1), B, 0, s>>>(start, nseg, n_ptr, waves, map);
    return hipGetLastError();
}
hipError_t launch_advance(StepState *st, hipStream_t s)
{
    k_advance<<<1, 1, 0, s>>>(st);
    return hipGetLastError();
}
hipError_t launch_status(const StatusSrc &src, int *out, hipStream_t s)
{
    k_status<<<1, 64, 0, s>>>(src, out);
    return hipGetLastError();
}

}  // namespace pppm_amd
