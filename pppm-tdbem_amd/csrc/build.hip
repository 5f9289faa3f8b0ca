// Triangle geometry, binning and neighbour lists on the GPU.
//
// Reference behaviour: update_triangles_kernel (src/algorithm/particle_grid.cu:40-88),
// getGridCoord / getGridBaseCoord (particle_grid.h:169-176), the per-cell lists and
// their compaction (particle_grid.cu:90-137), init_neigbor_list_size_kernel and
// fill_neighbor_list_{3,4}_kernel (particle_grid.cu:139-241).
//
// Layout: dense CSR over the N^3 cells (count -> exclusive scan -> fill) replaces the
// reference's fixed-capacity per-cell arrays (4112 B per cell). Lists are written with the
// triangles of every cell in ascending index order, which fixes the order the reference
// leaves to atomic_append. Non-empty cells are compacted in linear (x-major) order, as
// thrust::remove_if does in the reference.
//
// A mesh update is a short chain of launches; each launch does every piece of work whose
// inputs are complete, in block ranges of one grid:
//   k_tri_update   geometry, cell / base-cell counts, 3x3x3 marks, grid-cache dirty flags and
//                  the recompute inputs of the dirty triangles, interpolation weights (and the
//                  vertex copy)
//   [scan batch]   cell CSR offsets; compaction of base cells, marked cells, dirty triangles
//   k_lists_count  cell fill (atomic slots), 3x3x3 list lengths of the marked cells, face-list
//                  lengths (the 4x4x4 block around each triangle's base cell)
//   [scan batch]   3x3x3 and face-list offsets
//   k_lists_fill   3x3x3 lists and face lists (each cell's triangles sorted on the fly),
//                  corner map, both gather wave maps, and the reset of the counters
// The counters (cell and base counts, fill slots, marks) are zero when an update starts and
// are cleared again by k_lists_fill for the cells this update touched, so no N^3 clear is
// needed per update. The face lookup also adds up the per-tile counts of its dirty flags (one
// integer atomic per wave), so the face-pair scan is a single launch; k_tri_update clears them
// for the next update. Every launch has a host-known size (a capacity); the live counts stay
// in device memory and kernels stride over them, so an update needs no device-to-host round
// trip and can be captured in a HIP graph.
#include "internal.hpp"
#include "tri_device.hpp"

#include <algorithm>
#include <climits>

namespace pppm_amd
{

namespace
{
constexpr int B = 256;
constexpr int WPB = B / WAVE;

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

// lane l of a wave = neighbour cell l of the width^3 block around cell c (offset -1 ..
// width-2 per axis, x fastest): its [s, e) range in the cell CSR
template <int WIDTH>
__device__ inline int nbr_range(int c, int N, const int *__restrict__ cell_start, int lane, int &s, int &o)
{
    constexpr int NB = WIDTH * WIDTH * WIDTH;
    static_assert(NB <= 64, "one lane per neighbour cell");
    const int x = c / (N * N), y = c / N % N, z = c % N;
    s = 0;
    o = 0;
    int e = 0;
    if (lane < NB)
    {
        const int dx = lane % WIDTH - 1, dy = lane / WIDTH % WIDTH - 1, dz = lane / (WIDTH * WIDTH) - 1;
        const int nb = lin3(x + dx, y + dy, z + dz, N);
        s = cell_start[nb];
        e = cell_start[nb + 1];
        // 3x3x3 lists store tri*27 + weight offset (grid - tri_cell + 1, pppm_cache.h:36-40)
        o = (1 - dx) + 3 * (1 - dy) + 9 * (1 - dz);
    }
    return e - s;
}

// The segments of a width^3 list: the listed cells (cells[q]), or, for TRI, the base cells of
// the triangles (geom[q].base; the per-triangle face lists).
template <bool TRI>
__device__ inline int seg_cell(const int *__restrict__ cells, const TriGeom *__restrict__ geom, int q)
{
    if constexpr (TRI)
        return geom[q].base;
    else
        return cells[q];
}

// width^3 list lengths of the live segments, waves striding over them (wave index wv of nw);
// the capacity tail [n, cap) is zeroed by the range's threads (thread index tid of nth)
template <int WIDTH, bool TRI>
__device__ inline void nbr_count(int wv, int nw, long long tid, long long nth, const int *__restrict__ cells,
                                 const TriGeom *__restrict__ geom, int n, int cap,
                                 const int *__restrict__ cell_start, int N, int *cnt)
{
    const int lane = threadIdx.x & 63;
    for (long long q = n + tid; q < cap; q += nth)
        cnt[q] = 0;
    for (int q = wv; q < n; q += nw)
    {
        int s, o;
        const int len = nbr_range<WIDTH>(seg_cell<TRI>(cells, geom, q), N, cell_start, lane, s, o);
        const int incl = wave_incl_scan(len);
        if (lane == 63)
            cnt[q] = incl;
    }
}

// fills the width^3 lists in lane order; the triangles of a neighbour cell were placed by
// atomic slots, so each is written at its rank among them (ascending triangle index). With
// owner != null the segment index is written beside every entry.
template <int WIDTH, bool TRI>
__device__ inline void nbr_fill(int wv, int nw, const int *__restrict__ cells, const TriGeom *__restrict__ geom,
                                int n, const int *__restrict__ cell_start, const int *__restrict__ cell_items, int N,
                                const int *__restrict__ out_start, int out_cap, int *out, int *owner, int *err_flag)
{
    const int lane = threadIdx.x & 63;
    for (int q = wv; q < n; q += nw)
    {
        int s, o;
        const int len = nbr_range<WIDTH>(seg_cell<TRI>(cells, geom, q), N, cell_start, lane, s, o);
        const int incl = wave_incl_scan(len);
        const int base = out_start[q] + incl - len;
        const bool fits = out_start[q] + incl <= out_cap;
        if (lane == 63 && !fits)
            atomicOr(err_flag, ERR_CAPACITY);
        if (!fits)
            continue;
        if (len <= 8)
        {
            // up to 8 triangles: loaded together, sorted in registers (odd-even transposition)
            int v[8];
#pragma unroll
            for (int m = 0; m < 8; m++)
                v[m] = m < len ? cell_items[s + m] : INT_MAX;
#pragma unroll
            for (int r = 0; r < 8; r++)
#pragma unroll
                for (int m = r & 1; m + 1 < 8; m += 2)
                {
                    const int lo = min(v[m], v[m + 1]), hi = max(v[m], v[m + 1]);
                    v[m] = lo;
                    v[m + 1] = hi;
                }
#pragma unroll
            for (int m = 0; m < 8; m++)
                if (m < len)
                {
                    out[base + m] = WIDTH == 3 ? v[m] * 27 + o : v[m];
                    if (owner)
                        owner[base + m] = q;
                }
        }
        else
            for (int k = 0; k < len; k++)
            {
                const int tri = cell_items[s + k];
                int r = 0;
                for (int m = 0; m < len; m++)
                    r += cell_items[s + m] < tri ? 1 : 0;
                out[base + r] = WIDTH == 3 ? tri * 27 + o : tri;
                if (owner)
                    owner[base + r] = q;
            }
    }
}

// wave w of a segmented gather owns the segments whose first pair is in [64w, 64w+64):
// map[2w] = smallest s in [0, n] with start[s] >= 64 w, n = the live segment count, and
// map[2w+1] = start[map[2w]] (the wave's first pair). Searching only the live prefix matters:
// capacity-sized CSR arrays carry a tail of empty segments whose cell entries are not written,
// and a wave must never own them. Written without a search: segment s (s <= n) is the answer
// for exactly the waves with start[s-1] < 64 w <= start[s] (start[-1] = -1), and the waves
// with 64 w > start[n] get n. Thread i handles segment i and wave i.
__device__ inline void wave_map(int i, const int *__restrict__ start, int n, int waves, int *map)
{
    const int total = start[n];
    if (i <= n)
    {
        const int si = start[i];
        const int lo = i == 0 ? 0 : start[i - 1] / WAVE + 1;
        const int hi = min(si / WAVE, waves);
        for (int w = lo; w <= hi; w++)
        {
            map[2 * w] = i;
            map[2 * w + 1] = si;
        }
    }
    if (i <= waves && (long long)i * WAVE > total)
    {
        map[2 * i] = n;
        map[2 * i + 1] = total;
    }
}

// geometry, counts, 3x3x3 marks, grid-cache dirty flags and interpolation weights of every
// triangle; threads past nt copy the vertex array when the input is not the working copy
__global__ __launch_bounds__(B) void k_tri_update(const float *__restrict__ src, float *__restrict__ dst, int nv3,
                                                  const int *__restrict__ tris, TriGeom *__restrict__ geom,
                                                  BuildArgs ba, WeightArgs wa, const float *__restrict__ gmeta,
                                                  int first, int *__restrict__ gflag, GridTriIn *__restrict__ gin,
                                                  float *__restrict__ interp, UpdateCounters cn, int *err)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (dst && i < nv3)
        dst[i] = src[i];
    // the face-pair scan of the previous update has consumed its tile sums
    for (int q = i; q < cn.ts.n3; q += gridDim.x * B)
        cn.ts.b3[q] = 0;
    if (i >= ba.nt)
        return;
    TriGeom g;
    const int e = tri_geometry(src, tris, i, ba, g);
    if (e)
        atomicOr(err, e);
    geom[i] = g;
    atomicAdd(&cn.cell_cnt[g.cell], 1);
    atomicAdd(&cn.base_cnt[g.base], 1);
    const int N = ba.N;
    for (int o = 0; o < GC; o++)
        cn.mark[lin3(g.cell3.x + o % 3 - 1, g.cell3.y + o / 3 % 3 - 1, g.cell3.z + o / 9 - 1, N)] = 1;
    const int dirty = (first || grid_needs_recompute(g, gmeta + 7 * i, ba.h)) ? 1 : 0;
    gflag[i] = dirty;
    if (!(e & ERR_BAD_INDEX))
    {
        if (dirty)
        {
            // the grid-cache recompute's inputs (its one-point rule at the centroid)
            const TriD T = tri_d(src, tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]);
            const d3 p = l2g(TRI1_X, TRI1_X, T.v);
            GridTriIn r;
            r.px = p.x;
            r.py = p.y;
            r.pz = p.z;
            r.nx = T.n.x;
            r.ny = T.n.y;
            r.nz = T.n.z;
            r.jac = T.jac;
            r.cx = g.cell3.x;
            r.cy = g.cell3.y;
            r.cz = g.cell3.z;
            r.tri = i;
            r.meta[0] = g.center.x;
            r.meta[1] = g.center.y;
            r.meta[2] = g.center.z;
            r.meta[3] = g.normal.x;
            r.meta[4] = g.normal.y;
            r.meta[5] = g.normal.z;
            r.meta[6] = g.area;
            r.pad = 0;
            gin[i] = r;
        }
        interp_weights(src, tris, i, g.base3, wa, interp + 8 * i);
    }
}

// block ranges: [0, r1) cell fill, [r1, r2) 3x3x3 list lengths of the marked cells,
// [r2, grid) face-list lengths (the 4x4x4 block around each triangle's base cell)
__global__ __launch_bounds__(B) void k_lists_count(ListArgs a, int r1, int r2)
{
    const int b = blockIdx.x;
    if (b < r1)
    {
        const int i = b * B + threadIdx.x;
        if (i >= a.nt)
            return;
        const int c = a.geom[i].cell;
        const int r = atomicAdd(&a.cn.cell_fill[c], 1);
        a.cell_items[a.cell_start[c] + r] = i;
        a.cell_rank[i] = r;
    }
    else if (b < r2)
    {
        const int k = b - r1, nb = r2 - r1;
        nbr_count<3, false>(k * WPB + (threadIdx.x >> 6), nb * WPB, (long long)k * B + threadIdx.x, (long long)nb * B,
                     a.c3_cells, nullptr, min(*a.c3_n, a.cap_c3), a.cap_c3, a.cell_start, a.N, a.c3_cnt);
    }
    else
    {
        const int k = b - r2, nb = (int)gridDim.x - r2;
        nbr_count<4, true>(k * WPB + (threadIdx.x >> 6), nb * WPB, (long long)k * B + threadIdx.x, (long long)nb * B,
                     nullptr, a.geom, a.nt, a.nt, a.cell_start, a.N, a.face_cnt);
    }
}

// block ranges: [0, r1) 3x3x3 lists, [r1, r2) face lists with their owners, [r2, r3) per
// triangle (corner map, counter reset), [r3, r4) gather wave map of the 3x3x3 lists,
// [r4, grid) gather wave map of the face lists
__global__ __launch_bounds__(B) void k_lists_fill(ListArgs a, int r1, int r2, int r3, int r4)
{
    const int b = blockIdx.x;
    if (b < r1)
        nbr_fill<3, false>(b * WPB + (threadIdx.x >> 6), r1 * WPB, a.c3_cells, nullptr, min(*a.c3_n, a.cap_c3),
                    a.cell_start, a.cell_items, a.N, a.c3_start, 27 * a.nt, a.c3_pairs, nullptr, a.err);
    else if (b < r2)
        nbr_fill<4, true>((b - r1) * WPB + (threadIdx.x >> 6), (r2 - r1) * WPB, nullptr, a.geom, a.nt, a.cell_start,
                    a.cell_items, a.N, a.face_start, a.cap_p4, a.face_nbr, a.face_owner, a.err);
    else if (b < r3)
    {
        const int i = (b - r2) * B + threadIdx.x;
        if (i >= a.nt)
            return;
        const TriGeom g = a.geom[i];
        const int N = a.N;
        for (int w = 0; w < 8; w++)
            a.corner[8 * i + w] = a.c3_index[lin3(g.base3.x + (w >> 2), g.base3.y + ((w >> 1) & 1),
                                                   g.base3.z + (w & 1), N)];
        // every counter entry this update raised belongs to some triangle's cells: clear them
        a.cn.cell_cnt[g.cell] = 0;
        a.cn.cell_fill[g.cell] = 0;
        a.cn.base_cnt[g.base] = 0;
        for (int o = 0; o < GC; o++)
            a.cn.mark[lin3(g.cell3.x + o % 3 - 1, g.cell3.y + o / 3 % 3 - 1, g.cell3.z + o / 9 - 1, N)] = 0;
    }
    else if (b < r4)
        wave_map((b - r3) * B + threadIdx.x, a.c3_start, min(*a.c3_n, a.cap_c3), a.c3_waves, a.c3_wave);
    else
        wave_map((b - r4) * B + threadIdx.x, a.face_start, a.nt, a.face_waves, a.face_wave);
}

// 4x4x4 lists of the non-empty base cells (for inspection; the time step reads the
// per-triangle face lists): lengths, then, after their scan, the lists
__global__ __launch_bounds__(B) void k_b4_lists(ListArgs a, int fill)
{
    const int nb = (int)gridDim.x;
    const int n = min(*a.b4_n, a.nt);
    if (!fill)
        nbr_count<4, false>(blockIdx.x * WPB + (threadIdx.x >> 6), nb * WPB, (long long)blockIdx.x * B + threadIdx.x,
                     (long long)nb * B, a.b4_cells, nullptr, n, a.nt, a.cell_start, a.N, a.b4_cnt);
    else
        nbr_fill<4, false>(blockIdx.x * WPB + (threadIdx.x >> 6), nb * WPB, a.b4_cells, nullptr, n, a.cell_start,
                    a.cell_items, a.N, a.b4_start, a.cap_b4items, a.b4_tris, nullptr, a.err);
}

// skip (may be null): a mesh update's error word; a step behind a failed update does nothing
__global__ void k_advance(StepState *st, const int *skip)
{
    if (skip && *skip)
        return;
    st->t = st->t + 1;
}

inline int nblk(long long n) { return n > 0 ? (int)((n + B - 1) / B) : 1; }
inline int nblk_waves(long long waves) { return waves > 0 ? (int)((waves + WPB - 1) / WPB) : 1; }
// list kernels whose live count is on the device: a bounded grid that strides
inline int nblk_strided(long long waves) { return std::min(nblk_waves(waves), 1024); }
}  // namespace

__global__ void k_put_ints(int *__restrict__ a, int n, int v0, int v1, int v2, int v3)
{
    const int i = threadIdx.x;
    if (i < n)
        a[i] = i == 0 ? v0 : i == 1 ? v1 : i == 2 ? v2 : v3;
}
hipError_t launch_put_ints(int *a, int n, int v0, int v1, int v2, int v3, hipStream_t s)
{
    if (n < 1 || n > 4)
        return hipErrorInvalidValue;
    k_put_ints<<<1, WAVE, 0, s>>>(a, n, v0, v1, v2, v3);
    return hipGetLastError();
}
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
hipError_t launch_sort_cells(const TriGeom *geom, int nt, const int *start, int *items, const int *rank,
                             int use_base, hipStream_t s)
{
    k_sort_cells<<<nblk(nt), B, 0, s>>>(geom, nt, start, items, rank, use_base);
    return hipGetLastError();
}
hipError_t launch_tri_update(const float *src, float *dst, int nv, const int *tris, TriGeom *geom, const BuildArgs &ba,
                             const WeightArgs &wa, const float *gmeta, int first, int *gflag, GridTriIn *gin,
                             float *interp, const UpdateCounters &cn, int *err, hipStream_t s)
{
    const long long n = std::max<long long>(ba.nt, dst ? 3LL * nv : 0);
    k_tri_update<<<nblk(n), B, 0, s>>>(src, dst, 3 * nv, tris, geom, ba, wa, gmeta, first, gflag, gin, interp, cn,
                                       err);
    return hipGetLastError();
}
hipError_t launch_lists_count(const ListArgs &a, hipStream_t s)
{
    const int r1 = nblk(a.nt), n3 = nblk_strided(a.cap_c3), n4 = nblk_strided(a.nt);
    k_lists_count<<<r1 + n3 + n4, B, 0, s>>>(a, r1, r1 + n3);
    return hipGetLastError();
}
hipError_t launch_lists_fill(const ListArgs &a, hipStream_t s)
{
    const int n3 = nblk_strided(a.cap_c3), nf = nblk_strided(a.nt), nt = nblk(a.nt);
    const int m3 = nblk(std::max(a.c3_waves, a.cap_c3) + 1LL), mf = nblk(std::max(a.face_waves, a.nt) + 1LL);
    const int r1 = n3, r2 = r1 + nf, r3 = r2 + nt, r4 = r3 + m3;
    k_lists_fill<<<r4 + mf, B, 0, s>>>(a, r1, r2, r3, r4);
    return hipGetLastError();
}
hipError_t launch_b4_lists(const ListArgs &a, int fill, hipStream_t s)
{
    k_b4_lists<<<nblk_strided(a.nt), B, 0, s>>>(a, fill);
    return hipGetLastError();
}
hipError_t launch_advance(StepState *st, hipStream_t s, const int *skip)
{
    k_advance<<<1, 1, 0, s>>>(st, skip);
    return hipGetLastError();
}

}  // namespace pppm_amd
