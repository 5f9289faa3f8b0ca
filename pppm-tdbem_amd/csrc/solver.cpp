// C ABI of the MI355X PPPM library: state, orchestration of the build and step
// kernels, HIP-graph replay of the time step. Declarations: include/pppm_hip.h.
//
// Reference behaviour: PPPMSolver (src/algorithm/pppm.h:11-164), ParticleGrid::init /
// set_mesh / update_mesh / construct_neighbor_lists (src/algorithm/particle_grid.h:114-148,
// particle_grid.cu:133-241), GridCache / FaceCache::update_cache (pppm_cache.cu:125-134,
// 273-289), FDTD::init / step / reset (src/algorithm/fdtd.h:45-110).
//
// A mesh update is one asynchronous kernel sequence: every array has a host-known capacity
// and every live count stays in device memory, so no device-to-host round trip is needed
// and a moving-mesh time step (update + step) can be captured once and replayed. Counts
// and error bits are gathered into a status block that the host reads at sync points.
#include "pppm_hip.h"

#include <algorithm>
#include <utility>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "solver_state.hpp"

using namespace pppm_amd;

namespace
{

struct LWH72
{
    uint16_t s[STEPS];
    float ss;
    uint16_t d[STEPS];
    float sd;
};
static_assert(sizeof(LWH72) == 72, "LayerWeightHalf layout");

}  // namespace

namespace
{

int fail(pppm_solver *s, int code, const std::string &msg)
{
    if (s)
        s->err = msg;
    return code;
}

#define CK(expr)                                                                                 \
    do                                                                                           \
    {                                                                                            \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(s, PPPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

#define RC(expr)          \
    do                    \
    {                     \
        int rc_ = (expr); \
        if (rc_)          \
            return rc_;   \
    } while (0)

FdtdParams fdtd_params(const pppm_solver *s)
{
    FdtdParams p;
    p.N = s->N;
    const double c = s->c, dt = s->dt, h = s->h;
    p.coef = c * c * dt * dt;
    p.hh = h * h;
    p.c = s->c;
    p.dt = s->dt;
    p.h = s->h;
    p.pml = s->use_pml ? s->pml_width : 0;
    p.damp = 2.0f / s->h;
    return p;
}

WeightArgs weight_args(const pppm_solver *s)
{
    WeightArgs a;
    a.N = s->N;
    a.nt = s->nt;
    a.h = s->h;
    a.c = s->c;
    a.dt = s->dt;
    a.minx = s->minp[0];
    a.miny = s->minp[1];
    a.minz = s->minp[2];
    return a;
}

StepArgs step_args(const pppm_solver *s, int t0, long long stride, int n_listen)
{
    StepArgs a;
    a.N = s->N;
    a.nt = s->nt;
    a.rc = s->rc;
    a.t0 = t0;
    a.stride = stride;
    a.n_listen = n_listen;
    return a;
}

// ---------------------------------------------------------------- timing spans
hipEvent_t take_event(pppm_solver *s)
{
    if (!s->pool.empty())
    {
        hipEvent_t e = s->pool.back();
        s->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

int span_begin(pppm_solver *s, int phase)
{
    if (!s->timing)
        return -1;
    pppm_solver::Span sp{phase, take_event(s), take_event(s)};
    if (!sp.a || !sp.b || hipEventRecord(sp.a, s->stream) != hipSuccess)
        return -1;
    s->spans.push_back(sp);
    return (int)s->spans.size() - 1;
}

void span_end(pppm_solver *s, int idx)
{
    if (idx >= 0)
        (void)hipEventRecord(s->spans[idx].b, s->stream);
}

int collect_spans(pppm_solver *s)
{
    if (s->spans.empty())
        return PPPM_OK;
    CK(hipStreamSynchronize(s->stream));
    for (auto &sp : s->spans)
    {
        float ms = 0;
        if (hipEventElapsedTime(&ms, sp.a, sp.b) == hipSuccess)
        {
            s->phase_ms[sp.phase] += ms;
            s->phase_n[sp.phase]++;
        }
        s->pool.push_back(sp.a);
        s->pool.push_back(sp.b);
    }
    s->spans.clear();
    return PPPM_OK;
}

// ---------------------------------------------------------------- capacities
int scan(pppm_solver *s, const int *in, int *out, int n)
{
    CK(scan_exclusive(in, out, n, s->scan_tmp.p, s->stream));
    return PPPM_OK;
}

// (re)allocates every capacity-sized array for the current caps
int allocate(pppm_solver *s)
{
    const int N = s->N, nt = s->nt;
    const int NN = N * N * N;
    s->cap_c3 = std::min(27 * nt, NN);
    const int p3 = 27 * nt;
    s->c3_waves = p3 / WAVE + 1;
    s->face_waves = s->cap_p4 / WAVE + 1;
    const int batch_n[4] = {NN, NN, NN, std::max({nt, s->cap_c3, s->cap_p4})};
    if (s->counters.n < (size_t)4 * NN)
        s->counters_zero = false;
    CK(s->counters.ensure((size_t)4 * NN));
    CK(s->cell_start.ensure(NN + 1));
    CK(s->cell_items.ensure(nt));
    CK(s->cell_rank.ensure(nt));
    CK(s->scan_tmp.ensure(std::max(scan_batch_tmp_ints(batch_n, 4), scan_tmp_ints(NN))));
    CK(s->c3_cells.ensure(s->cap_c3));
    CK(s->c3_index.ensure(NN));
    CK(s->c3_cnt.ensure(s->cap_c3));
    CK(s->c3_start.ensure(s->cap_c3 + 1));
    CK(s->c3_pairs.ensure(p3));
    CK(s->c3_wave.ensure(2 * (size_t)(s->c3_waves + 1)));
    CK(s->far.ensure(s->cap_c3));
    CK(s->b4_cells.ensure(nt));
    CK(s->b4_cnt.ensure(nt));
    CK(s->b4_start.ensure(nt + 1));
    CK(s->b4_tris.ensure(s->cap_b4items));
    CK(s->face_cnt.ensure(nt));
    CK(s->face_start.ensure(nt + 1));
    CK(s->face_nbr.ensure(s->cap_p4));
    CK(s->face_slot.ensure(s->cap_p4));
    CK(s->prev_face_start.ensure(nt + 1));
    CK(s->prev_face_nbr.ensure(s->cap_p4));
    CK(s->prev_face_slot.ensure(s->cap_p4));
    s->faces_valid = false;  // buffers may have moved
    CK(s->face_owner.ensure(s->cap_p4));
    CK(s->face_flag.ensure(s->cap_p4));
    CK(s->dirty_list.ensure(s->cap_p4));
    CK(s->face_wave.ensure(2 * (size_t)(s->face_waves + 1)));
    CK(s->corner.ensure((size_t)nt * 8));
    CK(s->gflag.ensure(nt));
    CK(s->glist.ensure(nt));
    CK(s->ftmp.ensure(face_tmp_bytes(s->cap_p4)));
    CK(s->status.ensure(16));
    return PPPM_OK;
}

// grows the face-cache pool (keeping its contents) and the hash table (rehashing)
int ensure_pool(pppm_solver *s, int need)
{
    if (need > s->pool_cap)
    {
        const int ncap = std::max(need, s->pool_cap + s->pool_cap / 2);
        DBuf<uint32_t> nfw;
        DBuf<float2> nfsc, nfmeta;
        CK(nfw.ensure((size_t)ncap * STEPS));
        CK(nfsc.ensure(ncap));
        CK(nfmeta.ensure(ncap));
        if (s->pool_count > 0)
        {
            CK(hipMemcpyAsync(nfw.p, s->fw.p, (size_t)s->pool_count * STEPS * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s->stream));
            CK(hipMemcpyAsync(nfsc.p, s->fsc.p, (size_t)s->pool_count * sizeof(float2), hipMemcpyDeviceToDevice,
                              s->stream));
            CK(hipMemcpyAsync(nfmeta.p, s->fmeta.p, (size_t)s->pool_count * sizeof(float2),
                              hipMemcpyDeviceToDevice, s->stream));
        }
        CK(hipStreamSynchronize(s->stream));
        s->fw.take(nfw);
        s->fsc.take(nfsc);
        s->fmeta.take(nfmeta);
        s->pool_cap = ncap;
    }
    int want = 1024;
    while (want < 2 * s->pool_cap)
        want <<= 1;
    if (want > s->hcap)
    {
        DBuf<unsigned long long> nk;
        DBuf<int> nvv;
        CK(nk.ensure(want));
        CK(nvv.ensure(want));
        CK(launch_fill_u64(nk.p, want, ~0ull, s->stream));
        if (s->hcap > 0)
            CK(launch_rehash(s->hkeys.p, s->hvals.p, s->hcap, nk.p, nvv.p, want, s->stream));
        CK(hipStreamSynchronize(s->stream));
        s->hkeys.take(nk);
        s->hvals.take(nvv);
        s->hcap = want;
    }
    return PPPM_OK;
}

ListArgs list_args(pppm_solver *s, const UpdateCounters &cn)
{
    int *stp = s->status.p;
    ListArgs la{};
    la.N = s->N;
    la.nt = s->nt;
    la.geom = s->geom.p;
    la.cn = cn;
    la.cell_start = s->cell_start.p;
    la.cell_items = s->cell_items.p;
    la.cell_rank = s->cell_rank.p;
    la.c3_cells = s->c3_cells.p;
    la.c3_n = stp + ST_C3;
    la.c3_index = s->c3_index.p;
    la.c3_cnt = s->c3_cnt.p;
    la.c3_start = s->c3_start.p;
    la.c3_pairs = s->c3_pairs.p;
    la.c3_wave = s->c3_wave.p;
    la.cap_c3 = s->cap_c3;
    la.c3_waves = s->c3_waves;
    la.b4_cells = s->b4_cells.p;
    la.b4_n = stp + ST_B4;
    la.b4_cnt = s->b4_cnt.p;
    la.b4_start = s->b4_start.p;
    la.b4_tris = s->b4_tris.p;
    la.cap_b4items = s->cap_b4items;
    la.face_cnt = s->face_cnt.p;
    la.face_start = s->face_start.p;
    la.face_nbr = s->face_nbr.p;
    la.face_owner = s->face_owner.p;
    la.face_wave = s->face_wave.p;
    la.cap_p4 = s->cap_p4;
    la.face_waves = s->face_waves;
    la.corner = s->corner.p;
    la.err = stp + ST_ERR;
    return la;
}

// ---------------------------------------------------------------- the update sequence
// The grid-cache recompute goes to the update's second stream unless phases are being timed
// (spans on the handle's stream) or the stream is being captured into a graph.
bool fork_update(const pppm_solver *s)
{
    if (s->timing || !s->uf.aux)
        return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s->stream, &cs) != hipSuccess)
        return false;
    return cs == hipStreamCaptureStatusNone;
}

// Geometry, binning, lists and weight caches for the vertex positions in s->verts (or, with
// a series, row t+1-t0 of it). Asynchronous: only kernels are enqueued. The grid-cache
// recompute needs only the first scan batch (its dirty-triangle list); forked, it runs on the
// second stream while the lists and the face cache are built, and the handle's stream waits
// for it at the end (the step kernels read both caches).
int enqueue_rebuild(pppm_solver *s, bool first, const float *verts_series, long long vstride, int t0,
                    const float *vsrc = nullptr, bool use_prev = false)
{
    const int N = s->N, nt = s->nt;
    const int NN = N * N * N;
    hipStream_t st = s->stream;
    int *stp = s->status.p;
    int *err = stp + ST_ERR;

    if (verts_series)
        CK(launch_copy_verts(verts_series, vstride, t0, s->state.p, s->verts.p, s->nv * 3, st));
    UpdateCounters cn{s->counters.p, s->counters.p + NN, s->counters.p + 2 * NN, s->counters.p + 3 * NN, {}};

    // the update's three scan batches (capacity-sized, so their tile layouts are host-known);
    // the face lookup accumulates the tile sums of the third
    ScanBatch b1{}, b2{}, b3{};
    b1.njobs = 4;
    b1.job[0] = ScanJob{cn.cell_cnt, NN, 0, s->cell_start.p, nullptr, nullptr, nullptr, 0};
    b1.job[1] = ScanJob{cn.base_cnt, NN, 1, nullptr, s->b4_cells.p, nullptr, stp + ST_B4, 0};
    b1.job[2] = ScanJob{cn.mark, NN, 1, nullptr, s->c3_cells.p, s->c3_index.p, stp + ST_C3, 0};
    b1.job[3] = ScanJob{s->gflag.p, nt, 1, nullptr, s->glist.p, nullptr, stp + ST_GDIRTY, 0};
    b2.njobs = 2;
    b2.job[0] = ScanJob{s->c3_cnt.p, s->cap_c3, 0, s->c3_start.p, nullptr, nullptr, nullptr, 0};
    b2.job[1] = ScanJob{s->face_cnt.p, nt, 0, s->face_start.p, nullptr, nullptr, stp + ST_P4, 0};
    b3.njobs = 1;
    b3.job[0] = ScanJob{s->face_flag.p, s->cap_p4, 1, nullptr, s->dirty_list.p, nullptr, stp + ST_FDIRTY, 0};
    TileSums &ts = cn.ts;
    ts.n3 = scan_batch_layout(b3);
    const size_t nsum = (size_t)ts.n3;
    if (s->tsums.n < nsum)
    {
        CK(s->tsums.ensure(nsum));
        s->counters_zero = false;
    }
    ts.b3 = s->tsums.p;
    if (!s->counters_zero)
    {
        CK(launch_fill_int(s->counters.p, 4LL * NN, 0, st));
        CK(launch_fill_int(s->tsums.p, (long long)nsum, 0, st));
    }
    s->counters_zero = false;

    // geometry, counts, 3x3x3 marks, grid-cache dirty flags (GridCache::update_cache,
    // pppm_cache.cu:125-134), interpolation weights (update_interpolation_weight_kernel)
    const BuildArgs ba{N, nt, s->nv, s->h, s->minp[0], s->minp[1], s->minp[2]};
    const WeightArgs wa = weight_args(s);
    const float *src = vsrc ? vsrc : s->verts.p;
    int sp = span_begin(s, 7);
    CK(launch_tri_update(src, vsrc ? s->verts.p : nullptr, s->nv, s->tris.p, s->geom.p, ba, wa, s->gmeta.p,
                         first ? 1 : 0, s->gflag.p, s->gin.p, s->interp.p, cn, err, st));

    // cell CSR offsets; compaction of non-empty base cells, marked cells, dirty triangles
    CK(scan_batch(b1, s->scan_tmp.p, st));

    // grid cache: dirty triangles only (one block per CU, striding over the device-side
    // count). The face-cache launch sizes follow the previous update's recompute count (the
    // kernels stride, so a larger count is still covered); a first build sizes for everything.
    const long long fhint = first ? -1 : 2 * s->last_face_recomputes + 64;
    auto grid_weights = [&](hipStream_t gs) {
        return launch_grid_weights(s->gin.p, s->glist.p, stp + ST_GDIRTY, nt, wa, s->bem, s->wf.p, s->sf.p, s->wb.p,
                                   s->sb.p, s->gmeta.p, err, gs);
    };
    const bool fork = fork_update(s);
    if (fork)
    {
        CK(hipEventRecord(s->uf.fork, st));
        CK(hipStreamWaitEvent(s->uf.aux, s->uf.fork, 0));
        CK(grid_weights(s->uf.aux));
        CK(hipEventRecord(s->uf.join, s->uf.aux));
    }

    const ListArgs la = list_args(s, cn);

    // cell fill, 3x3x3 and face-list lengths (a triangle's face list is the 4x4x4 block around
    // its base cell), their offsets, the lists themselves, corner map, gather wave maps and the
    // reset of the counters
    CK(launch_lists_count(la, st));
    CK(scan_batch(b2, s->scan_tmp.p, st));
    CK(launch_lists_fill(la, st));
    s->counters_zero = true;
    span_end(s, sp);

    if (!fork)
    {
        sp = span_begin(s, 5);
        CK(grid_weights(st));
        span_end(s, sp);
    }
    sp = span_begin(s, 6);

    // face cache: persistent (c, j) slots, dirty pairs recomputed (FaceCache::update_cache)
    CK(launch_face_lookup(s->geom.p, s->tris.p, s->face_owner.p, s->face_nbr.p, s->face_start.p + nt, s->cap_p4,
                          s->hkeys.p, s->hvals.p, s->hcap, stp + ST_POOL, s->pool_cap, s->fmeta.p, s->face_slot.p,
                          s->face_flag.p, err, nt,
                          use_prev ? PrevFaces{s->prev_face_start.p, s->prev_face_nbr.p, s->prev_face_slot.p,
                                               s->face_start.p}
                                   : PrevFaces{nullptr, nullptr, nullptr, nullptr},
                          ts.b3, st));
    CK(scan_batch(b3, s->scan_tmp.p, st, ts.b3));
    CK(launch_face_weights(s->verts.p, s->tris.p, s->geom.p, s->dirty_list.p, stp + ST_FDIRTY, s->cap_p4,
                           s->face_owner.p, s->face_nbr.p, s->face_slot.p, s->interp.p, wa, s->bem, s->ftmp.p, s->fw.p,
                           s->fsc.p, s->fmeta.p, err, fhint, st));
    span_end(s, sp);
    if (fork)
        CK(hipStreamWaitEvent(st, s->uf.join, 0));

    return PPPM_OK;
}

int read_status(pppm_solver *s)
{
    CK(hipMemcpyAsync(s->st_h, s->status.p, ST_N * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    s->err_clean = s->st_h[ST_ERR] == 0;
    s->pool_count = s->st_h[ST_POOL];
    s->last_grid_recomputes = s->st_h[ST_GDIRTY];
    s->last_face_recomputes = s->st_h[ST_FDIRTY];
    return PPPM_OK;
}

// maps device error bits to a status (capacity problems are returned as PPPM_ERR_CAPACITY)
int status_error(pppm_solver *s)
{
    const int e = s->st_h[ST_ERR];
    if (e & ERR_BAD_INDEX)
        return fail(s, PPPM_ERR_ARG, "triangle references a vertex index out of range");
    if (e & ERR_OUT_OF_GRID)
        return fail(s, PPPM_ERR_OUT_OF_GRID, "a triangle centre lies outside grid cells [2, N-3]");
    if (e & ERR_CAPACITY)
        return fail(s, PPPM_ERR_CAPACITY, "an internal list or the face-cache pool ran out of capacity");
    return PPPM_OK;
}

bool needs_growth(const pppm_solver *s)
{
    return s->st_h[ST_P4] > s->cap_p4 || s->pool_count + s->cap_p4 > s->pool_cap;
}

// grows capacities to 1.25x the last observed needs (and the pool to hold one more update)
int grow(pppm_solver *s)
{
    bool realloc = false;
    if (s->st_h[ST_P4] > s->cap_p4 || s->cap_p4 == 0)
    {
        s->cap_p4 = std::max(s->st_h[ST_P4], 1) * 5 / 4 + 1024;
        realloc = true;
    }
    if (realloc)
        RC(allocate(s));
    return ensure_pool(s, s->pool_count + 2 * s->cap_p4);
}

// status bookkeeping after the status block has arrived in st_h
void take_status(pppm_solver *s)
{
    s->err_clean = s->st_h[ST_ERR] == 0;
    s->pool_count = s->st_h[ST_POOL];
    s->last_grid_recomputes = s->st_h[ST_GDIRTY];
    s->last_face_recomputes = s->st_h[ST_FDIRTY];
}

// a completed update: the face lists are valid, and its work counts while timing is on
void update_done(pppm_solver *s)
{
    s->faces_valid = true;
    if (s->timing)
    {
        s->work_updates++;
        s->work_grid += s->st_h[ST_GDIRTY];
        s->work_face += s->st_h[ST_FDIRTY];
    }
}

// the last update's face lists become the previous ones (the lookup reuses their slots) and this
// update writes the others; returns whether the previous lists are valid
bool begin_update(pppm_solver *s, bool first)
{
    const bool use_prev = !first && s->faces_valid;
    if (use_prev)
    {
        s->face_start.swap(s->prev_face_start);
        s->face_nbr.swap(s->prev_face_nbr);
        s->face_slot.swap(s->prev_face_slot);
    }
    s->faces_valid = false;
    return use_prev;
}

// enqueues one attempt of the update sequence (attempt 0 takes the vertices vsrc if given)
int enqueue_attempt(pppm_solver *s, bool first, const float *vsrc, bool use_prev)
{
    // the error word is cleared only when the last status read did not show it clean (no
    // kernel outside a mesh update sets it)
    if (!s->err_clean)
        CK(launch_fill_int(s->status.p + ST_ERR, 1, 0, s->stream));
    s->err_clean = false;
    if (first)
    {
        // fresh caches (GridCache::init / FaceCache::init, pppm_cache.h:29-33, 115-124)
        s->pool_count = 0;
        CK(launch_fill_u64(s->hkeys.p, s->hcap, ~0ull, s->stream));
        CK(launch_fill_int(s->status.p + ST_POOL, 1, 0, s->stream));
    }
    return enqueue_rebuild(s, first, nullptr, 0, 0, vsrc, use_prev);
}

// synchronous update attempts a0..2 (used by set_mesh / update_mesh): runs the sequence, reads
// the status, retries with grown capacities if a list or the pool overflowed
int rebuild_attempts_from(pppm_solver *s, bool first, const float *vsrc, int a0, bool use_prev)
{
    for (int attempt = a0; attempt < 3; attempt++)
    {
        RC(enqueue_attempt(s, first, attempt == 0 ? vsrc : nullptr, use_prev));
        RC(read_status(s));
        const int e = s->st_h[ST_ERR];
        if ((e & ~ERR_CAPACITY) != 0)
            return status_error(s);
        if (!(e & ERR_CAPACITY) && !needs_growth(s))
        {
            update_done(s);
            return PPPM_OK;
        }
        RC(grow(s));
        use_prev = false;  // grow may have moved the previous lists
    }
    return status_error(s);
}

// vsrc (device, may be null): vertex positions to take instead of s->verts; the first
// attempt copies them into s->verts
int rebuild_attempts(pppm_solver *s, bool first, const float *vsrc)
{
    const bool use_prev = begin_update(s, first);
    return rebuild_attempts_from(s, first, vsrc, 0, use_prev);
}

int rebuild_sync(pppm_solver *s, bool first, const float *vsrc = nullptr)
{
    const int rc = rebuild_attempts(s, first, vsrc);
    s->mesh_stale = rc != PPPM_OK;
    if (rc == PPPM_OK)
        s->mesh_version++;
    return rc;
}

// the step and its phases need a mesh whose last update succeeded
int check_ready(pppm_solver *s, const char *what)
{
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, std::string(what) + " before set_mesh");
    if (s->mesh_stale)
        return fail(s, PPPM_ERR_STATE,
                    std::string(what) + ": the last mesh update failed; update_mesh with a valid mesh first");
    return PPPM_OK;
}

int upload(pppm_solver *s, void *dst, const void *src, size_t bytes, int on_device)
{
    CK(hipMemcpyAsync(dst, src, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->stream));
    return PPPM_OK;
}

// per-step work buffers of the gathers: the handle's own, or (pppm_run_noalloc) slices of the
// per-device scratch array
struct StepBufs
{
    float *wn, *wd, *far;
    const int *lcell, *lc3;
};

StepBufs own_bufs(pppm_solver *s) { return StepBufs{s->wn.p, s->wd.p, s->far.p, s->lcell.p, s->lc3.p}; }

// the step kernels' view of the handle (neumann: series rows t - t0 or null; out: listener
// samples, n_listen per step, or null)
SceneStep scene_step(const pppm_solver *s, const StepBufs &b, const float *neumann, int t0, long long stride,
                     int n_listen, float *out)
{
    SceneStep d{};
    d.a = step_args(s, t0, stride, n_listen);
    d.st = s->state.p;
    d.grid3 = s->grid.p;
    d.c3_cells = s->c3_cells.p;
    d.c3_start = s->c3_start.p;
    d.c3_wave = s->c3_wave.p;
    d.c3_pairs = s->c3_pairs.p;
    d.c3_waves = s->c3_waves;
    d.face_waves = s->face_waves;
    d.face_start = s->face_start.p;
    d.face_wave = s->face_wave.p;
    d.face_nbr = s->face_nbr.p;
    d.face_slot = s->face_slot.p;
    d.wf = s->wf.p;
    d.wb = s->wb.p;
    d.fw = s->fw.p;
    d.sf = s->sf.p;
    d.sb = s->sb.p;
    d.fsc = s->fsc.p;
    if (s->ord_version == s->mesh_version)
    {
        d.owf = s->owf.p;
        d.owb = s->owb.p;
        d.ofw = s->ofw.p;
        d.osf = s->osf.p;
        d.osb = s->osb.p;
        d.ofsc = s->ofsc.p;
    }
    d.interp = s->interp.p;
    d.corner = s->corner.p;
    d.geom = s->geom.p;
    d.err = s->status.p + ST_ERR;
    d.hn = s->hn.p;
    d.hd = s->hd.p;
    d.wn = b.wn;
    d.wd = b.wd;
    d.far = b.far;
    d.neumann = neumann;
    d.lcell = b.lcell;
    d.lc3 = b.lc3;
    d.c3_index = s->c3_index.p;
    d.out = out;
    d.out_stride = n_listen;
    return d;
}

// one update_grid_and_face on the handle's stream; neumann is device memory (row t - t0).
// timed: phase spans when timing is on (events are created on first use, so the
// allocation-free path passes false)
// status_event != null: the FDTD launch also copies the mesh update's status block to st_h and
// the event is recorded right behind it (the caller has checked fused_status)
int enqueue_step(pppm_solver *s, const StepBufs &b, const float *neumann, int t0, long long stride, int n_listen,
                 float *out, bool timed = true, hipEvent_t status_event = nullptr)
{
    const FdtdParams fp = fdtd_params(s);
    const SceneStep d = scene_step(s, b, neumann, t0, stride, n_listen, out);
    // a step enqueued behind a mesh update does nothing if that update raises an error bit
    const int *skip = s->mesh_set ? d.err : nullptr;
    int sp = timed ? span_begin(s, 0) : -1;
    if (s->mesh_set && !s->use_pml)
    {
        // the ABC-shell launch also builds the gathers' history windows
        StepFuse fuse{s->hn.p, s->hd.p, d.wn, d.wd, s->nt, nullptr, nullptr, 0};
        if (status_event)
        {
            fuse.status = s->status.p;
            fuse.status_host = s->st_hdev;
            fuse.status_n = ST_N;
        }
        CK(launch_fdtd(s->grid.p, s->pmlbuf.p, fp, s->state.p, s->stream, skip, &fuse));
        if (status_event)
            CK(hipEventRecord(status_event, s->stream));
        span_end(s, sp);
    }
    else
    {
        CK(launch_fdtd(s->grid.p, s->pmlbuf.p, fp, s->state.p, s->stream, skip));
        span_end(s, sp);
        if (s->mesh_set)
            CK(launch_windows(d, 0, s->stream));
    }
    CK(launch_advance(s->state.p, s->stream, skip));
    if (!s->mesh_set)
    {
        if (out)
            CK(launch_listen_direct(d, s->stream));
        return PPPM_OK;
    }
    sp = timed ? span_begin(s, 1) : -1;
    CK(launch_far(d, s->stream));
    span_end(s, sp);
    sp = timed ? span_begin(s, 2) : -1;
    CK(launch_dirichlet(d, s->stream));
    span_end(s, sp);
    sp = timed ? span_begin(s, 3) : -1;
    CK(launch_near(d, s->stream));
    span_end(s, sp);
    if (timed && s->spans.size() > 4096)
        return collect_spans(s);
    return PPPM_OK;
}

int check_handle(pppm_solver *s) { return s ? PPPM_OK : PPPM_ERR_ARG; }

// update_mesh followed by update_grid_and_face (the demo loop's UpdateSound) with the status read
// overlapped with the step: the update, a copy of its status block and the step are enqueued
// together; the step's kernels do nothing if the update raised an error bit, and the host waits
// only for the status. Then, as the synchronous update: an invalid mesh returns its error (the
// step did nothing); an overflowed list or pool grows and the update runs again, followed by the
// step again; a pool short of headroom for the next update grows. The results are those of
// pppm_update_mesh + pppm_step. vsrc: device vertices (null: s->verts already holds them);
// neumann: device rows (row t - t0, stride; null: the ring slot is written elsewhere); out:
// listener samples by a direct read (may be null). The caller advances t_host when it returns OK.
int update_and_step(pppm_solver *s, const float *vsrc, const float *neumann, int t0, long long stride, int n_listen,
                    float *out)
{
    auto &c = s->stc;
    if (!c.ev_status)
        CK(hipEventCreateWithFlags(&c.ev_status, hipEventDisableTiming));
    const bool use_prev = begin_update(s, false);
    // the step's FDTD launch brings the status block to the host (no copy of its own on the way
    // to the step), except on the PML path, which launches no fused FDTD
    const bool fused_status = s->st_hdev && !s->use_pml;
    int sp = span_begin(s, 4);
    RC(enqueue_attempt(s, false, vsrc, use_prev));
    if (!fused_status)
    {
        CK(hipMemcpyAsync(s->st_h, s->status.p, ST_N * sizeof(int), hipMemcpyDeviceToHost, s->stream));
        CK(hipEventRecord(c.ev_status, s->stream));
    }
    span_end(s, sp);
    auto step = [&]() -> int {
        // the lists moved with this update: the near field looks the listeners' C3 indices up
        StepBufs b = own_bufs(s);
        b.lc3 = nullptr;
        return enqueue_step(s, b, neumann, t0, stride, n_listen, out, true, fused_status ? c.ev_status : nullptr);
    };
    RC(step());
    CK(hipEventSynchronize(c.ev_status));
    take_status(s);
    const int e = s->st_h[ST_ERR];
    if (e == 0)
    {
        update_done(s);
        s->mesh_stale = false;
        s->mesh_version++;
        if (needs_growth(s))
        {
            CK(hipStreamSynchronize(s->stream));  // the step's kernels use the buffers grow may move
            RC(grow(s));
        }
        return PPPM_OK;
    }
    CK(hipStreamSynchronize(s->stream));  // the step's kernels are done (they did nothing)
    if ((e & ~ERR_CAPACITY) != 0)
    {
        s->mesh_stale = true;
        return status_error(s);
    }
    RC(grow(s));
    const int rc = rebuild_attempts_from(s, false, nullptr, 1, false);
    s->mesh_stale = rc != PPPM_OK;
    if (rc)
        return rc;
    s->mesh_version++;
    return step();
}

// the weight records of the current mesh version in gather-list order (for static meshes; a
// stale copy is never used: scene_step compares the versions)
int ensure_ordered(pppm_solver *s)
{
    if (!s->mesh_set || s->mesh_stale || s->ord_version == s->mesh_version)
        return PPPM_OK;
    RC(read_status(s));
    const int n3 = GC * s->nt, n4 = s->st_h[ST_P4];
    CK(s->owf.ensure((size_t)n3 * STEPS));
    CK(s->owb.ensure((size_t)n3 * STEPS));
    CK(s->osf.ensure(n3));
    CK(s->osb.ensure(n3));
    CK(s->ofw.ensure((size_t)std::max(n4, 1) * STEPS));
    CK(s->ofsc.ensure(std::max(n4, 1)));
    CK(launch_order_records(s->c3_pairs.p, n3, s->wf.p, s->sf.p, s->wb.p, s->sb.p, s->owf.p, s->osf.p, s->owb.p,
                            s->osb.p, s->face_slot.p, n4, s->pool_cap, s->fw.p, s->fsc.p, s->ofw.p, s->ofsc.p,
                            s->stream));
    s->ord_version = s->mesh_version;
    return PPPM_OK;
}

int set_listeners(pppm_solver *s, const int *listen_xyz, int n_listen)
{
    const int N = s->N;
    std::vector<int> cells(n_listen), c3(n_listen, -1);
    for (int l = 0; l < n_listen; l++)
    {
        const int x = listen_xyz[3 * l], y = listen_xyz[3 * l + 1], z = listen_xyz[3 * l + 2];
        if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N)
            return fail(s, PPPM_ERR_ARG, "listener outside the grid");
        cells[l] = lin3(x, y, z, N);
    }
    // the same cells on the same mesh version: the device tables already hold them (the C3
    // lookups below wait for the stream, which a run of checkpoint segments would do per call)
    const long long version = s->mesh_set ? s->mesh_version : -2;
    if (n_listen > 0 && version == s->lis_version && cells == s->lis_cells)
        return PPPM_OK;
    s->lis_version = -1;
    CK(s->lcell.ensure(std::max(n_listen, 1)));
    CK(s->lc3.ensure(std::max(n_listen, 1)));
    if (n_listen > 0)
    {
        if (s->mesh_set)
        {
            CK(hipStreamSynchronize(s->stream));
            for (int l = 0; l < n_listen; l++)
                CK(hipMemcpy(&c3[l], s->c3_index.p + cells[l], sizeof(int), hipMemcpyDeviceToHost));
        }
        CK(hipMemcpyAsync(s->lcell.p, cells.data(), n_listen * sizeof(int), hipMemcpyHostToDevice, s->stream));
        CK(hipMemcpyAsync(s->lc3.p, c3.data(), n_listen * sizeof(int), hipMemcpyHostToDevice, s->stream));
        CK(hipStreamSynchronize(s->stream));
        s->lis_cells = cells;
        s->lis_version = version;
    }
    return PPPM_OK;
}

// Graph-cache key of a static-mesh run: the step's launch arguments with the per-call rows (the
// Neumann series and the listener output, which the kernels read and write only at
// p + (t - t0) * row) shifted to t = 0, so the segments of one series, run one after another
// (a checkpointed loss), replay one graph.
struct StepGraphKey
{
    SceneStep d;
    FdtdParams fp;
    long long nbase, obase;  // byte addresses of the rows of t = 0 (0: no series / no output)
    int chunk, pml;
};

StepGraphKey step_graph_key(pppm_solver *s, const float *neumann, int t0, long long stride, int n_listen, float *out,
                            int chunk)
{
    StepGraphKey k;
    std::memset(&k, 0, sizeof k);
    SceneStep d = scene_step(s, own_bufs(s), neumann, t0, stride, n_listen, out);
    canonical_bytes(d.a);
    k.nbase = d.neumann ? (long long)(uintptr_t)d.neumann - (long long)t0 * stride * (long long)sizeof(float) : 0;
    k.obase = d.out ? (long long)(uintptr_t)d.out - (long long)t0 * d.out_stride * (long long)sizeof(float) : 0;
    d.neumann = d.neumann ? s->grid.p : nullptr;  // presence only
    d.out = d.out ? s->grid.p : nullptr;
    d.a.t0 = 0;
    std::memcpy(&k.d, &d, sizeof d);
    FdtdParams fp = fdtd_params(s);
    canonical_bytes(fp);
    std::memcpy(&k.fp, &fp, sizeof fp);
    k.chunk = chunk;
    k.pml = s->use_pml ? 1 : 0;
    return k;
}

struct StaticChunk
{
    pppm_solver *s;
    const float *neumann;
    int t0;
    long long stride;
    int n_listen;
    float *out;
    int chunk;
};
int enqueue_static_chunk(void *p)
{
    const StaticChunk &c = *static_cast<const StaticChunk *>(p);
    for (int i = 0; i < c.chunk; i++)
        RC(enqueue_step(c.s, own_bufs(c.s), c.neumann, c.t0, c.stride, c.n_listen, c.out));
    return PPPM_OK;
}

// steps of (optional mesh update +) update_grid_and_face; graph-captured in chunks
int run_impl(pppm_solver *s, const float *verts_series, long long vstride, const float *neumann_series,
             long long stride, int steps, const int *listen_xyz, int n_listen, float *out_dev, int use_graph)
{
    if (steps < 0 || n_listen < 0 || n_listen > 64 || (n_listen > 0 && (!listen_xyz || !out_dev)))
        return fail(s, PPPM_ERR_ARG, "run: bad arguments (at most 64 listeners)");
    if (s->mesh_set && !neumann_series)
        return fail(s, PPPM_ERR_ARG, "run: neumann_series required");
    if (verts_series && !s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "run: moving mesh before set_mesh");
    if (s->mesh_set && !verts_series)
        RC(check_ready(s, "run"));
    if (steps == 0)
        return PPPM_OK;
    if (!verts_series)
        RC(ensure_ordered(s));
    // a listener's C3 index can change with the lists: moving runs read the grid directly
    RC(set_listeners(s, listen_xyz, n_listen));
    if (verts_series && n_listen > 0)
    {
        std::vector<int> none(n_listen, -1);
        CK(hipMemcpy(s->lc3.p, none.data(), n_listen * sizeof(int), hipMemcpyHostToDevice));
        s->lis_version = -1;  // the table no longer holds set_listeners' C3 indices
    }
    float *out = n_listen > 0 ? out_dev : nullptr;
    const int t0 = s->t_host + 1;
    // Moving meshes take the synchronous update of pppm_update_mesh before each step (row i
    // of the series, status read back, capacities grown when needed). The fully
    // asynchronous update (row picked on the device, capturable in a graph) stays off until
    // its graph replay has been verified on the device.
    constexpr bool kAsyncMovingStep = false;
    if (verts_series && !kAsyncMovingStep)
        use_graph = 0;
    auto one_step = [&](int i) -> int {
        if (verts_series && !kAsyncMovingStep)
            // the synchronous update with its status read overlapped with the step
            return update_and_step(s, verts_series + (size_t)i * vstride, neumann_series, t0, stride, n_listen, out);
        if (verts_series)
        {
            const int sp = span_begin(s, 4);
            RC(enqueue_rebuild(s, false, verts_series, vstride, t0));
            span_end(s, sp);
        }
        if (!verts_series)
            return enqueue_step(s, own_bufs(s), neumann_series, t0, stride, n_listen, out);
        // with moving lists the listener is read by its own launch after the near field
        RC(enqueue_step(s, own_bufs(s), neumann_series, t0, stride, n_listen, nullptr));
        if (out)
            CK(launch_listen_direct(scene_step(s, own_bufs(s), nullptr, t0, stride, n_listen, out), s->stream));
        return PPPM_OK;
    };
    if (verts_series)
    {
        CK(launch_fill_int(s->status.p + ST_ERR, 1, 0, s->stream));
        s->err_clean = false;
        s->faces_valid = false;
        RC(ensure_pool(s, s->pool_count + 2 * s->cap_p4));
    }
    int done = 0;
    if (use_graph && !verts_series)
    {
        // a short run (a checkpoint segment) in one graph, a long one in 32-step graphs; the
        // graph is kept for later runs with the same arguments up to the per-call rows
        const bool timing = s->timing;
        s->timing = false;
        const int chunk = steps <= 64 ? steps : 32;
        const StepGraphKey key = step_graph_key(s, neumann_series, t0, stride, n_listen, out, chunk);
        StaticChunk sc{s, neumann_series, t0, stride, n_listen, out, chunk};
        hipGraphExec_t ge = nullptr;
        int rc = solver_cached_graph(s, s->step_graphs, &key, sizeof key, enqueue_static_chunk, &sc, &ge);
        for (; rc == PPPM_OK && done + chunk <= steps; done += chunk)
        {
            const hipError_t le = hipGraphLaunch(ge, s->stream);
            if (le != hipSuccess)
                rc = fail(s, PPPM_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(le));
        }
        for (; rc == PPPM_OK && done < steps; done++)
            rc = one_step(done);
        s->timing = timing;
        if (rc)
        {
            s->t_host += done;
            return rc;
        }
    }
    else if (!use_graph)
    {
        for (; done < steps; done++)
        {
            const int rc = one_step(done);
            if (rc)
            {
                s->t_host += done;  // the steps before this one ran
                return rc;
            }
            if (verts_series && kAsyncMovingStep && (done % 64 == 63 || done == steps - 1))
            {
                RC(read_status(s));
                RC(status_error(s));
                if (needs_growth(s))
                    RC(grow(s));
            }
        }
    }
    else
    {
        // one captured graph holds `chunk` steps; kernels read t (and the input rows) from
        // device memory. Event timing is not captured.
        const bool timing = s->timing;
        s->timing = false;
        const int chunk = std::min(steps, 32);
        while (done < steps)
        {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            CK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            int rc = PPPM_OK;
            for (int i = 0; i < chunk && rc == PPPM_OK; i++)
                rc = one_step(done + i);
            hipError_t ee = hipStreamEndCapture(s->stream, &g);
            if (rc || ee != hipSuccess)
            {
                if (g)
                    (void)hipGraphDestroy(g);
                s->timing = timing;
                return rc ? rc : fail(s, PPPM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ee));
            }
            CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            bool regrow = false;
            while (done + chunk <= steps && !regrow)
            {
                CK(hipGraphLaunch(ge, s->stream));
                done += chunk;
                if (verts_series)
                {
                    RC(read_status(s));
                    RC(status_error(s));
                    regrow = needs_growth(s);
                }
            }
            CK(hipStreamSynchronize(s->stream));
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
            if (regrow)
                RC(grow(s));  // buffers moved: capture again
            else
                break;
        }
        for (; done < steps; done++)
            RC(one_step(done));
        s->timing = timing;
        if (verts_series)
        {
            RC(read_status(s));
            RC(status_error(s));
        }
    }
    s->t_host += steps;
    return PPPM_OK;
}

}  // namespace

namespace pppm_amd
{
int solver_fail(pppm_solver *s, int code, const std::string &msg) { return fail(s, code, msg); }
int solver_check_ready(pppm_solver *s, const char *what) { return check_ready(s, what); }
int solver_set_listeners(pppm_solver *s, const int *listen_xyz, int n_listen)
{
    return set_listeners(s, listen_xyz, n_listen);
}
FdtdParams solver_fdtd_params(const pppm_solver *s) { return fdtd_params(s); }
int solver_read_status(pppm_solver *s) { return read_status(s); }
int solver_ensure_ordered(pppm_solver *s) { return ensure_ordered(s); }
void canonical_bytes(StepArgs &a)
{
    StepArgs z;
    std::memset(&z, 0, sizeof z);
    z.N = a.N;
    z.nt = a.nt;
    z.rc = a.rc;
    z.t0 = a.t0;
    z.stride = a.stride;
    z.n_listen = a.n_listen;
    std::memcpy(&a, &z, sizeof z);
}
void canonical_bytes(FdtdParams &p)
{
    FdtdParams z;
    std::memset(&z, 0, sizeof z);
    z.N = p.N;
    z.coef = p.coef;
    z.hh = p.hh;
    z.c = p.c;
    z.dt = p.dt;
    z.h = p.h;
    z.pml = p.pml;
    z.damp = p.damp;
    std::memcpy(&p, &z, sizeof z);
}
int solver_cached_graph(pppm_solver *s, GraphCache &c, const void *key, size_t n, int (*enqueue)(void *), void *arg,
                        hipGraphExec_t *out)
{
    const unsigned char *kb = static_cast<const unsigned char *>(key);
    for (const auto &x : c.e)
        if (x.key.size() == n && std::memcmp(x.key.data(), kb, n) == 0)
        {
            *out = x.ge;
            return PPPM_OK;
        }
    hipGraph_t g = nullptr;
    CK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(arg);
    const hipError_t ee = hipStreamEndCapture(s->stream, &g);
    if (rc || ee != hipSuccess)
    {
        if (g)
            (void)hipGraphDestroy(g);
        return rc ? rc : fail(s, PPPM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ee));
    }
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CK(ie);
    if (c.e.size() >= GraphCache::kMax)
    {
        CK(hipStreamSynchronize(s->stream));  // the oldest graph may still be running
        (void)hipGraphExecDestroy(c.e.front().ge);
        c.e.erase(c.e.begin());
    }
    c.e.push_back(GraphCache::Entry{std::vector<unsigned char>(kb, kb + n), ge});
    *out = ge;
    return PPPM_OK;
}
SceneStep solver_scene_step(const pppm_solver *s, const float *neumann, int t0, long long stride, int n_listen,
                            float *out)
{
    return scene_step(s, StepBufs{s->wn.p, s->wd.p, s->far.p, s->lcell.p, s->lc3.p}, neumann, t0, stride, n_listen,
                      out);
}
}  // namespace pppm_amd

// ============================================================================ C ABI
extern "C" {

int pppm_create(pppm_solver **out, int res, float dl, float dt, const float *min_xyz, int pml_width,
                float reflect_coeff, float c, unsigned flags)
{
    if (!out)
        return PPPM_ERR_ARG;
    *out = nullptr;
    if (res < 8 || !(dl > 0) || !(dt > 0) || !(c > 0) || pml_width < 0 || 2 * pml_width >= res)
        return PPPM_ERR_ARG;
    pppm_solver *s = new pppm_solver();
    s->N = res;
    s->h = dl;
    s->dt = dt;
    s->c = c;
    s->rc = reflect_coeff;
    if (min_xyz)
        std::memcpy(s->minp, min_xyz, 3 * sizeof(float));
    s->pml_width = pml_width;
    s->flags = flags;
    s->use_pml = (flags & PPPM_FLAG_HONOR_PML) && pml_width > 0;
    tdbem_init(s->bem, dt, c, dl);
    hipError_t e = hipStreamCreate(&s->stream);
    if (e != hipSuccess)
    {
        delete s;
        return PPPM_ERR_HIP;
    }
    if (hipHostMalloc((void **)&s->st_pin.p, ST_N * sizeof(int), hipHostMallocDefault) != hipSuccess)
    {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return PPPM_ERR_HIP;
    }
    s->st_h = s->st_pin.p;
    if (hipHostGetDevicePointer((void **)&s->st_hdev, s->st_pin.p, 0) != hipSuccess)
        s->st_hdev = nullptr;
    std::memset(s->st_h, 0, ST_N * sizeof(int));
    s->own_stream = true;
    if (hipStreamCreateWithFlags(&s->uf.aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->uf.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->uf.join, hipEventDisableTiming) != hipSuccess)
    {
        pppm_destroy(s);
        return PPPM_ERR_HIP;
    }
    const size_t NN = (size_t)res * res * res;
    if (s->grid.ensure(3 * NN) != hipSuccess || s->state.ensure(1) != hipSuccess ||
        (s->use_pml && s->pmlbuf.ensure(3 * NN * 6) != hipSuccess))
    {
        pppm_destroy(s);
        return PPPM_ERR_HIP;
    }
    if (pppm_reset(s) != PPPM_OK)
    {
        pppm_destroy(s);
        return PPPM_ERR_HIP;
    }
    *out = s;
    return PPPM_OK;
}

int pppm_destroy(pppm_solver *s)
{
    if (!s)
        return PPPM_ERR_ARG;
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    for (auto &sp : s->spans)
    {
        (void)hipEventDestroy(sp.a);
        (void)hipEventDestroy(sp.b);
    }
    for (auto e : s->pool)
        (void)hipEventDestroy(e);
    if (s->own_stream && s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;  // device buffers are released by their owners
    return PPPM_OK;
}

int pppm_set_stream(pppm_solver *s, void *hip_stream)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    CK(hipStreamSynchronize(s->stream));
    if (s->own_stream)
        CK(hipStreamDestroy(s->stream));
    s->own_stream = false;
    s->stream = (hipStream_t)hip_stream;
    return PPPM_OK;
}

const char *pppm_last_error(const pppm_solver *s) { return s ? s->err.c_str() : "null handle"; }

int pppm_set_mesh(pppm_solver *s, const float *verts, int nv, const int *tris, int nt, int on_device)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!verts || !tris || nv < 3 || nt < 1)
        return fail(s, PPPM_ERR_ARG, "set_mesh: empty mesh");
    s->nv = nv;
    s->nt = nt;
    s->mesh_set = false;
    CK(s->verts.ensure((size_t)nv * 3));
    CK(s->tris.ensure((size_t)nt * 3));
    CK(s->geom.ensure(nt));
    RC(upload(s, s->verts.p, verts, (size_t)nv * 3 * sizeof(float), on_device));
    RC(upload(s, s->tris.p, tris, (size_t)nt * 3 * sizeof(int), on_device));
    CK(s->wf.ensure((size_t)nt * GC * STEPS));
    CK(s->wb.ensure((size_t)nt * GC * STEPS));
    CK(s->sf.ensure((size_t)nt * GC));
    CK(s->sb.ensure((size_t)nt * GC));
    CK(s->gmeta.ensure((size_t)nt * 7));
    CK(s->gin.ensure((size_t)nt));
    CK(s->interp.ensure((size_t)nt * 8));
    // first guess for the 4x4x4 list capacities; rebuild_sync grows them if needed
    s->cap_b4items = std::max(s->cap_b4items, 16 * nt + 1024);
    s->cap_p4 = std::max(s->cap_p4, 16 * nt + 1024);
    RC(allocate(s));
    s->pool_count = 0;
    RC(ensure_pool(s, 2 * s->cap_p4));
    RC(rebuild_sync(s, true));
    // dirichlet / neumann histories reset (pppm.h:88-91)
    CK(s->hn.ensure((size_t)nt * RING));
    CK(s->hd.ensure((size_t)nt * RING));
    CK(s->wn.ensure((size_t)nt * STEPS));
    CK(s->wd.ensure((size_t)nt * STEPS));
    CK(s->nstage.ensure(nt));
    CK(hipMemsetAsync(s->hn.p, 0, (size_t)nt * RING * sizeof(float), s->stream));
    CK(hipMemsetAsync(s->hd.p, 0, (size_t)nt * RING * sizeof(float), s->stream));
    CK(hipStreamSynchronize(s->stream));
    s->mesh_set = true;
    return PPPM_OK;
}

int pppm_update_mesh(pppm_solver *s, const float *verts, int on_device)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "update_mesh before set_mesh");
    if (!on_device)
        RC(upload(s, s->verts.p, verts, (size_t)s->nv * 3 * sizeof(float), 0));
    const int sp = span_begin(s, 4);
    const int rc = rebuild_sync(s, false, on_device ? verts : nullptr);
    span_end(s, sp);
    return rc;
}

int pppm_set_timing(pppm_solver *s, int enable)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    const int rc = collect_spans(s);
    s->timing = enable != 0;
    return rc;
}

int pppm_get_timing(pppm_solver *s, double *ms8, long long *count8)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    RC(collect_spans(s));
    for (int i = 0; i < 8; i++)
    {
        if (ms8)
            ms8[i] = s->phase_ms[i];
        if (count8)
            count8[i] = s->phase_n[i];
        s->phase_ms[i] = 0;
        s->phase_n[i] = 0;
    }
    return PPPM_OK;
}

int pppm_get_update_work(pppm_solver *s, long long *out3)
{
    if (check_handle(s) || !out3)
        return PPPM_ERR_ARG;
    out3[0] = s->work_updates;
    out3[1] = s->work_grid;
    out3[2] = s->work_face;
    s->work_updates = s->work_grid = s->work_face = 0;
    return PPPM_OK;
}

int pppm_step(pppm_solver *s, const float *neumann, int on_device)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    RC(check_ready(s, "step"));
    const float *nd = neumann;
    if (neumann && !on_device)
    {
        RC(upload(s, s->nstage.p, neumann, (size_t)s->nt * sizeof(float), 0));
        nd = s->nstage.p;
    }
    const int t_new = s->t_host + 1;
    RC(enqueue_step(s, own_bufs(s), nd, t_new, 0, 0, nullptr));
    s->t_host = t_new;
    return PPPM_OK;
}

int pppm_run(pppm_solver *s, const float *neumann_series, long long stride, int steps, const int *listen_xyz,
             int n_listen, float *out_dev, int use_graph)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    return run_impl(s, nullptr, 0, neumann_series, stride, steps, listen_xyz, n_listen, out_dev, use_graph);
}

int pppm_run_noalloc(pppm_solver *s, const float *neumann_series, long long stride, int steps, const int *listen_xyz,
                     int n_listen, float *out_dev)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    RC(check_ready(s, "run_noalloc"));
    if (steps < 0 || !neumann_series || stride < s->nt || n_listen < 0 || n_listen > 64 ||
        (n_listen > 0 && (!listen_xyz || !out_dev)))
        return fail(s, PPPM_ERR_ARG, "run_noalloc: bad arguments (at most 64 listeners, stride >= nt)");
    ListenTable lt{};
    lt.n = n_listen;
    for (int l = 0; l < n_listen; l++)
    {
        for (int d = 0; d < 3; d++)
        {
            const int v = listen_xyz[3 * l + d];
            if (v < 0 || v >= s->N)
                return fail(s, PPPM_ERR_ARG, "run_noalloc: listener outside the grid");
            lt.xyz[3 * l + d] = v;
        }
    }
    // scratch layout: listener cells and C3 indices (2 x 64 ints), the two history windows
    // (16 floats per triangle each), the far field of the C3 cells (capacity-sized)
    const size_t win = (size_t)s->nt * STEPS * sizeof(float);
    const size_t off_wn = 512, off_wd = off_wn + ((win + 255) & ~(size_t)255);
    const size_t off_far = off_wd + ((win + 255) & ~(size_t)255);
    const size_t need = off_far + (size_t)s->cap_c3 * sizeof(float);
    if (need > kScratchBytes)
        return fail(s, PPPM_ERR_CAPACITY,
                    "run_noalloc: the mesh needs " + std::to_string(need) + " bytes of scratch, the maximum is " +
                        std::to_string(kScratchBytes) + " (use pppm_run)");
    if (steps == 0)
        return PPPM_OK;
    void *base = nullptr;
    CK(scratch_ptr(&base));
    unsigned char *sb = (unsigned char *)base;
    int *lcell = (int *)sb, *lc3 = (int *)sb + 64;
    const StepBufs b{(float *)(sb + off_wn), (float *)(sb + off_wd), (float *)(sb + off_far), lcell, lc3};
    // every scratch entry a kernel reads is written earlier in the same call: the listener
    // table here, the windows by each step's window copy, the far field of the live C3 cells by
    // each step's far-field gather (the Dirichlet update and the near field read only those)
    if (n_listen > 0)
        CK(launch_listen_setup(lt, s->N, s->c3_index.p, lcell, lc3, s->stream));
    const int t0 = s->t_host + 1;
    float *out = n_listen > 0 ? out_dev : nullptr;
    for (int i = 0; i < steps; i++)
        RC(enqueue_step(s, b, neumann_series, t0, stride, n_listen, out, false));
    s->t_host += steps;
    return PPPM_OK;
}

int pppm_run_moving(pppm_solver *s, const float *verts_series, long long vstride, const float *neumann_series,
                    long long stride, int steps, const int *listen_xyz, int n_listen, float *out_dev, int use_graph)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!verts_series || vstride < (long long)s->nv * 3)
        return fail(s, PPPM_ERR_ARG, "run_moving: vertex series required (stride >= 3 nv)");
    return run_impl(s, verts_series, vstride, neumann_series, stride, steps, listen_xyz, n_listen, out_dev,
                    use_graph);
}

int pppm_update_mesh_and_step(pppm_solver *s, const float *verts, const float *neumann, int on_device)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "update_mesh before set_mesh");
    if (!verts)
        return fail(s, PPPM_ERR_ARG, "update_mesh_and_step: null vertices");
    const float *nd = neumann;
    if (!on_device)
    {
        RC(upload(s, s->verts.p, verts, (size_t)s->nv * 3 * sizeof(float), 0));
        if (neumann)
        {
            RC(upload(s, s->nstage.p, neumann, (size_t)s->nt * sizeof(float), 0));
            nd = s->nstage.p;
        }
    }
    RC(update_and_step(s, on_device ? verts : nullptr, nd, s->t_host + 1, 0, 0, nullptr));
    s->t_host++;
    return PPPM_OK;
}

int pppm_fdtd_step(pppm_solver *s)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    const FdtdParams fp = fdtd_params(s);
    CK(launch_fdtd(s->grid.p, s->pmlbuf.p, fp, s->state.p, s->stream));
    CK(launch_advance(s->state.p, s->stream));
    s->t_host++;
    return PPPM_OK;
}

int pppm_solve_far(pppm_solver *s)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    RC(check_ready(s, "solve_far"));
    const SceneStep d = scene_step(s, own_bufs(s), nullptr, s->t_host, 0, 0, nullptr);
    CK(launch_windows(d, 1, s->stream));
    CK(launch_far(d, s->stream));
    return PPPM_OK;
}

int pppm_set_neumann(pppm_solver *s, const float *neumann, int on_device)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!neumann)
        return fail(s, PPPM_ERR_ARG, "set_neumann: null input");
    RC(check_ready(s, "set_neumann"));
    const float *nd = neumann;
    if (!on_device)
    {
        RC(upload(s, s->nstage.p, neumann, (size_t)s->nt * sizeof(float), 0));
        nd = s->nstage.p;
    }
    CK(launch_set_neumann(scene_step(s, own_bufs(s), nd, s->t_host, 0, 0, nullptr), s->stream));
    return PPPM_OK;
}

int pppm_update_dirichlet(pppm_solver *s)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    RC(check_ready(s, "update_dirichlet"));
    const SceneStep d = scene_step(s, own_bufs(s), nullptr, s->t_host, 0, 0, nullptr);
    CK(launch_windows(d, 1, s->stream));
    CK(launch_dirichlet(d, s->stream));
    return PPPM_OK;
}

int pppm_solve_near(pppm_solver *s)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    RC(check_ready(s, "solve_near"));
    const SceneStep d = scene_step(s, own_bufs(s), nullptr, s->t_host, 0, 0, nullptr);
    CK(launch_windows(d, 1, s->stream));
    CK(launch_near(d, s->stream));
    return PPPM_OK;
}

int pppm_reset(pppm_solver *s)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    const size_t NN = (size_t)s->N * s->N * s->N;
    CK(hipMemsetAsync(s->grid.p, 0, 3 * NN * sizeof(float), s->stream));
    if (s->use_pml)
        CK(hipMemsetAsync(s->pmlbuf.p, 0, 3 * NN * 6 * sizeof(float), s->stream));
    if (s->mesh_set)
    {
        CK(hipMemsetAsync(s->hn.p, 0, (size_t)s->nt * RING * sizeof(float), s->stream));
        CK(hipMemsetAsync(s->hd.p, 0, (size_t)s->nt * RING * sizeof(float), s->stream));
    }
    static_assert(sizeof(StepState) == 2 * sizeof(int), "StepState is {t, pad}");
    CK(launch_put_ints(reinterpret_cast<int *>(s->state.p), 2, -1, 0, 0, 0, s->stream));
    s->t_host = -1;
    return PPPM_OK;
}

int pppm_time_idx(const pppm_solver *s, int *t)
{
    if (!s || !t)
        return PPPM_ERR_ARG;
    *t = s->t_host;
    return PPPM_OK;
}

int pppm_set_time_idx(pppm_solver *s, int t)
{
    if (check_handle(s) || t < -1)
        return PPPM_ERR_ARG;
    CK(launch_put_ints(reinterpret_cast<int *>(s->state.p), 2, t, 0, 0, 0, s->stream));
    s->t_host = t;
    return PPPM_OK;
}

int pppm_grid_ptr(pppm_solver *s, int rel, float **dev_ptr)
{
    if (check_handle(s) || !dev_ptr)
        return PPPM_ERR_ARG;
    const size_t NN = (size_t)s->N * s->N * s->N;
    const int l = (((s->t_host + rel) % 3) + 3) % 3;
    *dev_ptr = s->grid.p + (size_t)l * NN;
    return PPPM_OK;
}

int pppm_read_cells(pppm_solver *s, const int *xyz, int n, float *host_out)
{
    if (check_handle(s) || (n > 0 && (!xyz || !host_out)))
        return PPPM_ERR_ARG;
    float *g = nullptr;
    pppm_grid_ptr(s, 0, &g);
    const int N = s->N;
    for (int i = 0; i < n; i++)
    {
        const int x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N)
            return fail(s, PPPM_ERR_ARG, "read_cells: cell outside the grid");
        CK(hipMemcpyAsync(host_out + i, g + lin3(x, y, z, N), sizeof(float), hipMemcpyDeviceToHost, s->stream));
    }
    CK(hipStreamSynchronize(s->stream));
    return PPPM_OK;
}

int pppm_slice_image(pppm_solver *s, int axis, int index, float max_value, unsigned char *rgba_host, int *width)
{
    if (check_handle(s) || axis < 0 || axis > 2 || index < 0 || index >= s->N || !(max_value > 0.0f))
        return PPPM_ERR_ARG;
    if (s->N > 1080)
        return fail(s, PPPM_ERR_ARG, "slice_image: the image is 1080 pixels wide; the grid is wider");
    const int W = s->N * (1080 / s->N);
    if (width)
        *width = W;
    if (!rgba_host)
        return PPPM_OK;
    const size_t WW = (size_t)W * W;
    CK(s->img_cover.ensure(WW));
    CK(s->img_count.ensure(WW));
    CK(s->img_rgba.ensure(4 * WW));
    CK(launch_slice_image(s->N, s->h, s->minp, axis, index, max_value, s->state.p, s->grid.p,
                          s->mesh_set ? s->verts.p : nullptr, s->mesh_set ? s->tris.p : nullptr,
                          s->mesh_set ? s->nt : 0, s->img_cover.p, s->img_count.p, s->img_rgba.p, s->stream));
    CK(hipMemcpyAsync(rgba_host, s->img_rgba.p, 4 * WW, hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PPPM_OK;
}

int pppm_ffat_collect(pppm_solver *s, int w, float *map_dev)
{
    if (check_handle(s) || !map_dev)
        return PPPM_ERR_ARG;
    const int c = s->N / 2;
    if (w < 1 || c - w < 0 || c + w > s->N - 1)
        return fail(s, PPPM_ERR_ARG, "ffat_collect: the cube of half width w must lie inside the grid");
    CK(launch_ffat(s->state.p, s->grid.p, s->N, w, map_dev, s->stream));
    return PPPM_OK;
}

int pppm_history_ptrs(pppm_solver *s, float **neumann, float **dirichlet)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    if (neumann)
        *neumann = s->hn.p;
    if (dirichlet)
        *dirichlet = s->hd.p;
    return PPPM_OK;
}

int pppm_triangles(pppm_solver *s, float *host13)
{
    if (check_handle(s) || !host13)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    std::vector<TriGeom> g(s->nt);
    CK(hipMemcpyAsync(g.data(), s->geom.p, (size_t)s->nt * sizeof(TriGeom), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < s->nt; i++)
    {
        const TriGeom &t = g[i];
        float v[13] = {t.center.x, t.center.y, t.center.z, t.normal.x, t.normal.y, t.normal.z, t.area,
                       (float)t.cell3.x, (float)t.cell3.y, (float)t.cell3.z, (float)t.base3.x, (float)t.base3.y,
                       (float)t.base3.z};
        std::memcpy(host13 + 13 * i, v, sizeof(v));
    }
    return PPPM_OK;
}

int pppm_stats(pppm_solver *s, long long *o)
{
    if (check_handle(s) || !o)
        return PPPM_ERR_ARG;
    if (s->mesh_set)
        RC(read_status(s));
    o[0] = s->nt;
    o[1] = s->mesh_set ? s->st_h[ST_C3] : 0;
    o[2] = s->mesh_set ? 27LL * s->nt : 0;
    o[3] = s->mesh_set ? s->st_h[ST_B4] : 0;
    o[4] = s->mesh_set ? s->st_h[ST_P4] : 0;
    o[5] = s->last_grid_recomputes;
    o[6] = s->last_face_recomputes;
    o[7] = s->pool_count;
    return PPPM_OK;
}

static void to72(const uint32_t *w16, float2 sc, LWH72 &r)
{
    for (int k = 0; k < STEPS; k++)
    {
        r.s[k] = (uint16_t)(w16[k] & 0xffffu);
        r.d[k] = (uint16_t)(w16[k] >> 16);
    }
    r.ss = sc.x;
    r.sd = sc.y;
}
static void from72(const LWH72 &r, uint32_t *w16, float2 &sc)
{
    for (int k = 0; k < STEPS; k++)
        w16[k] = (uint32_t)r.s[k] | ((uint32_t)r.d[k] << 16);
    sc = make_float2(r.ss, r.sd);
}

int pppm_get_grid_weights(pppm_solver *s, void *host_recs72)
{
    if (check_handle(s) || !host_recs72)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    const size_t R = (size_t)s->nt * GC;
    std::vector<uint32_t> wf(R * STEPS), wb(R * STEPS);
    std::vector<float2> sf(R), sb(R);
    CK(hipMemcpyAsync(wf.data(), s->wf.p, R * STEPS * 4, hipMemcpyDeviceToHost, s->stream));
    CK(hipMemcpyAsync(wb.data(), s->wb.p, R * STEPS * 4, hipMemcpyDeviceToHost, s->stream));
    CK(hipMemcpyAsync(sf.data(), s->sf.p, R * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    CK(hipMemcpyAsync(sb.data(), s->sb.p, R * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    LWH72 *o = (LWH72 *)host_recs72;
    for (size_t r = 0; r < R; r++)
    {
        to72(&wf[r * STEPS], sf[r], o[2 * r]);
        to72(&wb[r * STEPS], sb[r], o[2 * r + 1]);
    }
    return PPPM_OK;
}

int pppm_set_grid_weights(pppm_solver *s, const void *host_recs72)
{
    if (s)
        s->ord_version = -1;  // the list-ordered copies no longer match
    if (check_handle(s) || !host_recs72)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    const size_t R = (size_t)s->nt * GC;
    std::vector<uint32_t> wf(R * STEPS), wb(R * STEPS);
    std::vector<float2> sf(R), sb(R);
    const LWH72 *in = (const LWH72 *)host_recs72;
    for (size_t r = 0; r < R; r++)
    {
        from72(in[2 * r], &wf[r * STEPS], sf[r]);
        from72(in[2 * r + 1], &wb[r * STEPS], sb[r]);
    }
    CK(hipMemcpyAsync(s->wf.p, wf.data(), R * STEPS * 4, hipMemcpyHostToDevice, s->stream));
    CK(hipMemcpyAsync(s->wb.p, wb.data(), R * STEPS * 4, hipMemcpyHostToDevice, s->stream));
    CK(hipMemcpyAsync(s->sf.p, sf.data(), R * sizeof(float2), hipMemcpyHostToDevice, s->stream));
    CK(hipMemcpyAsync(s->sb.p, sb.data(), R * sizeof(float2), hipMemcpyHostToDevice, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PPPM_OK;
}

int pppm_get_face_weights(pppm_solver *s, int *host_start, int *host_nbr, void *host_recs72)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    RC(read_status(s));
    const int P = s->st_h[ST_P4];
    std::vector<int> start(s->nt + 1), nbr(P), slot(P);
    CK(hipMemcpyAsync(start.data(), s->face_start.p, (s->nt + 1) * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    CK(hipMemcpyAsync(nbr.data(), s->face_nbr.p, P * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    CK(hipMemcpyAsync(slot.data(), s->face_slot.p, P * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    std::vector<uint32_t> fw((size_t)s->pool_count * STEPS);
    std::vector<float2> fsc(s->pool_count);
    if (s->pool_count > 0)
    {
        CK(hipMemcpyAsync(fw.data(), s->fw.p, fw.size() * 4, hipMemcpyDeviceToHost, s->stream));
        CK(hipMemcpyAsync(fsc.data(), s->fsc.p, fsc.size() * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    }
    CK(hipStreamSynchronize(s->stream));
    if (host_start)
        std::memcpy(host_start, start.data(), start.size() * sizeof(int));
    if (host_nbr)
        std::memcpy(host_nbr, nbr.data(), nbr.size() * sizeof(int));
    if (host_recs72)
    {
        LWH72 *o = (LWH72 *)host_recs72;
        for (int p = 0; p < P; p++)
            to72(&fw[(size_t)slot[p] * STEPS], fsc[slot[p]], o[p]);
    }
    return PPPM_OK;
}

int pppm_set_face_weights(pppm_solver *s, const void *host_recs72)
{
    if (s)
        s->ord_version = -1;  // the list-ordered copies no longer match
    if (check_handle(s) || !host_recs72)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    RC(read_status(s));
    const int P = s->st_h[ST_P4];
    std::vector<int> slot(P);
    CK(hipMemcpyAsync(slot.data(), s->face_slot.p, P * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    std::vector<uint32_t> fw((size_t)s->pool_count * STEPS);
    std::vector<float2> fsc(s->pool_count);
    CK(hipMemcpyAsync(fw.data(), s->fw.p, fw.size() * 4, hipMemcpyDeviceToHost, s->stream));
    CK(hipMemcpyAsync(fsc.data(), s->fsc.p, fsc.size() * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    const LWH72 *in = (const LWH72 *)host_recs72;
    for (int p = 0; p < P; p++)
        from72(in[p], &fw[(size_t)slot[p] * STEPS], fsc[slot[p]]);
    CK(hipMemcpyAsync(s->fw.p, fw.data(), fw.size() * 4, hipMemcpyHostToDevice, s->stream));
    CK(hipMemcpyAsync(s->fsc.p, fsc.data(), fsc.size() * sizeof(float2), hipMemcpyHostToDevice, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PPPM_OK;
}

int pppm_get_interp_weights(pppm_solver *s, float *host_out8)
{
    if (check_handle(s) || !host_out8)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    CK(hipMemcpyAsync(host_out8, s->interp.p, (size_t)s->nt * 8 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PPPM_OK;
}

int pppm_get_c3(pppm_solver *s, int *cells, int *start, int *pairs)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    RC(read_status(s));
    const int n = s->st_h[ST_C3];
    if (cells)
        CK(hipMemcpyAsync(cells, s->c3_cells.p, n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (start)
        CK(hipMemcpyAsync(start, s->c3_start.p, (n + 1) * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (pairs)
        CK(hipMemcpyAsync(pairs, s->c3_pairs.p, (size_t)27 * s->nt * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PPPM_OK;
}

int pppm_get_b4(pppm_solver *s, int *cells, int *start, int *tris)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "no mesh");
    RC(read_status(s));
    const int n = s->st_h[ST_B4], nt = s->nt;
    // the time step reads per-triangle face lists; the base-cell lists are built on request
    const int NN = s->N * s->N * s->N;
    const UpdateCounters cn{s->counters.p, s->counters.p + NN, s->counters.p + 2 * NN, s->counters.p + 3 * NN, {}};
    ListArgs la = list_args(s, cn);
    CK(launch_b4_lists(la, 0, s->stream));
    {
        ScanBatch b{};
        b.njobs = 1;
        b.job[0] = ScanJob{s->b4_cnt.p, nt, 0, s->b4_start.p, nullptr, nullptr, nullptr, 0};
        CK(scan_batch(b, s->scan_tmp.p, s->stream));
    }
    int items = 0;
    CK(hipMemcpyAsync(&items, s->b4_start.p + nt, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    if (items > s->cap_b4items)
    {
        s->cap_b4items = items + items / 4 + 1024;
        CK(s->b4_tris.ensure(s->cap_b4items));
        la.b4_tris = s->b4_tris.p;
        la.cap_b4items = s->cap_b4items;
    }
    CK(launch_b4_lists(la, 1, s->stream));
    if (cells)
        CK(hipMemcpyAsync(cells, s->b4_cells.p, n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (start)
        CK(hipMemcpyAsync(start, s->b4_start.p, (n + 1) * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (tris)
        CK(hipMemcpyAsync(tris, s->b4_tris.p, (size_t)items * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PPPM_OK;
}

// ---------------------------------------------------------------- cell classification
int pppm_classify(pppm_solver *s, int thin_shell, int *ghost_count)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return fail(s, PPPM_ERR_STATE, "classify before set_mesh");
    const int N = s->N, nt = s->nt, nv = s->nv;
    const long long NN = (long long)N * N * N;
    hipStream_t st = s->stream;
    CK(s->ctype.ensure(NN));
    CK(s->ctype_old.ensure(NN));
    CK(s->cghost.ensure(NN));
    CK(s->cnearest.ensure(NN));
    CK(s->ctyp2.ensure(NN));
    CK(s->cflag.ensure(NN));
    CK(s->cscan.ensure(NN + 1));
    CK(s->cpoint.ensure(3 * NN));
    CK(s->creflect.ensure(3 * NN));
    CK(s->cdist.ensure(NN));
    CK(s->vf_cnt.ensure(nv));
    CK(s->vf_fill.ensure(nv));
    CK(s->vf_start.ensure(nv + 1));
    CK(s->vf_items.ensure(3 * (size_t)nt));
    CK(s->vnormal.ensure(3 * (size_t)nv));
    CK(s->scan_tmp.ensure(scan_tmp_ints((int)std::max<long long>(NN, nv))));
    // the previous labels are kept for the fresh-cell test (cell_data_old.assign(cell_data),
    // ghost_cell.h:101); before the first classification every cell counts as AIR
    if (s->classified)
        CK(hipMemcpyAsync(s->ctype_old.p, s->ctype.p, NN * sizeof(int), hipMemcpyDeviceToDevice, st));
    else
        CK(launch_fill_int(s->ctype_old.p, NN, CELL_AIR, st));
    CK(launch_vertex_faces(s->tris.p, nt, nv, s->vf_cnt.p, s->vf_start.p, s->vf_fill.p, s->vf_items.p,
                           s->scan_tmp.p, st));
    CK(launch_vertex_normals(s->verts.p, s->tris.p, s->geom.p, nv, s->vf_start.p, s->vf_items.p, s->vnormal.p, st));
    // the mesh update leaves each cell's triangles in atomic-slot order; the nearest-triangle
    // search breaks ties by the first triangle in ascending order
    CK(launch_sort_cells(s->geom.p, nt, s->cell_start.p, s->cell_items.p, s->cell_rank.p, 0, st));
    const CellArgs a{N, s->h, s->minp[0], s->minp[1], s->minp[2], thin_shell ? 1 : 0};
    CK(launch_classify(a, s->verts.p, s->tris.p, s->cell_start.p, s->cell_items.p, s->vnormal.p, s->ctype.p,
                       s->cnearest.p, s->cpoint.p, s->creflect.p, s->cdist.p, s->ctyp2.p, s->cflag.p, s->cscan.p,
                       s->scan_tmp.p, s->cghost.p, st));
    int g = 0;
    CK(hipMemcpyAsync(&g, s->cscan.p + NN, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    s->ghost_count = g;
    s->classified = true;
    if (ghost_count)
        *ghost_count = g;
    return PPPM_OK;
}

int pppm_fill_fresh(pppm_solver *s, const float *acc_now, const float *acc_old, int on_device, int *fresh_count)
{
    if (check_handle(s) || !acc_now || !acc_old)
        return PPPM_ERR_ARG;
    if (!s->classified)
        return fail(s, PPPM_ERR_STATE, "fill_fresh before classify");
    const int N = s->N, nt = s->nt;
    const long long NN = (long long)N * N * N;
    hipStream_t st = s->stream;
    const float *an = acc_now, *ao = acc_old;
    if (!on_device)
    {
        CK(s->acc_stage.ensure(2 * (size_t)nt));
        RC(upload(s, s->acc_stage.p, acc_now, nt * sizeof(float), 0));
        RC(upload(s, s->acc_stage.p + nt, acc_old, nt * sizeof(float), 0));
        an = s->acc_stage.p;
        ao = s->acc_stage.p + nt;
    }
    CK(s->fresh_list.ensure(NN));
    CK(s->fresh_err.ensure(1));
    CK(launch_fill_int(s->fresh_err.p, 1, 0, st));
    CK(launch_fresh_flags(NN, nt, s->ctype_old.p, s->ctype.p, s->cnearest.p, s->cflag.p, s->fresh_err.p, st));
    {
        ScanBatch b{};
        b.njobs = 1;
        b.job[0] = ScanJob{s->cflag.p, (int)NN, 1, s->cscan.p, s->fresh_list.p, nullptr, nullptr, 0};
        CK(scan_batch(b, s->scan_tmp.p, st));
    }
    const CellArgs a{N, s->h, s->minp[0], s->minp[1], s->minp[2], 0};
    CK(launch_fresh_fill(a, s->state.p, s->fresh_list.p, s->cscan.p + NN, (int)NN, s->fresh_err.p, s->ctype_old.p,
                         s->ctype.p, s->cnearest.p, s->cpoint.p, an, ao, s->grid.p, st));
    int cnt = 0, err = 0;
    CK(hipMemcpyAsync(&cnt, s->cscan.p + NN, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(&err, s->fresh_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (fresh_count)
        *fresh_count = err ? -1 : cnt;
    return PPPM_OK;
}

int pppm_set_solid_zero(pppm_solver *s)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->classified)
        return fail(s, PPPM_ERR_STATE, "set_solid_zero before classify");
    const long long NN = (long long)s->N * s->N * s->N;
    CK(launch_solid_zero(NN, s->state.p, s->ctype.p, s->grid.p, s->stream));
    return PPPM_OK;
}

int pppm_ghost_setup(pppm_solver *s, float cond_threshold, int *second_order)
{
    if (check_handle(s) || !(cond_threshold >= 0.0f))
        return PPPM_ERR_ARG;
    if (!s->classified)
        return fail(s, PPPM_ERR_STATE, "ghost_setup before classify");
    const int n = s->ghost_count, nt = s->nt;
    hipStream_t st = s->stream;
    CK(s->gh_cell.ensure(n));
    CK(s->gh_order.ensure(n));
    CK(s->gh_pw.ensure(8 * (size_t)n));
    CK(s->gh_cond.ensure(n));
    CK(s->gh_mcol.ensure(9 * (size_t)n));
    CK(s->gh_mval.ensure(9 * (size_t)n));
    CK(s->gh_b.ensure(n));
    CK(s->gh_x.ensure(n));
    CK(s->gh_work.ensure(7 * (size_t)n));
    CK(s->gh_info.ensure(1));
    CK(s->gh_err.ensure(1));
    if (s->gh_nt != nt)
    {
        // the Neumann data start at zero (GhostCellSolver::set_mesh, ghost_cell.h:88-93)
        CK(s->gh_neu.ensure(nt));
        CK(s->gh_neu_old.ensure(nt));
        CK(launch_fill_int((int *)s->gh_neu.p, nt, 0, st));
        CK(launch_fill_int((int *)s->gh_neu_old.p, nt, 0, st));
        s->gh_nt = nt;
    }
    CK(launch_fill_int(s->gh_err.p, 1, 0, st));
    CK(launch_fill_int(s->gh_info.p, 1, 0, st));
    const CellArgs a{s->N, s->h, s->minp[0], s->minp[1], s->minp[2], 0};
    CK(launch_ghost_setup(a, n, s->ctype.p, s->cghost.p, s->cnearest.p, s->cpoint.p, s->creflect.p, cond_threshold,
                          s->gh_cell.p, s->gh_order.p, s->gh_pw.p, s->gh_cond.p, s->gh_mcol.p, s->gh_mval.p,
                          s->gh_err.p, st));
    int err = 0;
    std::vector<int> order(n > 0 ? n : 1);
    CK(hipMemcpyAsync(&err, s->gh_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    if (n > 0)
        CK(hipMemcpyAsync(order.data(), s->gh_order.p, n * sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    s->ghost_ready = false;
    if (err & ERR_OUT_OF_GRID)
        return fail(s, PPPM_ERR_OUT_OF_GRID, "ghost_setup: a ghost cell's stencil leaves the grid");
    if (err)
        return fail(s, PPPM_ERR_STATE, "ghost_setup: a ghost cell has no nearest triangle");
    s->gh_n = n;
    s->ghost_threshold = cond_threshold;
    s->ghost_linear = cond_threshold > 0.0f;
    s->ghost_ready = true;
    if (second_order)
    {
        int c = 0;
        for (int i = 0; i < n; i++)
            c += order[i];
        *second_order = c;
    }
    return PPPM_OK;
}

int pppm_ghost_set_neumann(pppm_solver *s, const float *neumann, int on_device)
{
    if (check_handle(s) || !neumann)
        return PPPM_ERR_ARG;
    if (!s->ghost_ready)
        return fail(s, PPPM_ERR_STATE, "ghost_set_neumann before ghost_setup");
    // previous <- current by swapping the two buffers, then the new data into current
    s->gh_neu.swap(s->gh_neu_old);
    RC(upload(s, s->gh_neu.p, neumann, s->nt * sizeof(float), on_device));
    return PPPM_OK;
}

int pppm_ghost_solve(pppm_solver *s, int *iterations)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->ghost_ready)
        return fail(s, PPPM_ERR_STATE, "ghost_solve before ghost_setup");
    const CellArgs a{s->N, s->h, s->minp[0], s->minp[1], s->minp[2], 0};
    CK(launch_ghost_solve(a, s->gh_n, s->state.p, s->grid.p, s->gh_cell.p, s->creflect.p, s->cdist.p, s->cnearest.p,
                          s->ctype.p, s->cghost.p, s->gh_order.p, s->gh_pw.p, s->gh_neu.p, s->gh_mcol.p,
                          s->gh_mval.p, s->ghost_linear, s->gh_b.p, s->gh_x.p, s->gh_work.p, s->gh_info.p,
                          s->stream));
    if (iterations)
    {
        int it = 0;
        if (s->ghost_linear && s->gh_n > 0)
        {
            CK(hipMemcpyAsync(&it, s->gh_info.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
            CK(hipStreamSynchronize(s->stream));
        }
        *iterations = it;
    }
    return PPPM_OK;
}

int pppm_ghost_step(pppm_solver *s, const float *neumann, int on_device)
{
    if (check_handle(s) || !neumann)
        return PPPM_ERR_ARG;
    if (!s->ghost_ready)
        return fail(s, PPPM_ERR_STATE, "ghost_step before ghost_setup");
    RC(pppm_fdtd_step(s));
    RC(pppm_ghost_set_neumann(s, neumann, on_device));
    RC(pppm_ghost_solve(s, nullptr));
    return pppm_set_solid_zero(s);
}

int pppm_ghost_update_mesh(pppm_solver *s, const float *verts, int on_device, int *fresh_count)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->ghost_ready)
        return fail(s, PPPM_ERR_STATE, "ghost_update_mesh before ghost_setup");
    RC(pppm_update_mesh(s, verts, on_device));
    RC(pppm_classify(s, s->ghost_threshold == 0.0f ? 1 : 0, nullptr));
    RC(pppm_fill_fresh(s, s->gh_neu.p, s->gh_neu_old.p, 1, fresh_count));
    return pppm_ghost_setup(s, s->ghost_threshold, nullptr);
}

int pppm_ghost_get(pppm_solver *s, int *cells, int *order, float *weights8, float *cond, float *rhs, float *x,
                   int *cols9, float *vals9)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->ghost_ready)
        return fail(s, PPPM_ERR_STATE, "ghost_get before ghost_setup");
    const size_t n = s->gh_n;
    hipStream_t st = s->stream;
    if (n > 0)
    {
        if (cells)
            CK(hipMemcpyAsync(cells, s->gh_cell.p, n * sizeof(int), hipMemcpyDeviceToHost, st));
        if (order)
            CK(hipMemcpyAsync(order, s->gh_order.p, n * sizeof(int), hipMemcpyDeviceToHost, st));
        if (weights8)
            CK(hipMemcpyAsync(weights8, s->gh_pw.p, 8 * n * sizeof(float), hipMemcpyDeviceToHost, st));
        if (cond)
            CK(hipMemcpyAsync(cond, s->gh_cond.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
        if (rhs)
            CK(hipMemcpyAsync(rhs, s->gh_b.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
        if (x)
            CK(hipMemcpyAsync(x, s->ghost_linear ? s->gh_x.p : s->gh_b.p, n * sizeof(float), hipMemcpyDeviceToHost,
                              st));
        if (cols9)
            CK(hipMemcpyAsync(cols9, s->gh_mcol.p, 9 * n * sizeof(int), hipMemcpyDeviceToHost, st));
        if (vals9)
            CK(hipMemcpyAsync(vals9, s->gh_mval.p, 9 * n * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    CK(hipStreamSynchronize(st));
    return PPPM_OK;
}

int pppm_get_cells(pppm_solver *s, int *type, int *ghost_idx, int *nearest, float *point3, float *reflect3,
                   float *dist)
{
    if (check_handle(s))
        return PPPM_ERR_ARG;
    if (!s->classified)
        return fail(s, PPPM_ERR_STATE, "get_cells before classify");
    const size_t NN = (size_t)s->N * s->N * s->N;
    hipStream_t st = s->stream;
    if (type)
        CK(hipMemcpyAsync(type, s->ctype.p, NN * sizeof(int), hipMemcpyDeviceToHost, st));
    if (ghost_idx)
        CK(hipMemcpyAsync(ghost_idx, s->cghost.p, NN * sizeof(int), hipMemcpyDeviceToHost, st));
    if (nearest)
        CK(hipMemcpyAsync(nearest, s->cnearest.p, NN * sizeof(int), hipMemcpyDeviceToHost, st));
    if (point3)
        CK(hipMemcpyAsync(point3, s->cpoint.p, 3 * NN * sizeof(float), hipMemcpyDeviceToHost, st));
    if (reflect3)
        CK(hipMemcpyAsync(reflect3, s->creflect.p, 3 * NN * sizeof(float), hipMemcpyDeviceToHost, st));
    if (dist)
        CK(hipMemcpyAsync(dist, s->cdist.p, NN * sizeof(float), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    return PPPM_OK;
}

}  // extern "C"
