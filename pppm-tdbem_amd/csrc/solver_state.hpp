// State of a solver handle (struct pppm_solver of the C ABI), shared by the translation units
// that implement the ABI (solver.cpp: one scene; batch.cpp: scenes stepped together).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace pppm_amd
{

// owning device buffer; grows on demand, contents are not preserved by ensure()
template <class T>
struct DBuf
{
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    hipError_t ensure(size_t m)
    {
        if (m == 0)
            m = 1;
        if (p && m <= n)
            return hipSuccess;
        release();
        hipError_t e = hipMalloc((void **)&p, m * sizeof(T));
        if (e != hipSuccess)
        {
            p = nullptr;
            n = 0;
            return e;
        }
        n = m;
        return hipSuccess;
    }
    void swap(DBuf &o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void take(DBuf &o)
    {
        release();
        p = o.p;
        n = o.n;
        o.p = nullptr;
        o.n = 0;
    }
};

// device status block layout: error bits, live counts written by the update's scans, the
// face-cache pool counter
enum
{
    ST_ERR = 0,
    ST_C3 = 1,      // marked (3x3x3) cells
    ST_B4 = 2,      // non-empty base cells
    ST_P4 = 3,      // face pairs
    ST_GDIRTY = 4,  // grid-cache triangles to recompute
    ST_FDIRTY = 5,  // face pairs to recompute
    ST_POOL = 6,    // face-cache slots in use
    ST_N = 8
};

// Instantiated step graphs of a handle, each with the bytes of the launch arguments it was
// captured for (kept until evicted or the handle is destroyed; at most kMax).
struct GraphCache
{
    static constexpr size_t kMax = 4;
    struct Entry
    {
        std::vector<unsigned char> key;
        hipGraphExec_t ge = nullptr;
    };
    std::vector<Entry> e;
    GraphCache() = default;
    GraphCache(const GraphCache &) = delete;
    GraphCache &operator=(const GraphCache &) = delete;
    ~GraphCache()
    {
        for (auto &x : e)
            if (x.ge)
                (void)hipGraphExecDestroy(x.ge);
    }
};

}  // namespace pppm_amd

struct pppm_solver
{
    template <class T>
    using DBuf = pppm_amd::DBuf<T>;
    using Tdbem = pppm_amd::Tdbem;
    using TriGeom = pppm_amd::TriGeom;
    using StepState = pppm_amd::StepState;
    int N = 0;
    float h = 0, dt = 0, c = 0, rc = 0;
    float minp[3] = {0, 0, 0};
    int pml_width = 0;
    unsigned flags = 0;
    bool use_pml = false;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    int t_host = -1;
    Tdbem bem;

    DBuf<float> grid, pmlbuf;
    DBuf<StepState> state;

    int nv = 0, nt = 0;
    bool mesh_set = false;
    // the last mesh update failed (out of grid, bad index, capacity): the lists and weights no
    // longer describe the mesh, so the step refuses to run until an update succeeds
    bool mesh_stale = false;
    DBuf<float> verts;
    DBuf<int> tris;
    DBuf<TriGeom> geom;

    // capacities (host-known) and derived launch sizes
    int cap_c3 = 0, cap_b4items = 0, cap_p4 = 0;
    int c3_waves = 0, face_waves = 0;

    // [cell_cnt | base_cnt | cell_fill | mark], 4 x N^3, all zero between updates
    DBuf<int> counters;
    DBuf<int> tsums;  // per-tile counts of the dirty face pairs, zero between updates
    bool counters_zero = false;
    DBuf<int> cell_start, cell_items, cell_rank;
    DBuf<int> scan_tmp;
    DBuf<int> c3_cells, c3_index, c3_cnt, c3_start, c3_pairs, c3_wave;
    DBuf<int> b4_cells, b4_cnt, b4_start, b4_tris;
    DBuf<int> face_cnt, face_start, face_nbr, face_slot, face_owner, face_flag, dirty_list;
    // face lists and slots of the previous update (swapped with the current ones per update):
    // the lookup takes a pair's slot from there when the pair sits at the same place
    DBuf<int> prev_face_start, prev_face_nbr, prev_face_slot;
    bool faces_valid = false;  // face_start / face_nbr / face_slot hold a completed update
    DBuf<int> face_wave, corner;
    DBuf<int> gflag, glist;
    DBuf<char> ftmp;
    // device status block (ST_*): error bits, the live counts the update's scans write, the
    // face-cache pool counter; the host reads it with one copy per update
    DBuf<int> status;
    bool err_clean = false;  // status[ST_ERR] known to be 0

    DBuf<uint32_t> wf, wb;
    DBuf<float2> sf, sb;
    DBuf<float> gmeta, interp;
    DBuf<pppm_amd::GridTriIn> gin;  // grid-cache recompute inputs of the dirty triangles (by triangle)

    int pool_cap = 0, hcap = 0, pool_count = 0;
    DBuf<uint32_t> fw;
    DBuf<float2> fsc, fmeta;
    DBuf<unsigned long long> hkeys;
    DBuf<int> hvals;

    // the step's weight records in gather-list order (C3 pair p, face pair p) for a static mesh:
    // built once per mesh version (ord_version) by the static-mesh entry points, so the gathers
    // fetch a pair's record with its index instead of after it; moving meshes read through
    // c3_pairs / face_slot
    long long ord_version = -1;
    DBuf<uint32_t> owf, owb, ofw;
    DBuf<float2> osf, osb, ofsc;

    DBuf<float> hn, hd, far, nstage;
    DBuf<float> wn, wd;  // per-step history windows, 16 floats per triangle and field
    DBuf<int> lcell, lc3;

    // cell classification and fresh cells (GhostCellSolver cell_data / cell_data_old,
    // ghost_cell.h:125-127), N^3 each
    bool classified = false;
    int ghost_count = 0;
    DBuf<int> ctype, ctype_old, cghost, cnearest, ctyp2, cflag, cscan, fresh_list, fresh_err;
    DBuf<float> cpoint, creflect, cdist, acc_stage;
    DBuf<int> vf_cnt, vf_start, vf_fill, vf_items;
    DBuf<float> vnormal;

    // ghost-cell boundary solve (GhostCellSolver, ghost_cell.h:72-193): ghost list, stencil
    // weights and order, 9-slot sparse rows, right-hand side and solution, the Neumann data of
    // the current and the previous step, FP64 solver work vectors
    bool ghost_ready = false, ghost_linear = false;
    float ghost_threshold = 25.0f;
    int gh_n = 0, gh_nt = 0, gh_iters = 0;
    DBuf<int> gh_cell, gh_order, gh_mcol, gh_info, gh_err;
    DBuf<float> gh_pw, gh_cond, gh_mval, gh_b, gh_x, gh_neu, gh_neu_old;
    DBuf<double> gh_work;

    // slice images: per-pixel mesh-line coverage and counts, the RGBA image
    DBuf<int> img_cover, img_count;
    DBuf<unsigned char> img_rgba;
    // host copy of the status block, in pinned memory so that the per-update read is a direct
    // DMA rather than a staged copy
    struct PinnedStatus
    {
        int *p = nullptr;
        ~PinnedStatus()
        {
            if (p)
                (void)hipHostFree(p);
        }
    } st_pin;
    int *st_h = nullptr;
    int *st_hdev = nullptr;  // the device's address of st_h (null: copies only)
    // update_and_step: recorded behind the step's FDTD launch, which writes the update's status
    // block into st_h (on the PML path, behind a copy); the step is enqueued before the status is
    // read (a copy on a second stream waiting for the update was slower: the step's kernels
    // queued behind the cross-stream wait)
    struct StatusCopy
    {
        hipEvent_t ev_status = nullptr;
        ~StatusCopy()
        {
            if (ev_status)
                (void)hipEventDestroy(ev_status);
        }
    } stc;
    long long last_grid_recomputes = 0, last_face_recomputes = 0;
    // second stream of a mesh update: the grid-cache recompute runs on it beside the lists and
    // the face cache (forked after the first scan batch, joined before the update ends)
    struct UpdateFork
    {
        hipStream_t aux = nullptr;
        hipEvent_t fork = nullptr, join = nullptr;
        ~UpdateFork()
        {
            if (aux)
                (void)hipStreamSynchronize(aux);
            if (fork)
                (void)hipEventDestroy(fork);
            if (join)
                (void)hipEventDestroy(join);
            if (aux)
                (void)hipStreamDestroy(aux);
        }
    } uf;

    // mesh version: counts successful mesh updates (the adjoint's transposed lists belong to one)
    long long mesh_version = 0;
    // adjoint (adjoint.cpp): transposed lists of the mesh version adj_version, the adjoint state
    // (grid levels, shell solutions, history rings, per-step buffers) and its reverse time index
    long long adj_version = -1;
    int adj_t = -1, adj_n_c3 = 0;
    DBuf<int> adj_tc3, adj_ft_start, adj_ft_c, adj_ft_slot, adj_ct_start, adj_ct_c;
    DBuf<float> adj_ct_w;
    DBuf<uint32_t> adj_tfw;  // face records in the transposed pair order (per j, increasing c)
    DBuf<float2> adj_tfsc;
    DBuf<double> adj_den;
    DBuf<float> adj_lam3, adj_ln, adj_ld, adj_unear, adj_lfar, adj_v;
    DBuf<float4> adj_bb;
    DBuf<pppm_amd::AdjState> adj_state;
    DBuf<unsigned char> state_trailer;
    // captured graphs of pppm_run (static mesh) and pppm_adjoint_run, replayed by later calls
    // whose launch arguments agree once the per-call rows are shifted to t = 0
    pppm_amd::GraphCache step_graphs, adj_graphs;
    // listener cells and the mesh version the C3 indices of set_listeners were read for
    std::vector<int> lis_cells;
    long long lis_version = -1;

    // per-phase GPU timing with HIP events on the handle's stream (the reference's
    // log_time / LOG_TIME per phase, src/core/macro.h:35-48)
    bool timing = false;
    struct Span
    {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<Span> spans;
    std::vector<hipEvent_t> pool;
    double phase_ms[8] = {0};
    long long phase_n[8] = {0};
    // work of the mesh updates completed while timing is on: updates, grid-cache triangles and
    // face-cache pairs recomputed
    long long work_updates = 0, work_grid = 0, work_face = 0;
};


namespace pppm_amd
{
// helpers of solver.cpp used by the batch
int solver_fail(pppm_solver *s, int code, const std::string &msg);
int solver_check_ready(pppm_solver *s, const char *what);
int solver_set_listeners(pppm_solver *s, const int *listen_xyz, int n_listen);
// reads the status block (live counts, pool count) into s->st_h
int solver_read_status(pppm_solver *s);
FdtdParams solver_fdtd_params(const pppm_solver *s);
// builds the list-ordered weight records for the current mesh version if they are stale
int solver_ensure_ordered(pppm_solver *s);
// the step kernels' view of the handle with its own per-step buffers
SceneStep solver_scene_step(const pppm_solver *s, const float *neumann, int t0, long long stride, int n_listen,
                            float *out);
// the adjoint kernels' view of the handle (adjoint.cpp): listener cotangents gout (rows
// gout_stride floats apart), gradient rows grad_stride floats apart, row 0 = step t_first
AdjScene solver_adj_scene(pppm_solver *s, const float *gout, long long gout_stride, int n_listen, float *grad,
                          long long grad_stride, int t_first);
// the graph of c whose key is the n bytes at key, or one captured now from what enqueue puts on
// the handle's stream (and then kept in c, evicting the oldest entry past GraphCache::kMax)
int solver_cached_graph(pppm_solver *s, GraphCache &c, const void *key, size_t n, int (*enqueue)(void *), void *arg,
                        hipGraphExec_t *out);
// zeroes the padding of a copy that is compared as bytes: the members are copied one by one
void canonical_bytes(StepArgs &a);
void canonical_bytes(FdtdParams &p);
}  // namespace pppm_amd
