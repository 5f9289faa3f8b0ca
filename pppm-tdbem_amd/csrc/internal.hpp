// Host-side declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace pppm_amd
{

// Device-resident step state: t follows the reference's fdtd.t (-1 before the first step,
// advanced right after the FDTD kernel, src/algorithm/fdtd.h:73-87).
struct StepState
{
    int t;
    int pad;
};

// device-side error bits (reported through the status block)
enum : int
{
    ERR_OUT_OF_GRID = 1,
    ERR_BAD_INDEX = 2,
    ERR_CAPACITY = 4,
};

struct FdtdParams
{
    int N;
    double coef;  // (c*c*dt*dt) in FP64, fdtd.cu:27
    double hh;    // h*h in FP64, fdtd.cu:16
    float c, dt, h;
    int pml;      // PML width (0 = ABC)
    float damp;   // 2/h, fdtd.h:61
};

// ---------------------------------------------------------------- fdtd.hip
// What the ABC-shell launch of a time step can do besides the FDTD: the gathers' history
// windows for the step, in extra blocks (wn / wd as in step.hip's windows_item with back = 0).
// wn == nullptr: nothing. (Advancing t in the same launch, by the last block to finish, cost more
// than the one-thread kernel it saves: every block counted itself done on one atomic.)
struct StepFuse
{
    const float *hn, *hd;
    float *wn, *wd;
    int nt;
    // status_host != null: the launch also copies the status_n ints of status (the mesh
    // update's status block) to status_host (pinned host memory), whether or not it steps
    const int *status;
    int *status_host;
    int status_n;
};
// skip (may be null): as for launch_advance
hipError_t launch_fdtd(float *grid3, float *pml3, const FdtdParams &p, const StepState *st, hipStream_t s,
                       const int *skip = nullptr, const StepFuse *fuse = nullptr);
// One scene of a batched FDTD step. nb_inner: blocks of its one-thread-per-cell interior in the
// batched shell launch (0: the interior is marched by a tile launch); xch: x planes per tile.
struct FdtdScene
{
    FdtdParams p;
    float *grid3;
    const StepState *st;
    int nb_inner, xch;
};
// FDTD interior kernels a scene can use: 0 one thread per cell, 1..4 LDS-staged 2.5D tiles of
// 8 x 64, 16 x 32, 4 x 64 and 8 x 32 (y, z) columns
constexpr int kFdtdVariants = 5;
int fdtd_shell_blocks(int N);
int fdtd_inner_blocks(int N);
// tiles of an N^3 grid's interior for a tile variant; sets the x planes per tile
int fdtd_tile_blocks(int N, int variant, int &xch);
struct BatchMap;
// shell of every scene plus the one-thread-per-cell interiors (nb_inner > 0)
hipError_t launch_batch_fdtd(const FdtdScene *ds, const BatchMap &m, hipStream_t s);
// interiors of the scenes using tile variant `variant` (1..4)
hipError_t launch_batch_tile(int variant, const FdtdScene *ds, const BatchMap &m, hipStream_t s);

// ---------------------------------------------------------------- scan.hip
// exclusive scan of n ints into out[0..n] (out[n] = total); tmp must hold scan_tmp_ints(n) ints
size_t scan_tmp_ints(int n);
hipError_t scan_exclusive(const int *in, int *out, int n, int *tmp, hipStream_t s);
// one array of a batched scan; every output pointer may be null
struct ScanJob
{
    const int *in;
    int n;
    int flag_mode;   // scan (in[q] != 0) instead of in[q]
    int *pos;        // exclusive positions, n + 1 entries (pos[n] = total)
    int *list;       // list[rank] = q for every selected q (flag mode)
    int *index_map;  // index_map[q] = rank if selected, else -1 (flag mode)
    int *total;      // receives the total
    int tile0;       // set by scan_batch
};
constexpr int SCAN_MAX_JOBS = 4;
struct ScanBatch
{
    ScanJob job[SCAN_MAX_JOBS];
    int njobs;
};
constexpr int SCAN_TILE = 1024;  // items per tile of a multi-tile scan
// tmp must hold scan_batch_tmp_ints(sizes) ints
size_t scan_batch_tmp_ints(const int *n, int njobs);
// sets job[k].tile0 (the tiles of all jobs laid out in job order); returns the tile count
int scan_batch_layout(ScanBatch &b);
// true when every job fits one tile of the single-launch scan (no tile sums are read)
bool scan_batch_single(const ScanBatch &b);
// presums (may be null): per-tile sums the producers of the inputs already accumulated, in the
// layout of scan_batch_layout (tile of item q of job k = tile0 + q / SCAN_TILE); null: computed
hipError_t scan_batch(ScanBatch b, int *tmp, hipStream_t s, const int *presums = nullptr);

// ---------------------------------------------------------------- build.hip (binning, lists)
struct WeightArgs
{
    int N, nt;
    float h, c, dt;
    float minx, miny, minz;
};
struct BuildArgs
{
    int N, nt, nv;
    float h;
    float minx, miny, minz;
};
hipError_t launch_fill_int(int *a, long long n, int value, hipStream_t s);
// a[0..n) = v[0..n) for n <= 4, from the launch arguments: a stream-ordered small write that
// needs no host buffer to outlive the call (so no wait for the stream)
hipError_t launch_put_ints(int *a, int n, int v0, int v1, int v2, int v3, hipStream_t s);
hipError_t launch_copy_verts(const float *series, long long stride, int t0, const StepState *st, float *verts, int n,
                             hipStream_t s);
// per-tile counts of the dirty face pairs, accumulated by the face lookup for the face-pair scan
// (layout of scan_batch_layout; zero between updates: k_tri_update clears them)
struct TileSums
{
    int *b3;
    int n3;
};
// counters of a mesh update, zero between updates (N^3 ints each)
struct UpdateCounters
{
    int *cell_cnt, *base_cnt, *cell_fill, *mark;
    TileSums ts;
};
struct ListArgs
{
    int N, nt;
    const TriGeom *geom;
    UpdateCounters cn;
    int *cell_start, *cell_items, *cell_rank;
    int *c3_cells, *c3_n, *c3_index, *c3_cnt, *c3_start, *c3_pairs, *c3_wave;
    int cap_c3, c3_waves;
    int *b4_cells, *b4_n, *b4_cnt, *b4_start, *b4_tris;
    int cap_b4items;
    int *face_cnt, *face_start, *face_nbr, *face_owner, *face_wave;
    int cap_p4, face_waves;
    int *corner;
    int *err;
};
// in-cell order of the cell CSR (classification needs ascending triangle indices)
hipError_t launch_sort_cells(const TriGeom *geom, int nt, const int *start, int *items, const int *rank,
                             int use_base, hipStream_t s);
// geometry, counts, marks, grid-cache dirty flags (and the recompute inputs of the dirty
// triangles), interpolation weights; dst != null copies the nv vertices of src into dst in the
// same launch
hipError_t launch_tri_update(const float *src, float *dst, int nv, const int *tris, TriGeom *geom, const BuildArgs &ba,
                             const WeightArgs &wa, const float *gmeta, int first, int *gflag, GridTriIn *gin,
                             float *interp, const UpdateCounters &cn, int *err, hipStream_t s);
hipError_t launch_lists_count(const ListArgs &a, hipStream_t s);
hipError_t launch_lists_fill(const ListArgs &a, hipStream_t s);
// 4x4x4 lists of the non-empty base cells: fill = 0 counts (b4_cnt), fill = 1 writes b4_tris
hipError_t launch_b4_lists(const ListArgs &a, int fill, hipStream_t s);
// skip (may be null): the error word of a mesh update the step was enqueued behind; if it is set
// when the kernel runs, the kernel does nothing (the step is run again after the update is)
hipError_t launch_advance(StepState *st, hipStream_t s, const int *skip = nullptr);

// ---------------------------------------------------------------- weights.hip (FP64 quadrature)
struct Tdbem
{
    double lambda;
    double wave_re[STEPS], wave_im[STEPS];
    double tw_re[STEPS / 2], tw_im[STEPS / 2];  // exp(2 pi_ref i j / 16), j = 0..7 (bem.h:25)
    double inv_pow[STEPS];                       // lambda^-n as the running quotient 1 / lambda / ... (bem.h:300-303)
    float threshold;
};
void tdbem_init(Tdbem &b, float dt, float c, float grid_size);

size_t face_tmp_bytes(int max_dirty);
// records tri*27+o of the *n_ptr triangles in list (cap bounds *n_ptr), from their inputs in
// gin (indexed by triangle, written by the binning pass)
hipError_t launch_grid_weights(const GridTriIn *gin, const int *list, const int *n_ptr, int cap, const WeightArgs &a,
                               const Tdbem &b, uint32_t *wf, float2 *sf, uint32_t *wb, float2 *sb, float *meta,
                               const int *err, hipStream_t s);
// face lists and slots of the previous update (start == null: none); cur_start = this update's
// face-list offsets
struct PrevFaces
{
    const int *start, *nbr, *slot, *cur_start;
};
// tsum (may be null): per-tile counts of the dirty flags (tile = pair / SCAN_TILE), accumulated
hipError_t launch_face_lookup(const TriGeom *geom, const int *tris, const int *owner, const int *face_nbr,
                              const int *n_ptr, int cap, unsigned long long *hkeys, int *hvals, int hcap,
                              int *pool_count, int pool_cap, float2 *fmeta, int *face_slot, int *dirty, int *err_flag,
                              int nt, const PrevFaces &prev, int *tsum, hipStream_t s);
hipError_t launch_face_weights(const float *verts, const int *tris, const TriGeom *geom, const int *dirty_pairs,
                               const int *n_ptr, int cap, const int *owner, const int *face_nbr, const int *face_slot,
                               const float *interp, const WeightArgs &a, const Tdbem &b, void *tmp, uint32_t *fw,
                               float2 *fsc, float2 *fmeta, const int *err, long long hint_dirty, hipStream_t s);
hipError_t launch_rehash(const unsigned long long *okeys, const int *ovals, int ocap, unsigned long long *nkeys,
                         int *nvals, int ncap, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *a, long long n, unsigned long long v, hipStream_t s);

// ---------------------------------------------------------------- cells.hip (classification)
// cell types of the ghost-cell solver (ghost_cell.h:11-17)
enum : int
{
    CELL_AIR = 0,
    CELL_SOLID = 1,
    CELL_GHOST = 2,
    CELL_UNKNOWN = 3,
};
struct CellArgs
{
    int N;
    float h;
    float minx, miny, minz;
    int thin_shell;
};
// vertex -> face CSR (cnt, fill: nv ints; start: nv+1; items: 3 nt)
hipError_t launch_vertex_faces(const int *tris, int nt, int nv, int *cnt, int *start, int *fill, int *items,
                               int *scan_tmp, hipStream_t s);
hipError_t launch_vertex_normals(const float *verts, const int *tris, const TriGeom *geom, int nv, const int *start,
                                 int *items, float *vnormal, hipStream_t s);
// types, nearest triangle / point / mirror point / distance, ghost indices of all N^3 cells
hipError_t launch_classify(const CellArgs &a, const float *verts, const int *tris, const int *cell_start,
                           const int *cell_items, const float *vnormal, int *ctype, int *cnearest, float *cpoint,
                           float *creflect, float *cdist, int *typ2, int *flag, int *scan, int *scan_tmp, int *cghost,
                           hipStream_t s);
hipError_t launch_fresh_flags(long long NN, int nt, const int *told, const int *tnew, const int *cnearest, int *flag,
                              int *err, hipStream_t s);
hipError_t launch_fresh_fill(const CellArgs &a, const StepState *st, const int *list, const int *n_ptr, int cap,
                             const int *err, const int *told, const int *tnew, const int *cnearest,
                             const float *cpoint, const float *acc_now, const float *acc_old, float *grid3,
                             hipStream_t s);
hipError_t launch_solid_zero(long long NN, const StepState *st, const int *ctype, float *grid3, hipStream_t s);

// ---------------------------------------------------------------- ghost.hip
// ghost list, stencil weights (order 1 = second order), 9-slot sparse rows of the ghost system
hipError_t launch_ghost_setup(const CellArgs &a, int n, const int *ctype, const int *cghost, const int *cnearest,
                              const float *cpoint, const float *creflect, float threshold, int *gcell, int *order,
                              float *pw, float *cond, int *mcol, float *mval, int *err, hipStream_t s);
// right-hand side at level t, the linear solve (if `linear`) and the ghost values into level t;
// work: 7 n doubles, info[0]: iterations of the solve
hipError_t launch_ghost_solve(const CellArgs &a, int n, const StepState *st, float *grid3, const int *gcell,
                              const float *creflect, const float *cdist, const int *cnearest, const int *ctype,
                              const int *cghost, const int *order, const float *pw, const float *neu,
                              const int *mcol, const float *mval, bool linear, float *b, float *x, double *work,
                              int *info, hipStream_t s);

// ---------------------------------------------------------------- adjoint.hip
// reverse time index of the adjoint (the forward step being reversed), in device memory
struct AdjState
{
    int t;
    int pad;
};
// Everything one reverse step reads and writes: the forward's mesh data (static), the
// transposed lists, the adjoint state, the listener cotangents in and the gradient out.
struct AdjScene
{
    FdtdParams fp;
    int N, nt, n_c3;
    AdjState *ast;
    float *lam3;   // adjoint grid levels (3 x N^3), slot = t mod 3 as the forward's
    float4 *bb;    // per shell cell: A^{-T} of its cotangents (N^3, shell entries used)
    float *ln, *ld;                  // adjoint history rings (nt x 32)
    float *unear, *lfar, *v;         // per C3 cell, per C3 cell, per triangle
    const int *c3_cells, *tc3;       // C3 cell -> grid index; (tri, o) -> C3 cell
    const uint32_t *wf, *wb, *fw;    // the forward's weight records
    const float2 *sf, *sb, *fsc;
    const int *ft_start, *ft_c, *ft_slot;  // per triangle j: face pairs (c, j) and their slots
    const uint32_t *tfw;                   // ... their face records and scales in that order
    const float2 *tfsc;
    const int *ct_start, *ct_c;            // per C3 cell: triangles with it as a base-cell corner
    const float *ct_w;                     // ... and rc x their interpolation weight
    const double *den;                     // A_c / 2 - sum_j D_0(c, j)
    const int *lcell, *lc3;
    int n_listen;
    const float *gout;   // listener cotangents, row t - t_first (n_listen used per row; null: none)
    long long gout_stride;  // floats between rows of gout
    float *grad;         // gradient of the Neumann inputs, row t - t_first
    long long grad_stride;
    int t_first;
};
// one reverse step (t = ast->t, then ast->t - 1)
hipError_t enqueue_adjoint_step(const AdjScene &d, hipStream_t s);
// kernels of a reverse step (read, near, Dirichlet divisor, Dirichlet, far, ABC, FDTD) and the
// block count of kernel `kind` for one scene
constexpr int kAdjKernels = 7;
int adj_blocks(const AdjScene &d, int kind);
// one reverse step of n scenes (descriptor table ds in device memory): maps[k] deals kernel k's
// blocks to the scenes
hipError_t enqueue_batch_adjoint_step(const AdjScene *ds, const BatchMap *maps, int n, hipStream_t s);

// ---------------------------------------------------------------- noalloc.hip
// bytes of the per-device scratch array of the allocation-free entry point
constexpr size_t kScratchBytes = 16u << 20;
// device address of that array on the current device (looked up once per device)
hipError_t scratch_ptr(void **out);

// ---------------------------------------------------------------- step.hip
struct StepArgs
{
    int N, nt;
    float rc;
    int t0;             // t of the first step of this launch sequence
    long long stride;   // Neumann row stride (floats)
    int n_listen;
};
// Everything the step kernels of one scene read and write: its sizes, lists, weight caches,
// histories and per-step buffers. Passed by value to the single-scene kernels; the batched
// kernels read an array of them from device memory.
struct SceneStep
{
    StepArgs a;
    StepState *st;
    float *grid3;
    // C3 gathers (far field, near field): cells, their pair lists, wave maps
    const int *c3_cells, *c3_start, *c3_wave, *c3_pairs;
    int c3_waves;
    // face gather (Dirichlet update)
    int face_waves;
    const int *face_start, *face_wave, *face_nbr, *face_slot;
    const uint32_t *wf, *wb, *fw;  // 16 x half2 {S_k, D_k} per record
    const float2 *sf, *sb, *fsc;   // scales
    // the same records in list order (record of C3 pair p / face pair p at p), or null: the
    // gathers then read wf / wb at c3_pairs[p] and fw at face_slot[p]
    const uint32_t *owf, *owb, *ofw;
    const float2 *osf, *osb, *ofsc;
    const float *interp;
    const int *corner;
    const TriGeom *geom;
    const int *err;
    // histories and per-step buffers
    float *hn, *hd, *wn, *wd, *far;
    const float *neumann;  // Neumann series (row t - t0), null: the ring slot is written elsewhere
    // listener recording: out[(t - t0) * out_stride + l] for l < a.n_listen (null: none); lc3
    // null: the listeners' C3 indices are read from c3_index (cell -> C3 index) per step
    const int *lcell, *lc3;
    const int *c3_index;
    float *out;
    long long out_stride;
};
// history windows of the gathers: wn/wd[tri*16 + k] = ring value at t - back - k, with t the
// device time index when the kernel runs (back = 0 before the advance of a step, 1 after)
hipError_t launch_windows(const SceneStep &d, int back, hipStream_t s);
hipError_t launch_far(const SceneStep &d, hipStream_t s);
hipError_t launch_dirichlet(const SceneStep &d, hipStream_t s);
hipError_t launch_near(const SceneStep &d, hipStream_t s);
hipError_t launch_listen_direct(const SceneStep &d, hipStream_t s);
hipError_t launch_set_neumann(const SceneStep &d, hipStream_t s);
// listener table on the device: cell[l] = linear index of xyz[l], c3[l] = c3_index[cell[l]]
// (-1: not a corrected cell); at most 64 listeners, passed by value
struct ListenTable
{
    int n;
    int xyz[3 * 64];
};
hipError_t launch_listen_setup(const ListenTable &lt, int N, const int *c3_index, int *cell, int *c3,
                               hipStream_t s);
// list-ordered copies of the weight records: o*[p] = *[c3_pairs[p]] for the n3 C3 pairs and
// ofw[p] = fw[face_slot[p]] for the n4 face pairs (slots below pool_cap)
hipError_t launch_order_records(const int *c3_pairs, int n3, const uint32_t *wf, const float2 *sf, const uint32_t *wb,
                                const float2 *sb, uint32_t *owf, float2 *osf, uint32_t *owb, float2 *osb,
                                const int *face_slot, int n4, int pool_cap, const uint32_t *fw, const float2 *fsc,
                                uint32_t *ofw, float2 *ofsc, hipStream_t s);

// Batched step kernels: one launch covers the blocks of every scene of a batch. scene_of[v] is
// the scene of virtual block v, first[scene] its first virtual block; total virtual blocks,
// per = ceil(total / 8) (the launch has 8 * per blocks, dealt to the XCDs in contiguous runs).
struct BatchMap
{
    const int *scene_of;
    const int *first;
    int total, per;
};
enum class BatchKernel
{
    Windows,
    Far,
    Dirichlet,
    Near
};
// blocks of one scene in a batched launch of kernel k
int batch_blocks(const SceneStep &d, BatchKernel k);
hipError_t launch_batch(BatchKernel k, const SceneStep *descs, const BatchMap &m, hipStream_t s);
hipError_t launch_batch_advance(const SceneStep *descs, int n, hipStream_t s);
// colour image of grid plane `index` normal to `axis` (0: x, 1: y, 2: z) at the current level
// with the mesh's cut drawn over it: W x W RGBA in PNG row order, W = N * (1080 / N);
// cover / count: W x W ints of scratch
hipError_t launch_slice_image(int N, float h, const float *mn3, int axis, int index, float vmax,
                              const StepState *st, const float *grid3, const float *verts, const int *tris, int nt,
                              int *cover, int *count, unsigned char *rgba, hipStream_t s);
// FFAT map accumulation at the current time level: map is 6 x 2w x 2w, 1 <= w <= N/2 - 1
hipError_t launch_ffat(const StepState *st, const float *grid3, int N, int w, float *map, hipStream_t s);

}  // namespace pppm_amd
