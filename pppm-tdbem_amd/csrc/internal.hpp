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
hipError_t launch_fdtd(float *grid3, float *pml3, const FdtdParams &p, const StepState *st, hipStream_t s);

// ---------------------------------------------------------------- scan.hip
// exclusive scan of n ints into out[0..n] (out[n] = total); tmp must hold scan_tmp_ints(n) ints
size_t scan_tmp_ints(int n);
hipError_t scan_exclusive(const int *in, int *out, int n, int *tmp, hipStream_t s);

// ---------------------------------------------------------------- build.hip (binning, lists)
struct BuildArgs
{
    int N, nt, nv;
    float h;
    float minx, miny, minz;
};
hipError_t launch_fill_int(int *a, long long n, int value, hipStream_t s);
hipError_t launch_copy_verts(const float *series, long long stride, int t0, const StepState *st, float *verts, int n,
                             hipStream_t s);
hipError_t launch_tri_geometry(const float *verts, const int *tris, TriGeom *geom, const BuildArgs &a, int *err_flag,
                               hipStream_t s);
hipError_t launch_count_cells(const TriGeom *geom, int nt, int *cell_cnt, int *base_cnt, hipStream_t s);
hipError_t launch_fill_cells(const TriGeom *geom, int nt, const int *cell_start, const int *base_start,
                             int *cell_fill, int *base_fill, int *cell_items, int *base_items, int *cell_rank,
                             int *base_rank, hipStream_t s);
hipError_t launch_sort_cells(const TriGeom *geom, int nt, const int *start, int *items, const int *rank,
                             int use_base, hipStream_t s);
hipError_t launch_mark27(const TriGeom *geom, int nt, int N, int *flag, hipStream_t s);
hipError_t launch_flag_base(const int *base_cnt, int N, int *flag, hipStream_t s);
hipError_t launch_scatter_selected(const int *flag, const int *pos, int n, int *out, int *index_map, hipStream_t s);
// width^3 neighbour blocks around compacted cells (count live ones, zero past *n_ptr; then fill)
hipError_t launch_nbr_count(const int *cells, const int *n_ptr, int cap, const int *cell_start, int N, int width,
                            int *cnt, hipStream_t s);
hipError_t launch_nbr_fill(const int *cells, const int *n_ptr, int cap, const int *cell_start, const int *cell_items,
                           int N, int width, const int *out_start, int out_cap, int *out, int *err_flag, hipStream_t s);
hipError_t launch_face_counts(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start, int *cnt,
                              hipStream_t s);
hipError_t launch_face_fill(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start,
                            const int *b4_tris, const int *face_start, int cap, int *face_nbr, int *face_owner,
                            int *err_flag, hipStream_t s);
hipError_t launch_flag_nonzero(const int *in, int n, int *flag01, hipStream_t s);
hipError_t launch_corner_map(const TriGeom *geom, int nt, int N, const int *c3_index, int *corner, hipStream_t s);
// n_ptr (device, may be null): live segment count; the map never extends past it
hipError_t launch_wave_map(const int *start, int nseg, const int *n_ptr, int waves, int *map, hipStream_t s);
hipError_t launch_advance(StepState *st, hipStream_t s);
struct StatusSrc
{
    const int *p[16];
    int n;
};
hipError_t launch_status(const StatusSrc &src, int *out, hipStream_t s);

// ---------------------------------------------------------------- weights.hip (FP64 quadrature)
struct Tdbem
{
    double lambda;
    double wave_re[STEPS], wave_im[STEPS];
    double tw_re[STEPS / 2], tw_im[STEPS / 2];  // exp(2 pi_ref i j / 16), j = 0..7 (bem.h:25)
    float threshold;
};
void tdbem_init(Tdbem &b, float dt, float c, float grid_size);

struct WeightArgs
{
    int N, nt;
    float h, c, dt;
    float minx, miny, minz;
};
size_t grid_tmp_bytes(int max_dirty);
size_t face_tmp_bytes(int max_dirty);
hipError_t launch_grid_dirty(const TriGeom *geom, float *meta, int nt, float h, int first, int *flag, hipStream_t s);
// records tri*27+o of the *n_ptr triangles in list (cap bounds *n_ptr). hint_dirty: host
// estimate of *n_ptr that sizes the grid-stride launch (-1: cap); any value is correct.
hipError_t launch_grid_weights(const float *verts, const int *tris, const TriGeom *geom, const int *list,
                               const int *n_ptr, int cap, const WeightArgs &a, const Tdbem &b, void *tmp,
                               uint32_t *wf, float2 *sf, uint32_t *wb, float2 *sb, float *meta, const int *err,
                               long long hint_dirty, hipStream_t s);
hipError_t launch_interp_weights(const float *verts, const int *tris, const TriGeom *geom, const WeightArgs &a,
                                 float *interp, const int *err, hipStream_t s);
hipError_t launch_face_lookup(const TriGeom *geom, const int *tris, const int *owner, const int *face_nbr,
                              const int *n_ptr, int cap, unsigned long long *hkeys, int *hvals, int hcap,
                              int *pool_count, int pool_cap, float2 *fmeta, int *face_slot, int *dirty, int *err_flag,
                              int nt, hipStream_t s);
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

// ---------------------------------------------------------------- step.hip
struct StepArgs
{
    int N, nt;
    float rc;
    int t0;             // t of the first step of this launch sequence
    long long stride;   // Neumann row stride (floats)
    int n_listen;
};
// history windows of the gathers: wn/wd[tri*16 + k] = ring value at t - back - k, with t the
// device time index when the kernel runs (back = 0 before the advance of a step, 1 after)
hipError_t launch_windows(int nt, int back, const StepState *st, const float *hn, const float *hd, float *wn,
                          float *wd, hipStream_t s);
// wave maps: map[w] = first segment owned by gather wave w (launch_wave_map)
hipError_t launch_far(const StepArgs &a, const StepState *st, const float *grid3, const int *c3_cells,
                      const int *c3_start, const int *c3_wave, int c3_waves, const int *c3_pairs, const uint32_t *wf,
                      const float2 *sf, float *hn, const float *hd, const float *wn, const float *wd, float *far,
                      const float *neumann, hipStream_t s);
hipError_t launch_dirichlet(const StepArgs &a, const StepState *st, const int *face_start, const int *face_wave,
                            int face_waves, const int *face_nbr, const int *face_slot, const uint32_t *fw,
                            const float2 *fsc, const float *hn, float *hd, const float *wn, const float *wd,
                            const float *interp, const int *corner, const float *far, const TriGeom *geom,
                            const int *err, hipStream_t s);
hipError_t launch_near(const StepArgs &a, const StepState *st, float *grid3, const int *c3_cells,
                       const int *c3_start, const int *c3_wave, int c3_waves, const int *c3_pairs, const uint32_t *wb,
                       const float2 *sb, const float *hn, const float *hd, const float *wn, const float *wd,
                       const float *far, const int *listen_cell, const int *listen_c3, float *out, hipStream_t s);
hipError_t launch_listen_direct(const StepArgs &a, const StepState *st, const float *grid3, const int *listen_cell,
                                float *out, hipStream_t s);
hipError_t launch_set_neumann(const StepArgs &a, const StepState *st, float *hn, const float *neumann,
                              hipStream_t s);

}  // namespace pppm_amd
