// Host-side declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "common.hpp"

namespace pppm_amd
{

// Device-resident step state. t follows the reference's fdtd.t (-1 before the
// first step, incremented by the FDTD kernel); done counts finished blocks of the
// launch that advances t.
struct StepState
{
    int t;
    unsigned done;
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
hipError_t launch_fdtd(float *grid3, float *pml3, const FdtdParams &p, StepState *st, hipStream_t s);

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
hipError_t launch_tri_geometry(const float *verts, const int *tris, TriGeom *geom, const BuildArgs &a, int *err_flag,
                               hipStream_t s);
hipError_t launch_count_cells(const TriGeom *geom, int nt, int *cell_cnt, int *base_cnt, hipStream_t s);
hipError_t launch_fill_cells(const TriGeom *geom, int nt, const int *cell_start, const int *base_start,
                             int *cell_fill, int *base_fill, int *cell_items, int *base_items, hipStream_t s);
hipError_t launch_sort_cells(const int *start, int *items, int ncell, hipStream_t s);
hipError_t launch_mark27(const TriGeom *geom, int nt, int N, int *flag, hipStream_t s);
hipError_t launch_flag_base(const int *base_cnt, int N, int *flag, hipStream_t s);
hipError_t launch_scatter_selected(const int *flag, const int *pos, int n, int *out, int *index_map, hipStream_t s);
hipError_t launch_count_nbr(const int *cells, int ncells, const int *cell_start, int N, int width, int *cnt,
                            hipStream_t s);
hipError_t launch_fill_c3(const int *cells, int ncells, const int *cell_start, const int *cell_items,
                          const int *c3_start, int N, int *pairs, hipStream_t s);
hipError_t launch_fill_b4(const int *cells, int ncells, const int *cell_start, const int *cell_items,
                          const int *b4_start, int N, int *tris, hipStream_t s);
hipError_t launch_face_counts(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start, int *cnt,
                              hipStream_t s);
hipError_t launch_face_fill(const TriGeom *geom, int nt, const int *b4_index, const int *b4_start,
                            const int *b4_tris, const int *face_start, int *face_nbr, hipStream_t s);
hipError_t launch_flag_nonzero(const int *in, int n, int *flag01, hipStream_t s);
hipError_t launch_corner_map(const TriGeom *geom, int nt, int N, const int *c3_index, int *corner, hipStream_t s);

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
// grid cache: dirty list of triangles -> fdtd/bem records at tri*27+o
hipError_t launch_grid_dirty(const TriGeom *geom, float *meta, int nt, float h, int first, int *flag, hipStream_t s);
hipError_t launch_grid_weights(const float *verts, const int *tris, const TriGeom *geom, const int *list, int n,
                               const WeightArgs &a, const Tdbem &b, uint32_t *wf, float2 *sf, uint32_t *wb,
                               float2 *sb, float *meta, hipStream_t s);
hipError_t launch_interp_weights(const float *verts, const int *tris, const TriGeom *geom, const WeightArgs &a,
                                 float *interp, hipStream_t s);
// face cache: hash lookup/insert of (c, j) pairs; dirty flags
hipError_t launch_face_lookup(const TriGeom *geom, const int *tris, const int *owner, const int *face_nbr, int n_p4,
                              unsigned long long *hkeys, int *hvals, int hcap, int *pool_count, float2 *fmeta,
                              int *face_slot, int *dirty, hipStream_t s);
hipError_t launch_face_weights(const float *verts, const int *tris, const TriGeom *geom, const int *dirty_pairs,
                               int n, const int *face_owner, const int *face_nbr, const int *face_slot,
                               const float *interp, const WeightArgs &a, const Tdbem &b, double *tmp,
                               uint32_t *fw, float2 *fsc, hipStream_t s);
hipError_t launch_face_owner(const int *face_start, int nt, int *owner, hipStream_t s);
hipError_t launch_face_denom(const TriGeom *geom, int nt, const int *face_start, const int *face_slot,
                             const uint32_t *fw, const float2 *fsc, double *denom, hipStream_t s);
hipError_t launch_rehash(const unsigned long long *okeys, const int *ovals, int ocap, unsigned long long *nkeys,
                         int *nvals, int ncap, hipStream_t s);

// ---------------------------------------------------------------- step.hip
struct StepArgs
{
    int N, nt, n_c3;
    float rc;
    int t0;             // t of the first step of this launch sequence
    long long stride;   // Neumann row stride (floats)
    int n_listen;
};
hipError_t launch_far(const StepArgs &a, const StepState *st, const float *grid3, const int *c3_cells,
                      const int *c3_start, const int *c3_pairs, const uint32_t *wf, const float2 *sf, float *hn,
                      const float *hd, float *far, const float *neumann, int n_pairs, hipStream_t s);
hipError_t launch_dirichlet(const StepArgs &a, const StepState *st, const int *face_start, const int *face_nbr,
                            const int *face_slot, const uint32_t *fw, const float2 *fsc, const float *hn, float *hd,
                            const float *interp, const int *corner, const float *far, const double *denom,
                            int n_pairs, hipStream_t s);
hipError_t launch_near(const StepArgs &a, const StepState *st, float *grid3, const int *c3_cells,
                       const int *c3_start, const int *c3_pairs, const uint32_t *wb, const float2 *sb,
                       const float *hn, const float *hd, const float *far, const int *listen_cell,
                       const int *listen_c3, float *out, int n_pairs, hipStream_t s);
hipError_t launch_listen_direct(const StepArgs &a, const StepState *st, const float *grid3, const int *listen_cell,
                                float *out, hipStream_t s);
hipError_t launch_set_neumann(const StepArgs &a, const StepState *st, float *hn, const float *neumann,
                              hipStream_t s);


}  // namespace pppm_amd
