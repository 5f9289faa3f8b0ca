/*
 * CPU oracle for the PPPM time step — TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch C++ restatement of the reference algorithm, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * The product path (libpppm_hip.so) never links, loads or calls this code.
 *
 * Two weight modes:
 *   ORC_REF32 : the reference's precisions (FP32 geometry and complex quadrature
 *               with the double promotions the reference's expressions make).
 *   ORC_FP64  : FP64 quadrature, wave numbers, lambda and FFT; the result is
 *               rounded to FP32 and then quantised to FP16+scale exactly as the
 *               reference does (bem.h:171-191).
 * The time step itself is identical in both modes: FP64 inner stencil on FP32
 * storage (fdtd.cu:7-33), FP32 ABC (fdtd.cu:36-137), FP16 weights with an FP32
 * scale, sequential (deterministic) sums in list order.
 */
#ifndef PPPM_ORACLE_H
#define PPPM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_REF32 = 0, ORC_FP64 = 1 };
enum { ORC_SINGLE_LAYER = 0, ORC_DOUBLE_LAYER = 1 };

typedef struct orc_scene orc_scene;

/* PPPMSolver(res, dl, dt, min_pos, pml_width, reflect_coeff)  (pppm.h:39-55);
 * c replaces the hard-coded AIR_WAVE_SPEED (macro.h:82). pml_width > 0 makes
 * the FDTD use the split-field PML instead of the ABC (fdtd.h:73-87). */
orc_scene *orc_create(int res, float dl, float dt, const float *min_xyz, float c, float reflect_coeff,
                      int weight_mode, int pml_width);
void orc_destroy(orc_scene *s);

/* pppm.h:75-94 / pppm.h:96-109. Returns 0 on success, <0 on error. */
int orc_set_mesh(orc_scene *s, const float *verts, int nv, const int *tris, int nt);
int orc_update_mesh(orc_scene *s, const float *verts);

/* update_grid_and_face (pppm.h:137-145) and its phases. */
int orc_step(orc_scene *s, const float *neumann);
int orc_fdtd_step(orc_scene *s);
int orc_solve_far(orc_scene *s);
int orc_set_neumann(orc_scene *s, const float *neumann);
int orc_update_dirichlet(orc_scene *s);
int orc_solve_near(orc_scene *s);
/* the brute-force verification path, pppm_direct.cu:6-48 (solve_fdtd_far_simple
 * includes the FDTD step, pppm.cu:227-233) */
int orc_solve_far_direct(orc_scene *s);
int orc_solve_near_direct(orc_scene *s);

int orc_time_idx(const orc_scene *s);
void orc_set_time_idx(orc_scene *s, int t); /* fdtd.t = t (experiments/demo/main.cu:255) */
/* grids[t + rel] (CircularArray semantics, array.h:39) */
int orc_get_grid(const orc_scene *s, int rel, float *out);
int orc_set_grid(orc_scene *s, int rel, const float *in);
int orc_get_far(const orc_scene *s, float *out);
float orc_read_cell(const orc_scene *s, int x, int y, int z);
/* History rings (T x 32 floats each), bem.h:55 */
int orc_get_history(const orc_scene *s, float *neumann, float *dirichlet);
int orc_set_history(orc_scene *s, const float *neumann, const float *dirichlet);
void orc_reset(orc_scene *s); /* fdtd.reset + neumann/dirichlet reset */

/* counts: [0]=T [1]=C3 cells [2]=P3 pairs [3]=base cells [4]=P4 pairs
 *         [5]=grid-cache recomputes at last update [6]=face-cache recomputes */
int orc_stats(const orc_scene *s, long long *out);
/* per-triangle geometry: center(3) normal(3) area cell(3) base(3) -> 13 floats */
int orc_get_triangles(const orc_scene *s, float *out13);
/* C3 cells (linear index) and their CSR pair lists (entries tri*27+offset) */
int orc_get_c3(const orc_scene *s, int *cells, int *start, int *pairs);
/* base cells with their L4 lists (triangle ids) */
int orc_get_b4(const orc_scene *s, int *cells, int *start, int *tris);
/* Grid-cache weights in the reference's LayerWeightHalf layout (72 B each):
 * rec[(tri*27+o)*2 + 0] = fdtd_near_weight[o], +1 = bem_near_weight[o] */
int orc_export_grid_weights(const orc_scene *s, void *recs72);
int orc_import_grid_weights(orc_scene *s, const void *recs72);
/* Face-cache weights, one 72 B record per (c, L4 entry) in CSR order of c. */
int orc_export_face_weights(const orc_scene *s, int *start, int *nbr, void *recs72);
int orc_import_face_weights(orc_scene *s, const void *recs72);
int orc_get_interp_weights(const orc_scene *s, float *out8);

/* Cell classification and fresh cells of the ghost-cell solver (ghost_cell_util.cu:7-230,
 * ghost_cell.cu:51-218, 547-560) on the scene's current mesh. Types: 0 AIR, 1 SOLID,
 * 2 GHOST, 3 UNKNOWN (ghost_cell.h:11-17). orc_classify keeps the previous types for the
 * fresh-cell test and returns the ghost-cell count. orc_fill_fresh fills the cells that
 * turned AIR at levels t and t-1 from per-triangle accelerations (now, previous) and
 * returns their number (-1: a fresh cell without a nearest triangle; 1e10 written). */
int orc_classify(orc_scene *s, int thin_shell);
int orc_get_cells(const orc_scene *s, int *type, int *ghost_idx, int *nearest, float *point3, float *reflect3,
                  float *dist);
/* test hook: sets the current and previous cell types, nearest triangles and nearest surface
 * points (N^3 entries each) directly, as if classify had produced them */
int orc_set_cells(orc_scene *s, const int *type, const int *type_old, const int *nearest, const float *point3);
int orc_fill_fresh(orc_scene *s, const float *acc_now, const float *acc_old);
int orc_set_solid_zero(orc_scene *s);
int orc_vertex_normals(const orc_scene *s, float *out3);
/* barycentric weights of the point of triangle tri9 closest to p3 (helper_math.h:1603-1800) */
void orc_closest_coeff(const float *p3, const float *tri9, float *out3);

/* The step as a linear map, in FP64 (adjoint_oracle.cpp), on the scene's current mesh and
 * caches (ABC only): forward of T steps from the zero state with Neumann inputs g (T x nt),
 * listener samples out (T x nl) at the grid cells lcell; the adjoint of that map
 * (grad = d<gout, out>/dg); central differences of J = sum (out - target)^2 at n entries
 * (step[e], tri[e]) with step eps, each J a full forward run. Return 0, or -1 without a mesh. */
int orc_linear_forward(const orc_scene *s, const double *g, int T, const int *lcell, int nl, double *out);
int orc_linear_adjoint(const orc_scene *s, const double *gout, int T, const int *lcell, int nl, double *grad);
int orc_linear_fd(const orc_scene *s, const double *g, const double *target, int T, const int *lcell, int nl,
                  int n_entries, const int *step, const int *tri, double eps, double *fd);

/* Primitives (for known-answer and closed-form tests) */
/* deterministic FP64 exp / sincos used by the FP64 quadrature (detmath.h) */
double orc_det_exp(double x);
void orc_det_sincos(double x, double *s, double *c);
unsigned short orc_f2h(float f);
float orc_h2f(unsigned short h);
/* TDBEM::init (bem.h:264-277): lambda and the 16 complex wave numbers */
void orc_tdbem_init(int mode, float dt, float c, double *lambda, double *wave_re_im);
/* TDBEM::scaledFFT (bem.h:292-304) and scaledIDFT (bem.h:279-290) */
void orc_scaled_fft(int mode, const double *in_re_im, double lambda, double *out16);
void orc_scaled_idft(const double *in_re_im, double lambda, double *out16);
/* face2PointIntegrand (integrand.h:261-267), face2FaceIntegrand (integrand.h:228-259) */
void orc_face2point(int mode, const float *tri9, const float *pt3, double kre, double kim, int type,
                    double *out_re_im);
void orc_face2face(int mode, const float *verts, const int *src3, const int *trg3, double kre, double kim,
                   int type, float threshold, double *out_re_im);
/* laplace_weight (bem.h:306-316) for a face->point / face->face pair */
void orc_laplace_weight_point(int mode, float dt, float c, float grid_size, const float *tri9, const float *pt3,
                              int type, double *out16);
/* same with a configurable Gauss order per Duffy dimension, FP64 (for convergence checks) */
void orc_singular_fp64(const float *verts, const int *src3, const int *trg3, double kre, double kim, int type,
                       int gauss_n, double *out_re_im);

#ifdef __cplusplus
}
#endif
#endif
