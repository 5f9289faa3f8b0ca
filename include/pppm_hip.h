/*
 * pppm_hip.h — C ABI of the MI355X-native PPPM time step (libpppm_hip.so).
 *
 * Drop-in for the reference's PPPMSolver / FDTD solver API (src/algorithm/pppm.h,
 * src/algorithm/fdtd.h). Plain pointers and sizes only; no C++ types cross the ABI.
 *
 * Conventions
 *  - Every entry point returns an int status (PPPM_OK == 0, negative on error);
 *    pppm_last_error() returns a message for the last failure on that handle.
 *    No exceptions cross the ABI (the reference throws std::runtime_error from
 *    cuSynchronize, src/core/macro.h:119-166).
 *  - One handle per host thread. All device work of a handle runs on its stream
 *    (pppm_set_stream; default: the legacy null stream, as the reference).
 *  - Arrays passed in are copied (host or device, selected by on_device); device
 *    pointers returned by *_ptr calls are borrowed and stay valid until the next
 *    set_mesh/update_mesh/destroy.
 *  - Grid layout is the reference's GArr3D: x-major, z contiguous,
 *    index = (x*N + y)*N + z (src/core/array3D.h:54-87). History rings are 32 floats
 *    per triangle, slot = t mod 32 (src/algorithm/bem.h:55, src/core/array.h:39).
 */
#ifndef PPPM_HIP_H
#define PPPM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pppm_solver pppm_solver;

enum pppm_status
{
    PPPM_OK = 0,
    PPPM_ERR_ARG = -1,         /* bad argument */
    PPPM_ERR_HIP = -2,         /* HIP runtime error */
    PPPM_ERR_OUT_OF_GRID = -3, /* a triangle lies outside cells [2, N-3] */
    PPPM_ERR_STATE = -4,       /* call order (e.g. step before set_mesh, or after a failed
                                  update_mesh until an update succeeds) */
    PPPM_ERR_CAPACITY = -5     /* an internal capacity was exceeded */
};

enum pppm_flags
{
    /* Honour pml_width in the FDTD. The reference's PPPMSolver drops it
     * (src/algorithm/particle_grid.h:115), so the default follows that. */
    PPPM_FLAG_HONOR_PML = 1u << 0,
};

/* PPPMSolver::PPPMSolver(res, dl, dt, min_pos, pml_width, reflect_coeff)
 * (src/algorithm/pppm.h:39-55) plus TDBEM::init (src/algorithm/bem.h:264-277).
 * c is the speed of sound, hard-coded as AIR_WAVE_SPEED 340.3 in the reference
 * (src/core/macro.h:82). min_xyz may be NULL (origin). */
int pppm_create(pppm_solver **out, int res, float dl, float dt, const float *min_xyz, int pml_width,
                float reflect_coeff, float c, unsigned flags);
/* PPPMSolver::clear (src/algorithm/pppm.h:111-125) + delete */
int pppm_destroy(pppm_solver *s);
/* all subsequent work of this handle is enqueued on stream (a hipStream_t) */
int pppm_set_stream(pppm_solver *s, void *hip_stream);
const char *pppm_last_error(const pppm_solver *s);

/* PPPMSolver::set_mesh (src/algorithm/pppm.h:75-94): verts nv x 3 floats, tris nt x 3 ints.
 * Bins triangles, builds the 3x3x3 / 4x4x4 lists and the FP64 weight caches. */
int pppm_set_mesh(pppm_solver *s, const float *verts, int nv, const int *tris, int nt, int on_device);
/* PPPMSolver::update_mesh (src/algorithm/pppm.h:96-109): new vertex positions, same topology.
 * Re-bins and recomputes the cache entries whose geometry moved past the thresholds
 * (src/algorithm/pppm_cache.h:64-69, 122-127). */
int pppm_update_mesh(pppm_solver *s, const float *verts, int on_device);

/* PPPMSolver::update_grid_and_face / update_step (src/algorithm/pppm.h:137-152):
 * FDTD step, far field, Neumann input (nt floats), Dirichlet update, near field. */
int pppm_step(pppm_solver *s, const float *neumann, int on_device);
/* pppm_update_mesh_and_step: pppm_update_mesh then pppm_step in one call: the demo's per-sample
 * UpdateSound with a moving mesh (experiments/demo/main.cu:170-205). Same inputs and errors as
 * the two calls; verts and neumann are both host or both device memory (on_device). */
int pppm_update_mesh_and_step(pppm_solver *s, const float *verts, const float *neumann, int on_device);
/* Runs `steps` consecutive update_steps. neumann_series is device memory holding
 * steps x stride floats (row i = Neumann input of step i, nt used). After each step
 * the pressure at the n_listen cells (xyz triples, host array) is written to
 * out_dev[i*n_listen + l] (device memory). use_graph != 0 replays the step as a
 * captured HIP graph (a run of up to 64 steps as one graph, longer runs 32 steps per graph),
 * which the handle keeps for later runs whose arguments differ only by the rows of the series
 * and of the output (the segments of one series). Replaces the demo loop's per-step one-float D2H read
 * (experiments/demo/main.cu:266, src/core/array.h:276-288). */
int pppm_run(pppm_solver *s, const float *neumann_series, long long stride, int steps, const int *listen_xyz,
             int n_listen, float *out_dev, int use_graph);

/* pppm_run_noalloc: pppm_run without the graph and without any device allocation — no
 * counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. Same arguments and
 * results as pppm_run(..., use_graph = 0) on a static mesh (bit for bit); the steps are
 * enqueued asynchronously on the handle's stream (synchronise it before reading out_dev).
 * The gathers' per-step buffers (history windows, far field, listener table) live in one
 * file-scope device array of 16 MiB per GPU, held until the process ends; a mesh needs
 * 512 + 2 * ceil256(64 nt) + 4 * min(27 nt, N^3) bytes of it (about 70 000 triangles at most),
 * PPPM_ERR_CAPACITY beyond. Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (two handles on concurrent streams) race; consecutive calls are independent (each
 * writes what it reads). Timing spans (pppm_set_timing) are not recorded. */
int pppm_run_noalloc(pppm_solver *s, const float *neumann_series, long long stride, int steps, const int *listen_xyz,
                     int n_listen, float *out_dev);

/* pppm_run for a moving mesh: before step i the vertices are replaced by row i of
 * verts_series (device memory, steps x vstride floats, vstride >= 3 nv) and the mesh is
 * updated as by pppm_update_mesh — the demo loop's UpdateSound (experiments/demo/main.cu:
 * 170-205: object transform, update_mesh, update_grid_and_face). Each update is
 * synchronous (its status is read back and capacities grow as needed), so use_graph is
 * ignored for moving meshes. */
int pppm_run_moving(pppm_solver *s, const float *verts_series, long long vstride, const float *neumann_series,
                    long long stride, int steps, const int *listen_xyz, int n_listen, float *out_dev, int use_graph);

/* Batched scenes: independent scenes stepped together, each kernel of the time step one
 * launch over every scene (no counterpart in the reference, which steps one PPPMSolver at a
 * time: src/algorithm/pppm.h:137-145 per scene). scenes[i] are handles with their meshes set
 * (static meshes, no PML); the batch references them, so they must outlive it, and its runs
 * advance their time indices as pppm_run would. Each scene's results equal its own pppm_run
 * bit for bit.
 * pppm_batch_run: per scene i, neumann_series[i] (device, steps x strides[i] floats, row k =
 *   the Neumann input of step k, strides[i] >= nt_i) and n_listen listener cells (1 to 64)
 *   listen_xyz[3 (i n_listen + l) ..]; out_dev (device, steps x n_scenes x n_listen floats) gets
 *   out_dev[(k n_scenes + i) n_listen + l] = scene i's listener l after step k. use_graph != 0:
 *   32 batch steps per captured HIP graph launch. Synchronous.
 * pppm_batch_autotune: times every FDTD interior variant (0: one thread per cell; 1..4:
 *   LDS-staged 2.5D tiles of 8x64, 16x32, 4x64, 8x32 (y, z) columns) for each size class of grid
 *   edge N present in the batch (classes [0,40) [40,56) [56,72) [72,88) [88,112) [112,160)
 *   [160,...)) over `reps` repetitions, keeps the fastest per class, and (ms_out optional,
 *   7 x 5 floats) reports the milliseconds per repetition [class][variant] (untimed entries are
 *   left as they were). The FDTD launches it times do not advance t and rewrite the same values,
 *   so the scenes' state is unchanged.
 * pppm_batch_variants / pppm_batch_set_variants: the FDTD variant of each scene (n_scenes ints).
 * pppm_batch_set_timing / pppm_batch_get_timing: HIP-event timing of each phase of eagerly
 *   launched batch steps (a timed run does not use the graph): 0 FDTD (shell and interiors),
 *   1 history windows and time advance, 2 far field, 3 Dirichlet update, 4 near field;
 *   milliseconds and launch counts accumulated since the last get (which resets them). */
typedef struct pppm_batch pppm_batch;
int pppm_batch_create(pppm_batch **out, pppm_solver *const *scenes, int n_scenes);
int pppm_batch_destroy(pppm_batch *b);
const char *pppm_batch_last_error(const pppm_batch *b);
int pppm_batch_set_stream(pppm_batch *b, void *hip_stream);
int pppm_batch_run(pppm_batch *b, const float *const *neumann_series, const long long *strides, int steps,
                   int n_listen, const int *listen_xyz, float *out_dev, int use_graph);
int pppm_batch_autotune(pppm_batch *b, int reps, float *ms_out);
int pppm_batch_variants(pppm_batch *b, int *variants);
int pppm_batch_set_variants(pppm_batch *b, const int *variants);
int pppm_batch_set_timing(pppm_batch *b, int enable);
int pppm_batch_get_timing(pppm_batch *b, double *ms8, long long *count8);

/* Adjoint: gradients of a listener loss with respect to the per-triangle Neumann inputs (no
 * counterpart in the reference). The step (src/algorithm/pppm.h:137-145) is linear in the grid
 * levels, the history rings and the inputs, so its adjoint — the transposed step run backwards,
 * ABC shell included (src/algorithm/fdtd.cu:20-150, pppm.cu:8-204) — needs no forward state,
 * only the cotangents of the listener samples. Static mesh: the mesh must not change between
 * the forward run and the adjoint (PPPM_ERR_STATE otherwise); the PML is not differentiated.
 * pppm_adjoint_reset: zero adjoint state; the next reverse step is the forward step t_last
 *   (the last step of the forward run, whose time index was t_last after it).
 * pppm_adjoint_run: reverses `steps` steps, t = T, T-1, ..., T-steps+1 (T the current reverse
 *   index). gout_dev (device, steps x n_listen): row r = the cotangent of the listener samples
 *   of forward step T-steps+1+r (the cells listen_xyz, host, n_listen <= 64); grad_dev (device,
 *   steps x grad_stride, grad_stride >= nt): row r = the gradient with respect to the Neumann
 *   input of step T-steps+1+r. Consecutive calls continue backwards, so a run can be reversed
 *   segment by segment. use_graph != 0: a run of up to 64 steps is one captured graph, a longer
 *   one 32 reverse steps per graph; the handle keeps the graph for later runs whose arguments
 *   differ only by the rows of gout and grad (segments of one reversal). Asynchronous on the
 *   handle's stream, as pppm_run.
 * pppm_state_bytes / pppm_save_state / pppm_load_state: a checkpoint of the forward state
 *   (grid levels t and t-1 with the ghost layer, both history rings, t) in caller-owned device
 *   memory; loading one restores the solver to that step (same grid and mesh only). */
int pppm_adjoint_reset(pppm_solver *s, int t_last);
int pppm_adjoint_run(pppm_solver *s, const float *gout_dev, int steps, const int *listen_xyz, int n_listen,
                     float *grad_dev, long long grad_stride, int use_graph);
int pppm_state_bytes(pppm_solver *s, long long *bytes);
int pppm_save_state(pppm_solver *s, void *dst_dev);
int pppm_load_state(pppm_solver *s, const void *src_dev);

/* Batched adjoint: the reverse steps of every scene of a batch, each reverse-step kernel one
 * launch over all scenes (no counterpart in the reference). Each scene's gradient equals its own
 * pppm_adjoint_run's bit for bit.
 * pppm_batch_adjoint_reset: pppm_adjoint_reset(scene, t_last) for every scene.
 * pppm_batch_adjoint_run: reverses `steps` steps of every scene (their reverse indices must be
 *   equal: reset together). gout_dev (device, steps x n_scenes x n_listen floats, the layout of
 *   pppm_batch_run's output): gout_dev[(r n_scenes + i) n_listen + l] = the cotangent of scene i's
 *   listener l (cell listen_xyz[3 (i n_listen + l) ..], host, n_listen <= 64) at forward step
 *   T-steps+1+r; grad_dev[i] (device, steps x grad_strides[i] floats, grad_strides[i] >= nt_i):
 *   scene i's gradient rows as in pppm_adjoint_run. use_graph != 0: 32 reverse batch steps per
 *   captured graph launch. Synchronous. */
int pppm_batch_adjoint_reset(pppm_batch *b, int t_last);
int pppm_batch_adjoint_run(pppm_batch *b, const float *gout_dev, int steps, int n_listen, const int *listen_xyz,
                           float *const *grad_dev, const long long *grad_strides, int use_graph);

/* Phase entry points used by the reference's tests (src/algorithm/pppm.h:127-135,
 * src/algorithm/fdtd.h:73-87). */
int pppm_fdtd_step(pppm_solver *s);                              /* FDTD::step */
int pppm_solve_far(pppm_solver *s);                              /* solve_fdtd_far */
int pppm_set_neumann(pppm_solver *s, const float *neumann, int on_device); /* set_neumann_condition */
int pppm_update_dirichlet(pppm_solver *s);                       /* update_dirichlet */
int pppm_solve_near(pppm_solver *s);                             /* solve_fdtd_near */
/* fdtd.reset(); neumann.reset(); dirichlet.reset() (experiments/demo/main.cu:251-254) */
int pppm_reset(pppm_solver *s);

/* time_idx() (src/algorithm/pppm.h:71); t starts at -1 (src/algorithm/fdtd.h:55) */
int pppm_time_idx(const pppm_solver *s, int *t);
/* fdtd.t = t, as the demo does after a reset (experiments/demo/main.cu:255) */
int pppm_set_time_idx(pppm_solver *s, int t);
/* fdtd.grids[t + rel] device pointer (N^3 floats); get_grid() is rel = 0 (pppm.h:154) */
int pppm_grid_ptr(pppm_solver *s, int rel, float **dev_ptr);
/* copies grids[t] at n cells (xyz triples, host) to host out — the listener read
 * without a full-grid transfer */
int pppm_read_cells(pppm_solver *s, const int *xyz, int n, float *host_out);
/* device pointers of the Neumann and Dirichlet rings (nt x 32 floats each, slot t mod 32) */
int pppm_history_ptrs(pppm_solver *s, float **neumann, float **dirichlet);
/* get_triangles() (pppm.h:156): 13 floats per triangle to host:
 * center(3) normal(3) area cell(3) base(3) */
int pppm_triangles(pppm_solver *s, float *host13);
/* counts: [0]=T [1]=C3 cells (cells within +-1 of a triangle cell) [2]=P3 pairs
 * [3]=non-empty base cells [4]=P4 pairs [5]=grid-cache recomputes at last update
 * [6]=face-cache recomputes at last update [7]=face-cache pool entries */
int pppm_stats(pppm_solver *s, long long *out8);

/* Per-phase GPU timing with HIP events on the handle's stream, the counterpart of the
 * reference's log_time / LOG_TIME phase timers (src/core/macro.h:35-48, pppm.cu:45-50).
 * Phases: 0 FDTD, 1 far field (+ Neumann input), 2 Dirichlet update, 3 near field,
 * 4 update_mesh, and within it 5 grid-cache recompute (k_grid_weights), 6 face-cache update
 * (lookup, dirty-pair scan, integrals), 7 geometry, binning and lists. Only eagerly launched
 * work is timed (graph replays are not).
 * pppm_get_timing returns accumulated milliseconds and launch counts, then resets them.
 * pppm_get_update_work: the mesh updates completed while timing was on, with the grid-cache
 * triangles and face-cache pairs they recomputed (out3 = updates, triangles, pairs), then
 * resets them. No counterpart in the reference (its log_time, src/core/macro.h:35-48, times
 * phases only). */
int pppm_set_timing(pppm_solver *s, int enable);
int pppm_get_timing(pppm_solver *s, double *ms8, long long *count8);
int pppm_get_update_work(pppm_solver *s, long long *out3);

/* Cache inspection / loading in the reference's LayerWeightHalf layout (72 B per
 * record, src/algorithm/bem.h:163-169). Grid cache: records (tri*27+o)*2 + {0: fdtd, 1: bem}
 * (GridCache::CacheElement, pppm_cache.h:13-20). Face cache: one record per
 * (triangle c, entry of the 4x4x4 list of c's base cell), CSR over c
 * (FaceCache::cache, pppm_cache.h:81-155); start has nt+1 entries. */
int pppm_get_grid_weights(pppm_solver *s, void *host_recs72);
int pppm_set_grid_weights(pppm_solver *s, const void *host_recs72);
int pppm_get_face_weights(pppm_solver *s, int *host_start, int *host_nbr, void *host_recs72);
int pppm_set_face_weights(pppm_solver *s, const void *host_recs72);
int pppm_get_interp_weights(pppm_solver *s, float *host_out8);
/* ParticleGrid lists (src/algorithm/particle_grid.h:100-112): C3 cells with their
 * 3x3x3 pair lists (entries tri*27 + weight offset), non-empty base cells with their
 * 4x4x4 triangle lists. Pass NULL to skip an array; sizes come from pppm_stats. */
int pppm_get_c3(pppm_solver *s, int *cells, int *start, int *pairs);
int pppm_get_b4(pppm_solver *s, int *cells, int *start, int *tris);

/* Cell classification and fresh cells for moving meshes — the ghost-cell solver's cell data
 * (GhostCellSolver::precompute_cell_data / fill_in_fresh_cell / set_solid_cell_zero,
 * src/algorithm/ghost_cell.h:69-212, ghost_cell_util.cu:7-230, ghost_cell.cu:51-218,547-560)
 * computed on the handle's current mesh and grid. Cell types: 0 AIR, 1 SOLID, 2 GHOST,
 * 3 UNKNOWN (ghost_cell.h:11-17).
 * pppm_classify: labels every cell (nearest triangle in the 7^3 window, side of the surface,
 *   ghost cells, column fill, ghost indices in linear order); the previous labels are kept
 *   for pppm_fill_fresh (all AIR before the first call). thin_shell: the reference's
 *   thin-shell test (condition_number_threshold == 0, ghost_cell.cu:236).
 * pppm_fill_fresh: cells that turned AIR get values at t and t-1 from the cells around their
 *   mirror point, with the per-triangle accelerations now / one step before (nt floats each).
 *   *fresh_count = their number, or -1 if one had no nearest triangle (then 1e10 is written
 *   at t, as the reference's error path).
 * pppm_set_solid_zero: SOLID cells of grid level t set to 0.
 * pppm_get_cells: host copies (N^3 entries, x-major; any pointer may be NULL). */
int pppm_classify(pppm_solver *s, int thin_shell, int *ghost_count);
int pppm_fill_fresh(pppm_solver *s, const float *acc_now, const float *acc_old, int on_device, int *fresh_count);
int pppm_set_solid_zero(pppm_solver *s);
int pppm_get_cells(pppm_solver *s, int *type, int *ghost_idx, int *nearest, float *point3, float *reflect3,
                   float *dist);

/* Ghost-cell boundary solve of the comparison solver (GhostCellSolver,
 * src/algorithm/ghost_cell.h:72-193), on the handle's grid and mesh. After pppm_classify:
 * pppm_ghost_setup: ghost list; per ghost the 8x8 stencil system of the trilinear-plus basis
 *   with the Neumann row at the nearest surface point (construct_phi_matrix_kernel,
 *   ghost_cell.cu:254-300), its SVD, condition number max S / min S and the stencil weights
 *   inv(Phi)^T phi(reflect) (precompute_p_weight_kernel, ghost_cell.cu:302-350,
 *   inverse_from_svd_kernel, svd.cu:129-148); rows with a condition number above cond_threshold
 *   use the first-order rule. cond_threshold = 0: first order everywhere and no linear solve
 *   (the reference's thin-shell mode, ghost_cell.cu:490-495). second_order (optional): the
 *   number of second-order rows. Synchronous. PPPM_ERR_OUT_OF_GRID if a stencil leaves the grid.
 * pppm_ghost_set_neumann: previous Neumann data <- current, current <- neumann (nt floats,
 *   set_boundary_condition, ghost_cell.h:118-123).
 * pppm_ghost_solve: right-hand side from the pressure at the current level and the Neumann
 *   data, the sparse solve, and the solution written into the ghost cells (solve_ghost_cell,
 *   ghost_cell.cu:522-540; BiCGSTAB_Solver::solve, sparse.cu:512-545). iterations (optional,
 *   synchronises): the solver's iteration count.
 * pppm_ghost_step: GhostCellSolver::update (ghost_cell.h:131-141): FDTD step, set_neumann,
 *   solve, SOLID cells to zero.
 * pppm_ghost_update_mesh: GhostCellSolver::update_mesh (ghost_cell.h:96-112): new vertices,
 *   classification (thin-shell when the threshold is 0), fresh cells from the current and
 *   previous Neumann data, ghost setup with the same threshold. fresh_count as pppm_fill_fresh.
 * pppm_ghost_get: host copies (any pointer may be NULL) of the ghost cells' linear indices,
 *   order (1 = second order), weights (8 per ghost), condition numbers, last right-hand side,
 *   last solution, and the sparse rows as 9 (column, value) slots per ghost (column -1: empty;
 *   slot 8: the unit diagonal). */
int pppm_ghost_setup(pppm_solver *s, float cond_threshold, int *second_order);
int pppm_ghost_set_neumann(pppm_solver *s, const float *neumann, int on_device);
int pppm_ghost_solve(pppm_solver *s, int *iterations);
int pppm_ghost_step(pppm_solver *s, const float *neumann, int on_device);
int pppm_ghost_update_mesh(pppm_solver *s, const float *verts, int on_device, int *fresh_count);
int pppm_ghost_get(pppm_solver *s, int *cells, int *order, float *weights8, float *cond, float *rhs, float *x,
                   int *cols9, float *vals9);

/* pppm_ffat_collect: far-field (FFAT) map accumulation at the current time level
 * (collect_ffat_map, experiments/static/main.cu:60-80 and experiments/acoustic/main.cu:46-74):
 * map_dev (device memory, 6 x 2w x 2w floats, row-major) gets map(b, x, y) += |p(cell)| with
 * face b = 2 i + (sign > 0) of axis i and cell = c + sign w e_i + (x - w) e_{i+1} + (y - w) e_{i+2}
 * (axes cyclic, c = (N/2, N/2, N/2)). Requires 1 <= w and the cube inside the grid
 * (w <= N/2 - 1 for even N); the reference does not check. Asynchronous on the handle's stream. */
int pppm_ffat_collect(pppm_solver *s, int w, float *map_dev);

/* pppm_slice_image: the experiments' picture of one grid plane at the current time level
 * (save_grid, src/utils/visualize.h:120-132, through CudaRender::setData / add_mesh_to_images,
 * src/ui/window.cu:28-330, and write_to_png, src/utils/array_writer.h:68-105): plane `index`
 * normal to `axis` (0: x, 1: y, 2: z; the reference's slice vector (0.5, 0, 0) is axis 0, index
 * N/2), upsampled to W = N * (1080 / N) pixels with bilinear interpolation, the jet colour map
 * over [-max_value, max_value] at half brightness, cell borders and the mesh's cut drawn in
 * white. rgba_host: W * W * 4 bytes, PNG row order (top row first), or NULL to query W only.
 * width (optional): W. Synchronous. */
int pppm_slice_image(pppm_solver *s, int axis, int index, float max_value, unsigned char *rgba_host, int *width);

/* Modal sound source of a rigid body (RigidBody, src/utils/RigidBody.h:53-166) on the GPU: the
 * surface modal matrix, the per-mode IIR oscillators and the surface accelerations that feed
 * the solver's Neumann input. A handle owns its device data and runs on its stream.
 * pppm_modal_create: eigvecs = the truncated eigenvector matrix, 3*nv rows (vertex dofs) x
 *   nmodes columns, row-major (RigidBody::LoadEigen_, RigidBody.cu:141-188); eigvals its
 *   eigenvalues; tris / normals the surface triangles and their unit normals. The IIR
 *   coefficients come from ModalInfo::SetCoeffs (RigidBody.cu:11-27) with the material's
 *   Rayleigh damping alpha, beta and timestep = 1 / sample_rate (RigidBody::InitIIR_); the
 *   surface matrix is MatrixVertToTri (RigidBody.cu:206-230): per triangle and mode the mean of
 *   the three vertex displacements dotted with the triangle normal. Oscillator state starts at 0.
 *   nt <= 65535 * 256.
 * pppm_modal_run: for each of nsamples samples (0 <= nsamples <= 2^30), the IIR step of every
 *   mode with the modal forces of that sample (forces: nsamples x nmodes, row-major) and the
 *   surface accelerations acc[s][t] = sum_i M(t, i) Q_i, summed in mode order with one fused
 *   multiply-add per term (CalculateIIR_ / update_surf_acc_kernel, RigidBody.cu:240-295),
 *   written to accs (nsamples x nt, device memory). on_device: where forces live.
 * pppm_modal_state / pppm_modal_set_state: the oscillators' (q1, q2, q3) per mode (nmodes x 3,
 *   host memory), for RigidBody::SaveState / LoadState (RigidBody.cu:60-88).
 * pppm_modal_surface_matrix: host copy of M (nt x nmodes, row-major).
 * pppm_modal_coeffs: host copy of the IIR coefficients (nmodes x 3: coeff1, coeff2, coeff3). */
typedef struct pppm_modal pppm_modal;
int pppm_modal_create(pppm_modal **out, const float *eigvecs, int nv, int nmodes, const float *eigvals,
                      const int *tris, int nt, const float *normals, float alpha, float beta, float sample_rate,
                      void *hip_stream);
int pppm_modal_destroy(pppm_modal *m);
int pppm_modal_run(pppm_modal *m, const float *forces, int nsamples, float *accs, int on_device);
int pppm_modal_state(pppm_modal *m, float *q3);
int pppm_modal_set_state(pppm_modal *m, const float *q3);
int pppm_modal_surface_matrix(pppm_modal *m, float *host_out);
int pppm_modal_coeffs(pppm_modal *m, float *host_out);

#ifdef __cplusplus
}
#endif
#endif
