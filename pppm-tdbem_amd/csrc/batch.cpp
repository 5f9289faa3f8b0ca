// Batched scenes: many independent PPPM scenes (solver handles with their meshes set) stepped
// together. Each kernel of the time step is one launch over every scene's blocks (a scene
// table in device memory gives each block its scene), and a captured HIP graph replays 32
// such batch steps per launch. The FDTD interior of each scene uses the kernel variant picked
// for its size class by the autotuner (one thread per cell, or LDS-staged 2.5D tiles); every
// variant gives the same bits, so the batch equals the scenes' single-scene runs bit for bit.
//
// Reference behaviour per scene: PPPMSolver::update_grid_and_face (src/algorithm/pppm.h:137-145)
// as in pppm_run; the reference steps one scene at a time.
#include "pppm_hip.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "solver_state.hpp"

using namespace pppm_amd;

namespace
{

// size classes of the autotuner table, by grid edge N: [lo, next lo)
constexpr int kClassLo[] = {0, 40, 56, 72, 88, 112, 160};
constexpr int kClasses = sizeof(kClassLo) / sizeof(kClassLo[0]);
int size_class(int N)
{
    int c = 0;
    while (c + 1 < kClasses && N >= kClassLo[c + 1])
        c++;
    return c;
}
// FDTD interior variant per size class before any tuning (0: one thread per cell)
constexpr int kDefaultVariant[kClasses] = {0, 0, 0, 0, 0, 0, 0};

struct Map
{
    DBuf<int> scene_of, first;
    int total = 0, per = 0;
    BatchMap view() const { return BatchMap{scene_of.p, first.p, total, per}; }
};

}  // namespace

struct pppm_batch
{
    std::vector<pppm_solver *> sc;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    std::vector<int> variant;  // FDTD interior variant per scene
    DBuf<SceneStep> descs;
    DBuf<FdtdScene> fdescs;
    Map m_fdtd, m_tile[kFdtdVariants], m_win, m_far, m_dir, m_near;
    // batched adjoint: scene table and one block map per reverse-step kernel
    DBuf<AdjScene> adescs;
    Map m_adj[kAdjKernels];
    // per-phase HIP event timing of eager batch steps: 0 FDTD (shell and interiors), 1 history
    // windows and time advance, 2 far field, 3 Dirichlet update, 4 near field
    bool timing = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> spans;
    std::vector<hipEvent_t> pool;
    double phase_ms[8] = {0};
    long long phase_n[8] = {0};
};

namespace
{

int bfail(pppm_batch *b, int code, const std::string &msg)
{
    if (b)
        b->err = msg;
    return code;
}

#define BCK(expr)                                                                                \
    do                                                                                           \
    {                                                                                            \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return bfail(b, PPPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

#define BRC(expr)         \
    do                    \
    {                     \
        int rc_ = (expr); \
        if (rc_)          \
            return rc_;   \
    } while (0)

// builds a block map from the per-scene block counts (0: the scene has no blocks in it)
int build_map(pppm_batch *b, Map &m, const std::vector<int> &blocks)
{
    const int n = (int)blocks.size();
    std::vector<int> first(n), scene_of;
    int total = 0;
    for (int i = 0; i < n; i++)
    {
        first[i] = total;
        total += blocks[i];
    }
    scene_of.resize(std::max(total, 1));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < blocks[i]; k++)
            scene_of[first[i] + k] = i;
    m.total = total;
    m.per = (total + 7) / 8;
    BCK(m.first.ensure(n));
    BCK(m.scene_of.ensure(scene_of.size()));
    BCK(hipMemcpyAsync(m.first.p, first.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BCK(hipMemcpyAsync(m.scene_of.p, scene_of.data(), scene_of.size() * sizeof(int), hipMemcpyHostToDevice,
                       b->stream));
    BCK(hipStreamSynchronize(b->stream));  // the host vectors go out of scope
    return PPPM_OK;
}

// FDTD scene table and maps for the current variants
int build_fdtd(pppm_batch *b)
{
    const int n = (int)b->sc.size();
    std::vector<FdtdScene> fd(n);
    std::vector<int> shell(n);
    std::vector<std::vector<int>> tiles(kFdtdVariants, std::vector<int>(n, 0));
    for (int i = 0; i < n; i++)
    {
        const pppm_solver *s = b->sc[i];
        FdtdScene &f = fd[i];
        f.p = solver_fdtd_params(s);
        f.grid3 = s->grid.p;
        f.st = s->state.p;
        f.xch = 0;
        const int v = b->variant[i];
        f.nb_inner = v == 0 ? fdtd_inner_blocks(s->N) : 0;
        shell[i] = fdtd_shell_blocks(s->N) + f.nb_inner;
        if (v > 0)
            tiles[v][i] = fdtd_tile_blocks(s->N, v, f.xch);
    }
    BCK(b->fdescs.ensure(n));
    BCK(hipMemcpyAsync(b->fdescs.p, fd.data(), n * sizeof(FdtdScene), hipMemcpyHostToDevice, b->stream));
    BRC(build_map(b, b->m_fdtd, shell));
    for (int v = 1; v < kFdtdVariants; v++)
        BRC(build_map(b, b->m_tile[v], tiles[v]));
    return PPPM_OK;
}

// FDTD of every scene (no advance): idempotent, the autotuner times it repeatedly
int enqueue_fdtd(pppm_batch *b)
{
    BCK(launch_batch_fdtd(b->fdescs.p, b->m_fdtd.view(), b->stream));
    for (int v = 1; v < kFdtdVariants; v++)
        if (b->m_tile[v].total > 0)
            BCK(launch_batch_tile(v, b->fdescs.p, b->m_tile[v].view(), b->stream));
    return PPPM_OK;
}

hipEvent_t take_event(pppm_batch *b)
{
    if (!b->pool.empty())
    {
        hipEvent_t e = b->pool.back();
        b->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

int span_begin(pppm_batch *b, int phase)
{
    if (!b->timing)
        return -1;
    hipEvent_t a = take_event(b), c = take_event(b);
    if (!a || !c || hipEventRecord(a, b->stream) != hipSuccess)
        return -1;
    b->spans.push_back({phase, {a, c}});
    return (int)b->spans.size() - 1;
}

void span_end(pppm_batch *b, int i)
{
    if (i >= 0)
        (void)hipEventRecord(b->spans[i].second.second, b->stream);
}

int collect_spans(pppm_batch *b)
{
    if (b->spans.empty())
        return PPPM_OK;
    BCK(hipStreamSynchronize(b->stream));
    for (auto &sp : b->spans)
    {
        float ms = 0;
        if (hipEventElapsedTime(&ms, sp.second.first, sp.second.second) == hipSuccess)
        {
            b->phase_ms[sp.first] += ms;
            b->phase_n[sp.first]++;
        }
        b->pool.push_back(sp.second.first);
        b->pool.push_back(sp.second.second);
    }
    b->spans.clear();
    return PPPM_OK;
}

int enqueue_batch_step(pppm_batch *b)
{
    int sp = span_begin(b, 0);
    BRC(enqueue_fdtd(b));
    span_end(b, sp);
    const int n = (int)b->sc.size();
    sp = span_begin(b, 1);
    BCK(launch_batch(BatchKernel::Windows, b->descs.p, b->m_win.view(), b->stream));
    BCK(launch_batch_advance(b->descs.p, n, b->stream));
    span_end(b, sp);
    sp = span_begin(b, 2);
    BCK(launch_batch(BatchKernel::Far, b->descs.p, b->m_far.view(), b->stream));
    span_end(b, sp);
    sp = span_begin(b, 3);
    BCK(launch_batch(BatchKernel::Dirichlet, b->descs.p, b->m_dir.view(), b->stream));
    span_end(b, sp);
    sp = span_begin(b, 4);
    BCK(launch_batch(BatchKernel::Near, b->descs.p, b->m_near.view(), b->stream));
    span_end(b, sp);
    if (b->spans.size() > 4096)
        return collect_spans(b);
    return PPPM_OK;
}

int check_scenes(pppm_batch *b)
{
    for (pppm_solver *s : b->sc)
    {
        const int rc = solver_check_ready(s, "batch");
        if (rc)
            return bfail(b, rc, s->err);
        if (s->use_pml)
            return bfail(b, PPPM_ERR_ARG, "batch: scenes with the PML are not batched (use pppm_run)");
    }
    return PPPM_OK;
}

}  // namespace

extern "C" {

int pppm_batch_create(pppm_batch **out, pppm_solver *const *scenes, int n_scenes)
{
    if (!out)
        return PPPM_ERR_ARG;
    *out = nullptr;
    if (!scenes || n_scenes < 1)
        return PPPM_ERR_ARG;
    for (int i = 0; i < n_scenes; i++)
        if (!scenes[i])
            return PPPM_ERR_ARG;
    pppm_batch *b = new pppm_batch();
    b->sc.assign(scenes, scenes + n_scenes);
    if (hipStreamCreate(&b->stream) != hipSuccess)
    {
        delete b;
        return PPPM_ERR_HIP;
    }
    b->own_stream = true;
    b->variant.resize(n_scenes);
    for (int i = 0; i < n_scenes; i++)
        b->variant[i] = kDefaultVariant[size_class(scenes[i]->N)];
    *out = b;
    return PPPM_OK;
}

int pppm_batch_destroy(pppm_batch *b)
{
    if (!b)
        return PPPM_ERR_ARG;
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    for (auto &sp : b->spans)
    {
        (void)hipEventDestroy(sp.second.first);
        (void)hipEventDestroy(sp.second.second);
    }
    for (hipEvent_t e : b->pool)
        (void)hipEventDestroy(e);
    if (b->own_stream && b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
    return PPPM_OK;
}

const char *pppm_batch_last_error(const pppm_batch *b) { return b ? b->err.c_str() : "null handle"; }

int pppm_batch_set_stream(pppm_batch *b, void *hip_stream)
{
    if (!b)
        return PPPM_ERR_ARG;
    BCK(hipStreamSynchronize(b->stream));
    if (b->own_stream)
        BCK(hipStreamDestroy(b->stream));
    b->own_stream = false;
    b->stream = (hipStream_t)hip_stream;
    return PPPM_OK;
}

int pppm_batch_run(pppm_batch *b, const float *const *neumann_series, const long long *strides, int steps,
                   int n_listen, const int *listen_xyz, float *out_dev, int use_graph)
{
    if (!b)
        return PPPM_ERR_ARG;
    const int n = (int)b->sc.size();
    if (steps < 0 || !neumann_series || !strides || !listen_xyz || !out_dev || n_listen < 1 || n_listen > 64)
        return bfail(b, PPPM_ERR_ARG, "batch_run: bad arguments (1 to 64 listeners per scene)");
    BRC(check_scenes(b));
    for (int i = 0; i < n; i++)
        if (!neumann_series[i] || strides[i] < b->sc[i]->nt)
            return bfail(b, PPPM_ERR_ARG, "batch_run: scene " + std::to_string(i) + ": Neumann series missing or stride < nt");
    if (steps == 0)
        return PPPM_OK;
    // per-scene set-up: list-ordered weight records, listener tables, pending work of the
    // scenes' own streams
    std::vector<SceneStep> hd(n);
    for (int i = 0; i < n; i++)
    {
        pppm_solver *s = b->sc[i];
        if (solver_ensure_ordered(s))
            return bfail(b, PPPM_ERR_HIP, "batch_run: scene " + std::to_string(i) + ": " + s->err);
        const int rc = solver_set_listeners(s, listen_xyz + 3 * (size_t)i * n_listen, n_listen);
        if (rc)
            return bfail(b, rc, "batch_run: scene " + std::to_string(i) + ": " + s->err);
        BCK(hipStreamSynchronize(s->stream));
        hd[i] = solver_scene_step(s, neumann_series[i], s->t_host + 1, strides[i], n_listen,
                                  out_dev + (size_t)i * n_listen);
        hd[i].out_stride = (long long)n * n_listen;
    }
    BCK(b->descs.ensure(n));
    BCK(hipMemcpyAsync(b->descs.p, hd.data(), n * sizeof(SceneStep), hipMemcpyHostToDevice, b->stream));
    std::vector<int> nw(n), nf(n), nd(n), nn(n);
    for (int i = 0; i < n; i++)
    {
        nw[i] = batch_blocks(hd[i], BatchKernel::Windows);
        nf[i] = batch_blocks(hd[i], BatchKernel::Far);
        nd[i] = batch_blocks(hd[i], BatchKernel::Dirichlet);
        nn[i] = batch_blocks(hd[i], BatchKernel::Near);
    }
    BRC(build_map(b, b->m_win, nw));
    BRC(build_map(b, b->m_far, nf));
    BRC(build_map(b, b->m_dir, nd));
    BRC(build_map(b, b->m_near, nn));
    BRC(build_fdtd(b));

    int done = 0;
    if (use_graph && !b->timing)
    {
        const int chunk = std::min(steps, 32);
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        BCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = PPPM_OK;
        for (int i = 0; i < chunk && rc == PPPM_OK; i++)
            rc = enqueue_batch_step(b);
        const hipError_t ee = hipStreamEndCapture(b->stream, &g);
        if (rc || ee != hipSuccess)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return rc ? rc : bfail(b, PPPM_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(ee));
        }
        if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
        {
            (void)hipGraphDestroy(g);
            return bfail(b, PPPM_ERR_HIP, "batch graph instantiate failed");
        }
        hipError_t le = hipSuccess;
        for (; done + chunk <= steps && le == hipSuccess; done += chunk)
            le = hipGraphLaunch(ge, b->stream);
        const hipError_t se = hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        BCK(le);
        BCK(se);
    }
    for (; done < steps; done++)
        BRC(enqueue_batch_step(b));
    BCK(hipStreamSynchronize(b->stream));
    for (pppm_solver *s : b->sc)
        s->t_host += steps;
    return PPPM_OK;
}

int pppm_batch_adjoint_reset(pppm_batch *b, int t_last)
{
    if (!b)
        return PPPM_ERR_ARG;
    BRC(check_scenes(b));
    for (size_t i = 0; i < b->sc.size(); i++)
    {
        const int rc = pppm_adjoint_reset(b->sc[i], t_last);
        if (rc)
            return bfail(b, rc, "batch_adjoint_reset: scene " + std::to_string(i) + ": " + b->sc[i]->err);
    }
    return PPPM_OK;
}

int pppm_batch_adjoint_run(pppm_batch *b, const float *gout_dev, int steps, int n_listen, const int *listen_xyz,
                           float *const *grad_dev, const long long *grad_strides, int use_graph)
{
    if (!b)
        return PPPM_ERR_ARG;
    const int n = (int)b->sc.size();
    if (steps < 0 || !grad_dev || !grad_strides || n_listen < 0 || n_listen > 64 ||
        (n_listen > 0 && (!listen_xyz || !gout_dev)))
        return bfail(b, PPPM_ERR_ARG, "batch_adjoint_run: bad arguments (at most 64 listeners per scene)");
    BRC(check_scenes(b));
    const int t_last = b->sc[0]->adj_t;
    for (int i = 0; i < n; i++)
    {
        pppm_solver *s = b->sc[i];
        const std::string sn = "batch_adjoint_run: scene " + std::to_string(i) + ": ";
        if (!grad_dev[i] || grad_strides[i] < s->nt)
            return bfail(b, PPPM_ERR_ARG, sn + "gradient missing or stride < nt");
        if (s->adj_version != s->mesh_version)
            return bfail(b, PPPM_ERR_STATE, sn + "no adjoint_reset since the mesh changed");
        if (s->adj_t != t_last)
            return bfail(b, PPPM_ERR_STATE, sn + "its reverse time differs from scene 0's (reset them together)");
    }
    if (steps > t_last + 1)
        return bfail(b, PPPM_ERR_ARG, "batch_adjoint_run: more steps than remain before t = 0");
    if (steps == 0)
        return PPPM_OK;
    // per-scene set-up: listener tables, pending work of the scenes' own streams
    const int t_first = t_last - steps + 1;
    std::vector<AdjScene> hd(n);
    for (int i = 0; i < n; i++)
    {
        pppm_solver *s = b->sc[i];
        const int rc = solver_set_listeners(s, n_listen > 0 ? listen_xyz + 3 * (size_t)i * n_listen : nullptr, n_listen);
        if (rc)
            return bfail(b, rc, "batch_adjoint_run: scene " + std::to_string(i) + ": " + s->err);
        BCK(hipStreamSynchronize(s->stream));
        // cotangent rows: (steps, n_scenes, n_listen), scene i at column block i
        hd[i] = solver_adj_scene(s, n_listen > 0 ? gout_dev + (size_t)i * n_listen : nullptr, (long long)n * n_listen,
                                 n_listen, grad_dev[i], grad_strides[i], t_first);
    }
    BCK(b->adescs.ensure(n));
    BCK(hipMemcpyAsync(b->adescs.p, hd.data(), n * sizeof(AdjScene), hipMemcpyHostToDevice, b->stream));
    BatchMap maps[kAdjKernels];
    for (int k = 0; k < kAdjKernels; k++)
    {
        std::vector<int> nb(n);
        for (int i = 0; i < n; i++)
            nb[i] = adj_blocks(hd[i], k);
        BRC(build_map(b, b->m_adj[k], nb));
        maps[k] = b->m_adj[k].view();
    }
    int done = 0;
    if (use_graph && steps >= 2)
    {
        const int chunk = std::min(steps, 32);
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        BCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int i = 0; i < chunk && e == hipSuccess; i++)
            e = enqueue_batch_adjoint_step(b->adescs.p, maps, n, b->stream);
        const hipError_t ee = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess || ee != hipSuccess)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return bfail(b, PPPM_ERR_HIP, "batch adjoint graph capture failed");
        }
        if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
        {
            (void)hipGraphDestroy(g);
            return bfail(b, PPPM_ERR_HIP, "batch adjoint graph instantiate failed");
        }
        hipError_t le = hipSuccess;
        for (; done + chunk <= steps && le == hipSuccess; done += chunk)
            le = hipGraphLaunch(ge, b->stream);
        const hipError_t se = hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        BCK(le);
        BCK(se);
    }
    for (; done < steps; done++)
        BCK(enqueue_batch_adjoint_step(b->adescs.p, maps, n, b->stream));
    BCK(hipStreamSynchronize(b->stream));
    for (pppm_solver *s : b->sc)
        s->adj_t -= steps;
    return PPPM_OK;
}

int pppm_batch_set_timing(pppm_batch *b, int enable)
{
    if (!b)
        return PPPM_ERR_ARG;
    BRC(collect_spans(b));
    b->timing = enable != 0;
    return PPPM_OK;
}

int pppm_batch_get_timing(pppm_batch *b, double *ms8, long long *count8)
{
    if (!b)
        return PPPM_ERR_ARG;
    BRC(collect_spans(b));
    for (int i = 0; i < 8; i++)
    {
        if (ms8)
            ms8[i] = b->phase_ms[i];
        if (count8)
            count8[i] = b->phase_n[i];
        b->phase_ms[i] = 0;
        b->phase_n[i] = 0;
    }
    return PPPM_OK;
}

int pppm_batch_variants(pppm_batch *b, int *variants)
{
    if (!b || !variants)
        return PPPM_ERR_ARG;
    std::memcpy(variants, b->variant.data(), b->variant.size() * sizeof(int));
    return PPPM_OK;
}

int pppm_batch_set_variants(pppm_batch *b, const int *variants)
{
    if (!b || !variants)
        return PPPM_ERR_ARG;
    for (size_t i = 0; i < b->sc.size(); i++)
        if (variants[i] < 0 || variants[i] >= kFdtdVariants)
            return bfail(b, PPPM_ERR_ARG, "batch_set_variants: variant out of range");
    b->variant.assign(variants, variants + b->sc.size());
    return PPPM_OK;
}

int pppm_batch_autotune(pppm_batch *b, int reps, float *ms_out)
{
    if (!b || reps < 1)
        return PPPM_ERR_ARG;
    BRC(check_scenes(b));
    const int n = (int)b->sc.size();
    for (pppm_solver *s : b->sc)
        BCK(hipStreamSynchronize(s->stream));
    hipEvent_t e0, e1;
    BCK(hipEventCreate(&e0));
    BCK(hipEventCreate(&e1));
    int rc = PPPM_OK;
    for (int c = 0; c < kClasses && rc == PPPM_OK; c++)
    {
        std::vector<int> members;
        for (int i = 0; i < n; i++)
            if (size_class(b->sc[i]->N) == c)
                members.push_back(i);
        if (members.empty())
            continue;
        float best = 1e30f;
        int best_v = b->variant[members[0]];
        for (int v = 0; v < kFdtdVariants && rc == PPPM_OK; v++)
        {
            for (int i : members)
                b->variant[i] = v;
            if ((rc = build_fdtd(b)) != PPPM_OK)
                break;
            // the FDTD launches do not advance t: repeating them rewrites the same values
            for (int w = 0; w < 2 && rc == PPPM_OK; w++)
                rc = enqueue_fdtd(b);
            if (hipEventRecord(e0, b->stream) != hipSuccess)
                rc = bfail(b, PPPM_ERR_HIP, "autotune: event record");
            for (int r = 0; r < reps && rc == PPPM_OK; r++)
                rc = enqueue_fdtd(b);
            if (rc == PPPM_OK && (hipEventRecord(e1, b->stream) != hipSuccess ||
                                  hipEventSynchronize(e1) != hipSuccess))
                rc = bfail(b, PPPM_ERR_HIP, "autotune: event wait");
            float ms = 0;
            if (rc == PPPM_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
            {
                if (ms_out)
                    ms_out[c * kFdtdVariants + v] = ms / reps;
                if (ms < best)
                {
                    best = ms;
                    best_v = v;
                }
            }
        }
        for (int i : members)
            b->variant[i] = best_v;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // extern "C"
