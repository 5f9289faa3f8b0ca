// C ABI of the adjoint (reverse-time transposed step, kernels in adjoint.hip) and of the solver
// state checkpoints that the checkpointed gradient uses. Declarations: include/pppm_hip.h.
//
// The transposed lists are built on the host from the forward's lists once per mesh version
// (the adjoint is defined for a static mesh over the reversed steps): per (triangle, offset) its
// C3 cell, per triangle j the face pairs (c, j) that read it, per C3 cell the triangles whose
// face far field interpolates it, and the Dirichlet divisors A_c / 2 - sum_j D_0(c, j) summed
// in the forward's order (update_dirichlet_kernel, src/algorithm/pppm.cu:172-193).
#include "pppm_hip.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "solver_state.hpp"

using namespace pppm_amd;

namespace
{

#define ACK(expr)                                                                                    \
    do                                                                                               \
    {                                                                                                \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return solver_fail(s, PPPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

#define ARC(expr)         \
    do                    \
    {                     \
        int rc_ = (expr); \
        if (rc_)          \
            return rc_;   \
    } while (0)

// IEEE half -> float (the forward decodes its FP16 weights with the hardware conversion)
float half_to_float(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff, bits;
    if (e == 0)
    {
        if (m == 0)
            bits = sign;
        else
        {
            e = 113;  // normalise the subnormal: value = m 2^-24
            while (!(m & 0x400))
            {
                m <<= 1;
                e--;
            }
            bits = sign | (e << 23) | ((m & 0x3ff) << 13);
        }
    }
    else if (e == 31)
        bits = sign | 0x7f800000u | (m << 13);
    else
        bits = sign | ((e + 112) << 23) | (m << 13);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

template <class T>
int download(pppm_solver *s, std::vector<T> &h, const T *d, size_t n)
{
    h.resize(std::max<size_t>(n, 1));
    if (n)
        ACK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s->stream));
    return PPPM_OK;
}

template <class T>
int upload_vec(pppm_solver *s, DBuf<T> &d, const std::vector<T> &h)
{
    ACK(d.ensure(std::max<size_t>(h.size(), 1)));
    if (!h.empty())
        ACK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s->stream));
    return PPPM_OK;
}

int adjoint_setup(pppm_solver *s)
{
    if (s->adj_version == s->mesh_version)
        return PPPM_OK;
    ARC(solver_read_status(s));
    const int N = s->N, nt = s->nt;
    const size_t NN = (size_t)N * N * N;
    const int n_c3 = s->st_h[ST_C3], P4 = s->st_h[ST_P4], pool = s->st_h[ST_POOL];
    std::vector<int> c3_index, fstart, fnbr, fslot, corner;
    std::vector<TriGeom> geom;
    std::vector<float> interp;
    std::vector<uint32_t> fw;
    std::vector<float2> fsc;
    ARC(download(s, c3_index, s->c3_index.p, NN));
    ARC(download(s, geom, s->geom.p, nt));
    ARC(download(s, fstart, s->face_start.p, (size_t)nt + 1));
    ARC(download(s, fnbr, s->face_nbr.p, P4));
    ARC(download(s, fslot, s->face_slot.p, P4));
    ARC(download(s, corner, s->corner.p, (size_t)8 * nt));
    ARC(download(s, interp, s->interp.p, (size_t)8 * nt));
    ARC(download(s, fw, s->fw.p, (size_t)pool * STEPS));
    ARC(download(s, fsc, s->fsc.p, pool));
    ACK(hipStreamSynchronize(s->stream));
    // (triangle, offset o) -> C3 cell of cell + (o % 3 - 1, o / 3 % 3 - 1, o / 9 - 1)
    std::vector<int> tc3((size_t)nt * GC);
    for (int i = 0; i < nt; i++)
        for (int o = 0; o < GC; o++)
        {
            const int3_ c = geom[i].cell3;
            const int q = c3_index[lin3(c.x + o % 3 - 1, c.y + o / 3 % 3 - 1, c.z + o / 9 - 1, N)];
            if (q < 0)
                return solver_fail(s, PPPM_ERR_STATE, "adjoint: a triangle's 3x3x3 cell is not in the C3 list");
            tc3[(size_t)i * GC + o] = q;
        }
    // face pairs transposed: per neighbour j, the pairs (c, j) in increasing c
    std::vector<int> ft_start(nt + 1, 0), ft_c(std::max(P4, 1)), ft_slot(std::max(P4, 1));
    for (int e = 0; e < P4; e++)
        ft_start[fnbr[e] + 1]++;
    for (int j = 0; j < nt; j++)
        ft_start[j + 1] += ft_start[j];
    {
        std::vector<int> fill(ft_start.begin(), ft_start.end() - 1);
        for (int c = 0; c < nt; c++)
            for (int e = fstart[c]; e < fstart[c + 1]; e++)
            {
                const int at = fill[fnbr[e]]++;
                ft_c[at] = c;
                ft_slot[at] = fslot[e];
            }
    }
    // their face records and scales in the same order, so the transposed gather streams them
    std::vector<uint32_t> tfw((size_t)std::max(P4, 1) * STEPS);
    std::vector<float2> tfsc(std::max(P4, 1));
    for (int e = 0; e < P4; e++)
    {
        const int sl = ft_slot[e];
        if (sl < 0 || sl >= pool)
            return solver_fail(s, PPPM_ERR_STATE, "adjoint: a face pair without a face-cache record");
        std::memcpy(&tfw[(size_t)e * STEPS], &fw[(size_t)sl * STEPS], STEPS * sizeof(uint32_t));
        tfsc[e] = fsc[sl];
    }
    // corners transposed: per C3 cell x, the (triangle, rc x interpolation weight) entries
    std::vector<int> ct_start(n_c3 + 1, 0), ct_c((size_t)8 * nt);
    std::vector<float> ct_w((size_t)8 * nt);
    for (int i = 0; i < 8 * nt; i++)
        if (corner[i] >= 0 && corner[i] < n_c3)  // corners of a base cell are C3 cells
            ct_start[corner[i] + 1]++;
    for (int x = 0; x < n_c3; x++)
        ct_start[x + 1] += ct_start[x];
    {
        std::vector<int> fill(ct_start.begin(), ct_start.end() - 1);
        for (int c = 0; c < nt; c++)
            for (int w = 0; w < 8; w++)
            {
                if (corner[8 * c + w] < 0 || corner[8 * c + w] >= n_c3)
                    continue;
                const int at = fill[corner[8 * c + w]]++;
                ct_c[at] = c;
                ct_w[at] = s->rc * interp[8 * c + w];
            }
    }
    // Dirichlet divisors, the tap-0 factors summed in list order in FP32 as the forward does
    std::vector<double> den(nt);
    for (int c = 0; c < nt; c++)
    {
        float fac = 0.0f;
        for (int e = fstart[c]; e < fstart[c + 1]; e++)
        {
            const int slot = fslot[e];
            fac += half_to_float((uint16_t)(fw[(size_t)slot * STEPS] >> 16)) * fsc[slot].y;
        }
        den[c] = 0.5 * (double)geom[c].area - (double)fac;
    }
    ARC(upload_vec(s, s->adj_tc3, tc3));
    ARC(upload_vec(s, s->adj_ft_start, ft_start));
    ARC(upload_vec(s, s->adj_ft_c, ft_c));
    ARC(upload_vec(s, s->adj_ft_slot, ft_slot));
    ARC(upload_vec(s, s->adj_tfw, tfw));
    ARC(upload_vec(s, s->adj_tfsc, tfsc));
    ARC(upload_vec(s, s->adj_ct_start, ct_start));
    ARC(upload_vec(s, s->adj_ct_c, ct_c));
    ARC(upload_vec(s, s->adj_ct_w, ct_w));
    ARC(upload_vec(s, s->adj_den, den));
    ACK(s->adj_lam3.ensure(3 * NN));
    ACK(s->adj_bb.ensure(NN));
    ACK(s->adj_ln.ensure((size_t)nt * RING));
    ACK(s->adj_ld.ensure((size_t)nt * RING));
    ACK(s->adj_unear.ensure(std::max(n_c3, 1)));
    ACK(s->adj_lfar.ensure(std::max(n_c3, 1)));
    ACK(s->adj_v.ensure(nt));
    ACK(s->adj_state.ensure(1));
    ACK(hipStreamSynchronize(s->stream));
    s->adj_n_c3 = n_c3;
    s->adj_version = s->mesh_version;
    return PPPM_OK;
}

AdjScene adj_scene(pppm_solver *s, const float *gout, long long gout_stride, int n_listen, float *grad,
                   long long grad_stride, int t_first)
{
    AdjScene d{};
    d.fp = solver_fdtd_params(s);
    d.N = s->N;
    d.nt = s->nt;
    d.n_c3 = s->adj_n_c3;
    d.ast = s->adj_state.p;
    d.lam3 = s->adj_lam3.p;
    d.bb = s->adj_bb.p;
    d.ln = s->adj_ln.p;
    d.ld = s->adj_ld.p;
    d.unear = s->adj_unear.p;
    d.lfar = s->adj_lfar.p;
    d.v = s->adj_v.p;
    d.c3_cells = s->c3_cells.p;
    d.tc3 = s->adj_tc3.p;
    d.wf = s->wf.p;
    d.wb = s->wb.p;
    d.fw = s->fw.p;
    d.sf = s->sf.p;
    d.sb = s->sb.p;
    d.fsc = s->fsc.p;
    d.ft_start = s->adj_ft_start.p;
    d.ft_c = s->adj_ft_c.p;
    d.ft_slot = s->adj_ft_slot.p;
    d.tfw = s->adj_tfw.p;
    d.tfsc = s->adj_tfsc.p;
    d.ct_start = s->adj_ct_start.p;
    d.ct_c = s->adj_ct_c.p;
    d.ct_w = s->adj_ct_w.p;
    d.den = s->adj_den.p;
    d.lcell = s->lcell.p;
    d.lc3 = s->lc3.p;
    d.n_listen = n_listen;
    d.gout = gout;
    d.gout_stride = gout_stride;
    d.grad = grad;
    d.grad_stride = grad_stride;
    d.t_first = t_first;
    return d;
}

size_t state_bytes(const pppm_solver *s)
{
    const size_t NN = (size_t)s->N * s->N * s->N;
    return 2 * NN * sizeof(float) + 2 * (size_t)s->nt * RING * sizeof(float) + 16;
}

float *level_ptr(pppm_solver *s, int t, int rel)
{
    const size_t NN = (size_t)s->N * s->N * s->N;
    return s->grid.p + (size_t)((((t + rel) % 3) + 3) % 3) * NN;
}

}  // namespace

namespace pppm_amd
{
AdjScene solver_adj_scene(pppm_solver *s, const float *gout, long long gout_stride, int n_listen, float *grad,
                          long long grad_stride, int t_first)
{
    return adj_scene(s, gout, gout_stride, n_listen, grad, grad_stride, t_first);
}
}  // namespace pppm_amd

namespace
{
// graph-cache key of an adjoint run: the reverse step's arguments with the per-call rows at t = 0
struct AdjGraphKey
{
    AdjScene d;
    long long gbase, rbase;  // byte addresses of the cotangent and gradient rows of t = 0
    int chunk, pad;
};
struct AdjChunk
{
    pppm_solver *s;
    const AdjScene *d;
    int chunk;
};
int enqueue_adjoint_chunk(void *p)
{
    const AdjChunk &c = *static_cast<const AdjChunk *>(p);
    for (int i = 0; i < c.chunk; i++)
        if (enqueue_adjoint_step(*c.d, c.s->stream) != hipSuccess)
            return solver_fail(c.s, PPPM_ERR_HIP, "adjoint graph capture failed");
    return PPPM_OK;
}
}  // namespace

extern "C" {

int pppm_adjoint_reset(pppm_solver *s, int t_last)
{
    if (!s)
        return PPPM_ERR_ARG;
    ARC(solver_check_ready(s, "adjoint_reset"));
    if (s->use_pml)
        return solver_fail(s, PPPM_ERR_ARG, "adjoint: the PML is not differentiated (ABC only)");
    if (t_last < -1)
        return solver_fail(s, PPPM_ERR_ARG, "adjoint_reset: t_last >= -1");
    if ((long long)s->N * s->N * s->N >= (1LL << 31) || s->N < 8)
        return solver_fail(s, PPPM_ERR_ARG, "adjoint: grid edges from 8 and grids below 2^31 cells");
    ARC(adjoint_setup(s));
    const size_t NN = (size_t)s->N * s->N * s->N;
    ACK(hipMemsetAsync(s->adj_lam3.p, 0, 3 * NN * sizeof(float), s->stream));
    ACK(hipMemsetAsync(s->adj_ln.p, 0, (size_t)s->nt * RING * sizeof(float), s->stream));
    ACK(hipMemsetAsync(s->adj_ld.p, 0, (size_t)s->nt * RING * sizeof(float), s->stream));
    AdjState st{t_last, 0};
    ACK(hipMemcpyAsync(s->adj_state.p, &st, sizeof(st), hipMemcpyHostToDevice, s->stream));
    ACK(hipStreamSynchronize(s->stream));
    s->adj_t = t_last;
    return PPPM_OK;
}


int pppm_adjoint_run(pppm_solver *s, const float *gout_dev, int steps, const int *listen_xyz, int n_listen,
                     float *grad_dev, long long grad_stride, int use_graph)
{
    if (!s)
        return PPPM_ERR_ARG;
    ARC(solver_check_ready(s, "adjoint_run"));
    if (steps < 0 || !grad_dev || grad_stride < s->nt || n_listen < 0 || n_listen > 64 ||
        (n_listen > 0 && (!listen_xyz || !gout_dev)))
        return solver_fail(s, PPPM_ERR_ARG, "adjoint_run: bad arguments (at most 64 listeners, grad_stride >= nt)");
    if (s->adj_version != s->mesh_version)
        return solver_fail(s, PPPM_ERR_STATE, "adjoint_run: the mesh changed since adjoint_reset");
    if (steps > s->adj_t + 1)
        return solver_fail(s, PPPM_ERR_ARG, "adjoint_run: more steps than remain before t = 0");
    if (steps == 0)
        return PPPM_OK;
    ARC(solver_set_listeners(s, listen_xyz, n_listen));
    const int t_first = s->adj_t - steps + 1;
    const AdjScene d =
        adj_scene(s, n_listen > 0 ? gout_dev : nullptr, n_listen, n_listen, grad_dev, grad_stride, t_first);
    int done = 0;
    if (use_graph && steps >= 2)
    {
        // a short run (a checkpoint segment) in one graph, a long one in 32-step graphs; the
        // graph is kept for later runs whose arguments agree once the cotangent and gradient
        // rows (read and written only at p + (t - t_first) * row) are shifted to t = 0
        const int chunk = steps <= 64 ? steps : 32;
        AdjGraphKey key;
        std::memset(&key, 0, sizeof key);
        AdjScene dk = d;
        canonical_bytes(dk.fp);
        key.gbase = dk.gout ? (long long)(uintptr_t)dk.gout - (long long)dk.t_first * dk.gout_stride * 4 : 0;
        key.rbase = (long long)(uintptr_t)dk.grad - (long long)dk.t_first * dk.grad_stride * 4;
        dk.gout = dk.gout ? s->adj_lam3.p : nullptr;  // presence only
        dk.grad = s->adj_lam3.p;
        dk.t_first = 0;
        std::memcpy(&key.d, &dk, sizeof dk);
        key.chunk = chunk;
        AdjChunk ac{s, &d, chunk};
        hipGraphExec_t ge = nullptr;
        ARC(solver_cached_graph(s, s->adj_graphs, &key, sizeof key, enqueue_adjoint_chunk, &ac, &ge));
        for (; done + chunk <= steps; done += chunk)
            ACK(hipGraphLaunch(ge, s->stream));
    }
    for (; done < steps; done++)
        ACK(enqueue_adjoint_step(d, s->stream));
    s->adj_t -= steps;
    return PPPM_OK;
}

int pppm_state_bytes(pppm_solver *s, long long *bytes)
{
    if (!s || !bytes)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return solver_fail(s, PPPM_ERR_STATE, "state_bytes before set_mesh");
    *bytes = (long long)state_bytes(s);
    return PPPM_OK;
}

int pppm_save_state(pppm_solver *s, void *dst_dev)
{
    if (!s || !dst_dev)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return solver_fail(s, PPPM_ERR_STATE, "save_state before set_mesh");
    const size_t NN = (size_t)s->N * s->N * s->N, R = (size_t)s->nt * RING;
    unsigned char *p = (unsigned char *)dst_dev;
    const int t = s->t_host;
    ACK(hipMemcpyAsync(p, level_ptr(s, t, 0), NN * 4, hipMemcpyDeviceToDevice, s->stream));
    ACK(hipMemcpyAsync(p + NN * 4, level_ptr(s, t, -1), NN * 4, hipMemcpyDeviceToDevice, s->stream));
    ACK(hipMemcpyAsync(p + 2 * NN * 4, s->hn.p, R * 4, hipMemcpyDeviceToDevice, s->stream));
    ACK(hipMemcpyAsync(p + 2 * NN * 4 + R * 4, s->hd.p, R * 4, hipMemcpyDeviceToDevice, s->stream));
    ACK(launch_put_ints(reinterpret_cast<int *>(p + 2 * NN * 4 + 2 * R * 4), 4, t, s->N, s->nt, 0x5157, s->stream));
    return PPPM_OK;
}

int pppm_load_state(pppm_solver *s, const void *src_dev)
{
    if (!s || !src_dev)
        return PPPM_ERR_ARG;
    if (!s->mesh_set)
        return solver_fail(s, PPPM_ERR_STATE, "load_state before set_mesh");
    const size_t NN = (size_t)s->N * s->N * s->N, R = (size_t)s->nt * RING;
    const unsigned char *p = (const unsigned char *)src_dev;
    int trailer[4];
    ACK(hipMemcpyAsync(trailer, p + 2 * NN * 4 + 2 * R * 4, sizeof(trailer), hipMemcpyDeviceToHost, s->stream));
    ACK(hipStreamSynchronize(s->stream));
    if (trailer[3] != 0x5157 || trailer[1] != s->N || trailer[2] != s->nt || trailer[0] < -1)
        return solver_fail(s, PPPM_ERR_ARG, "load_state: not a state of this solver's grid and mesh");
    const int t = trailer[0];
    ACK(hipMemcpyAsync(level_ptr(s, t, 0), p, NN * 4, hipMemcpyDeviceToDevice, s->stream));
    ACK(hipMemcpyAsync(level_ptr(s, t, -1), p + NN * 4, NN * 4, hipMemcpyDeviceToDevice, s->stream));
    ACK(hipMemcpyAsync(s->hn.p, p + 2 * NN * 4, R * 4, hipMemcpyDeviceToDevice, s->stream));
    ACK(hipMemcpyAsync(s->hd.p, p + 2 * NN * 4 + R * 4, R * 4, hipMemcpyDeviceToDevice, s->stream));
    return pppm_set_time_idx(s, t);
}

}  // extern "C"
