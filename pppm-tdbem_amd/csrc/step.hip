// Per-step near-field kernels: far-field extraction, Dirichlet update, near-field
// correction. Sparse and HBM-bound: no MFMA.
//
// Reference behaviour: solve_fdtd_far_kernel (src/algorithm/pppm.cu:8-51),
// set_neumann_condition_kernel (pppm.cu:206-225), solve_face_far_kernel,
// solve_face_near_kernel and update_dirichlet_kernel (pppm.cu:116-204),
// solve_fdtd_near_kernel (pppm.cu:53-114), LayerWeightHalf::convolution (bem.h:203-213).
//
// Segmented gather with a fixed summation order. A "segment" is a C3 cell (far / near)
// or a triangle (Dirichlet); its pairs are contiguous in a CSR list. Wave w owns the
// segments whose first pair lies in [64w, 64w+64) — at most 64, one leader lane each;
// the owned range comes from a wave map built with the lists, which also holds the first
// pair of that range, so a wave starts with two 8-byte loads. The wave walks its pair
// range in rounds of 64: lane i evaluates the 16-tap convolution of pair (round base + i)
// sequentially (k = 0..15, as the reference's loop), parks the value in LDS, and every
// leader adds the values of its own pairs in list order. Every sum therefore has the
// order of a sequential loop, whatever the segment lengths, and lanes stay busy across
// segment boundaries.
//
// Weights: 64 B record per pair (16 x half2 {S_k, D_k}) + float2 scales. History: the
// reference's 32-slot rings (slot = t mod 32), one 128 B line per triangle and field, plus
// per-step windows of the 16 latest values in time order (one 64 B row per triangle and
// field) from which the gathers read taps 1..15 as four 16 B loads.
//
// Time index: all kernels read t from device memory (StepState), so one captured HIP
// graph replays any step.
#include "internal.hpp"

namespace pppm_amd
{

namespace
{
constexpr int WPB = 4;            // waves per block
constexpr int TB = WPB * WAVE;    // threads per block

// Workgroups are dealt to the 8 XCDs in turn (block b to XCD b % 8). A gather launch covers
// 8 * per blocks and block b takes virtual block (b % 8) * per + b / 8, so each XCD works on
// one contiguous eighth of the wave range: neighbouring triangles and cells share history
// windows and grid values in that XCD's L2. Waves past a kernel's work do nothing.
__device__ inline int xcd_wave(int per, int wid)
{
    return ((int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3)) * WPB + wid;
}

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc plus the values vals[lo], ..., vals[hi - 1] of one round (64 values parked by the wave's
// lanes, 16-byte aligned), added one at a time in that order: the sequential loop's sum. Every
// lane sweeps all 64 positions with 16-byte LDS reads (the same addresses in every lane, so
// broadcasts) and keeps acc where a position is outside its range, so a round costs four
// waits on LDS instead of one per value, and lanes with several short ranges (or none) run
// in step. Positions outside [lo, hi) leave acc bit for bit as it was (a select, not + 0).
__device__ __forceinline__ float round_sum(float acc, const float *__restrict__ vals, int lo, int hi)
{
    if (lo <= 0 && hi >= WAVE)
    {
        // the whole round belongs to this lane's segment (the usual case of the long face lists)
#pragma unroll
        for (int c = 0; c < WAVE; c += 16)
        {
            float v[16];
#pragma unroll
            for (int i = 0; i < 4; i++)
                *reinterpret_cast<float4 *>(&v[4 * i]) = reinterpret_cast<const float4 *>(vals + c)[i];
#pragma unroll
            for (int i = 0; i < 16; i++)
                acc += v[i];
        }
        return acc;
    }
    if (lo >= hi)
        return acc;
#pragma unroll
    for (int c = 0; c < WAVE; c += 16)
    {
        float v[16];
#pragma unroll
        for (int i = 0; i < 4; i++)
            *reinterpret_cast<float4 *>(&v[4 * i]) = reinterpret_cast<const float4 *>(vals + c)[i];
#pragma unroll
        for (int i = 0; i < 16; i++)
        {
            const float sum = acc + v[i];
            acc = (c + i >= lo && c + i < hi) ? sum : acc;
        }
    }
    return acc;
}

// LayerWeightHalf::convolution (bem.h:203-213) for one pair, taps K0..15 in order, from the
// pair triangle's history windows (four 16 B loads per field): rows 0..14 hold the values at
// t-1 .. t-15 (taps 1..15), row 15 the value at t (tap 0; written into the window by the kernel
// that produces it, see windows_item). DZERO0: the Dirichlet value of step t is read as 0
// (pppm.cu:143).
template <int K0, bool DZERO0>
__device__ inline float conv_pair(const uint32_t *__restrict__ w16, float2 sc, const float *__restrict__ wn,
                                  const float *__restrict__ wd, uint32_t *tap0 = nullptr)
{
    uint4 q[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        q[i] = reinterpret_cast<const uint4 *>(w16)[i];
    float4 hn4[4], hd4[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        hn4[i] = reinterpret_cast<const float4 *>(wn)[i];
        hd4[i] = reinterpret_cast<const float4 *>(wd)[i];
    }
    const uint32_t *w = reinterpret_cast<const uint32_t *>(q);
    const float *hnw = reinterpret_cast<const float *>(hn4);
    const float *hdw = reinterpret_cast<const float *>(hd4);
    if (tap0)
        *tap0 = w[0];
    float r = 0.0f;
#pragma unroll
    for (int k = K0; k < STEPS; k++)
    {
        const float ws = f16_bits_to_f32(w[k]);
        const float wdk = f16_bits_to_f32(w[k] >> 16);
        const float n = k == 0 ? hnw[STEPS - 1] : hnw[k - 1];
        const float d = k == 0 ? (DZERO0 ? 0.0f : hdw[STEPS - 1]) : hdw[k - 1];
        const float a = -ws * sc.x * n;
        const float b = wdk * sc.y * d;
        r += a + b;
    }
    return r;
}

// History windows for the gathers of step t: row k < 15 of triangle i holds the ring value at
// t - 1 - k, so a gather reads tap k >= 1 at row k - 1; row 15 holds the value at t (tap 0).
// back = 0 when launched before t is advanced (the time step: the values at t do not exist yet;
// the far-field launch writes the Neumann value and the Dirichlet update the Dirichlet value
// into row 15 as into the ring), 1 after it (the phase entry points: the ring holds them).
__device__ __forceinline__ void windows_item(const SceneStep &d, int back, int i)
{
    if (i >= d.a.nt * STEPS || (d.err && *d.err))
        return;
    const int tri = i / STEPS, k = i % STEPS;
    const int tn = d.st->t - back + 1;  // t of the step the gathers serve
    const int slot = (k == STEPS - 1 ? tn : tn - 1 - k) & (RING - 1);
    d.wn[i] = d.hn[tri * RING + slot];
    d.wd[i] = d.hd[tri * RING + slot];
}

// far(c) = p(c) - sum_pairs W^fdtd (*) (N, D). W^fdtd has tap 0 == 0 (pppm_cache.cu:111),
// so taps 1..15 only (adding the zero tap leaves the sum bit-identical); the Neumann slot
// of step t is therefore free to be written by the extra waves of this launch.
__device__ __forceinline__ void far_wave(const SceneStep &d, int wave, int lane, float *__restrict__ vals)
{
    // the wave map is read before the error word is waited on (in bounds either way), so the
    // two loads take one memory round trip
    int2 w0 = {0, 0}, w1 = {0, 0};
    if (wave < d.c3_waves)
    {
        w0 = reinterpret_cast<const int2 *>(d.c3_wave)[wave];
        w1 = reinterpret_cast<const int2 *>(d.c3_wave)[wave + 1];
    }
    if (d.err && *d.err)  // behind a mesh update that failed: the step is run again after the update is
        return;
    const int t = d.st->t;
    if (wave < d.c3_waves)
    {
        const int c0 = w0.x, q0 = w0.y, c1 = w1.x, q1 = w1.y;
        if (c0 >= c1)
            return;
        const size_t nn = (size_t)d.a.N * d.a.N * d.a.N;
        const float *g = d.grid3 + (size_t)(((t % 3) + 3) % 3) * nn;
        const int cs = c0 + lane;
        const bool leader = cs < c1;
        // every index of the wave's first round is loaded at once, then what they point to
        const int m0 = leader ? d.c3_start[cs] : 0, m1 = leader ? d.c3_start[cs + 1] : 0;
        const unsigned cell = leader ? (unsigned)d.c3_cells[cs] : ~0u;
        int rec = q0 + lane < q1 ? d.c3_pairs[q0 + lane] : -1;
        const float gv = cell < nn ? g[cell] : 0.0f;
        float acc = 0.0f;
        for (int base = q0; base < q1; base += WAVE)
        {
            const int p = base + lane;
            // the next round's index is loaded now, beside this round's records
            const int rec_next = p + WAVE < q1 ? d.c3_pairs[p + WAVE] : -1;
            float v = 0.0f;
            if (p < q1)  // rec >= 0 exactly then; the list-ordered record does not wait for rec
            {
                const int tri = rec / GC;
                const uint32_t *w = d.owf ? d.owf + (size_t)p * STEPS : d.wf + (size_t)rec * STEPS;
                const float2 sc = d.owf ? d.osf[p] : d.sf[rec];
                v = conv_pair<1, false>(w, sc, d.wn + tri * STEPS, d.wd + tri * STEPS);
            }
            vals[lane] = v;
            wave_sync();
            acc = round_sum(acc, vals, max(m0, base) - base, min(m1, base + WAVE) - base);
            wave_sync();
            rec = rec_next;
        }
        if (leader)
            d.far[cs] = cell < nn ? gv - acc : 0.0f;
    }
    else
    {
        // Neumann input of step t into the ring (set_neumann_condition_kernel)
        const int i = (wave - d.c3_waves) * WAVE + lane;
        if (d.neumann && i < d.a.nt)
        {
            const float v = d.neumann[(size_t)(t - d.a.t0) * d.a.stride + i];
            d.hn[i * RING + (t & (RING - 1))] = v;
            d.wn[i * STEPS + STEPS - 1] = v;  // tap 0 of the Dirichlet update and the near field
        }
    }
}

// Face far field (8-corner interpolation, pppm.cu:116-144), face near field against the
// 4x4x4 list with D[t] = 0 and its tap-0 factors (pppm.cu:146-170), then
// D_c[t] = (far*rc + near) / (0.5*area - sum_j D_0(c, j)) (pppm.cu:172-193).
__device__ __forceinline__ void dirichlet_wave(const SceneStep &d, int wave, int lane, float *__restrict__ vals,
                                               float *__restrict__ facs)
{
    const int t = d.st->t;
    if (wave >= d.face_waves)
        return;
    // the wave map is read before the error word is waited on (one memory round trip)
    const int2 w0 = reinterpret_cast<const int2 *>(d.face_wave)[wave];
    const int2 w1 = reinterpret_cast<const int2 *>(d.face_wave)[wave + 1];
    // after a flagged mesh update the face lists may be incomplete: leave D[t] untouched
    if (*d.err != 0)
        return;
    const int c0 = w0.x, q0 = w0.y, c1 = w1.x, q1 = w1.y;
    if (c0 >= c1)
        return;
    const int cs = c0 + lane;
    const bool leader = cs < c1;
    // every index of the wave's first round is loaded at once, then what they point to; the
    // leaders' face far field (8-corner interpolation) does not depend on the gather
    const int m0 = leader ? d.face_start[cs] : 0, m1 = leader ? d.face_start[cs + 1] : 0;
    int j = -1, ws = 0;
    if (q0 + lane < q1)
    {
        j = d.face_nbr[q0 + lane];
        ws = d.face_slot[q0 + lane];
    }
    float ff = 0.0f, area = 0.0f;
    if (leader)
    {
        int ci[8];
        float iw[8];
#pragma unroll
        for (int w = 0; w < 8; w++)
        {
            ci[w] = d.corner[8 * cs + w];  // corners of a base cell are C3 cells (>= 0)
            iw[w] = d.interp[8 * cs + w];
        }
        area = d.geom[cs].area;
#pragma unroll
        for (int w = 0; w < 8; w++)
            ff += iw[w] * (ci[w] >= 0 ? d.far[ci[w]] : 0.0f);
    }
    float acc = 0.0f, fac = 0.0f;
    for (int base = q0; base < q1; base += WAVE)
    {
        const int p = base + lane;
        // the next round's indices are loaded now, beside this round's records
        int j_next = -1, ws_next = 0;
        if (p + WAVE < q1)
        {
            j_next = d.face_nbr[p + WAVE];
            ws_next = d.face_slot[p + WAVE];
        }
        float v = 0.0f, f = 0.0f;
        if (p < q1)  // j >= 0 exactly then; the list-ordered record does not wait for the slot
        {
            const float2 sc = d.ofw ? d.ofsc[p] : d.fsc[ws];
            const uint32_t *w = d.ofw ? d.ofw + (size_t)p * STEPS : d.fw + (size_t)ws * STEPS;
            uint32_t tap0;
            v = conv_pair<0, true>(w, sc, d.wn + j * STEPS, d.wd + j * STEPS, &tap0);
            f = f16_bits_to_f32(tap0 >> 16) * sc.y;  // face_factor = (float)D_0 * scale_D
        }
        vals[lane] = v;
        facs[lane] = f;
        wave_sync();
        {
            const int lo = max(m0, base) - base, hi = min(m1, base + WAVE) - base;
            acc = round_sum(acc, vals, lo, hi);
            fac = round_sum(fac, facs, lo, hi);
        }
        wave_sync();
        j = j_next;
        ws = ws_next;
    }
    if (leader)
    {
        const float num = ff * d.a.rc + acc;
        const double den = 0.5 * (double)area - (double)fac;
        const float dv = (float)((double)num / den);
        d.hd[cs * RING + (t & (RING - 1))] = dv;
        d.wd[cs * STEPS + STEPS - 1] = dv;  // tap 0 of the near field
    }
}

// p(c) = far + 0.98 * sum W^bem (*) (N, D) over all 16 taps; records listener samples.
__device__ __forceinline__ void near_wave(const SceneStep &d, int wave, int lane, float *__restrict__ vals)
{
    // the wave map is read before the error word is waited on (in bounds either way), and lane l
    // reads listener l's C3 index: the table set with the listeners, or, for lists that move with
    // the mesh (lc3 null), this update's map of the listener's cell
    int2 w0 = {0, 0}, w1 = {0, 0};
    if (wave < d.c3_waves)
    {
        w0 = reinterpret_cast<const int2 *>(d.c3_wave)[wave];
        w1 = reinterpret_cast<const int2 *>(d.c3_wave)[wave + 1];
    }
    int lc = -1;
    if (d.out && lane < d.a.n_listen)
        lc = d.lc3 ? d.lc3[lane] : d.c3_index[d.lcell[lane]];
    if (d.err && *d.err)
        return;
    const int t = d.st->t;
    const size_t nn = (size_t)d.a.N * d.a.N * d.a.N;
    float *g = d.grid3 + (size_t)(((t % 3) + 3) % 3) * nn;
    if (wave < d.c3_waves)
    {
        const int c0 = w0.x, q0 = w0.y, c1 = w1.x, q1 = w1.y;
        if (c0 >= c1)
            return;
        const int cs = c0 + lane;
        const bool leader = cs < c1;
        // every index of the wave's first round is loaded at once, then what they point to
        const int m0 = leader ? d.c3_start[cs] : 0, m1 = leader ? d.c3_start[cs + 1] : 0;
        const unsigned cell = leader ? (unsigned)d.c3_cells[cs] : ~0u;
        const float fv = leader ? d.far[cs] : 0.0f;
        int rec = q0 + lane < q1 ? d.c3_pairs[q0 + lane] : -1;
        float acc = 0.0f;
        for (int base = q0; base < q1; base += WAVE)
        {
            const int p = base + lane;
            // the next round's index is loaded now, beside this round's records
            const int rec_next = p + WAVE < q1 ? d.c3_pairs[p + WAVE] : -1;
            float v = 0.0f;
            if (p < q1)  // rec >= 0 exactly then; the list-ordered record does not wait for rec
            {
                const int tri = rec / GC;
                const uint32_t *w = d.owb ? d.owb + (size_t)p * STEPS : d.wb + (size_t)rec * STEPS;
                const float2 sc = d.owb ? d.osb[p] : d.sb[rec];
                v = conv_pair<0, false>(w, sc, d.wn + tri * STEPS, d.wd + tri * STEPS);
            }
            vals[lane] = v;
            wave_sync();
            acc = round_sum(acc, vals, max(m0, base) - base, min(m1, base + WAVE) - base);
            wave_sync();
            rec = rec_next;
        }
        const float v = (float)((double)fv + (double)acc * 0.98);
        if (leader && cell < nn)
            g[cell] = v;
        if (d.out)
            for (int l = 0; l < d.a.n_listen; l++)
            {
                const int c3l = __shfl(lc, l);  // every lane takes part
                if (leader && c3l == cs)
                    d.out[(size_t)(t - d.a.t0) * d.out_stride + l] = v;
            }
    }
    else if (d.out && wave == d.c3_waves)
    {
        // listeners outside the corrected region read the grid directly (no wave writes them)
        if (lane < d.a.n_listen && lc < 0)
            d.out[(size_t)(t - d.a.t0) * d.out_stride + lane] = g[d.lcell[lane]];
    }
}

// ---------------------------------------------------------------- single-scene kernels
__global__ __launch_bounds__(256) void k_windows(SceneStep d, int back)
{
    windows_item(d, back, blockIdx.x * 256 + threadIdx.x);
}

__global__ __launch_bounds__(TB) void k_far(SceneStep d, int xper)
{
    __shared__ __align__(16) float vals[WPB][WAVE];
    const int wid = threadIdx.x >> 6;
    far_wave(d, xcd_wave(xper, wid), threadIdx.x & 63, vals[wid]);
}

__global__ __launch_bounds__(TB) void k_dirichlet(SceneStep d)
{
    __shared__ __align__(16) float vals[WPB][WAVE];
    __shared__ __align__(16) float facs[WPB][WAVE];
    const int wid = threadIdx.x >> 6;
    dirichlet_wave(d, blockIdx.x * WPB + wid, threadIdx.x & 63, vals[wid], facs[wid]);
}

__global__ __launch_bounds__(TB) void k_near(SceneStep d, int xper)
{
    __shared__ __align__(16) float vals[WPB][WAVE];
    const int wid = threadIdx.x >> 6;
    near_wave(d, xcd_wave(xper, wid), threadIdx.x & 63, vals[wid]);
}

__global__ void k_listen_direct(SceneStep d)
{
    const int t = d.st->t;
    const int l = threadIdx.x;
    if (l >= d.a.n_listen || (d.err && *d.err))
        return;
    const size_t nn = (size_t)d.a.N * d.a.N * d.a.N;
    const float *g = d.grid3 + (size_t)(((t % 3) + 3) % 3) * nn;
    d.out[(size_t)(t - d.a.t0) * d.out_stride + l] = g[d.lcell[l]];
}

__global__ __launch_bounds__(256) void k_order_records(const int *__restrict__ c3_pairs, int n3,
                                                       const uint4 *__restrict__ wf, const float2 *__restrict__ sf,
                                                       const uint4 *__restrict__ wb, const float2 *__restrict__ sb,
                                                       uint4 *__restrict__ owf, float2 *__restrict__ osf,
                                                       uint4 *__restrict__ owb, float2 *__restrict__ osb,
                                                       const int *__restrict__ face_slot, int n4, int pool_cap,
                                                       const uint4 *__restrict__ fw, const float2 *__restrict__ fsc,
                                                       uint4 *__restrict__ ofw, float2 *__restrict__ ofsc)
{
    // one thread per 16-byte quarter of a record: consecutive threads write consecutive bytes
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < 4LL * n3)
    {
        const int pr = (int)(i >> 2), q = (int)(i & 3);
        const int r = c3_pairs[pr];
        const bool ok = (unsigned)r < (unsigned)n3;  // the C3 pairs are a permutation of [0, n3)
        owf[4 * (size_t)pr + q] = ok ? wf[4 * (size_t)r + q] : make_uint4(0, 0, 0, 0);
        owb[4 * (size_t)pr + q] = ok ? wb[4 * (size_t)r + q] : make_uint4(0, 0, 0, 0);
        if (q == 0)
        {
            osf[pr] = ok ? sf[r] : make_float2(0, 0);
            osb[pr] = ok ? sb[r] : make_float2(0, 0);
        }
        return;
    }
    i -= 4LL * n3;
    if (i < 4LL * n4)
    {
        const int pr = (int)(i >> 2), q = (int)(i & 3);
        const int sl = face_slot[pr];
        const bool ok = (unsigned)sl < (unsigned)pool_cap;
        ofw[4 * (size_t)pr + q] = ok ? fw[4 * (size_t)sl + q] : make_uint4(0, 0, 0, 0);
        if (q == 0)
            ofsc[pr] = ok ? fsc[sl] : make_float2(0, 0);
    }
}

__global__ void k_listen_setup(ListenTable lt, int N, const int *__restrict__ c3_index, int *__restrict__ cell,
                               int *__restrict__ c3)
{
    const int l = threadIdx.x;
    if (l >= lt.n)
        return;
    const int q = lin3(lt.xyz[3 * l], lt.xyz[3 * l + 1], lt.xyz[3 * l + 2], N);
    cell[l] = q;
    c3[l] = c3_index[q];
}

__global__ void k_set_neumann(SceneStep d)
{
    const int t = d.st->t;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.a.nt)
        d.hn[i * RING + (t & (RING - 1))] = d.neumann[(size_t)(t - d.a.t0) * d.a.stride + i];
}

// ---------------------------------------------------------------- batched kernels
// A batched launch covers the blocks of every scene: virtual block v belongs to scene
// m.scene_of[v] as its local block v - m.first[scene]. Blocks go to the 8 XCDs in turn and
// block b takes virtual block (b % 8) * m.per + b / 8, so each XCD works through one
// contiguous run of scenes (their lists, windows and grids stay in that XCD's L2).
__device__ __forceinline__ bool batch_block(const BatchMap &m, int &scene, int &local)
{
    const int v = (int)(blockIdx.x & 7) * m.per + (int)(blockIdx.x >> 3);
    if (v >= m.total)
        return false;
    scene = m.scene_of[v];
    local = v - m.first[scene];
    return true;
}

__global__ __launch_bounds__(256) void k_batch_windows(const SceneStep *__restrict__ ds, BatchMap m)
{
    int sc, lb;
    if (batch_block(m, sc, lb))
        windows_item(ds[sc], 0, lb * 256 + threadIdx.x);
}

__global__ __launch_bounds__(TB) void k_batch_far(const SceneStep *__restrict__ ds, BatchMap m)
{
    __shared__ __align__(16) float vals[WPB][WAVE];
    int sc, lb;
    if (!batch_block(m, sc, lb))
        return;
    const int wid = threadIdx.x >> 6;
    far_wave(ds[sc], lb * WPB + wid, threadIdx.x & 63, vals[wid]);
}

__global__ __launch_bounds__(TB) void k_batch_dirichlet(const SceneStep *__restrict__ ds, BatchMap m)
{
    __shared__ __align__(16) float vals[WPB][WAVE];
    __shared__ __align__(16) float facs[WPB][WAVE];
    int sc, lb;
    if (!batch_block(m, sc, lb))
        return;
    const int wid = threadIdx.x >> 6;
    dirichlet_wave(ds[sc], lb * WPB + wid, threadIdx.x & 63, vals[wid], facs[wid]);
}

__global__ __launch_bounds__(TB) void k_batch_near(const SceneStep *__restrict__ ds, BatchMap m)
{
    __shared__ __align__(16) float vals[WPB][WAVE];
    int sc, lb;
    if (!batch_block(m, sc, lb))
        return;
    const int wid = threadIdx.x >> 6;
    near_wave(ds[sc], lb * WPB + wid, threadIdx.x & 63, vals[wid]);
}

__global__ void k_batch_advance(const SceneStep *__restrict__ ds, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        ds[i].st->t = ds[i].st->t + 1;
}

inline int blocks_for_waves(long long waves) { return waves > 0 ? (int)((waves + WPB - 1) / WPB) : 1; }

// blocks per XCD of a gather launch over `waves` waves (the launch has 8 times as many)
inline int xcd_per(long long waves) { return (blocks_for_waves(waves) + 7) / 8; }

long long far_waves(const SceneStep &d) { return (long long)d.c3_waves + (d.neumann ? (d.a.nt + WAVE - 1) / WAVE : 0); }
long long near_waves(const SceneStep &d) { return (long long)d.c3_waves + (d.out ? 1 : 0); }
}  // namespace

int batch_blocks(const SceneStep &d, BatchKernel k)
{
    switch (k)
    {
    case BatchKernel::Windows:
    {
        const long long n = (long long)d.a.nt * STEPS;
        return n > 0 ? (int)((n + 255) / 256) : 1;
    }
    case BatchKernel::Far:
        return blocks_for_waves(far_waves(d));
    case BatchKernel::Dirichlet:
        return blocks_for_waves(d.face_waves);
    case BatchKernel::Near:
        return blocks_for_waves(near_waves(d));
    }
    return 1;
}

hipError_t launch_windows(const SceneStep &d, int back, hipStream_t s)
{
    const long long n = (long long)d.a.nt * STEPS;
    k_windows<<<n > 0 ? (int)((n + 255) / 256) : 1, 256, 0, s>>>(d, back);
    return hipGetLastError();
}

hipError_t launch_far(const SceneStep &d, hipStream_t s)
{
    const int per = xcd_per(far_waves(d));
    k_far<<<8 * per, TB, 0, s>>>(d, per);
    return hipGetLastError();
}

hipError_t launch_dirichlet(const SceneStep &d, hipStream_t s)
{
    // plain block order: the heavy face lists of finely meshed objects sit together in the wave
    // range, and round-robin spreads them over the XCDs (the contiguous mapping of the C3 gathers
    // cut this kernel's HBM traffic from 47.6 to 39.7 MB but raised its time from 19.1 to 21.6 us)
    k_dirichlet<<<blocks_for_waves(d.face_waves), TB, 0, s>>>(d);
    return hipGetLastError();
}

hipError_t launch_near(const SceneStep &d, hipStream_t s)
{
    const int per = xcd_per(near_waves(d));
    k_near<<<8 * per, TB, 0, s>>>(d, per);
    return hipGetLastError();
}

hipError_t launch_listen_direct(const SceneStep &d, hipStream_t s)
{
    k_listen_direct<<<1, 64, 0, s>>>(d);
    return hipGetLastError();
}

hipError_t launch_listen_setup(const ListenTable &lt, int N, const int *c3_index, int *cell, int *c3,
                               hipStream_t s)
{
    k_listen_setup<<<1, 64, 0, s>>>(lt, N, c3_index, cell, c3);
    return hipGetLastError();
}

hipError_t launch_order_records(const int *c3_pairs, int n3, const uint32_t *wf, const float2 *sf, const uint32_t *wb,
                                const float2 *sb, uint32_t *owf, float2 *osf, uint32_t *owb, float2 *osb,
                                const int *face_slot, int n4, int pool_cap, const uint32_t *fw, const float2 *fsc,
                                uint32_t *ofw, float2 *ofsc, hipStream_t s)
{
    const long long n = 4LL * n3 + 4LL * n4;
    if (n <= 0)
        return hipSuccess;
    k_order_records<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
        c3_pairs, n3, (const uint4 *)wf, sf, (const uint4 *)wb, sb, (uint4 *)owf, osf, (uint4 *)owb, osb, face_slot, n4,
        pool_cap, (const uint4 *)fw, fsc, (uint4 *)ofw, ofsc);
    return hipGetLastError();
}

hipError_t launch_set_neumann(const SceneStep &d, hipStream_t s)
{
    k_set_neumann<<<(d.a.nt + 255) / 256 > 0 ? (d.a.nt + 255) / 256 : 1, 256, 0, s>>>(d);
    return hipGetLastError();
}

hipError_t launch_batch(BatchKernel k, const SceneStep *descs, const BatchMap &m, hipStream_t s)
{
    const int grid = 8 * m.per;
    switch (k)
    {
    case BatchKernel::Windows:
        k_batch_windows<<<grid, 256, 0, s>>>(descs, m);
        break;
    case BatchKernel::Far:
        k_batch_far<<<grid, TB, 0, s>>>(descs, m);
        break;
    case BatchKernel::Dirichlet:
        k_batch_dirichlet<<<grid, TB, 0, s>>>(descs, m);
        break;
    case BatchKernel::Near:
        k_batch_near<<<grid, TB, 0, s>>>(descs, m);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_batch_advance(const SceneStep *descs, int n, hipStream_t s)
{
    k_batch_advance<<<(n + 255) / 256, 256, 0, s>>>(descs, n);
    return hipGetLastError();
}

}  // namespace pppm_amd
