// Modal sound source of a rigid body on the GPU (the C ABI's pppm_modal_*).
//
// Reference behaviour: ModalInfo::SetCoeffs (src/utils/RigidBody.cu:11-27), MatrixVertToTri
// (RigidBody.cu:206-230), the IIR recursion of CalculateIIR_ (RigidBody.cu:282-291) and
// update_surf_acc_kernel (RigidBody.cu:240-250). The reference steps the oscillators on the
// host one sample at a time and launches one surface-acceleration kernel per sample; here a
// call covers a block of samples: one thread per mode runs its oscillator recursion through the
// block, one thread per (mode, sample) forms the modal accelerations Q from the displacements,
// then one thread per (triangle, 32 samples) forms the surface accelerations, reading each surface-matrix
// entry once for the 32 samples. Every value keeps the reference's FP32 operations and order
// (host-side arithmetic of the oscillators uncontracted, the device sum as fused multiply-adds).
//
// Layout: the surface matrix is stored mode-major, Mt[i][t], so the accelerations stream it with
// coalesced loads; Q (modal accelerations) is mode-major too, Q[i][s] with a padded row stride,
// so a block reads the 32 values of its sample block as one contiguous, block-uniform run.
#include "internal.hpp"

#include <cmath>
#include <string>
#include <vector>

#include "pppm_hip.h"

namespace pppm_amd
{

namespace
{
constexpr int MB = 256;
constexpr int SPT = 32;  // samples per thread (acceleration kernel) and per prefetch (oscillators)

// MatrixVertToTri: M(t, i) = dot((u(v0) + u(v1) + u(v2)) / 3, n_t), u(v) = rows 3v..3v+2 of
// column i of the eigenvector matrix; written mode-major
__global__ __launch_bounds__(MB) void k_modal_surface(const float *__restrict__ eig, int nmodes,
                                                      const int *__restrict__ tris, const float *__restrict__ normals,
                                                      int nt, float *__restrict__ mt)
{
    const long long e = (long long)blockIdx.x * MB + threadIdx.x;
    if (e >= (long long)nt * nmodes)
        return;
    const int t = (int)(e % nt), i = (int)(e / nt);
    const int v[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
    float sx = eig[(size_t)(3 * v[0]) * nmodes + i] + eig[(size_t)(3 * v[1]) * nmodes + i];
    float sy = eig[(size_t)(3 * v[0] + 1) * nmodes + i] + eig[(size_t)(3 * v[1] + 1) * nmodes + i];
    float sz = eig[(size_t)(3 * v[0] + 2) * nmodes + i] + eig[(size_t)(3 * v[1] + 2) * nmodes + i];
    sx = (sx + eig[(size_t)(3 * v[2]) * nmodes + i]) / 3.0f;
    sy = (sy + eig[(size_t)(3 * v[2] + 1) * nmodes + i]) / 3.0f;
    sz = (sz + eig[(size_t)(3 * v[2] + 2) * nmodes + i]) / 3.0f;
    mt[(size_t)i * nt + t] = sx * normals[3 * t] + sy * normals[3 * t + 1] + sz * normals[3 * t + 2];
}

// One thread per mode: q0 = c1 q1 + c2 q2 + c3 f; (q3, q2, q1) <- (q2, q1, q0), for each sample
// of the block in order. Only the recursion is sequential, so this kernel keeps to it: a block
// holds 64 modes in two waves, a stepping wave and a loading wave. While the stepping wave runs
// window w of FD samples from LDS, the loading wave brings the forces of window w + 1 into the
// other LDS buffer, so the recursion never waits on memory. The displacements q0 go to the
// mode's row of qp four at a time. Row layout (stride ldp, a multiple of 4): [2] = q2 and
// [3] = q1 of the state at the start of the call, [4 + s] = q0 of sample s; entries of the last
// window past ns are written 0.
constexpr int FD = 64;

template <bool GUARD>
__device__ __forceinline__ void iir_window(const float (*__restrict__ sf)[64], int lane, int s0, int ns, float c1,
                                           float c2, float c3, float &q1, float &q2, float &q3,
                                           float *__restrict__ row)
{
#pragma unroll
    for (int k4 = 0; k4 < FD; k4 += 4)
    {
        float out[4];
#pragma unroll
        for (int k = k4; k < k4 + 4; k++)
        {
            const float f = sf[k][lane];
            out[k - k4] = 0.0f;
            if (!GUARD || s0 + k < ns)
            {
                const float q0 = c1 * q1 + c2 * q2 + c3 * f;
                q3 = q2;
                q2 = q1;
                q1 = q0;
                out[k - k4] = q0;
            }
        }
        if (!GUARD || s0 + k4 < ns)
            *(float4 *)(row + 4 + s0 + k4) = make_float4(out[0], out[1], out[2], out[3]);
    }
}

__global__ __launch_bounds__(128) void k_modal_iir(const float *__restrict__ coeff, float *__restrict__ state,
                                                   int nmodes, const float *__restrict__ forces, int ns,
                                                   float *__restrict__ qp, int ldp)
{
    __shared__ float sf[2][FD][64];
    const int lane = threadIdx.x & 63;
    const bool loader = threadIdx.x >= 64;
    const int i = blockIdx.x * 64 + lane;
    const bool live = i < nmodes;
    const int nw = (ns + FD - 1) / FD;
    auto load = [&](int w) {
        float v[FD];
#pragma unroll
        for (int k = 0; k < FD; k++)
        {
            const int s = w * FD + k;
            v[k] = (live && s < ns) ? forces[(size_t)s * nmodes + i] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < FD; k++)
            sf[w & 1][k][lane] = v[k];
    };
    float c1 = 0, c2 = 0, c3 = 0, q1 = 0, q2 = 0, q3 = 0;
    float *row = qp + (size_t)(live ? i : 0) * ldp;
    if (loader)
        load(0);
    else if (live)
    {
        c1 = coeff[3 * i], c2 = coeff[3 * i + 1], c3 = coeff[3 * i + 2];
        q1 = state[3 * i], q2 = state[3 * i + 1], q3 = state[3 * i + 2];
        row[2] = q2;
        row[3] = q1;
    }
    __syncthreads();
    for (int w = 0; w < nw; w++)
    {
        if (loader)
        {
            if (w + 1 < nw)
                load(w + 1);
        }
        else if (live)
        {
            if ((w + 1) * FD <= ns)
                iir_window<false>(sf[w & 1], lane, w * FD, ns, c1, c2, c3, q1, q2, q3, row);
            else
                iir_window<true>(sf[w & 1], lane, w * FD, ns, c1, c2, c3, q1, q2, q3, row);
        }
        __syncthreads();
    }
    if (!loader && live)
    {
        state[3 * i] = q1;
        state[3 * i + 1] = q2;
        state[3 * i + 2] = q3;
    }
}

// Modal accelerations Q = (q1 + q3 - 2 q2) / (dt dt) after each sample's shift, i.e. from the
// displacements of samples s, s - 2 and s - 1: thread per (mode, sample), Q stored mode-major
// (row stride ldq, a multiple of SPT; entries of the last sample block past ns are written 0).
__global__ __launch_bounds__(MB) void k_modal_q(const float *__restrict__ qp, int ldp, int nmodes, int ns,
                                                int ns_pad, float dt, float *__restrict__ q, int ldq)
{
    const long long e = (long long)blockIdx.x * MB + threadIdx.x;
    if (e >= (long long)nmodes * ns_pad)
        return;
    const int i = (int)(e / ns_pad), s = (int)(e % ns_pad);
    const float *row = qp + (size_t)i * ldp + 2 + s;  // row[0] = q3, row[1] = q2, row[2] = q1
    const float dt2 = dt * dt;
    q[(size_t)i * ldq + s] = s < ns ? (row[2] + row[0] - 2 * row[1]) / dt2 : 0.0f;
}

// acc[s][t] = sum_i Mt[i][t] Q[i][s] in mode order, each term added with one fused multiply-add
// (the reference's device loop `acc += M(t, i) * Q[i]` as its compiler contracts it). Block
// (x = sample block, y = triangle block): a thread holds SPT samples of one triangle; the SPT
// values of Q[i] are uniform over the block (scalar loads), and the blocks of one triangle range
// are launched together, so each surface-matrix row segment is fetched from HBM once per XCD.
__global__ __launch_bounds__(MB) void k_modal_accs(const float *__restrict__ mt, int nt, int nmodes,
                                                   const float *__restrict__ q, int ldq, int ns,
                                                   float *__restrict__ accs)
{
    const int t = blockIdx.y * MB + threadIdx.x;
    const int s0 = blockIdx.x * SPT;
    const bool live = t < nt;
    const float *mrow = mt + (live ? t : nt - 1);  // idle lanes read a live entry, store nothing
    float acc[SPT];
#pragma unroll
    for (int k = 0; k < SPT; k++)
        acc[k] = 0.0f;
    const float *qb = q + s0;
#pragma unroll 2
    for (int i = 0; i < nmodes; i++)
    {
        const float m = mrow[(size_t)i * nt];
        const float *qi = qb + (size_t)i * ldq;
#pragma unroll
        for (int k = 0; k < SPT; k++)
            acc[k] = __fmaf_rn(m, qi[k], acc[k]);
    }
    if (!live)
        return;
#pragma unroll
    for (int k = 0; k < SPT; k++)
        if (s0 + k < ns)
            accs[(size_t)(s0 + k) * nt + t] = acc[k];
}

inline unsigned nblocks(long long n) { return (unsigned)(n > 0 ? (n + MB - 1) / MB : 1); }

}  // namespace

// ModalInfo::SetCoeffs in FP32 with the host's libm, as the reference computes it on the host
void modal_coeffs(float eigval, float timestep, float alpha, float beta, float *c123)
{
    const float lambda = eigval;
    const float omega = std::sqrt(lambda);
    const float ksi = (alpha + beta * lambda) / (2 * omega);
    const float omega_prime = omega * std::sqrt(1 - ksi * ksi);
    const float epsilon = std::exp(-ksi * omega * timestep);
    const float sqr_epsilon = epsilon * epsilon;
    const float theta = omega_prime * timestep;
    const float gamma = std::asin(ksi);
    c123[0] = 2 * epsilon * std::cos(theta);
    c123[1] = -sqr_epsilon;
    const float item1 = epsilon * std::cos(theta + gamma);
    const float item2 = sqr_epsilon * std::cos(2 * theta + gamma);
    c123[2] = 2 * (item1 - item2) / (3 * omega * omega_prime);
}

}  // namespace pppm_amd

using namespace pppm_amd;

struct pppm_modal
{
    int nt = 0, nmodes = 0;
    float dt = 0;
    hipStream_t stream = nullptr;
    float *mt = nullptr;      // nmodes x nt
    float *coeff = nullptr;   // nmodes x 3
    float *state = nullptr;   // nmodes x 3 (q1, q2, q3)
    float *q = nullptr;       // nmodes x cap_samples (mode-major, row stride cap_samples)
    float *qp = nullptr;      // nmodes x (cap_samples + 4): displacements, two of them from before the call
    float *forces = nullptr;  // staging for host forces, cap_samples x nmodes
    int cap_samples = 0;      // a multiple of SPT
    std::vector<float> coeff_h;
};

namespace
{
int modal_release(pppm_modal *m)
{
    for (float *p : {m->mt, m->coeff, m->state, m->q, m->qp, m->forces})
        if (p)
            (void)hipFree(p);
    m->mt = m->coeff = m->state = m->q = m->qp = m->forces = nullptr;
    return PPPM_OK;
}

#define MCK(x)                                \
    do                                        \
    {                                         \
        if ((x) != hipSuccess)                \
            return PPPM_ERR_HIP;              \
    } while (0)
}  // namespace

extern "C" {

int pppm_modal_create(pppm_modal **out, const float *eigvecs, int nv, int nmodes, const float *eigvals,
                      const int *tris, int nt, const float *normals, float alpha, float beta, float sample_rate,
                      void *hip_stream)
{
    if (!out || !eigvecs || !eigvals || !tris || !normals || nv < 3 || nmodes < 1 || nt < 1 || !(sample_rate > 0) ||
        nt > 65535 * MB)
        return PPPM_ERR_ARG;
    for (int t = 0; t < 3 * nt; t++)
        if (tris[t] < 0 || tris[t] >= nv)
            return PPPM_ERR_ARG;
    *out = nullptr;
    pppm_modal *m = new pppm_modal();
    m->nt = nt;
    m->nmodes = nmodes;
    m->dt = 1.0f / sample_rate;  // RigidBody::InitIIR_ (RigidBody.cu:192)
    m->stream = (hipStream_t)hip_stream;
    m->coeff_h.resize((size_t)nmodes * 3);
    for (int i = 0; i < nmodes; i++)
        modal_coeffs(eigvals[i], m->dt, alpha, beta, &m->coeff_h[3 * (size_t)i]);
    float *eig = nullptr, *nrm = nullptr;
    int *tri = nullptr;
    auto fail = [&](int rc) {
        for (void *p : {(void *)eig, (void *)nrm, (void *)tri})
            if (p)
                (void)hipFree(p);
        modal_release(m);
        delete m;
        return rc;
    };
    const size_t neig = (size_t)3 * nv * nmodes;
    if (hipMalloc((void **)&m->mt, sizeof(float) * (size_t)nt * nmodes) != hipSuccess ||
        hipMalloc((void **)&m->coeff, sizeof(float) * 3 * (size_t)nmodes) != hipSuccess ||
        hipMalloc((void **)&m->state, sizeof(float) * 3 * (size_t)nmodes) != hipSuccess ||
        hipMalloc((void **)&eig, sizeof(float) * neig) != hipSuccess ||
        hipMalloc((void **)&nrm, sizeof(float) * 3 * (size_t)nt) != hipSuccess ||
        hipMalloc((void **)&tri, sizeof(int) * 3 * (size_t)nt) != hipSuccess)
        return fail(PPPM_ERR_HIP);
    if (hipMemcpyAsync(m->coeff, m->coeff_h.data(), sizeof(float) * 3 * nmodes, hipMemcpyHostToDevice, m->stream) ||
        hipMemsetAsync(m->state, 0, sizeof(float) * 3 * nmodes, m->stream) ||
        hipMemcpyAsync(eig, eigvecs, sizeof(float) * neig, hipMemcpyHostToDevice, m->stream) ||
        hipMemcpyAsync(nrm, normals, sizeof(float) * 3 * nt, hipMemcpyHostToDevice, m->stream) ||
        hipMemcpyAsync(tri, tris, sizeof(int) * 3 * nt, hipMemcpyHostToDevice, m->stream))
        return fail(PPPM_ERR_HIP);
    k_modal_surface<<<nblocks((long long)nt * nmodes), MB, 0, m->stream>>>(eig, nmodes, tri, nrm, nt, m->mt);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess)
        return fail(PPPM_ERR_HIP);
    (void)hipFree(eig);
    (void)hipFree(nrm);
    (void)hipFree(tri);
    *out = m;
    return PPPM_OK;
}

int pppm_modal_destroy(pppm_modal *m)
{
    if (!m)
        return PPPM_ERR_ARG;
    if (m->stream)
        (void)hipStreamSynchronize(m->stream);
    modal_release(m);
    delete m;
    return PPPM_OK;
}

int pppm_modal_run(pppm_modal *m, const float *forces, int nsamples, float *accs, int on_device)
{
    if (!m || !forces || !accs || nsamples < 0 || nsamples > (1 << 30))
        return PPPM_ERR_ARG;
    if (nsamples == 0)
        return PPPM_OK;
    if (nsamples > m->cap_samples)
    {
        MCK(hipStreamSynchronize(m->stream));
        for (float **p : {&m->q, &m->qp, &m->forces})
            if (*p)
            {
                (void)hipFree(*p);
                *p = nullptr;
            }
        m->cap_samples = 0;
        const int cap = (int)(((long long)nsamples + SPT - 1) / SPT * SPT);
        MCK(hipMalloc((void **)&m->q, sizeof(float) * (size_t)cap * m->nmodes));
        MCK(hipMalloc((void **)&m->qp, sizeof(float) * (size_t)(cap + 4) * m->nmodes));
        MCK(hipMalloc((void **)&m->forces, sizeof(float) * (size_t)cap * m->nmodes));
        m->cap_samples = cap;
    }
    const float *f = forces;
    if (!on_device)
    {
        MCK(hipMemcpyAsync(m->forces, forces, sizeof(float) * (size_t)nsamples * m->nmodes, hipMemcpyHostToDevice,
                           m->stream));
        f = m->forces;
    }
    const int ldp = m->cap_samples + 4;
    const int ns_pad = (nsamples + SPT - 1) / SPT * SPT;
    k_modal_iir<<<(m->nmodes + 63) / 64, 128, 0, m->stream>>>(m->coeff, m->state, m->nmodes, f, nsamples, m->qp, ldp);
    k_modal_q<<<nblocks((long long)m->nmodes * ns_pad), MB, 0, m->stream>>>(m->qp, ldp, m->nmodes, nsamples, ns_pad,
                                                                             m->dt, m->q, m->cap_samples);
    const dim3 grid((unsigned)((nsamples + SPT - 1) / SPT), (unsigned)((m->nt + MB - 1) / MB));
    k_modal_accs<<<grid, MB, 0, m->stream>>>(m->mt, m->nt, m->nmodes, m->q, m->cap_samples, nsamples, accs);
    MCK(hipGetLastError());
    if (!on_device)
        MCK(hipStreamSynchronize(m->stream));  // the host forces must outlive the copy
    return PPPM_OK;
}

int pppm_modal_state(pppm_modal *m, float *q3)
{
    if (!m || !q3)
        return PPPM_ERR_ARG;
    MCK(hipMemcpyAsync(q3, m->state, sizeof(float) * 3 * m->nmodes, hipMemcpyDeviceToHost, m->stream));
    MCK(hipStreamSynchronize(m->stream));
    return PPPM_OK;
}

int pppm_modal_set_state(pppm_modal *m, const float *q3)
{
    if (!m || !q3)
        return PPPM_ERR_ARG;
    MCK(hipMemcpyAsync(m->state, q3, sizeof(float) * 3 * m->nmodes, hipMemcpyHostToDevice, m->stream));
    MCK(hipStreamSynchronize(m->stream));
    return PPPM_OK;
}

int pppm_modal_surface_matrix(pppm_modal *m, float *host_out)
{
    if (!m || !host_out)
        return PPPM_ERR_ARG;
    std::vector<float> mt((size_t)m->nt * m->nmodes);
    MCK(hipMemcpyAsync(mt.data(), m->mt, sizeof(float) * mt.size(), hipMemcpyDeviceToHost, m->stream));
    MCK(hipStreamSynchronize(m->stream));
    for (int i = 0; i < m->nmodes; i++)
        for (int t = 0; t < m->nt; t++)
            host_out[(size_t)t * m->nmodes + i] = mt[(size_t)i * m->nt + t];
    return PPPM_OK;
}

int pppm_modal_coeffs(pppm_modal *m, float *host_out)
{
    if (!m || !host_out)
        return PPPM_ERR_ARG;
    for (size_t k = 0; k < m->coeff_h.size(); k++)
        host_out[k] = m->coeff_h[k];
    return PPPM_OK;
}

}  // extern "C"
