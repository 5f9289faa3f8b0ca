// Modal sound source of a rigid body on the GPU (the C ABI's pppm_modal_*).
//
// Reference behaviour: ModalInfo::SetCoeffs (src/utils/RigidBody.cu:11-27), MatrixVertToTri
// (RigidBody.cu:206-230), the IIR recursion of CalculateIIR_ (RigidBody.cu:282-291) and
// update_surf_acc_kernel (RigidBody.cu:240-250). The reference steps the oscillators on the
// host one sample at a time and launches one surface-acceleration kernel per sample; here a
// call covers a block of samples: one thread per mode runs its oscillator through the block,
// then one thread per (triangle, 8 samples) forms the accelerations, reading each surface-matrix
// entry once for the 8 samples. Every value keeps the reference's FP32 operations and order.
//
// Layout: the surface matrix is stored mode-major, Mt[i][t], so the accelerations stream it with
// coalesced loads; Q (modal accelerations) is [sample][mode].
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
constexpr int SPT = 8;  // samples per thread in the acceleration kernel

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

// one thread per mode: q0 = c1 q1 + c2 q2 + c3 f; (q3, q2, q1) <- (q2, q1, q0);
// Q = (q1 + q3 - 2 q2) / (dt dt), for each sample of the block in order
__global__ __launch_bounds__(MB) void k_modal_iir(const float *__restrict__ coeff, float *__restrict__ state,
                                                  int nmodes, const float *__restrict__ forces, int ns, float dt,
                                                  float *__restrict__ q)
{
    const int i = blockIdx.x * MB + threadIdx.x;
    if (i >= nmodes)
        return;
    const float c1 = coeff[3 * i], c2 = coeff[3 * i + 1], c3 = coeff[3 * i + 2];
    float q1 = state[3 * i], q2 = state[3 * i + 1], q3 = state[3 * i + 2];
    const float dt2 = dt * dt;
    for (int s = 0; s < ns; s++)
    {
        const float q0 = c1 * q1 + c2 * q2 + c3 * forces[(size_t)s * nmodes + i];
        q3 = q2;
        q2 = q1;
        q1 = q0;
        q[(size_t)s * nmodes + i] = (q1 + q3 - 2 * q2) / dt2;
    }
    state[3 * i] = q1;
    state[3 * i + 1] = q2;
    state[3 * i + 2] = q3;
}

// acc[s][t] = sum_i Mt[i][t] Q[s][i], in mode order; a thread holds SPT samples of one triangle
__global__ __launch_bounds__(MB) void k_modal_accs(const float *__restrict__ mt, int nt, int nmodes,
                                                   const float *__restrict__ q, int ns, float *__restrict__ accs)
{
    const long long e = (long long)blockIdx.x * MB + threadIdx.x;
    const int sblocks = (ns + SPT - 1) / SPT;
    if (e >= (long long)nt * sblocks)
        return;
    const int t = (int)(e % nt), s0 = (int)(e / nt) * SPT;
    float acc[SPT];
#pragma unroll
    for (int k = 0; k < SPT; k++)
        acc[k] = 0.0f;
    for (int i = 0; i < nmodes; i++)
    {
        const float m = mt[(size_t)i * nt + t];
#pragma unroll
        for (int k = 0; k < SPT; k++)
            if (s0 + k < ns)
                acc[k] += m * q[(size_t)(s0 + k) * nmodes + i];
    }
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
    float *q = nullptr;       // cap_samples x nmodes
    float *forces = nullptr;  // staging for host forces, cap_samples x nmodes
    int cap_samples = 0;
    std::vector<float> coeff_h;
};

namespace
{
int modal_release(pppm_modal *m)
{
    for (float *p : {m->mt, m->coeff, m->state, m->q, m->forces})
        if (p)
            (void)hipFree(p);
    m->mt = m->coeff = m->state = m->q = m->forces = nullptr;
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
    if (!out || !eigvecs || !eigvals || !tris || !normals || nv < 3 || nmodes < 1 || nt < 1 || !(sample_rate > 0))
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
    if (!m || !forces || !accs || nsamples < 0)
        return PPPM_ERR_ARG;
    if (nsamples == 0)
        return PPPM_OK;
    if (nsamples > m->cap_samples)
    {
        MCK(hipStreamSynchronize(m->stream));
        for (float **p : {&m->q, &m->forces})
            if (*p)
            {
                (void)hipFree(*p);
                *p = nullptr;
            }
        m->cap_samples = 0;
        MCK(hipMalloc((void **)&m->q, sizeof(float) * (size_t)nsamples * m->nmodes));
        MCK(hipMalloc((void **)&m->forces, sizeof(float) * (size_t)nsamples * m->nmodes));
        m->cap_samples = nsamples;
    }
    const float *f = forces;
    if (!on_device)
    {
        MCK(hipMemcpyAsync(m->forces, forces, sizeof(float) * (size_t)nsamples * m->nmodes, hipMemcpyHostToDevice,
                           m->stream));
        f = m->forces;
    }
    k_modal_iir<<<nblocks(m->nmodes), MB, 0, m->stream>>>(m->coeff, m->state, m->nmodes, f, nsamples, m->dt, m->q);
    const long long items = (long long)m->nt * ((nsamples + SPT - 1) / SPT);
    k_modal_accs<<<nblocks(items), MB, 0, m->stream>>>(m->mt, m->nt, m->nmodes, m->q, nsamples, accs);
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
