// CPU restatement of the reference's modal sound source — TEST INFRASTRUCTURE ONLY.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
//
//   orc_modal_coeffs          ModalInfo::SetCoeffs            src/utils/RigidBody.cu:11-27
//   orc_modal_surface_matrix  MatrixVertToTri                 src/utils/RigidBody.cu:206-224
//   orc_modal_iir             CalculateIIR_, recursion part   src/utils/RigidBody.cu:282-291
//   orc_modal_accs            update_surf_acc_kernel          src/utils/RigidBody.cu:240-250
//   orc_modal_forces          CalculateIIR_, impulse part     src/utils/RigidBody.cu:257-281,
//                             Impulse::amp / dead             src/utils/RigidBody.h:19-31
// FP32 throughout, in the reference's operation order (built with -ffp-contract=off). The
// reference's host code (coefficients, oscillators, forces) is uncontracted; its device sum
// `acc += M(t, i) * Q[i]` is contracted by its device compiler into a fused multiply-add, which
// orc_modal_accs states with std::fma.
#include <cmath>
#include <vector>

extern "C" {

void orc_modal_coeffs(const float *eigvals, int nmodes, float timestep, float alpha, float beta, float *c123)
{
    for (int i = 0; i < nmodes; i++)
    {
        float lam = eigvals[i];
        float om = std::sqrt(lam);
        float ksi = (alpha + beta * lam) / (2 * om);
        float omp = om * std::sqrt(1 - ksi * ksi);
        float eps = std::exp(-ksi * om * timestep);
        float eps2 = eps * eps;
        float th = omp * timestep;
        float gam = std::asin(ksi);
        c123[3 * i] = 2 * eps * std::cos(th);
        c123[3 * i + 1] = -eps2;
        float a = eps * std::cos(th + gam);
        float b = eps2 * std::cos(2 * th + gam);
        c123[3 * i + 2] = 2 * (a - b) / (3 * om * omp);
    }
}

// eig: 3 nv rows x nmodes columns, row-major; out: nt x nmodes, row-major
void orc_modal_surface_matrix(const float *eig, int nv, int nmodes, const int *tris, int nt, const float *normals,
                              float *out)
{
    (void)nv;
    for (int t = 0; t < nt; t++)
    {
        const int *v = tris + 3 * t;
        const float *n = normals + 3 * t;
        for (int i = 0; i < nmodes; i++)
        {
            float u[3];
            for (int d = 0; d < 3; d++)
            {
                float acc = eig[(size_t)(3 * v[0] + d) * nmodes + i];
                acc = acc + eig[(size_t)(3 * v[1] + d) * nmodes + i];
                acc = acc + eig[(size_t)(3 * v[2] + d) * nmodes + i];
                u[d] = acc / 3;
            }
            out[(size_t)t * nmodes + i] = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
        }
    }
}

// state: nmodes x (q1, q2, q3), updated; forces and q: nsamples x nmodes
void orc_modal_iir(const float *c123, int nmodes, float timestep, float *state, const float *forces, int nsamples,
                   float *q)
{
    for (int s = 0; s < nsamples; s++)
        for (int i = 0; i < nmodes; i++)
        {
            float *st = state + 3 * i;
            const float *c = c123 + 3 * i;
            float q0 = c[0] * st[0] + c[1] * st[1] + c[2] * forces[(size_t)s * nmodes + i];
            st[2] = st[1];
            st[1] = st[0];
            st[0] = q0;
            q[(size_t)s * nmodes + i] = (st[0] + st[2] - 2 * st[1]) / (timestep * timestep);
        }
}

// msurf: nt x nmodes; q: nsamples x nmodes; acc: nsamples x nt
void orc_modal_accs(const float *msurf, int nt, int nmodes, const float *q, int nsamples, float *acc)
{
    for (int s = 0; s < nsamples; s++)
        for (int t = 0; t < nt; t++)
        {
            float a = 0;
            for (int i = 0; i < nmodes; i++)
                a = std::fma(msurf[(size_t)t * nmodes + i], q[(size_t)s * nmodes + i], a);
            acc[(size_t)s * nt + t] = a;
        }
}

// Modal forces of one sample from the impulses active at `time`: each impulse is (start time,
// vertex, impulse vector); active[k] != 0 selects it. f_i += E(3v, i) I.x + E(3v+1, i) I.y +
// E(3v+2, i) I.z with I = impulse * amp, amp = sin(pi (time - start) / 0.0005) in double.
void orc_modal_forces(const float *eig, int nmodes, const float *imp_time, const int *imp_vert, const float *imp_vec,
                      const int *active, int nimp, float time, float *f)
{
    for (int i = 0; i < nmodes; i++)
    {
        float acc = 0;
        for (int j = 0; j < nimp; j++)
        {
            if (!active[j])
                continue;
            const float amp = (float)std::sin(M_PI * (time - imp_time[j]) / 0.0005f);
            const float ix = imp_vec[3 * j] * amp, iy = imp_vec[3 * j + 1] * amp, iz = imp_vec[3 * j + 2] * amp;
            const int v = imp_vert[j];
            acc += eig[(size_t)(3 * v) * nmodes + i] * ix + eig[(size_t)(3 * v + 1) * nmodes + i] * iy +
                   eig[(size_t)(3 * v + 2) * nmodes + i] * iz;
        }
        f[i] = acc;
    }
}

}  // extern "C"
