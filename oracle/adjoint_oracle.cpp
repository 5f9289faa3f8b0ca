// FP64 forward and adjoint of the PPPM step on its linear model — TEST INFRASTRUCTURE ONLY.
//
// The forward follows the reference's step with every operation in FP64 on the same
// coefficients (decoded FP16 weights x scales, interpolation weights, the ABC's FP32
// coefficients): fdtd_inner_kernel / ab_condition_solve (src/algorithm/fdtd.cu:7-137),
// solve_fdtd_far (pppm.cu:8-51), set_neumann_condition (pppm.cu:206-213), solve_face_far /
// solve_face_near / update_dirichlet (pppm.cu:116-193), solve_fdtd_near (pppm.cu:53-114).
// The adjoint is written as the transpose of each forward statement, y += a x  ->  x_bar += a y_bar,
// applied in reverse order (a scatter per statement, sequential): it shares no code with the GPU
// adjoint's gathers. Central finite differences of the FP64 forward pin it.
#include <cmath>
#include <cstring>
#include <vector>

#include "linear_model.h"

struct orc_scene;
extern "C" bool orc_linear_model(const orc_scene *s, LinearModel &m);

namespace
{

constexpr int ST = 16, RG = 32;

struct Lin
{
    const LinearModel &m;
    int N;
    size_t NN;
    Lin(const LinearModel &mm) : m(mm), N(mm.N), NN((size_t)mm.N * mm.N * mm.N) {}
    int lin(int x, int y, int z) const { return (x * N + y) * N + z; }
    static int lvl(int t) { return ((t % 3) + 3) % 3; }
    static int slot(int t) { return ((t % RG) + RG) % RG; }
};

// 4x4 solve in FP64 (Gaussian elimination, partial pivoting); trans: solve A^T x = b
void solve4d(const double A[4][4], const double b[4], double x[4], bool trans)
{
    double m[4][5];
    for (int r = 0; r < 4; r++)
    {
        for (int c = 0; c < 4; c++)
            m[r][c] = trans ? A[c][r] : A[r][c];
        m[r][4] = b[r];
    }
    for (int c = 0; c < 4; c++)
    {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(m[r][c]) > std::fabs(m[p][c]))
                p = r;
        for (int e = 0; e < 5; e++)
            std::swap(m[c][e], m[p][e]);
        for (int r = c + 1; r < 4; r++)
        {
            const double f = m[r][c] / m[c][c];
            for (int e = c; e < 5; e++)
                m[r][e] -= f * m[c][e];
        }
    }
    for (int r = 3; r >= 0; r--)
    {
        double s = m[r][4];
        for (int e = r + 1; e < 4; e++)
            s -= m[r][e] * x[e];
        x[r] = s / m[r][r];
    }
}

// ab_condition_solve's matrix (fdtd.cu:94-111) with FP32 entries
void abc_matrix(const LinearModel &m, const int *judge, double A[4][4])
{
    float a[4][4];
    a[3][3] = 1;
    for (int d = 0; d < 3; d++)
        a[3][d] = -m.acoef * (float)(judge[d] != 0);
    for (int idx = 0; idx < 3; idx++)
    {
        const int x = idx, y = (idx + 1) % 3, z = (idx + 2) % 3;
        a[idx][3] = m.coef1;
        a[idx][x] = m.coef2;
        a[idx][y] = m.coef3 * (float)(judge[y] != 0);
        a[idx][z] = m.coef3 * (float)(judge[z] != 0);
        (void)x;
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            A[r][c] = a[r][c];
}

struct Judge
{
    int j[3];
};
Judge judge_of(int N, int i, int j, int k)
{
    const int co[3] = {i, j, k};
    Judge J;
    for (int d = 0; d < 3; d++)
        J.j[d] = -((int)(co[d] == 1)) + ((int)(co[d] == N - 2));
    return J;
}

// The right-hand side of a shell cell as a list of terms (array 0: g, 1: gm; linear offset;
// row; coefficient), read off ab_condition_solve's expressions (fdtd.cu:96-118): every term of
// b[row] is coefficient x value.
struct Term
{
    int arr, off, row;
    double coef;
};
int abc_terms(const LinearModel &m, const Judge &J, int N, Term *t)
{
    const int strd[3] = {N * N, N, 1};
    const double coef = m.acoef, coef1 = m.coef1, coef2 = m.coef2, coef3 = m.coef3;
    int n = 0;
    // b[3] = -gm + 2 g - coef (6 g) + sum_d coef (g[+d] [judge <= 0] + g[-d] [judge >= 0])
    t[n++] = {1, 0, 3, -1.0};
    t[n++] = {0, 0, 3, 2.0};
    t[n++] = {0, 0, 3, -(double)(m.acoef * 6.0f)};
    for (int d = 0; d < 3; d++)
    {
        t[n++] = {0, strd[d], 3, coef * (J.j[d] <= 0)};
        t[n++] = {0, -strd[d], 3, coef * (J.j[d] >= 0)};
    }
    for (int idx = 0; idx < 3; idx++)
    {
        const int x = idx, y = (idx + 1) % 3, z = (idx + 2) % 3;
        // -coef1 (gm - 2 g)
        t[n++] = {1, 0, idx, -coef1};
        t[n++] = {0, 0, idx, 2.0 * coef1};
        // -coef2 (g[+x] (-[j<0]) + g[-x] (-[j>0]) + gm[+x] (-j) + gm[-x] j)
        t[n++] = {0, strd[x], idx, -coef2 * -(double)(J.j[x] < 0)};
        t[n++] = {0, -strd[x], idx, -coef2 * -(double)(J.j[x] > 0)};
        t[n++] = {1, strd[x], idx, -coef2 * (double)(-J.j[x])};
        t[n++] = {1, -strd[x], idx, -coef2 * (double)J.j[x]};
        // -coef3 (g[+y] [<=0] + g[-y] [>=0] + g[+z] [<=0] + g[-z] [>=0] - 4 g)
        t[n++] = {0, strd[y], idx, -coef3 * (J.j[y] <= 0)};
        t[n++] = {0, -strd[y], idx, -coef3 * (J.j[y] >= 0)};
        t[n++] = {0, strd[z], idx, -coef3 * (J.j[z] <= 0)};
        t[n++] = {0, -strd[z], idx, -coef3 * (J.j[z] >= 0)};
        t[n++] = {0, 0, idx, 4.0 * coef3};
    }
    return n;
}

bool is_shell(int N, int i, int j, int k)
{
    return i == 1 || j == 1 || k == 1 || i == N - 2 || j == N - 2 || k == N - 2;
}

// forward state: 3 grid levels (with ghosts), rings
struct FState
{
    std::vector<double> G[3], hn, hd, far;
    int t = -1;
};

void forward_step(const Lin &L, FState &S, const double *g_in)
{
    const LinearModel &m = L.m;
    const int N = L.N;
    const int t = S.t;  // the step writes level t + 1
    std::vector<double> &g = S.G[Lin::lvl(t)], &gm = S.G[Lin::lvl(t - 1)], &gp = S.G[Lin::lvl(t + 1)];
    for (int i = 2; i < N - 2; i++)
        for (int j = 2; j < N - 2; j++)
            for (int k = 2; k < N - 2; k++)
            {
                const int q = L.lin(i, j, k);
                double sum = g[q - N * N] + g[q + N * N] + g[q - N] + g[q + N] + g[q - 1] + g[q + 1] - 6 * g[q];
                gp[q] = 2 * g[q] + m.coef * (sum / m.hh) - gm[q];
            }
    // ABC shell: gp at the cell, ghost values of level t (all read before any ghost is written:
    // the terms never read g's ghosts)
    Term terms[64];
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++)
            for (int k = 1; k < N - 1; k++)
            {
                if (!is_shell(N, i, j, k))
                    continue;
                const Judge J = judge_of(N, i, j, k);
                double A[4][4], b[4] = {0, 0, 0, 0}, r[4];
                abc_matrix(m, J.j, A);
                const int me = L.lin(i, j, k);
                const int n = abc_terms(m, J, N, terms);
                for (int e = 0; e < n; e++)
                    if (terms[e].coef != 0)
                        b[terms[e].row] += terms[e].coef * (terms[e].arr ? gm : g)[me + terms[e].off];
                solve4d(A, b, r, false);
                gp[me] = r[3];
                const int strd[3] = {N * N, N, 1};
                for (int d = 0; d < 3; d++)
                    if (J.j[d] != 0)
                        g[me + J.j[d] * strd[d]] = r[d];
            }
    S.t = t + 1;
    const int tt = S.t, sl = Lin::slot(tt);
    std::vector<double> &P = S.G[Lin::lvl(tt)];
    const int nc3 = (int)m.c3_cells.size();
    // far field (taps 1..15; tap 0 of W^fdtd is zero)
    for (int c = 0; c < nc3; c++)
    {
        double acc = 0;
        for (int p = m.c3_start[c]; p < m.c3_start[c + 1]; p++)
        {
            const int rec = m.c3_pairs[p], tri = rec / 27;
            for (int k = 1; k < ST; k++)
            {
                const int s2 = Lin::slot(tt - k);
                acc += -(double)m.wfS[(size_t)rec * ST + k] * S.hn[tri * RG + s2] +
                       (double)m.wfD[(size_t)rec * ST + k] * S.hd[tri * RG + s2];
            }
        }
        S.far[c] = P[m.c3_cells[c]] - acc;
    }
    for (int i = 0; i < m.nt; i++)
        S.hn[i * RG + sl] = g_in[i];
    // Dirichlet update, D_j[t] read as 0 in the face near field
    std::vector<double> nd(m.nt);
    for (int i = 0; i < m.nt; i++)
        S.hd[i * RG + sl] = 0;
    for (int c = 0; c < m.nt; c++)
    {
        double ff = 0;
        for (int w = 0; w < 8; w++)
            ff += (double)m.interp[c * 8 + w] * S.far[m.corner[c * 8 + w]];
        double acc = 0;
        for (int e = m.f_start[c]; e < m.f_start[c + 1]; e++)
        {
            const int j = m.f_nbr[e];
            for (int k = 0; k < ST; k++)
            {
                const int s2 = Lin::slot(tt - k);
                acc += -(double)m.fS[(size_t)e * ST + k] * S.hn[j * RG + s2] +
                       (double)m.fD[(size_t)e * ST + k] * S.hd[j * RG + s2];
            }
        }
        nd[c] = (ff * m.rc + acc) / m.den[c];
    }
    for (int c = 0; c < m.nt; c++)
        S.hd[c * RG + sl] = nd[c];
    // near field
    for (int c = 0; c < nc3; c++)
    {
        double acc = 0;
        for (int p = m.c3_start[c]; p < m.c3_start[c + 1]; p++)
        {
            const int rec = m.c3_pairs[p], tri = rec / 27;
            for (int k = 0; k < ST; k++)
            {
                const int s2 = Lin::slot(tt - k);
                acc += -(double)m.wbS[(size_t)rec * ST + k] * S.hn[tri * RG + s2] +
                       (double)m.wbD[(size_t)rec * ST + k] * S.hd[tri * RG + s2];
            }
        }
        P[m.c3_cells[c]] = S.far[c] + acc * 0.98;
    }
}

void forward(const LinearModel &m, const double *g, int T, const int *lcell, int nl, double *out)
{
    Lin L(m);
    FState S;
    for (auto &v : S.G)
        v.assign(L.NN, 0.0);
    S.hn.assign((size_t)m.nt * RG, 0.0);
    S.hd.assign((size_t)m.nt * RG, 0.0);
    S.far.assign(m.c3_cells.size(), 0.0);
    for (int t = 0; t < T; t++)
    {
        forward_step(L, S, g + (size_t)t * m.nt);
        const std::vector<double> &P = S.G[Lin::lvl(S.t)];
        for (int l = 0; l < nl; l++)
            out[(size_t)t * nl + l] = P[lcell[l]];
    }
}

// adjoint: each forward statement transposed, in reverse order
void adjoint(const LinearModel &m, const double *gout, int T, const int *lcell, int nl, double *grad)
{
    Lin L(m);
    const int N = L.N;
    std::vector<double> lam[3], ln((size_t)m.nt * RG, 0.0), ld((size_t)m.nt * RG, 0.0);
    for (auto &v : lam)
        v.assign(L.NN, 0.0);
    const int nc3 = (int)m.c3_cells.size();
    std::vector<double> lfar(nc3), v(m.nt);
    Term terms[64];
    const int strd[3] = {N * N, N, 1};
    for (int tt = T - 1; tt >= 0; tt--)
    {
        std::vector<double> &lP = lam[Lin::lvl(tt)];
        const int sl = Lin::slot(tt);
        for (int l = 0; l < nl; l++)
            lP[lcell[l]] += gout[(size_t)tt * nl + l];
        // near: P[c] = far[c] + 0.98 sum(...)
        for (int c = 0; c < nc3; c++)
        {
            const double u = lP[m.c3_cells[c]];
            lP[m.c3_cells[c]] = 0;
            lfar[c] = u;
            for (int p = m.c3_start[c]; p < m.c3_start[c + 1]; p++)
            {
                const int rec = m.c3_pairs[p], tri = rec / 27;
                for (int k = 0; k < ST; k++)
                {
                    const int s2 = Lin::slot(tt - k);
                    ln[tri * RG + s2] += -(double)m.wbS[(size_t)rec * ST + k] * 0.98 * u;
                    ld[tri * RG + s2] += (double)m.wbD[(size_t)rec * ST + k] * 0.98 * u;
                }
            }
        }
        // Dirichlet: D_c[t] = (rc ff + acc) / den
        for (int c = 0; c < m.nt; c++)
        {
            v[c] = ld[c * RG + sl] / m.den[c];
            ld[c * RG + sl] = 0;
        }
        for (int c = 0; c < m.nt; c++)
        {
            for (int w = 0; w < 8; w++)
                lfar[m.corner[c * 8 + w]] += m.rc * (double)m.interp[c * 8 + w] * v[c];
            for (int e = m.f_start[c]; e < m.f_start[c + 1]; e++)
            {
                const int j = m.f_nbr[e];
                for (int k = 0; k < ST; k++)
                {
                    const int s2 = Lin::slot(tt - k);
                    ln[j * RG + s2] += -(double)m.fS[(size_t)e * ST + k] * v[c];
                    if (k > 0)
                        ld[j * RG + s2] += (double)m.fD[(size_t)e * ST + k] * v[c];
                }
            }
        }
        // Neumann input: N[t] = g[t] (D[t] = 0 before the face near field: consumed above)
        for (int i = 0; i < m.nt; i++)
        {
            grad[(size_t)tt * m.nt + i] = ln[i * RG + sl];
            ln[i * RG + sl] = 0;
        }
        // far: far[c] = P[c] - sum(...)
        for (int c = 0; c < nc3; c++)
        {
            lP[m.c3_cells[c]] += lfar[c];
            for (int p = m.c3_start[c]; p < m.c3_start[c + 1]; p++)
            {
                const int rec = m.c3_pairs[p], tri = rec / 27;
                for (int k = 1; k < ST; k++)
                {
                    const int s2 = Lin::slot(tt - k);
                    ln[tri * RG + s2] += (double)m.wfS[(size_t)rec * ST + k] * lfar[c];
                    ld[tri * RG + s2] += -(double)m.wfD[(size_t)rec * ST + k] * lfar[c];
                }
            }
        }
        // FDTD of step tt (levels tt-1 = g, tt-2 = gm -> tt): the ABC shell first (reverse of
        // the forward order), then the interior
        std::vector<double> &lg = lam[Lin::lvl(tt - 1)], &lgm = lam[Lin::lvl(tt - 2)];
        std::vector<double> lgm_new(L.NN, 0.0);
        for (int i = 1; i < N - 1; i++)
            for (int j = 1; j < N - 1; j++)
                for (int k = 1; k < N - 1; k++)
                {
                    if (!is_shell(N, i, j, k))
                        continue;
                    const Judge J = judge_of(N, i, j, k);
                    const int me = L.lin(i, j, k);
                    double A[4][4], rb[4], bb[4];
                    abc_matrix(m, J.j, A);
                    for (int d = 0; d < 3; d++)
                        rb[d] = J.j[d] != 0 ? lg[me + J.j[d] * strd[d]] : 0.0;
                    rb[3] = lP[me];
                    solve4d(A, rb, bb, true);
                    const int n = abc_terms(m, J, N, terms);
                    for (int e = 0; e < n; e++)
                        if (terms[e].coef != 0)
                            (terms[e].arr ? lgm_new : lg)[me + terms[e].off] += terms[e].coef * bb[terms[e].row];
                }
        for (int i = 2; i < N - 2; i++)
            for (int j = 2; j < N - 2; j++)
                for (int k = 2; k < N - 2; k++)
                {
                    const int q = L.lin(i, j, k);
                    const double a = lP[q];
                    lg[q] += 2 * a - 6 * (m.coef / m.hh) * a;
                    for (int d = 0; d < 3; d++)
                    {
                        lg[q + strd[d]] += (m.coef / m.hh) * a;
                        lg[q - strd[d]] += (m.coef / m.hh) * a;
                    }
                    lgm_new[q] += -a;
                }
        lgm.swap(lgm_new);
        // lP is the level of step tt+3 from here on (dead); its ghosts were the adjoint of the
        // values step tt+1 would have read, already consumed
        std::fill(lP.begin(), lP.end(), 0.0);
    }
}

}  // namespace

extern "C" {

// FP64 forward of T steps from the zero state: out[t * nl + l] = pressure at lcell[l] after step t
int orc_linear_forward(const orc_scene *s, const double *g, int T, const int *lcell, int nl, double *out)
{
    LinearModel m;
    if (!orc_linear_model(s, m))
        return -1;
    forward(m, g, T, lcell, nl, out);
    return 0;
}

// FP64 adjoint: grad[t * nt + i] = sum_{t', l} gout[t' * nl + l] d out[t', l] / d g[t, i]
int orc_linear_adjoint(const orc_scene *s, const double *gout, int T, const int *lcell, int nl, double *grad)
{
    LinearModel m;
    if (!orc_linear_model(s, m))
        return -1;
    adjoint(m, gout, T, lcell, nl, grad);
    return 0;
}

// Central differences of J(g) = sum (out - target)^2 at entries (step[e], tri[e]):
// fd[e] = (J(g + eps d_e) - J(g - eps d_e)) / (2 eps), each J a full FP64 forward run;
// entries in parallel (OpenMP)
int orc_linear_fd(const orc_scene *s, const double *g, const double *target, int T, const int *lcell, int nl,
                  int n_entries, const int *step, const int *tri, double eps, double *fd)
{
    LinearModel m;
    if (!orc_linear_model(s, m))
        return -1;
    const size_t G = (size_t)T * m.nt;
#pragma omp parallel for schedule(dynamic, 1)
    for (int e = 0; e < n_entries; e++)
    {
        std::vector<double> gg(g, g + G), out((size_t)T * nl);
        double J[2];
        for (int sgn = 0; sgn < 2; sgn++)
        {
            gg[(size_t)step[e] * m.nt + tri[e]] = g[(size_t)step[e] * m.nt + tri[e]] + (sgn ? -eps : eps);
            forward(m, gg.data(), T, lcell, nl, out.data());
            double acc = 0;
            for (size_t q = 0; q < out.size(); q++)
            {
                const double r = out[q] - target[q];
                acc += r * r;
            }
            J[sgn] = acc;
        }
        fd[e] = (J[0] - J[1]) / (2 * eps);
    }
    return 0;
}

}  // extern "C"
