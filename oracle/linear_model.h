/*
 * The PPPM step as a linear map (grid levels, history rings, Neumann input -> next state), in
 * flat arrays — TEST INFRASTRUCTURE ONLY (the oracle of the adjoint). Built from an oracle scene
 * by orc_linear_model (pppm_oracle.cpp); evaluated in FP64 by adjoint_oracle.cpp.
 */
#ifndef PPPM_LINEAR_MODEL_H
#define PPPM_LINEAR_MODEL_H

#include <vector>

struct LinearModel
{
    int N = 0, nt = 0;
    double rc = 0;
    double coef = 0, hh = 0;                      // leapfrog: c^2 dt^2 and h^2 in FP64 (fdtd.cu:7-33)
    float acoef = 0, coef1 = 0, coef2 = 0, coef3 = 0;  // ABC coefficients in FP32 (fdtd.cu:94-118)
    std::vector<int> c3_cells, c3_start, c3_pairs;  // C3 cells (grid index), CSR of tri*27+o
    std::vector<float> wfS, wfD, wbS, wbD;           // grid-cache weights [tri*27+o][16], decoded
    std::vector<int> f_start, f_nbr;                 // face lists (CSR over c)
    std::vector<float> fS, fD;                       // face weights [entry][16], decoded
    std::vector<int> corner;                         // [tri][8] C3 index of the base-cell corners
    std::vector<float> interp;                       // [tri][8]
    std::vector<double> den;                         // 0.5 A_c - sum_j D_0(c, j)
};

#endif
