// Shared device definitions for the MI355X PPPM library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pppm_amd
{

constexpr int STEPS = 16;                     // CQ taps per weight series (STEP_NUM, CMakeLists.txt:27)
constexpr int RING = 32;                      // history ring length (bem.h:55)
constexpr int GC = 27;                        // 3x3x3 offsets per triangle (pppm_cache.h:9)
constexpr double PI_R = 3.14159265359;        // the reference's PI (macro.h:80)
constexpr float EPS_F = 1e-8f;                // macro.h:99
constexpr double RECOMPUTE_THRESHOLD = 0.01;  // pppm_cache.h:8
constexpr int WAVE = 64;

struct int3_
{
    int x, y, z;
};

__host__ __device__ inline int lin3(int x, int y, int z, int N) { return (x * N + y) * N + z; }

// ------------------------------------------------------------------ FP16 <-> FP32 (RNE)
__device__ inline uint16_t f32_to_f16_bits(float f)
{
    _Float16 h = (_Float16)f;  // v_cvt_f16_f32, round to nearest even
    return __builtin_bit_cast(uint16_t, h);
}
__device__ inline float f16_bits_to_f32(uint32_t bits16)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(bits16 & 0xffffu));
}

// ------------------------------------------------------------------ wave reduction (fixed tree)
__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off, WAVE);
    return v;
}

// ------------------------------------------------------------------ geometry (FP32, as the reference)
struct f3
{
    float x, y, z;
};
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float dotf(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 crossf(f3 a, f3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float lengthf(f3 a) { return sqrtf(dotf(a, a)); }

struct d3
{
    double x, y, z;
};
__host__ __device__ inline d3 D(f3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dotd(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline d3 crossd(d3 a, d3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Per-triangle geometry produced by the binning pass (Triangle, particle_grid.h:68-80)
struct TriGeom
{
    f3 center;
    f3 normal;
    float area;
    int cell;  // linear cell index of the centre
    int base;  // linear index of the interpolation base cell
    int3_ cell3, base3;
};

// Inputs of a triangle's grid-cache recompute, written by the binning pass for the triangles it
// flags dirty: the one-point quadrature point (the centroid), the unit normal and twice the
// area in FP64, the centre's cell, and the centre / normal / area the recompute test compares
// against next time. 13 eight-byte words, so a wave copies one with a load per lane.
struct GridTriIn
{
    double px, py, pz;
    double nx, ny, nz;
    double jac;
    int cx, cy, cz;
    int tri;
    float meta[7];
    float pad;
};
static_assert(sizeof(GridTriIn) == 104 && sizeof(GridTriIn) % 8 == 0, "GridTriIn is 13 eight-byte words");

}  // namespace pppm_amd
