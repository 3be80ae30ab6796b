// b2r_heat_ops.h -- heat stencil arithmetic shared by the forward kernels and
// the adjoint: taps of fdm_heat_diffuse_delta1..4 and the per-axis second
// difference in a fixed evaluation order.
#pragma once

#include "b2r_common.h"

struct HeatK {
    float k[3];
    double adt;
    double h2[3];
};

// second-difference taps (heat-3D.cu:15-28; delta1 weights 1, -2)
__host__ __device__ constexpr float heat_c(int D, int j)
{
    return D == 1 ? (j == 0 ? -2.0f : 1.0f)
         : D == 2 ? (j == 0 ? -2.5f : j == 1 ? 4.f / 3.f : -1.f / 12.f)
         : D == 3 ? (j == 0 ? -49.f / 18.f : j == 1 ? 1.5f : j == 2 ? -3.f / 20.f : 1.f / 90.f)
                  : (j == 0 ? -205.f / 72.f : j == 1 ? 1.6f : j == 2 ? -0.2f
                     : j == 3 ? 8.f / 315.f : -1.f / 560.f);
}

// L_a from the two sides m[j] = u[-j], p[j] = u[+j] (j = 1..D) and the centre.
template <int D, int MODE>
__device__ __forceinline__ float lap(const float (&m)[D + 1], const float (&p)[D + 1], float uc)
{
    if (MODE == B2R_HEAT_REF && D == 1) return (m[1] - 2.0f * uc) + p[1];
    float acc = heat_c(D, D) * (m[D] + p[D]);
#pragma unroll
    for (int j = D - 1; j >= 1; --j) acc = acc + heat_c(D, j) * (m[j] + p[j]);
    return acc + heat_c(D, 0) * uc;
}

// F32 mode: L_a = c_D (m_D + p_D) + ... + c_1 (m_1 + p_1) + c_0 u (outermost tap
// first) by fused multiply-adds, then u' = u + t with t = k_x L_x + k_y L_y +
// k_z L_z, again fused: the coefficients c_j and k_a of the per-axis form,
// each product rounded once with its sum (6D+7 operations per cell instead
// of 10D+9, and no coefficient folded into another, which would add a
// rounding to the operator itself).
template <int D>
__device__ __forceinline__ float lap_fma(const float (&m)[D + 1], const float (&p)[D + 1], float uc)
{
    float acc = heat_c(D, D) * (m[D] + p[D]);
#pragma unroll
    for (int j = D - 1; j >= 1; --j) acc = __builtin_fmaf(heat_c(D, j), m[j] + p[j], acc);
    return __builtin_fmaf(heat_c(D, 0), uc, acc);
}

template <int D>
__device__ __forceinline__ float heat_f32(const float (&mx)[D + 1], const float (&px)[D + 1],
                                          const float (&my)[D + 1], const float (&py)[D + 1],
                                          const float (&mz)[D + 1], const float (&pz)[D + 1],
                                          float uc, const HeatK& hk)
{
    float t = hk.k[0] * lap_fma<D>(mx, px, uc);
    t = __builtin_fmaf(hk.k[1], lap_fma<D>(my, py, uc), t);
    t = __builtin_fmaf(hk.k[2], lap_fma<D>(mz, pz, uc), t);
    return uc + t;
}

// One step's update of a cell from its taps, in either mode.
template <int D, int MODE>
__device__ __forceinline__ float heat_update(const float (&mx)[D + 1], const float (&px)[D + 1],
                                             const float (&my)[D + 1], const float (&py)[D + 1],
                                             const float (&mz)[D + 1], const float (&pz)[D + 1],
                                             float uc, const HeatK& hk);

// The kernel arguments of one heat step.
static inline HeatK heat_k(const b2r_heat_params* p)
{
    HeatK hk;
    for (int a = 0; a < 3; ++a) {
        hk.k[a] = p->k[a];
        hk.h2[a] = p->h2[a];
    }
    hk.adt = p->adt;
    return hk;
}

template <int MODE>
__device__ __forceinline__ float heat_combine(float uc, float lx, float ly, float lz, const HeatK& hk)
{
    if (MODE == B2R_HEAT_F32) {
        float t = hk.k[0] * lx;
        t = t + hk.k[1] * ly;
        t = t + hk.k[2] * lz;
        return uc + t;
    } else {
        double t = (double)lx / hk.h2[0];
        t = t + (double)ly / hk.h2[1];
        t = t + (double)lz / hk.h2[2];
        return (float)((double)uc + hk.adt * t);
    }
}

template <int D, int MODE>
__device__ __forceinline__ float heat_update(const float (&mx)[D + 1], const float (&px)[D + 1],
                                             const float (&my)[D + 1], const float (&py)[D + 1],
                                             const float (&mz)[D + 1], const float (&pz)[D + 1],
                                             float uc, const HeatK& hk)
{
    if (MODE == B2R_HEAT_F32) return heat_f32<D>(mx, px, my, py, mz, pz, uc, hk);
    const float lx = lap<D, MODE>(mx, px, uc);
    const float ly = lap<D, MODE>(my, py, uc);
    const float lz = lap<D, MODE>(mz, pz, uc);
    return heat_combine<MODE>(uc, lx, ly, lz, hk);
}
