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
