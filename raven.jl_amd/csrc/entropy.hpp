// Entropy and entropy variables of the laws (euler.jl:70-91,
// advection.jl:20-21), evaluated in FP64 from state values of either
// precision; shared by the entropy passes and the fused relaxation stage.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace rv {

// s + c += s2 + c2 with the rounding error of s + s2 (Knuth's TwoSum, exact
// and independent of the operands' order) added to the compensation
__device__ inline void comp_add(double &s, double &c, double s2, double c2)
{
    const double t = s + s2;
    const double bb = t - s;
    const double err = (s - (t - bb)) + (s2 - bb);
    s = t;
    c = (c + c2) + err;
}

// entropy(law, q): advection q^2 / 2; Euler -rho log(p / rho^gamma) / (gamma - 1)
template <int LAW, int DIM>
__device__ inline double law_entropy(const double *q, double g)
{
    if constexpr (LAW == LAW_ADVECTION) {
        return q[0] * q[0] / 2;
    } else {
        double dot = q[1] * q[1];
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            dot = dot + q[1 + d] * q[1 + d];
        const double p = (g - 1) * (q[DIM + 1] - dot / (2 * q[0]));
        // log(p / rho^gamma) as two logarithms (no power)
        const double s = log(p) - g * log(q[0]);
        return -q[0] * s / (g - 1);
    }
}

// entropyvariables(law, q) . k: advection q k; Euler
// ((gamma - s)/(gamma - 1) - |u|^2 b, 2 b u, -2 b) . k with b = rho / 2p
template <int LAW, int DIM>
__device__ inline double law_entropy_dot(const double *q, const double *k, double g)
{
    if constexpr (LAW == LAW_ADVECTION) {
        return q[0] * k[0];
    } else {
        double dot = q[1] * q[1];
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            dot = dot + q[1 + d] * q[1 + d];
        const double p = (g - 1) * (q[DIM + 1] - dot / (2 * q[0]));
        const double s = log(p) - g * log(q[0]);
        const double b = q[0] / (2 * p);
        double u[DIM], uu;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            u[d] = q[1 + d] / q[0];
        uu = u[0] * u[0];
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            uu = uu + u[d] * u[d];
        double r = ((g - s) / (g - 1) - uu * b) * k[0];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            r = r + (2 * b * u[d]) * k[1 + d];
        return r + (-2 * b) * k[DIM + 1];
    }
}

// law_entropy(q) and law_entropy_dot(q, k) together, the logarithms once
template <int LAW, int DIM>
__device__ inline void law_entropy_and_dot(const double *q, const double *k, double g,
                                           double &eta, double &dot)
{
    if constexpr (LAW == LAW_ADVECTION) {
        eta = q[0] * q[0] / 2;
        dot = q[0] * k[0];
    } else {
        double uu2 = q[1] * q[1];
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            uu2 = uu2 + q[1 + d] * q[1 + d];
        const double p = (g - 1) * (q[DIM + 1] - uu2 / (2 * q[0]));
        const double s = log(p) - g * log(q[0]);
        eta = -q[0] * s / (g - 1);
        const double b = q[0] / (2 * p);
        double u[DIM], uu;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            u[d] = q[1 + d] / q[0];
        uu = u[0] * u[0];
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            uu = uu + u[d] * u[d];
        double r = ((g - s) / (g - 1) - uu * b) * k[0];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            r = r + (2 * b * u[d]) * k[1 + d];
        dot = r + (-2 * b) * k[DIM + 1];
    }
}

// law_entropy_dot for 3-D states, written for fewer registers (the column
// kernel's and the 3-D flux-differencing kernel's fused relaxation stages)
template <int LAW>
__device__ inline double law_entropy_dot_3d(const double *q, const double *k, double g)
{
    if constexpr (LAW == LAW_ADVECTION) {
        return q[0] * k[0];
    } else {
        const double rinv = 1.0 / q[0];
        const double u0 = q[1] * rinv, u1 = q[2] * rinv, u2 = q[3] * rinv;
        const double uu = u0 * u0 + u1 * u1 + u2 * u2;
        const double p = (g - 1) * (q[4] - 0.5 * q[0] * uu);
        const double s = log(p) - g * log(q[0]);
        const double b = 0.5 * q[0] / p;
        return ((g - s) / (g - 1) - uu * b) * k[0] + 2 * b * (u0 * k[1] + u1 * k[2] + u2 * k[3]) -
               2 * b * k[4];
    }
}

}  // namespace rv
