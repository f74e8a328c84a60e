// Pair potentials as device functors: each maps a pair's squared distance (below the cutoff)
// to its energy u and to fs = -(du/dr)/r, so that the force on i from j is fs * r_ij.
//
// LJ: u = 4 eps [(s2/r2)^6 - (s2/r2)^3] - u_c with u_c the same at rc
//     (potential/_lennard_jones/_potential.py:113-120,209-234).
// Table: Catmull-Rom on knots padded with a virtual knot at each end and the last knot read as 0;
//     below the first knot the row's tangent there (potential/_tabulated/_potential.py:59-91).
#pragma once

#include "common.hpp"

namespace ndn {

// 1 / value: the hardware reciprocal in fp32 (1 ulp), a division in fp64. For factors that only
// scale a force.
template <typename T>
__device__ __forceinline__ T reciprocal(T value);
template <>
__device__ __forceinline__ float reciprocal<float>(float value)
{
    return __builtin_amdgcn_rcpf(value);
}
template <>
__device__ __forceinline__ double reciprocal<double>(double value)
{
    return 1.0 / value;
}

// Row of the unordered type pair (a, b): a(2T - a - 1)/2 + b for a <= b
// (potential/_lennard_jones/_potential.py:143-151).
__device__ __forceinline__ int pair_row(std::int64_t a, std::int64_t b, int num_types)
{
    const std::int64_t low = a < b ? a : b;
    const std::int64_t high = a < b ? b : a;
    return static_cast<int>(low * (2 * num_types - low - 1) / 2 + high);
}

// Per-row LJ coefficients, built once per evaluation by a prologue kernel.
template <typename T>
struct LJRow
{
    T sigma_squared;
    T four_eps;
    T twentyfour_eps;
    T shift;
};

template <typename T>
struct LJ
{
    const LJRow<T>* __restrict__ rows;
    const std::int64_t* __restrict__ types;
    int num_types;

    __device__ __forceinline__ int row(std::int64_t i, std::int64_t j) const
    {
        return num_types == 1 ? 0 : pair_row(types[i], types[j], num_types);
    }

    __device__ __forceinline__ void eval(int row_index, T r2, T& u, T& fs) const
    {
        const LJRow<T> c = rows[row_index];
        // 1/r^2 by the hardware reciprocal (1 ulp in fp32), as the MD kernels take it
        const T inverse = reciprocal(r2);
        const T s = c.sigma_squared * inverse;
        const T p6 = s * s * s;
        u = c.four_eps * (p6 * p6 - p6) - c.shift;
        fs = c.twentyfour_eps * (T(2) * p6 * p6 - p6) * inverse;
    }
};

// The padded table of a row: [K + 2] knots, the last real knot 0, virtual knots at both ends.
template <typename T>
struct Table
{
    const T* __restrict__ padded;        // [P, K + 2]
    const T* __restrict__ start;         // [P]
    const T* __restrict__ spacing;       // [P]
    const std::int64_t* __restrict__ types;
    int num_types;
    int knots;                           // K

    __device__ __forceinline__ int row(std::int64_t i, std::int64_t j) const
    {
        return num_types == 1 ? 0 : pair_row(types[i], types[j], num_types);
    }

    __device__ __forceinline__ void eval(int row_index, T r2, T& u, T& fs) const
    {
        // r and the knot coordinate with the reference's own operations, each correctly
        // rounded (sqrt, subtract, divide), so t and the segment are the reference's bits
        const T r = sqrt(r2);
        const T h = spacing[row_index];
        const T x = (r - start[row_index]) / h;
        T segment = floor(x);
        segment = segment < T(0) ? T(0) : (segment > T(knots - 2) ? T(knots - 2) : segment);
        const T t = x - segment;
        const T* k = padded + static_cast<std::int64_t>(row_index) * (knots + 2) +
                     static_cast<int>(segment);
        const T u0 = k[0], u1 = k[1], u2 = k[2], u3 = k[3];
        const T a = u2 - u0;
        const T b = T(2) * u0 - T(5) * u1 + T(4) * u2 - u3;
        const T c = T(3) * (u1 - u2) + u3 - u0;
        const T s = t > T(0) ? t : T(0);
        const T cubic = u1 + T(0.5) * s * (a + s * (b + s * c));
        const bool below = t < T(0);
        u = cubic + (below ? T(0.5) * a * t : T(0));
        const T du_dt = below ? T(0.5) * a : T(0.5) * (a + s * (T(2) * b + T(3) * s * c));
        // 1/h and 1/r only scale the force: hardware reciprocals
        fs = -(du_dt * reciprocal(h)) * reciprocal(r);
    }
};

}  // namespace ndn
