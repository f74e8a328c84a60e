// Device helpers shared by the pair-force kernels: exact geometry, the potential view, and the
// per-block partial sums of energy and virial.
#pragma once

#include "common.hpp"
#include "potentials.hpp"

namespace ndn {

// Partial slots every workspace reserves at least: enough for any fixed-grid launch here.
constexpr int kPartialSlots = 4096;

// center - other by minimum image with the near-face coordinate shifted first
// (space.py:11-19): (left - L) - right when raw > L/2, left - (right - L) when raw < -L/2.
template <typename T>
__device__ __forceinline__ T displacement(T left, T right, T length)
{
#pragma clang fp contract(off)
    const T raw = left - right;
    const T half = T(0.5) * length;
    const T left_shift = raw > half ? length : T(0);
    const T right_shift = raw < -half ? length : T(0);
    return (left - left_shift) - (right - right_shift);
}

// (dx dx + dy dy) + dz dz, each product and sum rounded on its own.
template <typename T>
__device__ __forceinline__ T exact_squared(T dx, T dy, T dz)
{
#pragma clang fp contract(off)
    const T xx = dx * dx;
    const T yy = dy * dy;
    const T zz = dz * dz;
    return (xx + yy) + zz;
}

template <typename T>
struct PotentialView
{
    int kind;
    T cutoff_squared;
    LJ<T> lj;
    Table<T> table;
};

// Writes this block's energy and virial sums to partials[8 * blockIdx.x ...].
template <bool kVirial, typename T>
__device__ __forceinline__ void store_partials(double energy, const T* w, double* shared,
                                               double* partials)
{
    const double e = block_sum(energy, shared);
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (kVirial)
        for (int k = 0; k < 6; ++k) v[k] = block_sum(static_cast<double>(w[k]), shared);
    if (threadIdx.x == 0)
    {
        double* out = partials + 8LL * blockIdx.x;
        out[0] = e;
        for (int k = 0; k < 6; ++k) out[1 + k] = v[k];
        out[7] = 0.0;
    }
}

std::int64_t coefficient_bytes(const ndn_potential* pot, int element);
int check_potential(const ndn_potential* pot, std::int64_t num_atoms);

template <typename T>
int prepare_potential(const ndn_potential* pot, void* workspace, hipStream_t stream,
                      PotentialView<T>* view);

template <typename T>
int finish(const double* partials, int blocks, void* energy, void* virial, double scale,
           hipStream_t stream);

}  // namespace ndn
