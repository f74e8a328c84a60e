// Shared pieces of the neighbor list for other translation units: the exact geometry helpers,
// the reuse-rule preparation as a device function, and the launch sequences of a build.
#pragma once

#include <limits>

#include "common.hpp"
#include "scan.hpp"

namespace ndn {

template <typename T>
struct Epsilon
{
    static constexpr T value = std::numeric_limits<T>::epsilon();
};

// center - other by minimum image, the near-face coordinate shifted first
// (neighbors/csrc/common/traversal.cuh:31-38).
template <typename T>
__device__ __forceinline__ T minimum_image(T center, T other, T length)
{
#pragma clang fp contract(off)
    const T raw = center - other;
    return raw > T(0.5) * length ? (center - length) - other
         : raw < T(-0.5) * length ? center - (other - length)
         : raw;
}

// The update's preparation (update/update.cuh:27-75): resets the rebuild flag, checks the
// cell, chooses the grid (at most `limits` bins per axis, each wider than the radius) and sets
// the squared reuse threshold and per-axis scales; records the previous pair total for the
// padding. One thread. `reset_moved` clears the check's record of the largest displacement.
template <typename T>
__device__ __forceinline__ void prepare_update(const T* __restrict__ cell,
                                               const T* __restrict__ reference_cell, T cutoff,
                                               T radius, const index_t* __restrict__ limits,
                                               index_t* dims, T* widths, T* reuse,
                                               index_t* rebuild, index_t* status,
                                               const index_t* pair_total, ListHeader* header,
                                               bool reset_moved = true)
{
#pragma clang fp contract(off)
    const T epsilon = Epsilon<T>::value;
    header->old_total = pair_total[0];
    if (reset_moved)
        for (int k = 0; k < kMovedSlots; ++k) *moved_slot(header, k) = 0ull;
    rebuild[0] = 0;
    if (status[0] != 0) return;
    T smallest_scale = cell[0] / reference_cell[0];
    T longest = cell[0];
    for (int axis = 0; axis < 3; ++axis)
    {
        const T length = cell[axis];
        if (!(isfinite(length) && length > 2 * cutoff))
        {
            status[0] |= kStatusInvalidCell;
            return;
        }
        const T scale = length / reference_cell[axis];
        reuse[1 + axis] = scale;
        smallest_scale = scale < smallest_scale ? scale : smallest_scale;
        longest = length > longest ? length : longest;
        const T fit = floor(length / (radius + 4 * epsilon * length));
        const index_t count =
            fit < 1 ? 1 : (fit > limits[axis] ? limits[axis] : static_cast<index_t>(fit));
        dims[axis] = count;
        widths[axis] = length / count;
    }
    const T allowance = radius * smallest_scale - cutoff - 32 * epsilon * longest;
    if (allowance <= 0)
    {
        rebuild[0] = 1;
        return;
    }
    reuse[0] = allowance * allowance / 4;
}

// Squared displacement of one atom from its position at the last rebuild, scaled with the
// cell: ((0 + dx^2) + dy^2) + dz^2 (update/update.cuh:104-110).
template <typename T>
__device__ __forceinline__ T moved_squared(T x, T y, T z, const T* __restrict__ x0,
                                           const T* __restrict__ reuse, const T* __restrict__ cell)
{
#pragma clang fp contract(off)
    const T dx = minimum_image(x, reuse[1] * x0[0], cell[0]);
    const T dy = minimum_image(y, reuse[2] * x0[1], cell[1]);
    const T dz = minimum_image(z, reuse[3] * x0[2], cell[2]);
    T moved = T(0) + dx * dx;
    moved = moved + dy * dy;
    return moved + dz * dz;
}

template <typename T>
__device__ __forceinline__ unsigned long long order_bits(T value);
template <>
__device__ __forceinline__ unsigned long long order_bits<float>(float value)
{
    return static_cast<unsigned long long>(__float_as_uint(value));
}
template <>
__device__ __forceinline__ unsigned long long order_bits<double>(double value)
{
    return static_cast<unsigned long long>(__double_as_longlong(value));
}

// The rebuild after the reuse decision: grid, count, scans, fill, total and padding, commit.
// Every kernel is gated on the device flag. Instantiated for float and double in
// neighbors.hip.
template <typename T>
int launch_rebuild(const T* positions, const T* cell, double radius, const ndn_neighbor_buffers& b,
                   T* reference_positions, T* reference_cell, index_t* rebuild_count,
                   const index_t* rebuild, hipStream_t stream);

}  // namespace ndn
