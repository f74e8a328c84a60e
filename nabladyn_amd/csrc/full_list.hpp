// The full-list force schedule: one wave per 64 rows of a sliced-ELL list, a lane per atom.
// Each lane sums its atom's force over all of its neighbors (both orientations are listed), so
// no force is written twice and nothing is atomic; energy and virial count every pair twice
// and are halved by the caller's finish.
//
// Periodic images: a wave whose atoms all sit farther than the cutoff (plus a margin) from
// every face takes x_i - x_j directly. That is the reference's minimum-image vector for every
// pair inside the cutoff, and any other image it might pick is longer than L/2 > rc, so the
// mask drops it exactly as the reference does. Other waves use the reference's displacement
// (space.py:11-19) per pair.
#pragma once

#include "common.hpp"
#include "pair_forces.hpp"
#include "potentials.hpp"

namespace ndn {

// Blocks in flight are dealt to the 8 XCDs round-robin; renumbering them so that consecutive
// logical blocks (neighboring rows, sharing neighbor positions) run on one XCD keeps those
// gathers in that XCD's L2. grid is a multiple of 8.
__device__ __forceinline__ unsigned xcd_block(unsigned block, unsigned grid)
{
    const unsigned per = grid / 8;
    return (block % 8) * per + block / 8;
}

inline unsigned xcd_grid(std::int64_t blocks)
{
    return static_cast<unsigned>((blocks + 7) / 8 * 8);
}

struct FullList
{
    const index_t* __restrict__ rows;    // row -> atom id, or null for the identity
    const index_t* __restrict__ counts;  // [N] neighbors per row
    const index_t* __restrict__ slices;  // [slices + 1] start of each slice (64 * width)
    const index_t* __restrict__ index;   // neighbor atom ids
    index_t num_atoms;
};

template <typename T>
struct RowResult
{
    T fx, fy, fz;
    T energy;
    T w[6];
};

// Sums one row's pair terms. `slice` is the row's slice (uniform over the wave).
template <typename T, class Pot, bool kVirial>
__device__ __forceinline__ RowResult<T> full_row(const T* __restrict__ positions, T lx, T ly,
                                                 T lz, T cutoff, T cutoff_squared,
                                                 const FullList& list, const Pot& pot,
                                                 std::int64_t slice, index_t row, index_t atom,
                                                 bool live)
{
    RowResult<T> out{};
    T xi = 0, yi = 0, zi = 0;
    index_t count = 0;
    if (live)
    {
        xi = positions[3LL * atom];
        yi = positions[3LL * atom + 1];
        zi = positions[3LL * atom + 2];
        count = list.counts[row];
    }
    const std::int64_t begin = list.slices[slice];
    const index_t width = static_cast<index_t>((list.slices[slice + 1] - begin) / kWave);
    const int lane = row % kWave;
    const T margin = T(1e-3);
    const T low = cutoff + margin;
    const bool interior = !live || (xi > low && xi < lx - low && yi > low && yi < ly - low &&
                                    zi > low && zi < lz - low);
    const bool fast = ballot(!interior) == 0ull;
    const index_t* __restrict__ column = list.index + begin + lane;
    if (fast)
    {
        for (index_t k = 0; k < width; ++k)
        {
            if (k >= count) continue;
            const index_t j = column[static_cast<std::int64_t>(k) * kWave];
            const T dx = xi - positions[3LL * j];
            const T dy = yi - positions[3LL * j + 1];
            const T dz = zi - positions[3LL * j + 2];
            const T r2 = exact_squared(dx, dy, dz);
            if (!(r2 < cutoff_squared)) continue;
            T u, fs;
            pot.eval(pot.row(atom, j), r2, u, fs);
            out.energy += u;
            out.fx += fs * dx;
            out.fy += fs * dy;
            out.fz += fs * dz;
            if (kVirial)
            {
                out.w[0] += dx * (fs * dx);
                out.w[1] += dy * (fs * dy);
                out.w[2] += dz * (fs * dz);
                out.w[3] += dx * (fs * dy);
                out.w[4] += dx * (fs * dz);
                out.w[5] += dy * (fs * dz);
            }
        }
    }
    else
    {
        for (index_t k = 0; k < width; ++k)
        {
            if (k >= count) continue;
            const index_t j = column[static_cast<std::int64_t>(k) * kWave];
            const T dx = displacement(xi, positions[3LL * j], lx);
            const T dy = displacement(yi, positions[3LL * j + 1], ly);
            const T dz = displacement(zi, positions[3LL * j + 2], lz);
            const T r2 = exact_squared(dx, dy, dz);
            if (!(r2 < cutoff_squared)) continue;
            T u, fs;
            pot.eval(pot.row(atom, j), r2, u, fs);
            out.energy += u;
            out.fx += fs * dx;
            out.fy += fs * dy;
            out.fz += fs * dz;
            if (kVirial)
            {
                out.w[0] += dx * (fs * dx);
                out.w[1] += dy * (fs * dy);
                out.w[2] += dz * (fs * dz);
                out.w[3] += dx * (fs * dy);
                out.w[4] += dx * (fs * dz);
                out.w[5] += dy * (fs * dz);
            }
        }
    }
    return out;
}

}  // namespace ndn
