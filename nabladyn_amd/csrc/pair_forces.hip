// Pair energy, forces and virial on gfx950, for LJ and tabulated potentials in fp32 and fp64.
//
// Two schedules share the potential functors and the exact cutoff mask:
// - half list [2, C] (any source, the reference's NeighborPairs.index): a lane per column; the
//   second atom's force added with hardware float atomics, the first atom's summed across each
//   run of consecutive columns with the same first atom (the drop-in list writes each atom's
//   pairs together) and added once per run. It serves quantity.* and the plain per-system path.
// - full list (sliced ELL, rows of 64 in one wave): a lane per atom, forces summed in registers
//   with no atomics and no re-reading; used by the MD step and the batched path.
// Energy and virial are summed per block into fp64 partials, then by one block in a fixed
// order, so they do not depend on scheduling.
//
// Mask and geometry follow the reference exactly: valid = (i != j) & (r^2 < rc^2), r^2 the
// rounded (dx dx + dy dy) + dz dz of the minimum-image vector with the near-face coordinate
// shifted first (space.py:11-19; _lennard_jones/_potential.py:218-223). Only the distance
// test is kept free of contraction; the energy and force arithmetic may fuse.

#include <algorithm>
#include <climits>

#include "common.hpp"
#include "pair_forces.hpp"

#include <mutex>
#include <string>
#include "potentials.hpp"

namespace ndn {

// ---- prologues ------------------------------------------------------------------------------

template <typename T>
__global__ void lj_rows_kernel(const T* __restrict__ sigma, const T* __restrict__ epsilon,
                               int rows, T cutoff_squared, LJRow<T>* __restrict__ out)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const T s2 = sigma[row] * sigma[row];
    const T e = epsilon[row];
    const T base = s2 / cutoff_squared;
    const T p6 = base * base * base;
    out[row] = LJRow<T>{s2, T(4) * e, T(24) * e, T(4) * e * (p6 * p6 - p6)};
}

// [P, K + 2] padded knots and the spacing (potential/_tabulated/_potential.py:79-82).
template <typename T>
__global__ void table_pad_kernel(const T* __restrict__ table, const T* __restrict__ spacing,
                                 int rows, int knots, T* __restrict__ padded,
                                 T* __restrict__ spacing_out)
{
    const std::int64_t index = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t width = knots + 2;
    if (index < rows) spacing_out[index] = spacing[index];
    if (index >= rows * width) return;
    const std::int64_t row = index / width;
    const std::int64_t column = index % width;
    const T* v = table + row * knots;
    auto value = [&](std::int64_t k) { return k == knots - 1 ? T(0) : v[k]; };
    T out;
    if (column == 0)
        out = T(3) * value(0) - T(3) * value(1) + value(2);
    else if (column == width - 1)
        out = T(3) * value(knots - 1) - T(3) * value(knots - 2) + value(knots - 3);
    else
        out = value(column - 1);
    padded[index] = out;
}

template <typename T>
__global__ void zero_kernel(T* __restrict__ data, std::int64_t count)
{
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < count; i += stride)
        data[i] = T(0);
}

// ---- half-list schedule ---------------------------------------------------------------------

// The sum of `value` over each run of lanes with the same key (a run starts where the key
// changes), left in the run's last lane; `start` is the run's first lane. Equal keys apart are
// separate runs.
template <typename T>
__device__ __forceinline__ T run_sum(T value, int lane, int start)
{
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1)
    {
        const T other = __shfl_up(value, offset, kWave);
        if (lane - offset >= start) value += other;
    }
    return value;
}

//
// Frames of one system (training stacks) take blockIdx.y as the frame: its positions [N, 3],
// cell [3], pairs [2, C], forces [N, 3] and the grid's partial slots follow the previous frame's.
template <typename T, class Pot, bool kForces, bool kVirial>
__global__ void __launch_bounds__(kBlock) half_list_kernel(
    const T* __restrict__ positions, const T* __restrict__ cell,
    const index_t* __restrict__ pairs, std::int64_t capacity, index_t num_atoms,
    T cutoff_squared, Pot pot, T* __restrict__ forces, double* __restrict__ partials)
{
    __shared__ double shared[16];
    const std::int64_t frame = blockIdx.y;
    positions += 3 * frame * num_atoms;
    cell += 3 * frame;
    pairs += 2 * frame * capacity;
    if (kForces) forces += 3 * frame * num_atoms;
    partials += 8 * frame * gridDim.x;
    const T lx = cell[0], ly = cell[1], lz = cell[2];
    const int lane = lane_id();
    T energy = 0;
    T w[6] = {0, 0, 0, 0, 0, 0};
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    // the wave walks consecutive columns together (every lane takes part in the run sums); a
    // list that keeps each atom's pairs in consecutive columns, as the drop-in list does, has
    // long runs of the same first atom, whose force is summed across the run first and added
    // with one atomic per run
    for (std::int64_t first =
             static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x - lane;
         first < capacity; first += stride)
    {
        const std::int64_t column = first + lane;
        index_t i = -1, j = -1;
        if (column < capacity)
        {
            i = pairs[column];
            j = pairs[capacity + column];
        }
        // self-pairs are padding; an index outside [0, N) is skipped rather than read
        const bool listed = i != j && static_cast<std::uint32_t>(i) < static_cast<std::uint32_t>(num_atoms) &&
                            static_cast<std::uint32_t>(j) < static_cast<std::uint32_t>(num_atoms);
        T fx = 0, fy = 0, fz = 0;
        bool inside = false;
        if (listed)
        {
            const T* p = positions + 3LL * i;
            const T* q = positions + 3LL * j;
            const T dx = displacement(p[0], q[0], lx);
            const T dy = displacement(p[1], q[1], ly);
            const T dz = displacement(p[2], q[2], lz);
            const T r2 = exact_squared(dx, dy, dz);
            inside = r2 < cutoff_squared;
            if (inside)
            {
                T u, fs;
                pot.eval(pot.row(i, j), r2, u, fs);
                energy += u;
                fx = fs * dx;
                fy = fs * dy;
                fz = fs * dz;
                if (kVirial)
                {
                    w[0] += dx * fx;
                    w[1] += dy * fy;
                    w[2] += dz * fz;
                    w[3] += dx * fy;
                    w[4] += dx * fz;
                    w[5] += dy * fz;
                }
            }
        }
        if (kForces)
        {
            if (inside)
            {
                unsafeAtomicAdd(forces + 3LL * j, -fx);
                unsafeAtomicAdd(forces + 3LL * j + 1, -fy);
                unsafeAtomicAdd(forces + 3LL * j + 2, -fz);
            }
            // runs of the same first atom (padding and unlisted columns key -1)
            const index_t key = listed ? i : -1;
            const index_t before = __shfl_up(key, 1, kWave);
            const bool head = lane == 0 || before != key;
            int start = head ? lane : 0;
#pragma unroll
            for (int offset = 1; offset < kWave; offset <<= 1)
            {
                const int other = __shfl_up(start, offset, kWave);
                if (lane >= offset && other > start) start = other;
            }
            const T sx = run_sum(fx, lane, start);
            const T sy = run_sum(fy, lane, start);
            const T sz = run_sum(fz, lane, start);
            const int next_start = __shfl_down(start, 1, kWave);
            const bool tail = lane == kWave - 1 || next_start != start;
            if (tail && key >= 0)
            {
                unsafeAtomicAdd(forces + 3LL * key, sx);
                unsafeAtomicAdd(forces + 3LL * key + 1, sy);
                unsafeAtomicAdd(forces + 3LL * key + 2, sz);
            }
        }
    }
    store_partials<kVirial>(static_cast<double>(energy), w, shared, partials);
}

// ---- full-row schedule ----------------------------------------------------------------------

// A full list: every pair within the list radius in both orientations, sliced by 64 rows: row
// r's k-th neighbor (an atom id) at index[slices[r / 64] + r % 64 + 64 k], its count counts[r];
// row r is atom rows[r] (rows null: atom r). A lane per row walks it, sums the row's force in
// registers and writes it once: no atomics. Each pair is met twice, so E and W take half of
// each meeting. A slot at or past the capacity (an overflowed build) and an id outside [0, N)
// are skipped rather than read.
//
// The walk takes kAhead entries per turn: their ids, then their positions, are all read before
// any is used (every lane reads, a lane past its row from a harmless address), so a turn waits
// for memory once rather than once per entry; each entry's terms are then computed on a safe
// distance and added only for a pair inside the cutoff, in the row's order.
template <typename T, class Pot>
constexpr int full_rows_ahead()
{
    return sizeof(T) == 4 ? 8 : 4;
}

template <typename T, class Pot, bool kForces, bool kVirial>
__global__ void __launch_bounds__(kBlock) full_rows_kernel(
    const T* __restrict__ positions, const T* __restrict__ cell, const index_t* __restrict__ rows,
    const index_t* __restrict__ counts, const index_t* __restrict__ slices,
    const index_t* __restrict__ index, std::int64_t capacity, index_t num_atoms,
    T cutoff_squared, Pot pot, T* __restrict__ forces, double* __restrict__ partials)
{
    constexpr int kAhead = full_rows_ahead<T, Pot>();
    __shared__ double shared[16];
    const std::int64_t slices_total = (static_cast<std::int64_t>(num_atoms) + kWave - 1) / kWave;
    // frames of one system (blockIdx.y): each frame's arrays follow the previous frame's (rows
    // [N], counts [N], slices [N / 64 + 1], `capacity` index entries)
    const std::int64_t frame = blockIdx.y;
    positions += 3 * frame * num_atoms;
    cell += 3 * frame;
    if (rows != nullptr) rows += frame * num_atoms;
    counts += frame * num_atoms;
    slices += frame * (slices_total + 1);
    index += frame * capacity;
    if (kForces) forces += 3 * frame * num_atoms;
    partials += 8 * frame * gridDim.x;
    const T lx = cell[0], ly = cell[1], lz = cell[2];
    // a distance inside every potential's domain, for the terms of entries that are not pairs
    const T safe_r2 = T(0.25) * cutoff_squared;
    const int lane = lane_id();
    T energy = 0;
    T w[6] = {0, 0, 0, 0, 0, 0};
    const std::int64_t waves = static_cast<std::int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (std::int64_t slice =
             static_cast<std::int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
         slice < slices_total; slice += waves)
    {
        const std::int64_t row = slice * kWave + lane;
        const bool live = row < num_atoms;
        const index_t i = live ? (rows != nullptr ? rows[row] : static_cast<index_t>(row)) : 0;
        const index_t count = live ? counts[row] : 0;
        const std::int64_t base = live ? static_cast<std::int64_t>(slices[slice]) + lane : 0;
        const bool own = live && static_cast<std::uint32_t>(i) < static_cast<std::uint32_t>(num_atoms);
        const index_t me = own ? i : 0;
        // the entries this lane walks: its count, cut where its slots reach the capacity
        std::int64_t walk = own && base < capacity ? count : 0;
        if (walk > 0)
        {
            const std::int64_t fit = (capacity - base + kWave - 1) / kWave;
            walk = walk < fit ? walk : fit;
        }
        const T* p = positions + 3LL * me;
        const T px = p[0], py = p[1], pz = p[2];
        T fx = 0, fy = 0, fz = 0;
        for (std::int64_t k0 = 0; ballot(k0 < walk) != 0; k0 += kAhead)
        {
            index_t j[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; ++u)
            {
                const bool take = k0 + u < walk;
                const index_t v = index[take ? base + (k0 + u) * kWave : 0];
                j[u] = take ? v : me;
            }
            bool valid[kAhead];
            T qx[kAhead], qy[kAhead], qz[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; ++u)
            {
                valid[u] = j[u] != i &&
                           static_cast<std::uint32_t>(j[u]) < static_cast<std::uint32_t>(num_atoms);
                const T* q = positions + 3LL * (valid[u] ? j[u] : me);
                qx[u] = q[0];
                qy[u] = q[1];
                qz[u] = q[2];
            }
#pragma unroll
            for (int u = 0; u < kAhead; ++u)
            {
                const T dx = displacement(px, qx[u], lx);
                const T dy = displacement(py, qy[u], ly);
                const T dz = displacement(pz, qz[u], lz);
                const T r2 = exact_squared(dx, dy, dz);
                const bool inside = valid[u] && r2 < cutoff_squared;
                T e, fs;
                pot.eval(pot.row(me, valid[u] ? j[u] : me), inside ? r2 : safe_r2, e, fs);
                const T gx = fs * dx, gy = fs * dy, gz = fs * dz;
                energy = inside ? energy + T(0.5) * e : energy;
                fx = inside ? fx + gx : fx;
                fy = inside ? fy + gy : fy;
                fz = inside ? fz + gz : fz;
                if (kVirial)
                {
                    w[0] = inside ? w[0] + T(0.5) * dx * gx : w[0];
                    w[1] = inside ? w[1] + T(0.5) * dy * gy : w[1];
                    w[2] = inside ? w[2] + T(0.5) * dz * gz : w[2];
                    w[3] = inside ? w[3] + T(0.5) * dx * gy : w[3];
                    w[4] = inside ? w[4] + T(0.5) * dx * gz : w[4];
                    w[5] = inside ? w[5] + T(0.5) * dy * gz : w[5];
                }
            }
        }
        if (kForces && own)
        {
            forces[3LL * i] = fx;
            forces[3LL * i + 1] = fy;
            forces[3LL * i + 2] = fz;
        }
    }
    store_partials<kVirial>(static_cast<double>(energy), w, shared, partials);
}

// ---- final reduction ------------------------------------------------------------------------

// Sums the per-block partials (8 doubles each: energy, then xx yy zz xy xz yz) in a fixed order
// and writes E [1] and W [3, 3] (symmetric) in T, each when its pointer is given. Block f sums
// frame f's `blocks` slots into E [f] and W [f].
template <typename T>
__global__ void __launch_bounds__(kBlock) finish_kernel(const double* __restrict__ partials,
                                                        int blocks, T* energy, T* virial,
                                                        double scale)
{
    __shared__ double shared[16];
    const std::int64_t frame = blockIdx.x;
    partials += 8 * frame * blocks;
    if (energy) energy += frame;
    if (virial) virial += 9 * frame;
    double sums[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < blocks; b += blockDim.x)
        for (int k = 0; k < 7; ++k) sums[k] += partials[8LL * b + k];
    for (int k = 0; k < 7; ++k) sums[k] = block_sum(sums[k], shared);
    if (threadIdx.x == 0)
    {
        if (energy) energy[0] = static_cast<T>(scale * sums[0]);
        if (virial)
        {
            const double xx = scale * sums[1], yy = scale * sums[2], zz = scale * sums[3];
            const double xy = scale * sums[4], xz = scale * sums[5], yz = scale * sums[6];
            virial[0] = static_cast<T>(xx);
            virial[1] = static_cast<T>(xy);
            virial[2] = static_cast<T>(xz);
            virial[3] = static_cast<T>(xy);
            virial[4] = static_cast<T>(yy);
            virial[5] = static_cast<T>(yz);
            virial[6] = static_cast<T>(xz);
            virial[7] = static_cast<T>(yz);
            virial[8] = static_cast<T>(zz);
        }
    }
}

// ---- host side ------------------------------------------------------------------------------

int check_potential(const ndn_potential* pot, std::int64_t num_atoms)
{
    NDN_REQUIRE(pot != nullptr, "the potential must not be null");
    NDN_REQUIRE(pot->next == nullptr,
                "a sum of potentials (next) is taken only by ndn_md_workspace, ndn_md_step, "
                "ndn_md_evaluate, ndn_md_refresh and ndn_md_force");
    NDN_REQUIRE(pot->num_types >= 1, "num_types must be at least 1");
    NDN_REQUIRE(pot->num_types == 1 || pot->atom_types != nullptr,
                "atom_types is required with more than one type");
    NDN_REQUIRE(pot->cutoff > 0, "the cutoff must be positive");
    if (pot->kind == NDN_LJ)
    {
        NDN_REQUIRE(pot->sigma && pot->epsilon, "LJ needs sigma and epsilon");
    }
    else
    {
        NDN_REQUIRE(pot->kind == NDN_TABLE, "kind must be NDN_LJ or NDN_TABLE");
        NDN_REQUIRE(pot->table && pot->start && pot->spacing && pot->knots >= 4,
                    "a table needs table, start, spacing and at least 4 knots");
    }
    (void)num_atoms;
    return NDN_OK;
}

std::int64_t coefficient_bytes(const ndn_potential* pot, int element)
{
    const std::int64_t rows = static_cast<std::int64_t>(pot->num_types) * (pot->num_types + 1) / 2;
    std::int64_t bytes;
    if (pot->kind == NDN_LJ)
        bytes = rows * 4 * element;
    else
        bytes = rows * (pot->knots + 2) * element + rows * element;
    return (bytes + 255) / 256 * 256;
}

template <typename T>
int prepare_potential(const ndn_potential* pot, void* workspace, hipStream_t stream,
                      PotentialView<T>* view)
{
    const int rows = pot->num_types * (pot->num_types + 1) / 2;
    const T cutoff_squared = static_cast<T>(pot->cutoff * pot->cutoff);
    view->kind = pot->kind;
    view->cutoff_squared = cutoff_squared;
    if (pot->kind == NDN_LJ)
    {
        auto* out = static_cast<LJRow<T>*>(workspace);
        lj_rows_kernel<T><<<blocks_for(rows), kBlock, 0, stream>>>(
            static_cast<const T*>(pot->sigma), static_cast<const T*>(pot->epsilon), rows,
            cutoff_squared, out);
        NDN_LAUNCHED("lj_rows");
        view->lj = LJ<T>{out, pot->atom_types, pot->num_types};
    }
    else
    {
        const int knots = static_cast<int>(pot->knots);
        T* padded = static_cast<T*>(workspace);
        T* spacing = padded + static_cast<std::int64_t>(rows) * (knots + 2);
        table_pad_kernel<T><<<blocks_for(static_cast<std::int64_t>(rows) * (knots + 2)), kBlock, 0,
                              stream>>>(static_cast<const T*>(pot->table),
                                        static_cast<const T*>(pot->spacing), rows, knots, padded,
                                        spacing);
        NDN_LAUNCHED("table_pad");
        view->table = Table<T>{padded, static_cast<const T*>(pot->start), spacing,
                               pot->atom_types, pot->num_types, knots};
    }
    return NDN_OK;
}

template int prepare_potential<float>(const ndn_potential*, void*, hipStream_t,
                                      PotentialView<float>*);
template int prepare_potential<double>(const ndn_potential*, void*, hipStream_t,
                                       PotentialView<double>*);

constexpr int kHalfBlocks = 2048;
constexpr std::int64_t kMaxFrames = 65535;  // the grid's y extent

int half_blocks(std::int64_t capacity)
{
    const std::int64_t blocks = (capacity + kBlock - 1) / kBlock;
    return static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(blocks, kHalfBlocks)));
}

template <typename T>
int finish(const double* partials, int blocks, void* energy, void* virial, double scale,
           hipStream_t stream)
{
    if (energy == nullptr && virial == nullptr) return NDN_OK;
    finish_kernel<T><<<1, kBlock, 0, stream>>>(partials, blocks, static_cast<T*>(energy),
                                                static_cast<T*>(virial), scale);
    NDN_LAUNCHED("finish");
    return NDN_OK;
}

template int finish<float>(const double*, int, void*, void*, double, hipStream_t);
template int finish<double>(const double*, int, void*, void*, double, hipStream_t);

template <typename T, class Pot>
void launch_half(const T* x, const T* cell, const index_t* pairs, std::int64_t capacity,
                 index_t num_atoms, T cutoff_squared, const Pot& pot, T* forces, bool virial,
                 double* partials, int columns_blocks, hipStream_t stream, int frames = 1)
{
    const dim3 blocks(static_cast<unsigned>(columns_blocks), static_cast<unsigned>(frames));
    if (forces && virial)
        half_list_kernel<T, Pot, true, true><<<blocks, kBlock, 0, stream>>>(
            x, cell, pairs, capacity, num_atoms, cutoff_squared, pot, forces, partials);
    else if (forces)
        half_list_kernel<T, Pot, true, false><<<blocks, kBlock, 0, stream>>>(
            x, cell, pairs, capacity, num_atoms, cutoff_squared, pot, forces, partials);
    else if (virial)
        half_list_kernel<T, Pot, false, true><<<blocks, kBlock, 0, stream>>>(
            x, cell, pairs, capacity, num_atoms, cutoff_squared, pot, forces, partials);
    else
        half_list_kernel<T, Pot, false, false><<<blocks, kBlock, 0, stream>>>(
            x, cell, pairs, capacity, num_atoms, cutoff_squared, pot, forces, partials);
}

template <typename T, class Pot>
void launch_full(const T* x, const T* cell, const index_t* rows, const index_t* counts,
                 const index_t* slices, const index_t* index, std::int64_t capacity,
                 index_t num_atoms, T cutoff_squared, const Pot& pot, T* forces, bool virial,
                 double* partials, int row_blocks, hipStream_t stream, int frames = 1)
{
    const dim3 blocks(static_cast<unsigned>(row_blocks), static_cast<unsigned>(frames));
    if (forces && virial)
        full_rows_kernel<T, Pot, true, true><<<blocks, kBlock, 0, stream>>>(
            x, cell, rows, counts, slices, index, capacity, num_atoms, cutoff_squared, pot,
            forces, partials);
    else if (forces)
        full_rows_kernel<T, Pot, true, false><<<blocks, kBlock, 0, stream>>>(
            x, cell, rows, counts, slices, index, capacity, num_atoms, cutoff_squared, pot,
            forces, partials);
    else if (virial)
        full_rows_kernel<T, Pot, false, true><<<blocks, kBlock, 0, stream>>>(
            x, cell, rows, counts, slices, index, capacity, num_atoms, cutoff_squared, pot,
            forces, partials);
    else
        full_rows_kernel<T, Pot, false, false><<<blocks, kBlock, 0, stream>>>(
            x, cell, rows, counts, slices, index, capacity, num_atoms, cutoff_squared, pot,
            forces, partials);
}

// Blocks of the full-row launch: a wave per 64 rows, at most kHalfBlocks blocks striding.
int full_blocks(std::int64_t num_atoms)
{
    const std::int64_t slices = (num_atoms + kWave - 1) / kWave;
    const std::int64_t blocks = (slices + kBlock / kWave - 1) / (kBlock / kWave);
    return static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(blocks, kHalfBlocks)));
}

template <typename T>
int pair_efw_full_impl(const void* positions, const void* cell, std::int64_t num_atoms,
                       const index_t* rows, const index_t* counts, const index_t* slices,
                       const index_t* index, std::int64_t capacity, const ndn_potential* pot,
                       void* energy, void* forces, void* virial, void* workspace,
                       hipStream_t stream)
{
    PotentialView<T> view;
    const int prepared = prepare_potential<T>(pot, workspace, stream, &view);
    if (prepared != NDN_OK) return prepared;
    const int blocks = full_blocks(num_atoms);
    auto* partials = reinterpret_cast<double*>(static_cast<std::uint8_t*>(workspace) +
                                               coefficient_bytes(pot, sizeof(T)));
    T* f = static_cast<T*>(forces);
    const T* x = static_cast<const T*>(positions);
    const T* L = static_cast<const T*>(cell);
    const index_t n = static_cast<index_t>(num_atoms);
    if (view.kind == NDN_LJ)
        launch_full<T>(x, L, rows, counts, slices, index, capacity, n, view.cutoff_squared,
                       view.lj, f, virial != nullptr, partials, blocks, stream);
    else
        launch_full<T>(x, L, rows, counts, slices, index, capacity, n, view.cutoff_squared,
                       view.table, f, virial != nullptr, partials, blocks, stream);
    NDN_LAUNCHED("full_rows");
    return finish<T>(partials, blocks, energy, virial, 1.0, stream);
}

// Frames of one system over full rows built from their pair stacks: one launch of F x blocks,
// one finish block per frame.
template <typename T>
int pair_efw_full_frames_impl(const void* positions, const void* cells, std::int64_t frames,
                              std::int64_t num_atoms, const index_t* counts,
                              const index_t* slices, const index_t* index,
                              std::int64_t capacity, const ndn_potential* pot, void* energy,
                              void* forces, void* virial, void* workspace, hipStream_t stream)
{
    PotentialView<T> view;
    const int prepared = prepare_potential<T>(pot, workspace, stream, &view);
    if (prepared != NDN_OK) return prepared;
    const int blocks = full_blocks(num_atoms);
    auto* partials = reinterpret_cast<double*>(static_cast<std::uint8_t*>(workspace) +
                                               coefficient_bytes(pot, sizeof(T)));
    T* f = static_cast<T*>(forces);
    const T* x = static_cast<const T*>(positions);
    const T* L = static_cast<const T*>(cells);
    const index_t n = static_cast<index_t>(num_atoms);
    const int count = static_cast<int>(frames);
    if (view.kind == NDN_LJ)
        launch_full<T>(x, L, nullptr, counts, slices, index, capacity, n, view.cutoff_squared,
                       view.lj, f, virial != nullptr, partials, blocks, stream, count);
    else
        launch_full<T>(x, L, nullptr, counts, slices, index, capacity, n, view.cutoff_squared,
                       view.table, f, virial != nullptr, partials, blocks, stream, count);
    NDN_LAUNCHED("full_rows_frames");
    if (energy == nullptr && virial == nullptr) return NDN_OK;
    finish_kernel<T><<<count, kBlock, 0, stream>>>(partials, blocks, static_cast<T*>(energy),
                                                   static_cast<T*>(virial), 1.0);
    NDN_LAUNCHED("finish_frames");
    return NDN_OK;
}

template <typename T>
int pair_efw_impl(const void* positions, const void* cell, std::int64_t num_atoms,
                  const index_t* pairs, std::int64_t capacity, const ndn_potential* pot,
                  void* energy, void* forces, void* virial, void* workspace, hipStream_t stream)
{
    PotentialView<T> view;
    const int prepared = prepare_potential<T>(pot, workspace, stream, &view);
    if (prepared != NDN_OK) return prepared;
    const int blocks = half_blocks(capacity);
    auto* partials = reinterpret_cast<double*>(static_cast<std::uint8_t*>(workspace) +
                                               coefficient_bytes(pot, sizeof(T)));
    T* f = static_cast<T*>(forces);
    if (f != nullptr)
    {
        zero_kernel<T><<<std::min<unsigned>(blocks_for(3 * num_atoms), 1024u), kBlock, 0,
                         stream>>>(f, 3 * num_atoms);
        NDN_LAUNCHED("zero_forces");
    }
    const T* x = static_cast<const T*>(positions);
    const T* L = static_cast<const T*>(cell);
    if (view.kind == NDN_LJ)
        launch_half<T>(x, L, pairs, capacity, static_cast<index_t>(num_atoms), view.cutoff_squared,
                       view.lj, f, virial != nullptr, partials, blocks, stream);
    else
        launch_half<T>(x, L, pairs, capacity, static_cast<index_t>(num_atoms), view.cutoff_squared,
                       view.table, f, virial != nullptr, partials, blocks, stream);
    NDN_LAUNCHED("half_list");
    return finish<T>(partials, blocks, energy, virial, 1.0, stream);
}

// Frames of one system over their pair stacks [F, 2, C]: one launch of F x blocks, one finish
// block per frame.
template <typename T>
int pair_efw_frames_impl(const void* positions, const void* cells, std::int64_t frames,
                         std::int64_t num_atoms, const index_t* stacks, std::int64_t capacity,
                         const ndn_potential* pot, void* energy, void* forces, void* virial,
                         void* workspace, hipStream_t stream)
{
    PotentialView<T> view;
    const int prepared = prepare_potential<T>(pot, workspace, stream, &view);
    if (prepared != NDN_OK) return prepared;
    const int blocks = half_blocks(capacity);
    auto* partials = reinterpret_cast<double*>(static_cast<std::uint8_t*>(workspace) +
                                               coefficient_bytes(pot, sizeof(T)));
    T* f = static_cast<T*>(forces);
    if (f != nullptr)
    {
        zero_kernel<T><<<std::min<unsigned>(blocks_for(3 * frames * num_atoms), 4096u), kBlock,
                         0, stream>>>(f, 3 * frames * num_atoms);
        NDN_LAUNCHED("zero_forces");
    }
    const T* x = static_cast<const T*>(positions);
    const T* L = static_cast<const T*>(cells);
    const index_t n = static_cast<index_t>(num_atoms);
    const int count = static_cast<int>(frames);
    if (view.kind == NDN_LJ)
        launch_half<T>(x, L, stacks, capacity, n, view.cutoff_squared, view.lj, f,
                       virial != nullptr, partials, blocks, stream, count);
    else
        launch_half<T>(x, L, stacks, capacity, n, view.cutoff_squared, view.table, f,
                       virial != nullptr, partials, blocks, stream, count);
    NDN_LAUNCHED("half_list_frames");
    if (energy == nullptr && virial == nullptr) return NDN_OK;
    finish_kernel<T><<<count, kBlock, 0, stream>>>(partials, blocks, static_cast<T*>(energy),
                                                   static_cast<T*>(virial), 1.0);
    NDN_LAUNCHED("finish_frames");
    return NDN_OK;
}

// ---- the entry point that allocates nothing --------------------------------------------------

// The scratch of ndn_pair_efw_noalloc: the potential's coefficients from byte 0, then one
// partial slot (8 doubles) per block of the pair launch. File-scope device memory: every GPU the
// library runs on holds its own copy until the process ends. Host code never passes this name
// to a kernel (on the host it is a stand-in, not a device address); it passes the address
// hipGetSymbolAddress returns for the device, looked up once per device and kept.
static __device__ __align__(256) unsigned char scratch[NDN_PAIR_EFW_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void* scratch_address[kMaxDevices] = {};

// The device address of `scratch` on the stream's device, looked up on the first call there.
int scratch_for(hipStream_t stream, void** address)
{
    hipDevice_t device = 0;
    hipError_t status = hipStreamGetDevice(stream, &device);
    if (status != hipSuccess)
    {
        set_error(std::string("pair_efw_noalloc: hipStreamGetDevice: ") + hipGetErrorString(status));
        return NDN_ERR_HIP;
    }
    NDN_REQUIRE(device >= 0 && device < kMaxDevices, "pair_efw_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_address[device] == nullptr)
    {
        int current = 0;
        status = hipGetDevice(&current);
        if (status == hipSuccess && current != device) status = hipSetDevice(device);
        void* found = nullptr;
        if (status == hipSuccess) status = hipGetSymbolAddress(&found, HIP_SYMBOL(scratch));
        if (current != device) (void)hipSetDevice(current);
        if (status != hipSuccess || found == nullptr)
        {
            set_error(std::string("pair_efw_noalloc: hipGetSymbolAddress: ") +
                      hipGetErrorString(status));
            return NDN_ERR_HIP;
        }
        scratch_address[device] = found;
    }
    *address = scratch_address[device];
    return NDN_OK;
}

}  // namespace ndn

using namespace ndn;

extern "C" {

int ndn_pair_workspace_bytes(int dtype, const ndn_potential* pot, int64_t columns,
                             int64_t* bytes)
{
    const int checked = check_potential(pot, 0);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    NDN_REQUIRE(columns >= 0 && bytes != nullptr, "columns must be nonnegative");
    const int element = dtype == NDN_F32 ? 4 : 8;
    const std::int64_t blocks = std::max<std::int64_t>(half_blocks(columns), kPartialSlots);
    *bytes = coefficient_bytes(pot, element) + blocks * 8 * static_cast<std::int64_t>(sizeof(double));
    return NDN_OK;
}

int ndn_pair_frames_workspace_bytes(int dtype, const ndn_potential* pot, int64_t frames,
                                    int64_t columns, int64_t* bytes)
{
    const int checked = check_potential(pot, 0);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    NDN_REQUIRE(frames >= 1 && frames <= kMaxFrames && columns >= 0 && bytes != nullptr,
                "frames must lie in [1, 65535] and columns be nonnegative");
    const int element = dtype == NDN_F32 ? 4 : 8;
    const std::int64_t slots = frames * half_blocks(columns);
    *bytes = coefficient_bytes(pot, element) +
             std::max<std::int64_t>(slots, kPartialSlots) * 8 *
                 static_cast<std::int64_t>(sizeof(double));
    return NDN_OK;
}

int ndn_pair_efw_full_frames(int dtype, const void* positions, const void* cells, int64_t frames,
                             int64_t num_atoms, const int32_t* counts, const int32_t* slices,
                             const int32_t* index, int64_t capacity, const ndn_potential* pot,
                             void* energy, void* forces, void* virial, void* workspace,
                             int64_t workspace_bytes, ndn_stream_t stream)
{
    const int checked = check_potential(pot, num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(positions && cells && counts && slices && index && workspace, "a buffer is null");
    NDN_REQUIRE(num_atoms >= 1 && num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(capacity >= 1, "each frame's full list must have a positive capacity");
    std::int64_t needed = 0;
    const int sized = ndn_pair_frames_workspace_bytes(dtype, pot, frames, num_atoms, &needed);
    if (sized != NDN_OK) return sized;
    NDN_REQUIRE(workspace_bytes >= needed, "the workspace is too small");
    if (dtype == NDN_F32)
        return pair_efw_full_frames_impl<float>(positions, cells, frames, num_atoms, counts, slices,
                                                index, capacity, pot, energy, forces, virial,
                                                workspace, stream);
    return pair_efw_full_frames_impl<double>(positions, cells, frames, num_atoms, counts, slices,
                                             index, capacity, pot, energy, forces, virial,
                                             workspace, stream);
}

int ndn_pair_efw_frames(int dtype, const void* positions, const void* cells, int64_t frames,
                        int64_t num_atoms, const int32_t* stacks, int64_t capacity,
                        const ndn_potential* pot, void* energy, void* forces, void* virial,
                        void* workspace, int64_t workspace_bytes, ndn_stream_t stream)
{
    const int checked = check_potential(pot, num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(positions && cells && stacks && workspace, "a buffer is null");
    NDN_REQUIRE(num_atoms >= 1 && num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(capacity >= 1 && capacity <= INT32_MAX, "stacks must be [F, 2, C] with int32 C");
    std::int64_t needed = 0;
    const int sized = ndn_pair_frames_workspace_bytes(dtype, pot, frames, capacity, &needed);
    if (sized != NDN_OK) return sized;
    NDN_REQUIRE(workspace_bytes >= needed, "the workspace is too small");
    if (dtype == NDN_F32)
        return pair_efw_frames_impl<float>(positions, cells, frames, num_atoms, stacks, capacity,
                                           pot, energy, forces, virial, workspace, stream);
    return pair_efw_frames_impl<double>(positions, cells, frames, num_atoms, stacks, capacity,
                                        pot, energy, forces, virial, workspace, stream);
}

int ndn_pair_efw(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                 const int32_t* pairs, int64_t capacity, const ndn_potential* pot, void* energy,
                 void* forces, void* virial, void* workspace, int64_t workspace_bytes,
                 ndn_stream_t stream)
{
    const int checked = check_potential(pot, num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(positions && cell && pairs && workspace, "a buffer is null");
    NDN_REQUIRE(num_atoms >= 1 && num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(capacity >= 1 && capacity <= INT32_MAX, "pairs must be [2, C] with int32 C");
    std::int64_t needed = 0;
    const int sized = ndn_pair_workspace_bytes(dtype, pot, capacity, &needed);
    if (sized != NDN_OK) return sized;
    NDN_REQUIRE(workspace_bytes >= needed, "the workspace is too small");
    if (dtype == NDN_F32)
        return pair_efw_impl<float>(positions, cell, num_atoms, pairs, capacity, pot, energy,
                                    forces, virial, workspace, stream);
    return pair_efw_impl<double>(positions, cell, num_atoms, pairs, capacity, pot, energy, forces,
                                 virial, workspace, stream);
}

int ndn_pair_efw_full(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                      const int32_t* rows, const int32_t* counts, const int32_t* slices,
                      const int32_t* index, int64_t capacity, const ndn_potential* pot,
                      void* energy, void* forces, void* virial, void* workspace,
                      int64_t workspace_bytes, ndn_stream_t stream)
{
    const int checked = check_potential(pot, num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    NDN_REQUIRE(positions && cell && counts && slices && index && workspace, "a buffer is null");
    NDN_REQUIRE(num_atoms >= 1 && num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(capacity >= 1, "the full list must have a positive capacity");
    std::int64_t needed = 0;
    const int sized = ndn_pair_workspace_bytes(dtype, pot, 1, &needed);
    if (sized != NDN_OK) return sized;
    NDN_REQUIRE(workspace_bytes >= needed, "the workspace is too small");
    if (dtype == NDN_F32)
        return pair_efw_full_impl<float>(positions, cell, num_atoms, rows, counts, slices, index,
                                         capacity, pot, energy, forces, virial, workspace, stream);
    return pair_efw_full_impl<double>(positions, cell, num_atoms, rows, counts, slices, index,
                                      capacity, pot, energy, forces, virial, workspace, stream);
}

}  // extern "C"

extern "C" int ndn_pair_efw_noalloc(int dtype, const void* positions, const void* cell,
                                    int64_t num_atoms, const int32_t* pairs, int64_t capacity,
                                    const ndn_potential* pot, void* energy, void* forces,
                                    void* virial, ndn_stream_t stream)
{
    const int checked = check_potential(pot, num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    NDN_REQUIRE(positions && cell && pairs, "a buffer is null");
    NDN_REQUIRE(num_atoms >= 1 && num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(capacity >= 1 && capacity <= INT32_MAX, "pairs must be [2, C] with int32 C");
    const int element = dtype == NDN_F32 ? 4 : 8;
    const std::int64_t needed = coefficient_bytes(pot, element) +
                                static_cast<std::int64_t>(half_blocks(capacity)) * 8 *
                                    static_cast<std::int64_t>(sizeof(double));
    if (needed > NDN_PAIR_EFW_NOALLOC_SCRATCH_BYTES)
    {
        set_error("pair_efw_noalloc: the potential's coefficients and partial sums need " +
                  std::to_string(needed) + " bytes of scratch, more than its " +
                  std::to_string(NDN_PAIR_EFW_NOALLOC_SCRATCH_BYTES) +
                  " (use ndn_pair_efw with a workspace)");
        return NDN_ERR_ARGUMENT;
    }
    void* workspace = nullptr;
    const int found = scratch_for(stream, &workspace);
    if (found != NDN_OK) return found;
    if (dtype == NDN_F32)
        return pair_efw_impl<float>(positions, cell, num_atoms, pairs, capacity, pot, energy,
                                    forces, virial, workspace, stream);
    return pair_efw_impl<double>(positions, cell, num_atoms, pairs, capacity, pot, energy, forces,
                                 virial, workspace, stream);
}
