// The neighbor list on gfx950: reuse check, cell binning, pair count and fill (the reference's
// half list, in its column order, plus an optional full list for the fused force kernels),
// total, padding and commit. Built with -ffp-contract=off: every distance is the source-text
// IEEE expression (dx dx + dy dy) + dz dz, rounded step by step, so the pair set is bit-exact
// with a scalar CPU evaluation of the same expression.
//
// Reference semantics followed (paths under src/nabladyn/neighbors/csrc/):
//   update/update.cuh:27-75 prepare, :79-116 reuse check, :119-145 commit;
//   grid/grid.cuh:10-118 binning and (bin, atom id) order; pairs/pairs.cuh:43-157 count, fill,
//   total, pad; pairs/all_pairs.cuh:29-114 all-pairs path; common/traversal.cuh:12-147 visit
//   order and periodic images; runtime/entry_points.cu:127-508 launch order and gating.

#include <climits>
#include <limits>
#include <mutex>
#include <string>

#include "common.hpp"
#include "neighbors.hpp"
#include "scan.hpp"

namespace ndn {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& message)
{
    g_last_error = message;
}

const char* last_error()
{
    return g_last_error.c_str();
}

// ---- geometry -------------------------------------------------------------------------------

__device__ __forceinline__ index_t wrap_bin(index_t home, int step, index_t count)
{
    const index_t bin = home + step;
    return bin < 0 ? bin + count : (bin >= count ? bin - count : bin);
}

template <typename T>
__device__ __forceinline__ T bin_shift(index_t home, int step, index_t count, T length)
{
    const index_t bin = home + step;
    return bin < 0 ? -length : (bin >= count ? length : T(0));
}

template <typename T>
__device__ __forceinline__ index_t axis_bin(T x, T width, index_t count)
{
    const index_t bin = static_cast<index_t>(floor(x / width));
    return bin < 0 ? 0 : (bin >= count ? count - 1 : bin);
}

template <typename T>
__device__ __forceinline__ T squared(T dx, T dy, T dz)
{
    return dx * dx + dy * dy + dz * dz;
}

// ---- prepare and reuse ----------------------------------------------------------------------

template <typename T>
__global__ void prepare_update_kernel(const T* __restrict__ cell,
                                      const T* __restrict__ reference_cell, T cutoff, T radius,
                                      const index_t* __restrict__ limits, index_t* dims, T* widths,
                                      T* reuse, index_t* rebuild, index_t* status,
                                      const index_t* pair_total, ListHeader* header)
{
    prepare_update(cell, reference_cell, cutoff, radius, limits, dims, widths, reuse, rebuild,
                   status, pair_total, header);
}

// One atom per lane: flags a non-finite coordinate, and reduces the squared displacement since
// the last rebuild over the wave. Its maximum decides the rebuild: max >= threshold is the
// reference's "any atom moved at least the threshold", bit for bit.
template <typename T>
__global__ void __launch_bounds__(kBlock) check_reuse_kernel(
    const T* __restrict__ positions, const T* __restrict__ reference_positions,
    const T* __restrict__ cell, const T* __restrict__ reuse, index_t num_atoms, index_t* rebuild,
    index_t* status, ListHeader* header)
{
    __shared__ T block_max[kBlock / kWave];
    // every thread reaches the barrier below, whatever status a wave reads
    const bool active = status[0] == 0;
    const std::int64_t atom = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool finite = true;
    T moved = 0;
    if (active && atom < num_atoms)
    {
        const T x = positions[3 * atom], y = positions[3 * atom + 1], z = positions[3 * atom + 2];
        finite = isfinite(x) && isfinite(y) && isfinite(z);
        if (finite) moved = moved_squared(x, y, z, reference_positions + 3 * atom, reuse, cell);
    }
    if (!finite) atomicOr(status, kStatusNonFinite);
    const T wave_moved = wave_max(finite ? moved : T(0));
    if (lane_id() == 0)
    {
        block_max[threadIdx.x / kWave] = wave_moved;
        if (active && wave_moved >= reuse[0] && rebuild[0] == 0) rebuild[0] = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        T m = block_max[0];
        for (int w = 1; w < kBlock / kWave; ++w) m = block_max[w] > m ? block_max[w] : m;
        if (m > T(0)) atomicMax(moved_slot(header, blockIdx.x), order_bits(m));
    }
}

// ---- grid: bins, counting sort, deterministic in-bin order ----------------------------------

template <typename T>
__global__ void __launch_bounds__(kBlock) assign_bins_kernel(
    const T* __restrict__ positions, const index_t* __restrict__ dims,
    const T* __restrict__ widths, index_t num_atoms, index_t* __restrict__ atom_bin,
    index_t* __restrict__ bin_counts, const index_t* rebuild, const index_t* status)
{
    const std::int64_t atom = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (atom >= num_atoms || !rebuilding(rebuild, status)) return;
    const T* p = positions + 3 * atom;
    const index_t nx = dims[0], ny = dims[1];
    const index_t bin = axis_bin(p[0], widths[0], nx) +
                        nx * (axis_bin(p[1], widths[1], ny) + ny * axis_bin(p[2], widths[2], dims[2]));
    atom_bin[atom] = bin;
    // one atomic per run of a bin's atoms in neighbouring lanes
    const LaneRun run = equal_runs(bin);
    if (run.head) atomicAdd(bin_counts + bin, run.length);
}

// Places each atom in its bin's slice in the order of the atomics; the counts return to zero.
__global__ void __launch_bounds__(kBlock) scatter_atoms_kernel(
    const index_t* __restrict__ atom_bin, const index_t* __restrict__ bin_offsets,
    index_t num_atoms, index_t* __restrict__ bin_counts, index_t* __restrict__ unordered_ids,
    const index_t* rebuild, const index_t* status)
{
    const std::int64_t atom = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (atom >= num_atoms || !rebuilding(rebuild, status)) return;
    const index_t bin = atom_bin[atom];
    // one atomic per run of a bin's atoms: the run's lanes take the places below the count it
    // returns, one each (order_bins_kernel ranks a bin's atoms by id after)
    const LaneRun run = equal_runs(bin);
    index_t count = 0;
    if (run.head) count = atomicSub(bin_counts + bin, run.length);
    count = __shfl(count, run.first, kWave);
    unordered_ids[bin_offsets[bin] + count - 1 - (lane_id() - run.first)] =
        static_cast<index_t>(atom);
}

// One wave per bin: ranks each atom among its bin's ids (readlane broadcasts, no atomics), so the
// order inside a bin is by atom id whatever the atomics did, then writes the sorted ids and
// gathers the sorted positions.
template <typename T>
__global__ void __launch_bounds__(kBlock) order_bins_kernel(
    const T* __restrict__ positions, const index_t* __restrict__ unordered_ids,
    const index_t* __restrict__ bin_offsets, const index_t* __restrict__ dims,
    std::int64_t bin_slots, index_t* __restrict__ sorted_ids, T* __restrict__ sorted_positions,
    const index_t* rebuild, const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    const std::int64_t bin =
        static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
    const std::int64_t bins = static_cast<std::int64_t>(dims[0]) * dims[1] * dims[2];
    if (bin >= bins || bin >= bin_slots - 1) return;
    const int lane = lane_id();
    const index_t begin = bin_offsets[bin];
    const index_t end = bin_offsets[bin + 1];
    for (index_t chunk = begin; chunk < end; chunk += kWave)
    {
        const bool mine = chunk + lane < end;
        const index_t id = mine ? unordered_ids[chunk + lane] : INT_MAX;
        index_t rank = 0;
        for (index_t other = begin; other < end; other += kWave)
        {
            const index_t theirs = other + lane < end ? unordered_ids[other + lane] : INT_MAX;
            const int live = end - other < kWave ? end - other : kWave;
            for (int m = 0; m < live; ++m)
            {
                rank += __shfl(theirs, m, kWave) < id;
            }
        }
        if (mine)
        {
            const index_t slot = begin + rank;
            sorted_ids[slot] = id;
            const T* p = positions + 3LL * id;
            T* q = sorted_positions + 3LL * slot;
            q[0] = p[0];
            q[1] = p[1];
            q[2] = p[2];
        }
    }
}

// ---- pair visit -----------------------------------------------------------------------------

// Visits, for sorted center `center`, every candidate in the reference's order (neighbor bins
// z, y, x from -1 to +1; candidates ascending) and calls visit(neighbor, after) for each one
// closer than the list radius, after = neighbor > center. Half lists visit only candidates
// after the center, as the reference does; full lists visit all other candidates.
template <typename T, bool kFull, typename Visit>
__device__ __forceinline__ void visit_grid(const T* __restrict__ sp,
                                           const index_t* __restrict__ bin_offsets, index_t home,
                                           index_t nx, index_t ny, index_t nz, T lx, T ly, T lz,
                                           T radius_squared, index_t center, Visit& visit)
{
    const T cx = sp[3LL * center], cy = sp[3LL * center + 1], cz = sp[3LL * center + 2];
    const index_t hx = home % nx, hy = (home / nx) % ny, hz = home / (nx * ny);
    if (nx >= 3 && ny >= 3 && nz >= 3)
    {
        for (int sz = -1; sz <= 1; ++sz)
        {
            const T shz = bin_shift(hz, sz, nz, lz);
            const T az = shz > T(0) ? cz - shz : cz;
            const T bz_shift = shz < T(0) ? shz : T(0);
            const index_t bz = wrap_bin(hz, sz, nz);
            for (int sy = -1; sy <= 1; ++sy)
            {
                const T shy = bin_shift(hy, sy, ny, ly);
                const T ay = shy > T(0) ? cy - shy : cy;
                const T by_shift = shy < T(0) ? shy : T(0);
                const index_t by = wrap_bin(hy, sy, ny);
                for (int sx = -1; sx <= 1; ++sx)
                {
                    const T shx = bin_shift(hx, sx, nx, lx);
                    const T ax = shx > T(0) ? cx - shx : cx;
                    const T bx_shift = shx < T(0) ? shx : T(0);
                    const index_t bin = wrap_bin(hx, sx, nx) + nx * (by + ny * bz);
                    const index_t first = bin_offsets[bin];
                    const index_t end = bin_offsets[bin + 1];
                    const index_t start = kFull ? first : (first > center ? first : center + 1);
                    for (index_t nb = start; nb < end; ++nb)
                    {
                        const T* p = sp + 3LL * nb;
                        const T dx = ax - (p[0] + bx_shift);
                        const T dy = ay - (p[1] + by_shift);
                        const T dz = az - (p[2] + bz_shift);
                        if (squared(dx, dy, dz) < radius_squared && nb != center)
                        {
                            visit(nb, nb > center);
                        }
                    }
                }
            }
        }
        return;
    }
    // A short axis (fewer than three bins): every bin along it, minimum image per candidate.
    const int span_x = nx >= 3 ? 3 : nx, span_y = ny >= 3 ? 3 : ny, span_z = nz >= 3 ? 3 : nz;
    for (int iz = 0; iz < span_z; ++iz)
    {
        const index_t bz = nz >= 3 ? wrap_bin(hz, iz - 1, nz) : iz;
        for (int iy = 0; iy < span_y; ++iy)
        {
            const index_t by = ny >= 3 ? wrap_bin(hy, iy - 1, ny) : iy;
            for (int ix = 0; ix < span_x; ++ix)
            {
                const index_t bin = (nx >= 3 ? wrap_bin(hx, ix - 1, nx) : ix) + nx * (by + ny * bz);
                const index_t first = bin_offsets[bin];
                const index_t end = bin_offsets[bin + 1];
                const index_t start = kFull ? first : (first > center ? first : center + 1);
                for (index_t nb = start; nb < end; ++nb)
                {
                    const T* p = sp + 3LL * nb;
                    const T dx = minimum_image(cx, p[0], lx);
                    const T dy = minimum_image(cy, p[1], ly);
                    const T dz = minimum_image(cz, p[2], lz);
                    if (squared(dx, dy, dz) < radius_squared && nb != center)
                    {
                        visit(nb, nb > center);
                    }
                }
            }
        }
    }
}

struct CountVisit
{
    index_t half = 0;
    index_t full = 0;
    __device__ __forceinline__ void operator()(index_t, bool after)
    {
        half += after;
        ++full;
    }
};

struct FillVisit
{
    const index_t* __restrict__ sorted_ids;
    index_t center_id;
    std::int64_t column;
    std::int64_t capacity;
    index_t* __restrict__ pairs;
    // full list
    index_t* __restrict__ full_index;
    std::int64_t full_slot;     // slice base + lane within slice
    std::int64_t full_capacity;
    __device__ __forceinline__ void operator()(index_t nb, bool after)
    {
        const index_t id = sorted_ids[nb];
        if (after)
        {
            if (column < capacity)
            {
                pairs[column] = center_id;
                pairs[capacity + column] = id;
            }
            ++column;
        }
        if (full_index != nullptr)
        {
            if (full_slot < full_capacity) full_index[full_slot] = id;
            full_slot += kWave;
        }
    }
};

// Thread per sorted center. Writes the half count (and with a full list the full count, and
// the slice widths: 64 * the largest full count of each 64 rows) for the scans that follow.
template <typename T, bool kFull>
__global__ void __launch_bounds__(kBlock) count_pairs_kernel(
    const T* __restrict__ sp, const index_t* __restrict__ sorted_ids,
    const index_t* __restrict__ atom_bin, const index_t* __restrict__ bin_offsets,
    const index_t* __restrict__ dims, const T* __restrict__ cell, T radius_squared,
    index_t num_atoms, index_t* __restrict__ pair_counts, index_t* __restrict__ full_counts,
    index_t* __restrict__ full_slices, const index_t* rebuild, const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    const std::int64_t center = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    CountVisit visit;
    if (center < num_atoms)
    {
        visit_grid<T, kFull>(sp, bin_offsets, atom_bin[sorted_ids[center]], dims[0], dims[1],
                             dims[2], cell[0], cell[1], cell[2], radius_squared,
                             static_cast<index_t>(center), visit);
        pair_counts[center] = visit.half;
        if (kFull) full_counts[center] = visit.full;
    }
    if (kFull)
    {
        const index_t width = wave_max(center < num_atoms ? visit.full : 0);
        const std::int64_t slice = center / kWave;
        if (lane_id() == 0 && slice * kWave < num_atoms) full_slices[slice] = width * kWave;
    }
}

template <typename T, bool kFull>
__global__ void __launch_bounds__(kBlock) fill_pairs_kernel(
    const T* __restrict__ sp, const index_t* __restrict__ sorted_ids,
    const index_t* __restrict__ atom_bin, const index_t* __restrict__ bin_offsets,
    const index_t* __restrict__ dims, const T* __restrict__ cell,
    const index_t* __restrict__ pair_offsets, T radius_squared, index_t num_atoms,
    std::int64_t capacity, index_t* __restrict__ pairs, const index_t* __restrict__ full_slices,
    index_t* __restrict__ full_index, std::int64_t full_capacity, const index_t* rebuild,
    const index_t* status)
{
    const std::int64_t center = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (center >= num_atoms || !rebuilding(rebuild, status)) return;
    FillVisit visit{sorted_ids, sorted_ids[center], pair_offsets[center], capacity, pairs,
                    kFull ? full_index : nullptr,
                    kFull ? full_slices[center / kWave] + center % kWave : 0, full_capacity};
    visit_grid<T, kFull>(sp, bin_offsets, atom_bin[sorted_ids[center]], dims[0], dims[1], dims[2],
                         cell[0], cell[1], cell[2], radius_squared, static_cast<index_t>(center),
                         visit);
}

// ---- all-pairs path -------------------------------------------------------------------------

// One wave per atom i: 64 candidates per ballot, each 64-bit hit mask stored as two words of
// i's bit row. A half list tests j > i only; a full list tests every j != i.
template <typename T, bool kFull>
__global__ void __launch_bounds__(kBlock) test_all_pairs_kernel(
    const T* __restrict__ positions, const T* __restrict__ cell, T radius_squared,
    index_t num_atoms, std::uint32_t* __restrict__ bits, std::int64_t row_words,
    index_t* __restrict__ pair_counts, index_t* __restrict__ full_counts,
    index_t* __restrict__ full_slices, const index_t* rebuild, const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    const std::int64_t center =
        static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
    if (center >= num_atoms) return;
    const int lane = lane_id();
    const T lx = cell[0], ly = cell[1], lz = cell[2];
    const T cx = positions[3 * center], cy = positions[3 * center + 1],
            cz = positions[3 * center + 2];
    const std::int64_t chunks = (num_atoms + kWave - 1) / kWave;
    index_t half = 0, full = 0;
    for (std::int64_t chunk = kFull ? 0 : (center + 1) / kWave; chunk < chunks; ++chunk)
    {
        const std::int64_t j = chunk * kWave + lane;
        bool hit = false;
        if (j < num_atoms && j != center && (kFull || j > center))
        {
            const T dx = minimum_image(cx, positions[3 * j], lx);
            const T dy = minimum_image(cy, positions[3 * j + 1], ly);
            const T dz = minimum_image(cz, positions[3 * j + 2], lz);
            hit = squared(dx, dy, dz) < radius_squared;
        }
        const std::uint64_t mask = ballot(hit);
        const std::uint64_t after = ballot(hit && j > center);
        half += __popcll(after);
        full += __popcll(mask);
        if (lane == 0)
        {
            bits[center * row_words + 2 * chunk] = static_cast<std::uint32_t>(mask);
            if (2 * chunk + 1 < row_words)
                bits[center * row_words + 2 * chunk + 1] = static_cast<std::uint32_t>(mask >> 32);
        }
    }
    if (lane == 0)
    {
        pair_counts[center] = half;
        if (kFull) full_counts[center] = full;
    }
}

// Slice widths of the full list on the all-pairs path, from the full counts.
__global__ void __launch_bounds__(kBlock) slice_widths_kernel(
    const index_t* __restrict__ full_counts, index_t num_atoms, index_t* __restrict__ full_slices,
    const index_t* rebuild, const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    const std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const index_t width = wave_max(row < num_atoms ? full_counts[row] : 0);
    if (lane_id() == 0 && row < num_atoms) full_slices[row / kWave] = width * kWave;
}

template <bool kFull>
__global__ void __launch_bounds__(kBlock) fill_all_pairs_kernel(
    const std::uint32_t* __restrict__ bits, std::int64_t row_words,
    const index_t* __restrict__ pair_offsets, index_t num_atoms, std::int64_t capacity,
    index_t* __restrict__ pairs, const index_t* __restrict__ full_slices,
    index_t* __restrict__ full_index, std::int64_t full_capacity, const index_t* rebuild,
    const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    const std::int64_t center =
        static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
    if (center >= num_atoms) return;
    const int lane = lane_id();
    const std::int64_t chunks = (num_atoms + kWave - 1) / kWave;
    std::int64_t column = pair_offsets[center];
    std::int64_t full_column = 0;
    const std::int64_t slice_base = kFull ? full_slices[center / kWave] + center % kWave : 0;
    for (std::int64_t chunk = kFull ? 0 : (center + 1) / kWave; chunk < chunks; ++chunk)
    {
        const std::uint32_t low = bits[center * row_words + 2 * chunk];
        const std::uint32_t high =
            2 * chunk + 1 < row_words ? bits[center * row_words + 2 * chunk + 1] : 0u;
        const std::uint64_t mask = (static_cast<std::uint64_t>(high) << 32) | low;
        const std::int64_t j = chunk * kWave + lane;
        const bool hit = (mask >> lane) & 1ull;
        const std::uint64_t after = ballot(hit && j > center);
        if (hit && j > center)
        {
            const std::int64_t slot = column + bits_below(after);
            if (slot < capacity)
            {
                pairs[slot] = static_cast<index_t>(center);
                pairs[capacity + slot] = static_cast<index_t>(j);
            }
        }
        column += __popcll(after);
        if (kFull)
        {
            if (hit)
            {
                const std::int64_t slot = slice_base + (full_column + bits_below(mask)) * kWave;
                if (slot < full_capacity) full_index[slot] = static_cast<index_t>(j);
            }
            full_column += __popcll(mask);
        }
    }
}

// ---- total, pad, commit ---------------------------------------------------------------------

// The pair total and, above the capacity, the overflow bit with the pairs needed; then the
// padding. Column k past the pairs holds (k mod N, k mod N) whatever the build, so after the
// first frame only the columns between this total and the previous one need writing.
__global__ void __launch_bounds__(kBlock) total_pad_kernel(
    const index_t* __restrict__ pair_counts, const index_t* __restrict__ pair_offsets,
    index_t num_atoms, std::int64_t capacity, index_t* __restrict__ pair_total, index_t* status,
    index_t* __restrict__ pairs, const ListHeader* __restrict__ header, int first_frame,
    const index_t* __restrict__ full_slices, std::int64_t num_slices, std::int64_t full_capacity,
    const index_t* rebuild)
{
    if (!rebuilding(rebuild, status)) return;
    const index_t total = pair_offsets[num_atoms - 1] + pair_counts[num_atoms - 1];
    const std::int64_t gid = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool full_over = full_slices != nullptr && full_slices[num_slices] > full_capacity;
    if (gid == 0)
    {
        pair_total[0] = total;
        index_t flags = 0;
        if (total > capacity)
        {
            flags |= kStatusOverflow;
            status[1] = total > status[1] ? total : status[1];
        }
        if (full_over) flags |= kStatusFullOverflow;
        if (flags) status[0] |= flags;
    }
    if (total > capacity || full_over) return;
    const std::int64_t end =
        first_frame ? capacity : (header->old_total < capacity ? header->old_total : capacity);
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t column = total + gid; column < end; column += stride)
    {
        const index_t atom = static_cast<index_t>(column % num_atoms);
        pairs[column] = atom;
        pairs[capacity + column] = atom;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) commit_rebuild_kernel(
    const T* __restrict__ positions, const T* __restrict__ cell, index_t num_atoms,
    T* __restrict__ reference_positions, T* __restrict__ reference_cell,
    index_t* __restrict__ rebuild_count, const index_t* rebuild, const index_t* status)
{
    const std::int64_t index = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (index >= 3LL * num_atoms || !rebuilding(rebuild, status)) return;
    reference_positions[index] = positions[index];
    if (index < 3) reference_cell[index] = cell[index];
    if (index == 0) ++rebuild_count[0];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) wrap_kernel(const T* positions, const T* __restrict__ cell,
                                                      std::int64_t count, T* out,
                                                      std::int64_t* __restrict__ images)
{
    const std::int64_t index = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (index >= count) return;
    const T length = cell[index % 3];
    const T v = positions[index];
    T k = floor(v / length);
    T w = v - k * length;
    if (w < T(0))
    {
        w = w + length;
        k = k - T(1);
    }
    if (w >= length)
    {
        w = w - length;
        k = k + T(1);
    }
    out[index] = w;
    if (images != nullptr) images[index] += static_cast<std::int64_t>(k);
}

// ---- launch sequences -----------------------------------------------------------------------

constexpr unsigned kPadBlocks = 1024;

template <typename T>
int launch_grid(const T* positions, const ndn_neighbor_buffers& b, const index_t* rebuild,
                hipStream_t stream)
{
    const index_t n = static_cast<index_t>(b.num_atoms);
    assign_bins_kernel<T><<<blocks_for(n), kBlock, 0, stream>>>(
        positions, b.dims, static_cast<const T*>(b.widths), n, b.atom_bin, b.bin_counts, rebuild,
        b.status);
    NDN_LAUNCHED("assign_bins");
    const int scanned = exclusive_scan(b.bin_counts, b.bin_offsets, b.bin_slots, b.scan_temp,
                                       rebuild, b.status, stream);
    if (scanned != NDN_OK) return scanned;
    scatter_atoms_kernel<<<blocks_for(n), kBlock, 0, stream>>>(
        b.atom_bin, b.bin_offsets, n, b.bin_counts, b.unordered_ids, rebuild, b.status);
    NDN_LAUNCHED("scatter_atoms");
    const std::int64_t bins = b.bin_slots - 1;
    order_bins_kernel<T><<<blocks_for(bins, kBlock / kWave), kBlock, 0, stream>>>(
        positions, b.unordered_ids, b.bin_offsets, b.dims, b.bin_slots, b.sorted_ids,
        static_cast<T*>(b.sorted_positions), rebuild, b.status);
    NDN_LAUNCHED("order_bins");
    return NDN_OK;
}

template <typename T>
int launch_count(const T* positions, const T* cell, double radius, const ndn_neighbor_buffers& b,
                 const index_t* rebuild, hipStream_t stream)
{
    const index_t n = static_cast<index_t>(b.num_atoms);
    const T r = static_cast<T>(radius);
    const T r2 = r * r;
    const bool full = b.full_index != nullptr;
    const std::int64_t slices = (b.num_atoms + kWave - 1) / kWave;
    if (b.bits_words > 0)
    {
        const std::int64_t row_words = (b.num_atoms + 31) / 32;
        auto* bits = reinterpret_cast<std::uint32_t*>(b.bits);
        const unsigned blocks = blocks_for(n, kBlock / kWave);
        if (full)
        {
            test_all_pairs_kernel<T, true><<<blocks, kBlock, 0, stream>>>(
                positions, cell, r2, n, bits, row_words, b.pair_counts, b.full_counts,
                b.full_slices, rebuild, b.status);
            NDN_LAUNCHED("test_all_pairs");
            slice_widths_kernel<<<blocks_for(n), kBlock, 0, stream>>>(b.full_counts, n,
                                                                      b.full_slices, rebuild,
                                                                      b.status);
            NDN_LAUNCHED("slice_widths");
        }
        else
        {
            test_all_pairs_kernel<T, false><<<blocks, kBlock, 0, stream>>>(
                positions, cell, r2, n, bits, row_words, b.pair_counts, nullptr, nullptr, rebuild,
                b.status);
            NDN_LAUNCHED("test_all_pairs");
        }
    }
    else
    {
        const T* sp = static_cast<const T*>(b.sorted_positions);
        if (full)
        {
            count_pairs_kernel<T, true><<<blocks_for(n), kBlock, 0, stream>>>(
                sp, b.sorted_ids, b.atom_bin, b.bin_offsets, b.dims, cell, r2, n, b.pair_counts,
                b.full_counts, b.full_slices, rebuild, b.status);
        }
        else
        {
            count_pairs_kernel<T, false><<<blocks_for(n), kBlock, 0, stream>>>(
                sp, b.sorted_ids, b.atom_bin, b.bin_offsets, b.dims, cell, r2, n, b.pair_counts,
                nullptr, nullptr, rebuild, b.status);
        }
        NDN_LAUNCHED("count_pairs");
    }
    int scanned = exclusive_scan(b.pair_counts, b.pair_offsets, b.num_atoms, b.scan_temp, rebuild,
                                 b.status, stream);
    if (scanned != NDN_OK) return scanned;
    if (full)
    {
        // slices + 1 entries: the last (kept 0 by the counts) becomes the total.
        scanned = exclusive_scan(b.full_slices, b.full_slices, slices + 1, b.scan_temp, rebuild,
                                 b.status, stream);
        if (scanned != NDN_OK) return scanned;
    }
    return NDN_OK;
}

template <typename T>
int launch_fill(const T* cell, double radius, const ndn_neighbor_buffers& b,
                const index_t* rebuild, bool first_frame, hipStream_t stream)
{
    const index_t n = static_cast<index_t>(b.num_atoms);
    const T r = static_cast<T>(radius);
    const T r2 = r * r;
    const bool full = b.full_index != nullptr;
    const std::int64_t slices = (b.num_atoms + kWave - 1) / kWave;
    if (b.bits_words > 0)
    {
        const std::int64_t row_words = (b.num_atoms + 31) / 32;
        const auto* bits = reinterpret_cast<const std::uint32_t*>(b.bits);
        const unsigned blocks = blocks_for(n, kBlock / kWave);
        if (full)
            fill_all_pairs_kernel<true><<<blocks, kBlock, 0, stream>>>(
                bits, row_words, b.pair_offsets, n, b.capacity, b.pairs, b.full_slices,
                b.full_index, b.full_capacity, rebuild, b.status);
        else
            fill_all_pairs_kernel<false><<<blocks, kBlock, 0, stream>>>(
                bits, row_words, b.pair_offsets, n, b.capacity, b.pairs, nullptr, nullptr, 0,
                rebuild, b.status);
        NDN_LAUNCHED("fill_all_pairs");
    }
    else
    {
        const T* sp = static_cast<const T*>(b.sorted_positions);
        if (full)
            fill_pairs_kernel<T, true><<<blocks_for(n), kBlock, 0, stream>>>(
                sp, b.sorted_ids, b.atom_bin, b.bin_offsets, b.dims, cell, b.pair_offsets, r2, n,
                b.capacity, b.pairs, b.full_slices, b.full_index, b.full_capacity, rebuild,
                b.status);
        else
            fill_pairs_kernel<T, false><<<blocks_for(n), kBlock, 0, stream>>>(
                sp, b.sorted_ids, b.atom_bin, b.bin_offsets, b.dims, cell, b.pair_offsets, r2, n,
                b.capacity, b.pairs, nullptr, nullptr, 0, rebuild, b.status);
        NDN_LAUNCHED("fill_pairs");
    }
    total_pad_kernel<<<kPadBlocks, kBlock, 0, stream>>>(
        b.pair_counts, b.pair_offsets, n, b.capacity, b.pair_total, b.status, b.pairs,
        reinterpret_cast<const ListHeader*>(b.scan_temp), first_frame ? 1 : 0,
        full ? b.full_slices : nullptr, slices, b.full_capacity, rebuild);
    NDN_LAUNCHED("total_pad");
    return NDN_OK;
}

template <typename T>
int launch_rebuild(const T* positions, const T* cell, double radius, const ndn_neighbor_buffers& b,
                   T* reference_positions, T* reference_cell, index_t* rebuild_count,
                   const index_t* rebuild, hipStream_t stream)
{
    const index_t n = static_cast<index_t>(b.num_atoms);
    if (b.bits_words == 0)
    {
        const int grid = launch_grid<T>(positions, b, rebuild, stream);
        if (grid != NDN_OK) return grid;
    }
    int step = launch_count<T>(positions, cell, radius, b, rebuild, stream);
    if (step != NDN_OK) return step;
    step = launch_fill<T>(cell, radius, b, rebuild, false, stream);
    if (step != NDN_OK) return step;
    commit_rebuild_kernel<T><<<blocks_for(3LL * n), kBlock, 0, stream>>>(
        positions, cell, n, reference_positions, reference_cell, rebuild_count, rebuild, b.status);
    NDN_LAUNCHED("commit_rebuild");
    return NDN_OK;
}

template int launch_rebuild<float>(const float*, const float*, double, const ndn_neighbor_buffers&,
                                   float*, float*, index_t*, const index_t*, hipStream_t);
template int launch_rebuild<double>(const double*, const double*, double,
                                    const ndn_neighbor_buffers&, double*, double*, index_t*,
                                    const index_t*, hipStream_t);

int check_buffers(const ndn_neighbor_buffers* b)
{
    NDN_REQUIRE(b != nullptr, "buffers must not be null");
    NDN_REQUIRE(b->num_atoms >= 1 && b->num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(b->dims && b->widths && b->atom_bin && b->bin_counts && b->bin_offsets &&
                    b->unordered_ids && b->sorted_ids && b->sorted_positions && b->scan_temp &&
                    b->pair_counts && b->pair_offsets && b->pair_total && b->status && b->pairs,
                "a list buffer is null");
    NDN_REQUIRE(b->bin_slots >= 2, "bin_counts must hold at least 2 entries");
    NDN_REQUIRE(b->capacity >= 1 && b->capacity <= INT32_MAX, "pairs must be [2, C] with int32 C");
    const std::int64_t words = b->num_atoms * ((b->num_atoms + 31) / 32);
    NDN_REQUIRE(b->bits_words == 0 || (b->bits_words == words && b->bits != nullptr),
                "bits must be empty or hold N rows of ceil(N / 32) words");
    const std::int64_t slices = (b->num_atoms + kWave - 1) / kWave;
    const std::int64_t items = std::max(std::max(b->bin_slots, b->num_atoms), slices + 1);
    NDN_REQUIRE(b->scan_temp_bytes >= scan_bytes(items), "scan_temp is too small");
    if (b->full_index != nullptr || b->full_counts != nullptr || b->full_slices != nullptr)
    {
        NDN_REQUIRE(b->full_index && b->full_counts && b->full_slices && b->full_capacity >= 1 &&
                        b->full_capacity <= INT32_MAX,
                    "a full list needs counts, slices and an int32-sized index");
    }
    return NDN_OK;
}

template <typename T>
int count_first_frame_impl(const void* positions, const void* cell, double radius,
                           const ndn_neighbor_buffers& b, hipStream_t stream)
{
    const T* x = static_cast<const T*>(positions);
    if (b.bits_words == 0)
    {
        const int grid = launch_grid<T>(x, b, nullptr, stream);
        if (grid != NDN_OK) return grid;
    }
    return launch_count<T>(x, static_cast<const T*>(cell), radius, b, nullptr, stream);
}

template <typename T>
int fill_first_frame_impl(const void* cell, double radius, const ndn_neighbor_buffers& b,
                          hipStream_t stream)
{
    return launch_fill<T>(static_cast<const T*>(cell), radius, b, nullptr, true, stream);
}

template <typename T>
int update_impl(const void* positions, const void* cell, double cutoff, double radius,
                void* reference_positions, void* reference_cell, const index_t* limits,
                void* reuse, index_t* rebuild, index_t* rebuild_count,
                const ndn_neighbor_buffers& b, hipStream_t stream)
{
    const T* x = static_cast<const T*>(positions);
    const T* L = static_cast<const T*>(cell);
    const index_t n = static_cast<index_t>(b.num_atoms);
    auto* header = reinterpret_cast<ListHeader*>(b.scan_temp);
    prepare_update_kernel<T><<<1, 1, 0, stream>>>(
        L, static_cast<const T*>(reference_cell), static_cast<T>(cutoff), static_cast<T>(radius),
        limits, b.dims, static_cast<T*>(b.widths), static_cast<T*>(reuse), rebuild, b.status,
        b.pair_total, header);
    NDN_LAUNCHED("prepare_update");
    check_reuse_kernel<T><<<blocks_for(n), kBlock, 0, stream>>>(
        x, static_cast<const T*>(reference_positions), L, static_cast<const T*>(reuse), n, rebuild,
        b.status, header);
    NDN_LAUNCHED("check_reuse");
    return launch_rebuild<T>(x, L, radius, b, static_cast<T*>(reference_positions),
                             static_cast<T*>(reference_cell), rebuild_count, rebuild, stream);
}

}  // namespace ndn

using namespace ndn;

extern "C" {

const char* ndn_last_error(void)
{
    return ndn::last_error();
}

int ndn_scan_temp_bytes(int64_t items, int64_t* bytes)
{
    NDN_REQUIRE(items > 0 && items <= INT32_MAX, "the scan length must fit int32");
    NDN_REQUIRE(bytes != nullptr, "bytes must not be null");
    // Room for the largest scan a list runs: max(bins + 1, N) and, with a full list, the
    // ceil(N / 64) + 1 slice widths, which never exceed N + 1.
    *bytes = scan_bytes(items + 1);
    return NDN_OK;
}

int ndn_count_first_frame(int dtype, const void* positions, const void* cell, double radius,
                          const ndn_neighbor_buffers* buffers, ndn_stream_t stream)
{
    const int checked = check_buffers(buffers);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(positions && cell, "positions and cell must not be null");
    NDN_REQUIRE(radius > 0, "the radius must be positive");
    if (dtype == NDN_F32)
        return count_first_frame_impl<float>(positions, cell, radius, *buffers, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return count_first_frame_impl<double>(positions, cell, radius, *buffers, stream);
}

int ndn_fill_first_frame(int dtype, const void* positions, const void* cell, double radius,
                         const ndn_neighbor_buffers* buffers, ndn_stream_t stream)
{
    const int checked = check_buffers(buffers);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(positions && cell, "positions and cell must not be null");
    NDN_REQUIRE(radius > 0, "the radius must be positive");
    if (dtype == NDN_F32) return fill_first_frame_impl<float>(cell, radius, *buffers, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return fill_first_frame_impl<double>(cell, radius, *buffers, stream);
}

int ndn_update(int dtype, const void* positions, const void* cell, double cutoff, double radius,
               void* reference_positions, void* reference_cell, const int32_t* limits,
               void* reuse, int32_t* rebuild, int32_t* rebuild_count,
               const ndn_neighbor_buffers* buffers, ndn_stream_t stream)
{
    const int checked = check_buffers(buffers);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(positions && cell && reference_positions && reference_cell && limits && reuse &&
                    rebuild && rebuild_count,
                "an update buffer is null");
    NDN_REQUIRE(cutoff > 0 && radius >= cutoff, "need 0 < cutoff <= radius");
    if (dtype == NDN_F32)
        return update_impl<float>(positions, cell, cutoff, radius, reference_positions,
                                  reference_cell, limits, reuse, rebuild, rebuild_count,
                                  *buffers, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return update_impl<double>(positions, cell, cutoff, radius, reference_positions,
                               reference_cell, limits, reuse, rebuild, rebuild_count, *buffers,
                               stream);
}

int ndn_wrap(int dtype, const void* positions, const void* cell, int64_t num_atoms, void* out,
             int64_t* images, ndn_stream_t stream)
{
    NDN_REQUIRE(positions && cell && out, "positions, cell and out must not be null");
    NDN_REQUIRE(num_atoms >= 0, "N must be nonnegative");
    if (num_atoms == 0) return NDN_OK;
    const std::int64_t count = 3 * num_atoms;
    if (dtype == NDN_F32)
        wrap_kernel<float><<<blocks_for(count), kBlock, 0, stream>>>(
            static_cast<const float*>(positions), static_cast<const float*>(cell), count,
            static_cast<float*>(out), images);
    else
    {
        NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
        wrap_kernel<double><<<blocks_for(count), kBlock, 0, stream>>>(
            static_cast<const double*>(positions), static_cast<const double*>(cell), count,
            static_cast<double*>(out), images);
    }
    NDN_LAUNCHED("wrap");
    return NDN_OK;
}

}  // extern "C"
