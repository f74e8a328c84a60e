// The molecular-dynamics step on gfx950: velocity Verlet with an optional Nose-Hoover chain of
// three, the neighbor update and the pair forces, replayed as one HIP graph per step.
//
// The resident frame is kept in cell order ("slots"): at every rebuild the atoms are counting-
// sorted into sub-bins at least half the list radius wide, ordered by atom id inside a sub-bin
// (so every run is deterministic), and the whole frame is permuted with them. Every per-atom
// access of a step is then coalesced, and a neighbor's position is a short gather from nearby
// memory. The frame is read out in atom order only on demand.
//
//   prologue     one block: the chain's opening half step (fp64), the reuse rule's preparation
//                (update/update.cuh:27-75), the potential's coefficient rows
//   first_half   per slot: chain scale, half kick, drift, wrap, image count, the float4 copy the
//                force kernel gathers from, and the squared displacement since the last rebuild
//                reduced to the rebuild flag (update/update.cuh:79-116)
//   rebuild      gated on that flag: sub-bin counting sort, permutation of the frame, then one
//                wave per atom tests the candidates of its culled 5 x 5 x 5 sub-bin stencil 64 at
//                a time and ballots the hits into its row of a fixed-stride full list
//   prune        gated on a build, or on an atom having moved half the pruned list's skin since
//                its last refresh: keeps the full rows' pairs within cutoff + that skin, so the
//                force kernel reads about two thirds of the entries
//   force        a lane group per row reads its pruned row coalesced, sums the force in
//                registers, then the second half kick and the partial sums of E, W and K
//   finish       one block: E, W, K in fixed order and the chain's closing half step; its
//                velocity factor is applied by the next first_half (or by settle)
//
// Reference step order (md/_simulation.py:343-348): first_half, NeighborList._update_wrapped,
// energy_and_force + second_half, closing; ensembles in md/_ensembles/_verlet.py:8-31,
// _nve.py:39-58, _nvt.py:88-129, _chain.py:29-105. The integrator's arithmetic is the
// reference's eager expressions, each product and sum rounded on its own; the list's distance
// test is the exact minimum-image expression, so its pair set equals the reference's.

#include <algorithm>
#include <climits>

#include "common.hpp"
#include "neighbors.hpp"
#include "pair_forces.hpp"
#include "potentials.hpp"

namespace ndn {

// Thermostat slots (fp64): chain positions [0..2], momenta [3..5], the pending closing factor
// [6], the opening factor of the current step [7].
constexpr int kChainQ = 0;
constexpr int kChainP = 3;
constexpr int kPending = 6;
constexpr int kOpening = 7;
constexpr int kStencil = 5;          // sub-bins per axis a stencil spans

// The chain's half step against the kinetic energy K it thermostats (_chain.py:29-76), in the
// reference's operation order. Updates q, p and K; returns the velocity factor.
__device__ __forceinline__ double chain_half_step(double* kinetic, double* q, double* p,
                                                  const double* mass, double g, double kT,
                                                  double dt)
{
#pragma clang fp contract(off)
    const double quarter = 0.25 * dt;
    const double eighth = 0.125 * dt;
    double damping[2];
    double K = *kinetic;
    auto driving = [&](int index) {
        return index == 0 ? 2.0 * K - g * kT : p[index - 1] * p[index - 1] / mass[index - 1] - kT;
    };
    for (int index = 2; index >= 0; --index)
    {
        if (index == 2)
            p[2] = p[2] + quarter * driving(2);
        else
        {
            const double kick = quarter * driving(index);
            p[index] = damping[index] * (damping[index] * p[index] + kick);
        }
        if (index > 0) damping[index - 1] = exp(-eighth * p[index] / mass[index]);
    }
    const double factor = exp(-0.5 * dt * p[0] / mass[0]);
    K = K * (factor * factor);
    for (int k = 0; k < 3; ++k) q[k] = q[k] + 0.5 * dt * p[k] / mass[k];
    for (int index = 0; index <= 2; ++index)
    {
        if (index == 2)
            p[2] = p[2] + quarter * driving(2);
        else
        {
            const double kick = quarter * driving(index);
            p[index] = damping[index] * (damping[index] * p[index] + kick);
        }
    }
    *kinetic = K;
    return factor;
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float>
{
    using type = float4;
};
template <>
struct Vec4<double>
{
    using type = double4;
};

// The sentinel that pads the rows of the MD list: slot N of the float4 copy holds
// (kFarAway, kFarAway, kFarAway), far beyond any cutoff from every atom, yet its squared
// distances stay finite in fp32.
template <typename T>
constexpr T kFarAway = T(1e18);

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

// The force kernel's potentials. A pair's two types arrive as the w of the two gathered
// float4s (small integers, exact); bind() runs once per kernel and returns what the pair loop
// evaluates: eval(type_i, type_j, r2, u, fs) with fs = -(du/dr)/r.
//
// LJ with one type: the row's coefficients held in registers for the whole kernel; the hardware
// reciprocal (1 ulp in fp32).
template <typename T>
struct OneTypeLJ
{
    template <typename E>
    static constexpr int waves = sizeof(E) == 4 ? 8 : 1;

    const LJRow<T>* __restrict__ rows;

    struct Bound
    {
        T sigma_squared, four_eps, twentyfour_eps, shift;

        __device__ __forceinline__ void eval(T, T, T r2, T& u, T& fs) const
        {
            const T inverse = reciprocal(r2);
            const T s = sigma_squared * inverse;
            const T p6 = s * s * s;
            const T p12 = p6 * p6;
            u = four_eps * (p12 - p6) - shift;
            fs = twentyfour_eps * ((p12 + p12) - p6) * inverse;
        }
    };

    __device__ __forceinline__ Bound bind() const
    {
        const LJRow<T> c = rows[0];
        return Bound{c.sigma_squared, c.four_eps, c.twentyfour_eps, c.shift};
    }
};

// LJ with several types: the pair's row read from the small, cache-resident coefficient table.
template <typename T>
struct TypedLJ
{
    template <typename E>
    static constexpr int waves = sizeof(E) == 4 ? 5 : 1;

    const LJRow<T>* __restrict__ rows;
    int num_types;

    __device__ __forceinline__ const TypedLJ& bind() const { return *this; }

    __device__ __forceinline__ void eval(T type_i, T type_j, T r2, T& u, T& fs) const
    {
        const LJRow<T> c = rows[pair_row(static_cast<int>(type_i), static_cast<int>(type_j),
                                         num_types)];
        const T inverse = reciprocal(r2);
        const T s = c.sigma_squared * inverse;
        const T p6 = s * s * s;
        const T p12 = p6 * p6;
        u = c.four_eps * (p12 - p6) - c.shift;
        fs = c.twentyfour_eps * ((p12 + p12) - p6) * inverse;
    }
};

// The Catmull-Rom table, its row from the carried types.
template <typename T>
struct TypedTable
{
    template <typename E>
    static constexpr int waves = sizeof(E) == 4 ? 4 : 1;

    Table<T> table;

    __device__ __forceinline__ const TypedTable& bind() const { return *this; }

    __device__ __forceinline__ void eval(T type_i, T type_j, T r2, T& u, T& fs) const
    {
        const int row = table.num_types == 1
                            ? 0
                            : pair_row(static_cast<int>(type_i), static_cast<int>(type_j),
                                       table.num_types);
        table.eval(row, r2, u, fs);
    }
};

template <typename T>
struct StepArgs
{
    // frame, in slot order
    T* positions;
    T* velocities;
    T* forces;
    std::int64_t* images;
    T* masses;
    T* inverse_masses;
    const T* cell;
    T* energy;
    T* virial;
    index_t num_atoms;
    // thermostat
    double* thermostat;
    const double* chain_masses;
    double* kinetic;
    double kT;
    double degrees;
    double timestep;
    int nvt;
    // reuse rule
    T cutoff;
    T radius;
    T* reference_positions;
    T* reference_cell;
    const index_t* limits;
    T* reuse;
    index_t* rebuild;
    index_t* rebuild_count;
    // list
    index_t* status;
    index_t* dims_scratch;
    T* widths_scratch;
    index_t* total_scratch;
    ListHeader* header;
    index_t nx, ny, nz;
    T wx, wy, wz;
    std::int64_t bins;
    std::int64_t stride;
    index_t* bin_counts;
    index_t* bin_offsets;
    index_t* bin_of;
    index_t* unordered;
    index_t* order;
    index_t* counts;
    index_t* index;
    index_t* ids;
    std::int64_t* types;
    T* positions4;
    std::uint8_t* scratch;
    std::uint8_t* scan_temp;
    // the pruned list (inner_index null: none)
    T inner_radius;
    index_t* inner_counts;
    index_t* inner_index;
    T* prune_positions;
};

// The pruned list's slots in the list header (after the 64 max slots): its reuse threshold (the
// squared half allowance, set by the prologue), the refresh flag of this step and a count.
struct PruneState
{
    double threshold;
    index_t refresh;
    index_t refreshes;
};

__device__ __forceinline__ PruneState* prune_state(ListHeader* header)
{
    return reinterpret_cast<PruneState*>(reinterpret_cast<std::uint8_t*>(header) + 640);
}

// Whether the gated refresh of the pruned list runs: after a build (the full rows changed) or
// when the check asked for it, and no status bit is set.
template <typename T>
__device__ __forceinline__ bool refreshing(const StepArgs<T>& a)
{
    return a.status[0] == 0 &&
           (a.rebuild[0] != 0 || prune_state(a.header)->refresh != 0);
}

// ---- prologue -------------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(kBlock) prologue_kernel(StepArgs<T> a, const ndn_potential pot,
                                                          void* coefficients)
{
    if (threadIdx.x == 0)
    {
        if (a.nvt)
        {
            double* t = a.thermostat;
            t[kOpening] = chain_half_step(a.kinetic, t + kChainQ, t + kChainP, a.chain_masses,
                                          a.degrees, a.kT, a.timestep);
        }
        prepare_update(a.cell, a.reference_cell, a.cutoff, a.radius, a.limits, a.dims_scratch,
                       a.widths_scratch, a.reuse, a.rebuild, a.status, a.total_scratch, a.header);
        if (a.inner_index)
        {
            // the outer rule's allowance with the pruned list's radius (the cell is fixed
            // between refreshes); none left means a refresh every step
#pragma clang fp contract(off)
            const T epsilon = Epsilon<T>::value;
            T longest = a.cell[0];
            for (int axis = 1; axis < 3; ++axis) longest = a.cell[axis] > longest ? a.cell[axis] : longest;
            const T allowance = a.inner_radius - a.cutoff - 32 * epsilon * longest;
            PruneState* prune = prune_state(a.header);
            prune->threshold = allowance > 0 ? static_cast<double>(allowance * allowance / 4) : -1.0;
            prune->refresh = 0;
        }
    }
    const int rows = pot.num_types * (pot.num_types + 1) / 2;
    if (pot.kind == NDN_LJ)
    {
        auto* out = static_cast<LJRow<T>*>(coefficients);
        const T* sigma = static_cast<const T*>(pot.sigma);
        const T* epsilon = static_cast<const T*>(pot.epsilon);
        const T cutoff_squared = static_cast<T>(pot.cutoff * pot.cutoff);
        for (int row = threadIdx.x; row < rows; row += blockDim.x)
        {
            const T s2 = sigma[row] * sigma[row];
            const T e = epsilon[row];
            const T base = s2 / cutoff_squared;
            const T p6 = base * base * base;
            out[row] = LJRow<T>{s2, T(4) * e, T(24) * e, T(4) * e * (p6 * p6 - p6)};
        }
    }
    else
    {
        const int knots = static_cast<int>(pot.knots);
        const std::int64_t width = knots + 2;
        T* padded = static_cast<T*>(coefficients);
        T* inverse = padded + rows * width;
        const T* table = static_cast<const T*>(pot.table);
        const T* spacing = static_cast<const T*>(pot.spacing);
        for (std::int64_t i = threadIdx.x; i < rows * width; i += blockDim.x)
        {
            const std::int64_t row = i / width;
            const std::int64_t column = i % width;
            const T* v = table + row * knots;
            auto value = [&](std::int64_t k) { return k == knots - 1 ? T(0) : v[k]; };
            T out;
            if (column == 0)
                out = T(3) * value(0) - T(3) * value(1) + value(2);
            else if (column == width - 1)
                out = T(3) * value(knots - 1) - T(3) * value(knots - 2) + value(knots - 3);
            else
                out = value(column - 1);
            padded[i] = out;
        }
        for (int row = threadIdx.x; row < rows; row += blockDim.x) inverse[row] = T(1) / spacing[row];
    }
}

// ---- first half -----------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(kBlock) first_half_kernel(StepArgs<T> a, T timestep,
                                                            T half_timestep)
{
#pragma clang fp contract(off)
    using V4 = typename Vec4<T>::type;
    __shared__ T block_max[kBlock / kWave];
    const std::int64_t atom = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool check = a.status[0] == 0;
    bool finite = true;
    T moved = 0, moved_inner = 0;
    if (atom < a.num_atoms)
    {
        T* v = a.velocities + 3 * atom;
        const T* f = a.forces + 3 * atom;
        T* x = a.positions + 3 * atom;
        const T inverse_mass = a.inverse_masses[atom];
        T vx = v[0], vy = v[1], vz = v[2];
        if (a.nvt)
        {
            // the previous step's closing factor, then this step's opening factor
            const T closing = static_cast<T>(a.thermostat[kPending]);
            const T opening = static_cast<T>(a.thermostat[kOpening]);
            vx = vx * closing;
            vy = vy * closing;
            vz = vz * closing;
            vx = vx * opening;
            vy = vy * opening;
            vz = vz * opening;
        }
        vx = vx + half_timestep * f[0] * inverse_mass;
        vy = vy + half_timestep * f[1] * inverse_mass;
        vz = vz + half_timestep * f[2] * inverse_mass;
        T out[3];
        const T vel[3] = {vx, vy, vz};
        std::int64_t crossings[3];
        bool crossed = false;
        for (int axis = 0; axis < 3; ++axis)
        {
            const T length = a.cell[axis];
            const T drifted = x[axis] + timestep * vel[axis];
            T k = floor(drifted / length);
            T w = drifted - k * length;
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
            out[axis] = w;
            crossings[axis] = static_cast<std::int64_t>(k);
            crossed = crossed || crossings[axis] != 0;
        }
        v[0] = vx;
        v[1] = vy;
        v[2] = vz;
        x[0] = out[0];
        x[1] = out[1];
        x[2] = out[2];
        V4 p4;
        p4.x = out[0];
        p4.y = out[1];
        p4.z = out[2];
        p4.w = a.types ? static_cast<T>(a.types[atom]) : T(0);
        reinterpret_cast<V4*>(a.positions4)[atom] = p4;
        if (crossed)
        {
            std::int64_t* image = a.images + 3 * atom;
            image[0] += crossings[0];
            image[1] += crossings[1];
            image[2] += crossings[2];
        }
        if (check)
        {
            finite = isfinite(out[0]) && isfinite(out[1]) && isfinite(out[2]);
            if (finite)
            {
                moved = moved_squared(out[0], out[1], out[2], a.reference_positions + 3 * atom,
                                      a.reuse, a.cell);
                if (a.inner_index)
                {
                    const T* x_prune = a.prune_positions + 3 * atom;
                    const T dx = minimum_image(out[0], x_prune[0], a.cell[0]);
                    const T dy = minimum_image(out[1], x_prune[1], a.cell[1]);
                    const T dz = minimum_image(out[2], x_prune[2], a.cell[2]);
                    moved_inner = ((dx * dx) + dy * dy) + dz * dz;
                }
            }
        }
    }
    // every thread reaches the barrier below, whatever status a wave read
    if (!finite) atomicOr(a.status, kStatusNonFinite);
    const T wave_moved = wave_max(finite ? moved : T(0));
    const T wave_moved_inner = a.inner_index ? wave_max(finite ? moved_inner : T(0)) : T(0);
    const int wave = threadIdx.x / kWave;
    if (lane_id() == 0)
    {
        block_max[wave] = wave_moved;
        if (check && wave_moved >= a.reuse[0] && a.rebuild[0] == 0) a.rebuild[0] = 1;
        if (check && a.inner_index)
        {
            PruneState* prune = prune_state(a.header);
            if (static_cast<double>(wave_moved_inner) >= prune->threshold && prune->refresh == 0)
                prune->refresh = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        T m = block_max[0];
        for (int w = 1; w < kBlock / kWave; ++w) m = block_max[w] > m ? block_max[w] : m;
        if (m > T(0)) atomicMax(moved_slot(a.header, blockIdx.x), order_bits(m));
    }
}

// ---- rebuild --------------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ index_t sub_bin(T x, T width, index_t count)
{
    const index_t bin = static_cast<index_t>(floor(x / width));
    return bin < 0 ? 0 : (bin >= count ? count - 1 : bin);
}

// Sub-bin numbering: x fastest along each (y, z) row of sub-bins, so every row is one contiguous
// run of slots; the rows in blocks of kLayerBlock z-layers, y outer and z inner within a block.
// The force kernel walks the slots in this order, so the rows it meets within a short stretch
// span a few y-rows of all the block's layers: the positions a row's neighbors need were
// gathered moments before by the rows beside it, and stay in L2 (x-y-z order would revisit each
// position once per layer of the stencil, 5 times, a whole layer of work apart).
constexpr index_t kLayerBlock = 8;

__device__ __forceinline__ std::int64_t row_first_bin(index_t y, index_t z, index_t nx,
                                                      index_t ny, index_t nz)
{
    const index_t block = z / kLayerBlock;
    const index_t layer = z - block * kLayerBlock;
    const index_t depth =
        nz - block * kLayerBlock < kLayerBlock ? nz - block * kLayerBlock : kLayerBlock;
    return static_cast<std::int64_t>(nx) *
           (static_cast<std::int64_t>(ny) * kLayerBlock * block +
            static_cast<std::int64_t>(y) * depth + layer);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) md_assign_kernel(StepArgs<T> a, int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t slot = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         slot < a.num_atoms; slot += static_cast<std::int64_t>(gridDim.x) * blockDim.x)
    {
        const T* p = a.positions + 3 * slot;
        const index_t bin = static_cast<index_t>(
            sub_bin(p[0], a.wx, a.nx) + row_first_bin(sub_bin(p[1], a.wy, a.ny),
                                                      sub_bin(p[2], a.wz, a.nz), a.nx, a.ny, a.nz));
        a.bin_of[slot] = bin;
        atomicAdd(a.bin_counts + bin, 1);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) md_scatter_kernel(StepArgs<T> a, int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t slot = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         slot < a.num_atoms; slot += static_cast<std::int64_t>(gridDim.x) * blockDim.x)
    {
        const index_t bin = a.bin_of[slot];
        a.unordered[a.bin_offsets[bin] + atomicSub(a.bin_counts + bin, 1) - 1] =
            static_cast<index_t>(slot);
    }
}

// One atom of the frame, as the permutation moves it.
template <typename T>
struct Packed
{
    T x[3], v[3], f[3];
    std::int64_t image[3];
    T mass, inverse_mass;
    std::int64_t type;
    index_t id;
};

// One thread per atom, in the sub-bin-sorted (unordered within a sub-bin) array: ranks the atom
// by atom id among its sub-bin's atoms (a handful), so the new order is (sub-bin, atom id)
// whatever the atomics did, and gathers its whole frame entry into scratch at its new slot;
// md_unpack_kernel puts the frame home.
template <typename T>
__global__ void __launch_bounds__(kBlock) md_order_kernel(StepArgs<T> a, int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < a.num_atoms; i += static_cast<std::int64_t>(gridDim.x) * blockDim.x)
    {
        const std::int64_t old = a.unordered[i];
        const index_t bin = a.bin_of[old];
        const index_t begin = a.bin_offsets[bin];
        const index_t end = a.bin_offsets[bin + 1];
        const index_t id = a.ids[old];
        index_t rank = 0;
        for (index_t k = begin; k < end; ++k) rank += a.ids[a.unordered[k]] < id;
        const std::int64_t slot = begin + rank;
        a.order[slot] = static_cast<index_t>(old);
        Packed<T> p;
        for (int c = 0; c < 3; ++c)
        {
            p.x[c] = a.positions[3 * old + c];
            p.v[c] = a.velocities[3 * old + c];
            p.f[c] = a.forces[3 * old + c];
            p.image[c] = a.images[3 * old + c];
        }
        p.mass = a.masses[old];
        p.inverse_mass = a.inverse_masses[old];
        p.type = a.types ? a.types[old] : 0;
        p.id = id;
        reinterpret_cast<Packed<T>*>(a.scratch)[slot] = p;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) md_unpack_kernel(StepArgs<T> a, int forced)
{
    using V4 = typename Vec4<T>::type;
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t slot = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         slot < a.num_atoms; slot += static_cast<std::int64_t>(gridDim.x) * blockDim.x)
    {
        const Packed<T> p = reinterpret_cast<const Packed<T>*>(a.scratch)[slot];
        for (int k = 0; k < 3; ++k)
        {
            a.positions[3 * slot + k] = p.x[k];
            // the rebuild's commit (update.cuh:119-145): the reference frame is the new frame
            a.reference_positions[3 * slot + k] = p.x[k];
            a.velocities[3 * slot + k] = p.v[k];
            a.forces[3 * slot + k] = p.f[k];
            a.images[3 * slot + k] = p.image[k];
        }
        a.masses[slot] = p.mass;
        a.inverse_masses[slot] = p.inverse_mass;
        if (a.types) a.types[slot] = p.type;
        a.ids[slot] = p.id;
        V4 p4;
        p4.x = p.x[0];
        p4.y = p.x[1];
        p4.z = p.x[2];
        p4.w = static_cast<T>(p.type);
        reinterpret_cast<V4*>(a.positions4)[slot] = p4;
        if (slot == 0)
        {
            for (int k = 0; k < 3; ++k) a.reference_cell[k] = a.cell[k];
            if (a.rebuild_count) ++a.rebuild_count[0];
            // the sentinel that pads the list's rows
            V4 far;
            far.x = far.y = far.z = kFarAway<T>;
            far.w = T(0);
            reinterpret_cast<V4*>(a.positions4)[a.num_atoms] = far;
        }
    }
}

// Distance from coordinate c to the slab [lo, lo + w): 0 inside.
template <typename T>
__device__ __forceinline__ T slab_distance(T c, T lo, T w)
{
    const T below = lo - c;
    const T above = c - (lo + w);
    return below > T(0) ? below : (above > T(0) ? above : T(0));
}

// The sub-bin numbered `bin` (the inverse of row_first_bin plus x).
__device__ __forceinline__ void bin_coordinates(std::int64_t bin, index_t nx, index_t ny,
                                                index_t nz, index_t* x, index_t* y, index_t* z)
{
    const std::int64_t per_block = static_cast<std::int64_t>(nx) * ny * kLayerBlock;
    const index_t block = static_cast<index_t>(bin / per_block);
    const std::int64_t rest = bin - per_block * block;
    const index_t depth =
        nz - block * kLayerBlock < kLayerBlock ? nz - block * kLayerBlock : kLayerBlock;
    const index_t row = static_cast<index_t>(rest / nx);
    *x = static_cast<index_t>(rest - static_cast<std::int64_t>(row) * nx);
    *y = row / depth;
    *z = block * kLayerBlock + row % depth;
}

// Distance between the sub-bin a stencil offset o (-2..2) away and the center's sub-bin, along
// one axis: the gap of |o| - 1 sub-bins between them.
template <typename T>
__device__ __forceinline__ T bin_gap(int offset, T width)
{
    const int steps = (offset < 0 ? -offset : offset) - 1;
    return steps > 0 ? static_cast<T>(steps) * width : T(0);
}

// One wave per sub-bin. Its atoms (the centers) share one stencil: lanes 0..49 hold the slot
// ranges of the 25 stencil rows (runs of x-consecutive sub-bins; two parts where a run wraps
// around x), culled by the distance between sub-bins, so every sub-bin a pair within R could
// touch is kept. The candidates are packed 64 at a time; each chunk is gathered once and tested
// against every center in turn with the exact minimum-image expression, and each center's hits
// are ballot-compacted into its row. Rows are then padded to the stride with the sentinel slot
// N (a far-away point, never within any cutoff), so the force kernel reads whole rows unmasked.
// With stride 0 it only counts (the first frame, to size the rows).
template <typename T>
__global__ void __launch_bounds__(kBlock) md_fill_kernel(StepArgs<T> a, T radius_squared,
                                                         T cull_squared, int forced)
{
#pragma clang fp contract(off)
    using V4 = typename Vec4<T>::type;
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    const int lane = lane_id();
    const V4* pos = reinterpret_cast<const V4*>(a.positions4);
    const T lx = a.cell[0], ly = a.cell[1], lz = a.cell[2];
    const index_t nx = a.nx, ny = a.ny, nz = a.nz;
    const bool full_x = nx < kStencil, full_y = ny < kStencil, full_z = nz < kStencil;
    const int span_x = full_x ? nx : kStencil;
    const int span_y = full_y ? ny : kStencil;
    const int span_z = full_z ? nz : kStencil;
    const index_t sentinel = a.num_atoms;
    // one wave per sub-bin, the waves of the grid striding over the sub-bins
    const std::int64_t waves = static_cast<std::int64_t>(gridDim.x) * (kBlock / kWave);
    for (std::int64_t bin =
             static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
         bin < a.bins; bin += waves)
    {
        const index_t begin = a.bin_offsets[bin];
        const index_t end = a.bin_offsets[bin + 1];
        if (begin == end) continue;
        index_t hx, hy, hz;
        bin_coordinates(bin, nx, ny, nz, &hx, &hy, &hz);
        // this lane's range: stencil row lane / 2, part lane % 2 of its run of x sub-bins
        index_t lo = 0, length = 0;
        const int row = lane / 2, part = lane % 2;
        if (row < span_y * span_z)
        {
            const int oy = row % span_y, oz = row / span_y;
            const T dy = full_y ? T(0) : bin_gap(oy - 2, a.wy);
            const T dz = full_z ? T(0) : bin_gap(oz - 2, a.wz);
            const T dyz = dy * dy + dz * dz;
            if (dyz < cull_squared)
            {
                int first = span_x, last = -1;
                for (int ox = 0; ox < span_x; ++ox)
                {
                    const T dx = full_x ? T(0) : bin_gap(ox - 2, a.wx);
                    if (dx * dx + dyz < cull_squared)
                    {
                        first = ox < first ? ox : first;
                        last = ox;
                    }
                }
                if (last >= first)
                {
                    const index_t by = full_y ? oy : hy + oy - 2;
                    const index_t bz = full_z ? oz : hz + oz - 2;
                    const index_t wy = by < 0 ? by + ny : (by >= ny ? by - ny : by);
                    const index_t wz = bz < 0 ? bz + nz : (bz >= nz ? bz - nz : bz);
                    const std::int64_t base = row_first_bin(wy, wz, nx, ny, nz);
                    const index_t x0 = full_x ? first : hx + first - 2;
                    const index_t x1 = full_x ? last : hx + last - 2;
                    // a run that wraps around the x faces splits into two contiguous runs
                    index_t run_lo = 0, run_hi = -1;
                    if (x0 < 0)
                    {
                        run_lo = part == 0 ? x0 + nx : 0;
                        run_hi = part == 0 ? nx - 1 : x1;
                    }
                    else if (x1 >= nx)
                    {
                        run_lo = part == 0 ? x0 : 0;
                        run_hi = part == 0 ? nx - 1 : x1 - nx;
                    }
                    else if (part == 0)
                    {
                        run_lo = x0;
                        run_hi = x1;
                    }
                    if (run_hi >= run_lo)
                    {
                        lo = a.bin_offsets[base + run_lo];
                        length = a.bin_offsets[base + run_hi + 1] - lo;
                    }
                }
            }
        }
        const index_t inclusive = wave_inclusive_sum(length);
        const index_t start = inclusive - length;
        const index_t total = __shfl(inclusive, kWave - 1, kWave);
        // near a face (or with a short axis) the images come from the minimum image per pair
        const bool wraps = full_x || full_y || full_z || hx < 2 || hx + 2 >= nx || hy < 2 ||
                           hy + 2 >= ny || hz < 2 || hz + 2 >= nz;
        // the centers, 64 at a time: lane k holds center first + k and its count
        for (index_t first = begin; first < end; first += kWave)
        {
            const int centers = end - first < kWave ? end - first : kWave;
            const V4 mine = pos[lane < centers ? first + lane : first];
            index_t mine_count = 0;
            for (index_t chunk = 0; chunk < total; chunk += kWave)
            {
                const index_t t = chunk + lane;
                // the range holding t: the last lane whose start is <= t
                int r = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1)
                {
                    const index_t s = __shfl(start, r + step, kWave);
                    if (s <= t) r += step;
                }
                const bool valid = t < total;
                const index_t slot = __shfl(lo, r, kWave) + (t - __shfl(start, r, kWave));
                const V4 q = pos[valid ? slot : first];
                for (int k = 0; k < centers; ++k)
                {
                    const index_t center = first + k;
                    const T cx = read_lane(mine.x, k);
                    const T cy = read_lane(mine.y, k);
                    const T cz = read_lane(mine.z, k);
                    T dx, dy, dz;
                    if (wraps)
                    {
                        dx = minimum_image(cx, q.x, lx);
                        dy = minimum_image(cy, q.y, ly);
                        dz = minimum_image(cz, q.z, lz);
                    }
                    else
                    {
                        dx = cx - q.x;
                        dy = cy - q.y;
                        dz = cz - q.z;
                    }
                    const T xx = dx * dx;
                    const T yy = dy * dy;
                    const T zz = dz * dz;
                    const bool hit = valid && slot != center && (xx + yy) + zz < radius_squared;
                    const std::uint64_t mask = ballot(hit);
                    const index_t count = read_lane(mine_count, k);
                    if (hit && a.stride > 0)
                    {
                        const index_t position = count + bits_below(mask);
                        if (position < a.stride) a.index[center * a.stride + position] = slot;
                    }
                    if (lane == k) mine_count += static_cast<index_t>(__popcll(mask));
                }
            }
            // the centers' counts, the overflow flag, and the sentinel padding of each row
            if (lane < centers)
            {
                a.counts[first + lane] = mine_count;
                if (a.stride > 0 && mine_count > a.stride)
                {
                    atomicOr(a.status, kStatusFullOverflow);
                    atomicMax(a.status + 1, mine_count);
                }
            }
            if (a.stride > 0)
                for (int k = 0; k < centers; ++k)
                {
                    const index_t count = read_lane(mine_count, k);
                    index_t* row_out = a.index + static_cast<std::int64_t>(first + k) * a.stride;
                    for (index_t p = count + lane; p < a.stride; p += kWave) row_out[p] = sentinel;
                }
        }
    }
}

// Four list entries, read once per step: a non-temporal load, so the list streaming through L2
// does not evict the positions the gathers reuse.
typedef index_t Quad __attribute__((ext_vector_type(4)));

__device__ __forceinline__ Quad stream_quad(const index_t* entries)
{
    return __builtin_nontemporal_load(reinterpret_cast<const Quad*>(entries));
}

// The position of slot j: a 32-bit byte offset from the float4 copy's base (a uniform pointer),
// which a gather takes in one instruction (the host keeps the copy under 4 GiB).
template <typename V4>
__device__ __forceinline__ V4 gather(const V4* __restrict__ pos, index_t j)
{
    const std::uint32_t offset =
        static_cast<std::uint32_t>(j) * static_cast<std::uint32_t>(sizeof(V4));
    return *reinterpret_cast<const V4*>(reinterpret_cast<const char*>(pos) + offset);
}

// Refreshes the pruned list from the full rows: four lanes per row, each reading 4 consecutive
// entries per pass (the group 16), keeping those within the pruned radius (the exact
// minimum-image test) in row order by ballot compaction. Since every pair kept is in the full
// row, a pruned row never outgrows the stride. Rows come in aligned groups of 64 (one block); a
// row is padded with the sentinel slot N up to the group's largest count rounded to 4, and every
// force schedule's wave (64 / G consecutive rows, aligned) lies inside one such group, so the
// force kernel never reads past the padding. The frame of the refresh is kept for the check.
template <typename T>
__global__ void __launch_bounds__(kBlock) md_prune_kernel(StepArgs<T> a, T inner_radius,
                                                          T inner_squared, int forced)
{
#pragma clang fp contract(off)
    using V4 = typename Vec4<T>::type;
    if (!forced && !refreshing(a)) return;
    constexpr int G = 4;
    constexpr int kRows = kBlock / G;
    __shared__ index_t group_max[kBlock / kWave];
    const int lane = lane_id();
    const int sub = lane % G;
    const std::uint64_t mine = 0xFull << (lane - sub);  // the lanes of this row
    const std::uint64_t below = (1ull << lane) - 1ull;
    const V4* __restrict__ pos = reinterpret_cast<const V4*>(a.positions4);
    const T lx = a.cell[0], ly = a.cell[1], lz = a.cell[2];
    const index_t sentinel = a.num_atoms;
    const std::int64_t groups = (static_cast<std::int64_t>(a.num_atoms) + kRows - 1) / kRows;
    if (blockIdx.x == 0 && threadIdx.x == 0) ++prune_state(a.header)->refreshes;
    for (std::int64_t g = blockIdx.x; g < groups; g += gridDim.x)
    {
        const std::int64_t row = g * kRows + threadIdx.x / G;
        const bool live = row < a.num_atoms;
        const V4 c = pos[live ? row : 0];
        // the interior fast path of the force kernel: a wave whose atoms all sit farther than
        // the pruned radius from every face takes x_i - x_j directly (the minimum image of every
        // pair that can be kept)
        const T low = inner_radius + T(1e-3);
        const bool interior = !live || (c.x > low && c.x < lx - low && c.y > low &&
                                        c.y < ly - low && c.z > low && c.z < lz - low);
        const bool fast = ballot(!interior) == 0ull;
        index_t kept = 0;
        if (live)
        {
            const index_t listed = a.counts[row];
            const index_t count = listed < a.stride ? listed : static_cast<index_t>(a.stride);
            const index_t* __restrict__ in_row = a.index + row * a.stride;
            index_t* __restrict__ out_row = a.inner_index + row * a.stride;
            // every lane of the row takes the same number of passes (its ballots stay whole)
            for (index_t pass = 0; pass < count; pass += 4 * G)
            {
                const index_t base = pass + 4 * sub;
                const bool any = base < count;
                // entries in [count, stride) are the sentinel; base + 3 < stride when base < count
                const Quad quad = stream_quad(in_row + (any ? base : 0));
                const index_t js[4] = {quad.x, quad.y, quad.z, quad.w};
                // all four gathers before any store
                const V4 q[4] = {gather(pos, js[0]), gather(pos, js[1]), gather(pos, js[2]),
                                 gather(pos, js[3])};
                bool hits[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                {
                    T dx, dy, dz;
                    if (fast)
                    {
                        dx = c.x - q[u].x;
                        dy = c.y - q[u].y;
                        dz = c.z - q[u].z;
                    }
                    else
                    {
                        dx = displacement(c.x, q[u].x, lx);
                        dy = displacement(c.y, q[u].y, ly);
                        dz = displacement(c.z, q[u].z, lz);
                    }
                    hits[u] = any && exact_squared(dx, dy, dz) < inner_squared;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                {
                    const std::uint64_t mask = ballot(hits[u]) & mine;
                    if (hits[u]) out_row[kept + __popcll(mask & below)] = js[u];
                    kept += static_cast<index_t>(__popcll(mask));
                }
            }
            if (sub < 3) a.prune_positions[3 * row + sub] = a.positions[3 * row + sub];
            if (sub == 0) a.inner_counts[row] = kept;
        }
        // the group's largest count, rounded to 4: the padding every row of the group gets
        const index_t wave_largest = wave_max(kept);
        if (lane == 0) group_max[threadIdx.x / kWave] = wave_largest;
        __syncthreads();
        index_t largest = group_max[0];
        for (int w = 1; w < kBlock / kWave; ++w) largest = group_max[w] > largest ? group_max[w] : largest;
        __syncthreads();
        const index_t padded = (largest + 3) / 4 * 4;
        if (live)
        {
            index_t* out_row = a.inner_index + row * a.stride;
            for (index_t p = kept + sub; p < padded; p += G) out_row[p] = sentinel;
        }
    }
}

// ---- forces ---------------------------------------------------------------------------------

inline unsigned xcd_grid(std::int64_t blocks)
{
    return static_cast<unsigned>((blocks + 7) / 8 * 8);
}

template <int G, typename V>
__device__ __forceinline__ V group_sum(V value)
{
#pragma unroll
    for (int offset = G / 2; offset > 0; offset >>= 1) value += __shfl_xor(value, offset, kWave);
    return value;
}



// One group's walk over its row: each lane takes 4 consecutive entries (one 16-byte load) per
// pass, the group 4G entries, and the lane's four position gathers are issued together. A row is
// padded to the stride with the sentinel slot N, a point far from every atom, so entries past
// the row's count fail the cutoff test by themselves and need no mask. The row start is 16-byte
// aligned and `width` <= stride, a multiple of 8, so no read leaves the row.
//
// In refresh mode (the step that refreshes the pruned list) the walk reads a full row and also
// keeps its pairs within the pruned radius: each lane's four are ballot-compacted, in row order,
// into the pruned row `kept_row`, and `kept` counts them (the same on every lane of the group:
// every lane takes the same passes).
template <int G, bool kFast, bool kVirial, bool kRefresh = false, typename T, typename V4,
          typename Ev>
__device__ __forceinline__ void walk_row(const Ev& ev, const V4& c, const V4* __restrict__ pos,
                                         const index_t* __restrict__ entries, index_t width,
                                         T lx, T ly, T lz, T cutoff_squared, T& fx, T& fy, T& fz,
                                         T& energy, T* w, T inner_squared = T(0),
                                         index_t* __restrict__ kept_row = nullptr,
                                         index_t* kept = nullptr)
{
    const int lane = lane_id();
    const std::uint64_t group = ((1ull << G) - 1ull) << (lane - lane % G);
    const std::uint64_t below = (1ull << lane) - 1ull;
    // A plain walk lets each lane stop at its own last pass. A refresh carries the compaction
    // count across the lanes of a row, so there every lane takes the same passes: a lane whose
    // entries start at or past `width` reads the row's first entries (valid slots) and keeps
    // and counts nothing.
    const index_t start = kRefresh ? 0 : 4 * (lane % G);
    const index_t offset = kRefresh ? 4 * (lane % G) : 0;
    for (index_t pass = start; pass < width; pass += 4 * G)
    {
        const index_t base = pass + offset;
        const bool any = !kRefresh || base < width;
        const Quad quad = stream_quad(entries + (any ? base : 0));
        const V4 q[4] = {gather(pos, quad.x), gather(pos, quad.y), gather(pos, quad.z),
                         gather(pos, quad.w)};
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            T dx, dy, dz;
            if (kFast)
            {
                dx = c.x - q[u].x;
                dy = c.y - q[u].y;
                dz = c.z - q[u].z;
            }
            else
            {
                dx = displacement(c.x, q[u].x, lx);
                dy = displacement(c.y, q[u].y, ly);
                dz = displacement(c.z, q[u].z, lz);
            }
            const T r2 = exact_squared(dx, dy, dz);
            const bool inside = any && r2 < cutoff_squared;
            if (kRefresh)
            {
                const bool keep = any && r2 < inner_squared;
                const std::uint64_t mask = ballot(keep) & group;
                if (keep) kept_row[*kept + __popcll(mask & below)] = (u == 0 ? quad.x : u == 1 ? quad.y : u == 2 ? quad.z : quad.w);
                *kept += static_cast<index_t>(__popcll(mask));
            }
            T pu, fs;
            ev.eval(c.w, q[u].w, r2, pu, fs);
            // a select, not a product: a masked pair's fs may be inf or NaN
            pu = inside ? pu : T(0);
            fs = inside ? fs : T(0);
            energy += pu;
            fx += fs * dx;
            fy += fs * dy;
            fz += fs * dz;
            if (kVirial)
            {
                w[0] += dx * (fs * dx);
                w[1] += dy * (fs * dy);
                w[2] += dz * (fs * dz);
                w[3] += dx * (fs * dy);
                w[4] += dx * (fs * dz);
                w[5] += dy * (fs * dz);
            }
        }
    }
}

// Waves per SIMD the force kernel is compiled for: each potential's `waves`, the most its fp32
// kernels without the virial reach without spilling (one-type LJ 8, typed LJ 5, table 4). The virial's six more sums, and fp64, need more
// registers: those are left to the compiler.
template <typename T, class Pot, bool kVirial, int kMode>
constexpr int kForceWaves = kVirial || kMode == 2 ? 1 : Pot::template waves<T>;

// The full list's rows in tiles of kBlock / G rows, G lanes per row. The grid is persistent: the
// blocks of one XCD (blockIdx % 8; blocks are dealt to the 8 XCDs round-robin) walk a contiguous
// eighth of the tiles, so neighboring rows, which gather the same positions, share that XCD's
// L2. Each block sums its E, W and K in a fixed order into its partial slot, so every run gives
// the same bits.
//
// Images are chosen once per wave: a wave whose atoms all sit farther than the cutoff (plus a
// margin) from every face takes x_i - x_j directly, which is the minimum image for every pair
// within the cutoff; any other image it may pick is longer than L/2 > rc, so the mask drops it
// exactly as the reference does. Other waves use the reference's displacement (space.py:11-19)
// per pair.
// kMode: 0 reads the pruned rows (the full rows when there are none); in a step with a pruned
// list two launches share the work, each gated on the device's refresh flag: mode 1 reads the
// pruned rows and runs when there is no refresh, mode 2 runs when there is one, reading the full
// rows and writing the pruned rows, their counts (each row padded with the sentinel up to the
// tile's largest count rounded to 4; tiles are at least 32 aligned rows, and so hold every wave
// of every schedule of at most 8 lanes per row) and the frame of the refresh.
//
// The partial slots past this launch's grid, up to `slots`, are zeroed: the two launches of a
// step may have different grids and the finish reads the larger.
template <typename T, class Pot, int G, bool kVirial, bool kKick, int kMode = 0>
__global__ void __launch_bounds__(kBlock, (kForceWaves<T, Pot, kVirial, kMode>)) force_kernel(StepArgs<T> a, Pot pot, T cutoff,
                                                       T cutoff_squared, T half_timestep,
                                                       double* __restrict__ partials,
                                                       index_t tiles, index_t slots)
{
    using V4 = typename Vec4<T>::type;
    constexpr bool kRefresh = kMode == 2;
    if (kMode == 1 && refreshing(a)) return;
    if (kMode == 2 && !refreshing(a)) return;
    __shared__ double shared[16];
    __shared__ index_t tile_largest[kBlock / kWave];
    constexpr int kRows = kBlock / G;
    const auto ev = pot.bind();
    const int sub = lane_id() % G;
    const V4* __restrict__ pos = reinterpret_cast<const V4*>(a.positions4);
    const T lx = a.cell[0], ly = a.cell[1], lz = a.cell[2];
    // a refresh keeps pairs out to the pruned radius: the direct difference must be their
    // minimum image too
    const T low = (kRefresh ? a.inner_radius : cutoff) + T(1e-3);
    const T inner_squared = kRefresh ? a.inner_radius * a.inner_radius : T(0);
    if (kRefresh && blockIdx.x == 0 && threadIdx.x == 0) ++prune_state(a.header)->refreshes;
    const unsigned xcd = blockIdx.x % 8;
    const index_t first = static_cast<index_t>(static_cast<std::int64_t>(tiles) * xcd / 8);
    const index_t last = static_cast<index_t>(static_cast<std::int64_t>(tiles) * (xcd + 1) / 8);
    const index_t step = static_cast<index_t>(gridDim.x / 8);
    double energy_sum = 0.0, kinetic = 0.0;
    double w_sum[6] = {0, 0, 0, 0, 0, 0};
    for (index_t tile = first + static_cast<index_t>(blockIdx.x / 8); tile < last; tile += step)
    {
        const index_t row = tile * kRows + static_cast<index_t>(threadIdx.x) / G;
        const bool live = row < a.num_atoms;
        // a row past the last atom walks row 0's entries from a far point of its own (apart from
        // the sentinel too), so none of its pairs is within the cutoff
        V4 c = pos[live ? row : 0];
        if (!live) c.x = c.y = c.z = -kFarAway<T>;
        // the pruned rows when there are any (not on a refresh); a full row that overflowed
        // holds only `stride` entries (the status bit reports it)
        const bool pruned = !kRefresh && a.inner_index != nullptr;
        const index_t* counts = pruned ? a.inner_counts : a.counts;
        const index_t listed = live ? counts[row] : 0;
        const index_t count = listed < a.stride ? listed : static_cast<index_t>(a.stride);
        const bool interior = live && c.x > low && c.x < lx - low && c.y > low &&
                              c.y < ly - low && c.z > low && c.z < lz - low;
        const bool fast = ballot(!interior) == 0ull;
        const index_t width = wave_max(count);
        const std::int64_t row_start = static_cast<std::int64_t>(live ? row : 0) * a.stride;
        const index_t* entries = (pruned ? a.inner_index : a.index) + row_start;
        T fx = 0, fy = 0, fz = 0, energy = 0;
        T w[6] = {0, 0, 0, 0, 0, 0};
        index_t kept = 0;
        index_t* kept_row = kRefresh ? a.inner_index + row_start : nullptr;
        if (fast)
            walk_row<G, true, kVirial, kRefresh>(ev, c, pos, entries, width, lx, ly, lz,
                                                 cutoff_squared, fx, fy, fz, energy, w,
                                                 inner_squared, kept_row, &kept);
        else
            walk_row<G, false, kVirial, kRefresh>(ev, c, pos, entries, width, lx, ly, lz,
                                                  cutoff_squared, fx, fy, fz, energy, w,
                                                  inner_squared, kept_row, &kept);
        if (kRefresh)
        {
            // the tile's largest pruned count, rounded to 4: every row's padding
            const index_t wave_largest = wave_max(live ? kept : 0);
            if (lane_id() == 0) tile_largest[threadIdx.x / kWave] = wave_largest;
            __syncthreads();
            index_t largest = tile_largest[0];
            for (int k = 1; k < kBlock / kWave; ++k)
                largest = tile_largest[k] > largest ? tile_largest[k] : largest;
            __syncthreads();
            const index_t padded = (largest + 3) / 4 * 4;
            if (live)
            {
                for (index_t p = kept + sub; p < padded; p += G) kept_row[p] = a.num_atoms;
                if (sub == 0) a.inner_counts[row] = kept;
                if (sub < 3) a.prune_positions[3LL * row + sub] = a.positions[3LL * row + sub];
            }
        }
        fx = group_sum<G>(fx);
        fy = group_sum<G>(fy);
        fz = group_sum<G>(fz);
        energy_sum += static_cast<double>(energy);
        if (kVirial)
            for (int k = 0; k < 6; ++k) w_sum[k] += static_cast<double>(w[k]);
        if (live && sub == 0)
        {
            T* f = a.forces + 3LL * row;
            f[0] = fx;
            f[1] = fy;
            f[2] = fz;
            if (kKick)
            {
#pragma clang fp contract(off)
                T* v = a.velocities + 3LL * row;
                const T inverse_mass = a.inverse_masses[row];
                const T vx = v[0] + half_timestep * fx * inverse_mass;
                const T vy = v[1] + half_timestep * fy * inverse_mass;
                const T vz = v[2] + half_timestep * fz * inverse_mass;
                v[0] = vx;
                v[1] = vy;
                v[2] = vz;
                kinetic += 0.5 * static_cast<double>(a.masses[row]) *
                           (static_cast<double>(vx) * vx + static_cast<double>(vy) * vy +
                            static_cast<double>(vz) * vz);
            }
        }
    }
    double* out = partials + 8LL * blockIdx.x;
    const double e = block_sum(energy_sum, shared);
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (kVirial)
        for (int k = 0; k < 6; ++k) v[k] = block_sum(w_sum[k], shared);
    const double k_sum = kKick ? block_sum(kinetic, shared) : 0.0;
    if (threadIdx.x == 0)
    {
        out[0] = e;
        for (int k = 0; k < 6; ++k) out[1 + k] = v[k];
        out[7] = k_sum;
    }
    if (threadIdx.x < 8)
        for (index_t slot = blockIdx.x + gridDim.x; slot < slots; slot += gridDim.x)
            partials[8LL * slot + threadIdx.x] = 0.0;
}

// One block of 1024 threads reduces the force kernel's partials.
constexpr int kFinishThreads = 1024;

// E, W and K from the partials (pairs counted twice, so E and W halve); then the chain's
// closing half step, whose velocity factor waits in the thermostat for the next step.
template <typename T>
__global__ void __launch_bounds__(kFinishThreads) finish_step_kernel(StepArgs<T> a,
                                                                     const double* __restrict__ partials,
                                                                     int blocks,
                                                                     int kinetic_from_partials)
{
    __shared__ double shared[16];
    double sums[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // a block's 8 partials are 64 contiguous bytes: two 32-byte loads per block, all of a
    // thread's loads in flight together (the grid is at most a few thousand blocks)
    for (int b = threadIdx.x; b < blocks; b += kFinishThreads)
    {
        const double4* p = reinterpret_cast<const double4*>(partials + 8LL * b);
        const double4 lo = p[0], hi = p[1];
        sums[0] += lo.x;
        sums[1] += lo.y;
        sums[2] += lo.z;
        sums[3] += lo.w;
        sums[4] += hi.x;
        sums[5] += hi.y;
        sums[6] += hi.z;
        sums[7] += hi.w;
    }
    for (int k = 0; k < 8; ++k) sums[k] = block_sum(sums[k], shared);
    if (threadIdx.x != 0) return;
    if (a.energy) a.energy[0] = static_cast<T>(0.5 * sums[0]);
    if (a.virial)
    {
        const double xx = 0.5 * sums[1], yy = 0.5 * sums[2], zz = 0.5 * sums[3];
        const double xy = 0.5 * sums[4], xz = 0.5 * sums[5], yz = 0.5 * sums[6];
        const double w[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
        for (int k = 0; k < 9; ++k) a.virial[k] = static_cast<T>(w[k]);
    }
    if (kinetic_from_partials && a.kinetic) a.kinetic[0] = sums[7];
    if (a.nvt && kinetic_from_partials)
    {
        double* t = a.thermostat;
        t[kPending] = chain_half_step(a.kinetic, t + kChainQ, t + kChainP, a.chain_masses,
                                      a.degrees, a.kT, a.timestep);
    }
}

// Applies the pending closing factor to the velocities (before they are read out).
template <typename T>
__global__ void __launch_bounds__(kBlock) settle_kernel(T* velocities, std::int64_t count,
                                                        const double* thermostat)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) velocities[i] = velocities[i] * static_cast<T>(thermostat[kPending]);
}

__global__ void settle_reset_kernel(double* thermostat)
{
    thermostat[kPending] = 1.0;
}

// Slot order -> atom order for readout: out[ids[slot]] = in[slot], `width` values per atom.
template <typename E>
__global__ void __launch_bounds__(kBlock) to_atom_order_kernel(const E* __restrict__ in,
                                                               const index_t* __restrict__ ids,
                                                               std::int64_t num_atoms, int width,
                                                               E* __restrict__ out)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= num_atoms * width) return;
    const std::int64_t slot = i / width;
    out[static_cast<std::int64_t>(ids[slot]) * width + i % width] = in[i];
}

// ---- host -----------------------------------------------------------------------------------

template <typename T>
StepArgs<T> step_args(const ndn_md_system& s, const ndn_md_settings* settings,
                      const ndn_list_state* list, const ndn_md_list& m)
{
    StepArgs<T> a{};
    a.positions = static_cast<T*>(s.positions);
    a.velocities = static_cast<T*>(s.velocities);
    a.forces = static_cast<T*>(s.forces);
    a.images = s.images;
    a.masses = static_cast<T*>(s.masses);
    a.inverse_masses = static_cast<T*>(s.inverse_masses);
    a.cell = static_cast<const T*>(s.cell);
    a.energy = static_cast<T*>(s.energy);
    a.virial = static_cast<T*>(s.virial);
    a.num_atoms = static_cast<index_t>(s.num_atoms);
    a.thermostat = s.thermostat;
    a.chain_masses = s.chain_masses;
    a.kinetic = s.kinetic;
    if (settings)
    {
        a.kT = settings->kT;
        a.degrees = static_cast<double>(settings->degrees_of_freedom);
        a.timestep = settings->timestep;
        a.nvt = settings->thermostat;
    }
    if (list)
    {
        a.cutoff = static_cast<T>(list->cutoff);
        a.radius = static_cast<T>(list->radius);
        a.reference_positions = static_cast<T*>(list->reference_positions);
        a.reference_cell = static_cast<T*>(list->reference_cell);
        a.limits = list->limits;
        a.reuse = static_cast<T*>(list->reuse);
        a.rebuild = list->rebuild;
        a.rebuild_count = list->rebuild_count;
    }
    a.status = m.status;
    a.dims_scratch = m.prepare_dims;
    a.widths_scratch = static_cast<T*>(m.prepare_widths);
    a.total_scratch = m.prepare_total;
    a.header = reinterpret_cast<ListHeader*>(m.scan_temp);
    a.nx = m.dims[0];
    a.ny = m.dims[1];
    a.nz = m.dims[2];
    a.wx = static_cast<T>(m.widths[0]);
    a.wy = static_cast<T>(m.widths[1]);
    a.wz = static_cast<T>(m.widths[2]);
    a.bins = static_cast<std::int64_t>(m.dims[0]) * m.dims[1] * m.dims[2];
    a.stride = m.stride;
    a.bin_counts = m.bin_counts;
    a.bin_offsets = m.bin_offsets;
    a.bin_of = m.bin_of;
    a.unordered = m.unordered;
    a.order = m.order;
    a.counts = m.counts;
    a.index = m.index;
    a.ids = m.ids;
    a.types = m.types;
    a.positions4 = static_cast<T*>(m.positions4);
    a.scratch = static_cast<std::uint8_t*>(m.scratch);
    a.scan_temp = m.scan_temp;
    if (m.inner_index != nullptr)
    {
        a.inner_radius = static_cast<T>(m.inner_radius);
        a.inner_counts = m.inner_counts;
        a.inner_index = m.inner_index;
        a.prune_positions = static_cast<T*>(m.prune_positions);
    }
    return a;
}

// Grid of a rebuild kernel gated on the device flag. On the many steps without a rebuild every
// block exits at once, so the grid is capped (its blocks stride over the items) to keep those
// idle launches short; 4096 blocks of 256 threads still fill the chip when it does rebuild.
constexpr std::int64_t kGatedBlocks = 4096;

inline unsigned gated_grid(std::int64_t items, int per_block = kBlock)
{
    return static_cast<unsigned>(
        std::max<std::int64_t>(1, std::min<std::int64_t>(blocks_for(items, per_block), kGatedBlocks)));
}

template <typename T>
int launch_md_rebuild(const StepArgs<T>& a, double radius, int forced, hipStream_t stream,
                      bool prune = true)
{
    const std::int64_t n = a.num_atoms;
    md_assign_kernel<T><<<gated_grid(n), kBlock, 0, stream>>>(a, forced);
    NDN_LAUNCHED("md_assign");
    const index_t* gate = forced ? nullptr : a.rebuild;
    const int scanned = exclusive_scan(a.bin_counts, a.bin_offsets, a.bins + 1, a.scan_temp, gate,
                                       a.status, stream);
    if (scanned != NDN_OK) return scanned;
    md_scatter_kernel<T><<<gated_grid(n), kBlock, 0, stream>>>(a, forced);
    NDN_LAUNCHED("md_scatter");
    md_order_kernel<T><<<gated_grid(n), kBlock, 0, stream>>>(a, forced);
    NDN_LAUNCHED("md_order");
    md_unpack_kernel<T><<<gated_grid(n), kBlock, 0, stream>>>(a, forced);
    NDN_LAUNCHED("md_unpack");
    const T r = static_cast<T>(radius);
    const T r2 = r * r;
    // culling keeps every sub-bin a pair within R could sit in, with room for rounding
    const T cull = static_cast<T>(radius * (1.0 + 1e-4) + 1e-6);
    md_fill_kernel<T><<<gated_grid(n, kBlock / kWave), kBlock, 0, stream>>>(a, r2, cull * cull,
                                                                            forced);
    NDN_LAUNCHED("md_fill");
    if (prune && a.inner_index != nullptr && a.stride > 0)
    {
        const T inner = a.inner_radius;
        md_prune_kernel<T><<<gated_grid(n, kBlock / 4), kBlock, 0, stream>>>(a, inner,
                                                                             inner * inner, forced);
        NDN_LAUNCHED("md_prune");
    }
    return NDN_OK;
}

// Tiles of kBlock / G rows.
std::int64_t force_tiles(std::int64_t num_atoms, int group)
{
    const std::int64_t rows_per_tile = kBlock / group;
    return (num_atoms + rows_per_tile - 1) / rows_per_tile;
}

// The persistent grid: as many blocks as are resident on the device at once, a multiple of 8
// (one share per XCD), and no more than the tiles need (so never more than xcd_grid(tiles)).
template <typename Kernel>
unsigned persistent_grid(Kernel kernel, std::int64_t tiles)
{
    int device = 0, units = 0, per_unit = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_unit, kernel, kBlock, 0) != hipSuccess)
        return xcd_grid(tiles);
    const std::int64_t resident =
        static_cast<std::int64_t>(std::max(units, 1)) * std::max(per_unit, 1);
    return xcd_grid(std::min(resident, tiles));
}

template <typename T, class Pot, int G, bool kVirial, bool kKick>
unsigned launch_force_kernel(const StepArgs<T>& a, const Pot& pot, T cutoff, T cutoff_squared,
                             T half_timestep, double* partials, bool refresh_pair,
                             hipStream_t stream)
{
    const std::int64_t tiles = force_tiles(a.num_atoms, G);
    if constexpr (kKick && G <= 8)
    {
        if (refresh_pair)
        {
            // a step with a pruned list: the two gated launches share one grid (and so one set
            // of partial slots for the finish)
            auto plain = force_kernel<T, Pot, G, kVirial, kKick, 1>;
            auto refresh = force_kernel<T, Pot, G, kVirial, kKick, 2>;
            const unsigned plain_grid = persistent_grid(plain, tiles);
            const unsigned refresh_grid = persistent_grid(refresh, tiles);
            const index_t slots = static_cast<index_t>(std::max(plain_grid, refresh_grid));
            plain<<<plain_grid, kBlock, 0, stream>>>(a, pot, cutoff, cutoff_squared,
                                                     half_timestep, partials,
                                                     static_cast<index_t>(tiles), slots);
            refresh<<<refresh_grid, kBlock, 0, stream>>>(a, pot, cutoff, cutoff_squared,
                                                         half_timestep, partials,
                                                         static_cast<index_t>(tiles), slots);
            return static_cast<unsigned>(slots);
        }
    }
    auto kernel = force_kernel<T, Pot, G, kVirial, kKick, 0>;
    const unsigned grid = persistent_grid(kernel, tiles);
    kernel<<<grid, kBlock, 0, stream>>>(a, pot, cutoff, cutoff_squared, half_timestep, partials,
                                        static_cast<index_t>(tiles), static_cast<index_t>(grid));
    return grid;
}

template <typename T, class Pot, int G>
unsigned launch_force_g(const StepArgs<T>& a, const Pot& pot, T cutoff, T cutoff_squared,
                        T half_timestep, bool virial, bool kick, double* partials,
                        bool refresh_pair, hipStream_t stream)
{
    if (virial && kick)
        return launch_force_kernel<T, Pot, G, true, true>(a, pot, cutoff, cutoff_squared,
                                                          half_timestep, partials, refresh_pair,
                                                          stream);
    if (kick)
        return launch_force_kernel<T, Pot, G, false, true>(a, pot, cutoff, cutoff_squared,
                                                           half_timestep, partials, refresh_pair,
                                                           stream);
    // evaluations without the kick never refresh
    if (virial)
        return launch_force_kernel<T, Pot, G, true, false>(a, pot, cutoff, cutoff_squared,
                                                           half_timestep, partials, false, stream);
    return launch_force_kernel<T, Pot, G, false, false>(a, pot, cutoff, cutoff_squared,
                                                        half_timestep, partials, false, stream);
}

template <typename T, class Pot>
unsigned launch_force(const StepArgs<T>& a, const Pot& pot, int group, T cutoff,
                      T cutoff_squared, T half_timestep, bool virial, bool kick,
                      double* partials, bool refresh_pair, hipStream_t stream)
{
    switch (group)
    {
        case 2:
            return launch_force_g<T, Pot, 2>(a, pot, cutoff, cutoff_squared, half_timestep,
                                             virial, kick, partials, refresh_pair, stream);
        case 4:
            return launch_force_g<T, Pot, 4>(a, pot, cutoff, cutoff_squared, half_timestep,
                                             virial, kick, partials, refresh_pair, stream);
        case 8:
            return launch_force_g<T, Pot, 8>(a, pot, cutoff, cutoff_squared, half_timestep,
                                             virial, kick, partials, refresh_pair, stream);
        default:
            return launch_force_g<T, Pot, 16>(a, pot, cutoff, cutoff_squared, half_timestep,
                                              virial, kick, partials, refresh_pair, stream);
    }
}

// The widest lane group. A tile holds kBlock / G rows, so the widest group has the most tiles;
// the persistent grid never exceeds xcd_grid(tiles), so xcd_grid(force_tiles(N, kMaxGroup))
// partial slots suffice for every schedule.
constexpr int kMaxGroup = 16;

int valid_group(int group)
{
    return group == 2 || group == 4 || group == 8 || group == 16;
}

template <typename T>
int launch_forces(const StepArgs<T>& a, const ndn_potential* pot, void* coefficients, int group,
                  bool virial, bool kick, double* partials, std::int64_t partial_slots,
                  hipStream_t stream, int* grid_out, bool refresh_pair = false)
{
    NDN_REQUIRE(valid_group(group), "group must be 2, 4, 8 or 16");
    // pruned rows are padded for waves of at most 32 rows, i.e. at least 2 lanes per row and
    // tiles of at least 32 rows: at most 8 lanes per row
    NDN_REQUIRE(a.inner_index == nullptr || group <= 8,
                "with a pruned list the force schedule takes at most 8 lanes per row");
    NDN_REQUIRE(static_cast<std::int64_t>(xcd_grid(force_tiles(a.num_atoms, group))) <=
                    partial_slots,
                "the force launch may have more blocks than partial slots");
    NDN_REQUIRE(a.stride >= 8 && a.stride % 4 == 0,
                "the list stride must be a multiple of 4 and at least 8");
    // gathers take 32-bit byte offsets into the float4 copy of N + 1 slots
    NDN_REQUIRE((static_cast<std::int64_t>(a.num_atoms) + 1) * 4 * sizeof(T) < (1LL << 32),
                "too many atoms for the force kernel's 32-bit gather offsets");
    const T cutoff = static_cast<T>(pot->cutoff);
    const T cutoff_squared = static_cast<T>(pot->cutoff * pot->cutoff);
    const T half_timestep = static_cast<T>(0.5 * a.timestep);
    const int rows = pot->num_types * (pot->num_types + 1) / 2;
    unsigned grid;
    if (pot->kind == NDN_LJ && pot->num_types == 1)
    {
        const OneTypeLJ<T> lj{static_cast<const LJRow<T>*>(coefficients)};
        grid = launch_force<T>(a, lj, group, cutoff, cutoff_squared, half_timestep, virial, kick,
                               partials, refresh_pair, stream);
    }
    else if (pot->kind == NDN_LJ)
    {
        const TypedLJ<T> lj{static_cast<const LJRow<T>*>(coefficients), pot->num_types};
        grid = launch_force<T>(a, lj, group, cutoff, cutoff_squared, half_timestep, virial, kick,
                               partials, refresh_pair, stream);
    }
    else
    {
        T* padded = static_cast<T*>(coefficients);
        const TypedTable<T> table{Table<T>{
            padded, static_cast<const T*>(pot->start),
            padded + static_cast<std::int64_t>(rows) * (pot->knots + 2), nullptr,
            pot->num_types, static_cast<int>(pot->knots)}};
        grid = launch_force<T>(a, table, group, cutoff, cutoff_squared, half_timestep, virial,
                               kick, partials, refresh_pair, stream);
    }
    NDN_LAUNCHED("md_force");
    *grid_out = static_cast<int>(grid);
    return NDN_OK;
}

template <typename T>
int md_step_impl(const ndn_md_system& s, const ndn_md_settings* settings, const ndn_potential* pot,
                 const ndn_list_state* list, const ndn_md_list& m, hipStream_t stream)
{
    StepArgs<T> a = step_args<T>(s, settings, list, m);
    const T timestep = static_cast<T>(settings->timestep);
    const T half_timestep = static_cast<T>(0.5 * settings->timestep);
    prologue_kernel<T><<<1, kBlock, 0, stream>>>(a, *pot, s.coefficients);
    NDN_LAUNCHED("md_prologue");
    first_half_kernel<T><<<blocks_for(s.num_atoms), kBlock, 0, stream>>>(a, timestep, half_timestep);
    NDN_LAUNCHED("md_first_half");
    // the step's force launches refresh the pruned list themselves
    const int rebuilt = launch_md_rebuild<T>(a, list->radius, 0, stream, false);
    if (rebuilt != NDN_OK) return rebuilt;
    int blocks = 0;
    const int forces = launch_forces<T>(a, pot, s.coefficients, m.group, settings->virial != 0,
                                        true, s.partials, s.partial_slots, stream, &blocks,
                                        a.inner_index != nullptr);
    if (forces != NDN_OK) return forces;
    finish_step_kernel<T><<<1, kFinishThreads, 0, stream>>>(a, s.partials, blocks, 1);
    NDN_LAUNCHED("md_finish");
    return NDN_OK;
}

template <typename T>
int md_evaluate_impl(const ndn_md_system& s, const ndn_potential* pot, const ndn_md_list& m,
                     int virial, int kick, double timestep, hipStream_t stream)
{
    StepArgs<T> a = step_args<T>(s, nullptr, nullptr, m);
    a.timestep = timestep;
    PotentialView<T> view;
    const int prepared = prepare_potential<T>(pot, s.coefficients, stream, &view);
    if (prepared != NDN_OK) return prepared;
    int blocks = 0;
    const int forces = launch_forces<T>(a, pot, s.coefficients, m.group, virial != 0, kick != 0,
                                        s.partials, s.partial_slots, stream, &blocks);
    if (forces != NDN_OK) return forces;
    finish_step_kernel<T><<<1, kFinishThreads, 0, stream>>>(a, s.partials, blocks, 0);
    NDN_LAUNCHED("md_evaluate_finish");
    return NDN_OK;
}

int check_system(const ndn_md_system* s, const ndn_md_list* m)
{
    NDN_REQUIRE(s != nullptr && m != nullptr, "system and list must not be null");
    NDN_REQUIRE(s->num_atoms >= 1 && s->num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(s->positions && s->velocities && s->forces && s->images && s->masses &&
                    s->inverse_masses && s->cell && s->energy && s->thermostat && s->kinetic &&
                    s->partials && s->coefficients,
                "a system buffer is null");
    NDN_REQUIRE(m->bin_counts && m->bin_offsets && m->bin_of && m->unordered && m->order &&
                    m->counts && m->ids && m->positions4 && m->scratch && m->scan_temp &&
                    m->status && m->prepare_dims && m->prepare_widths && m->prepare_total,
                "a list buffer is null");
    NDN_REQUIRE(m->dims[0] >= 1 && m->dims[1] >= 1 && m->dims[2] >= 1, "bad sub-bin dims");
    const std::int64_t bins = static_cast<std::int64_t>(m->dims[0]) * m->dims[1] * m->dims[2];
    NDN_REQUIRE(bins < INT32_MAX, "too many sub-bins");
    NDN_REQUIRE(m->scan_temp_bytes >= scan_bytes(bins + 1) && m->scan_temp_bytes >= 1024,
                "scan_temp is too small");
    NDN_REQUIRE(m->stride == 0 || m->index != nullptr, "a list with rows needs its index");
    NDN_REQUIRE(m->stride >= 0 && m->stride * s->num_atoms <= (1LL << 40), "bad stride");
    NDN_REQUIRE(valid_group(m->group), "group must be 2, 4, 8 or 16");
    NDN_REQUIRE(s->partial_slots >=
                    static_cast<std::int64_t>(xcd_grid(force_tiles(s->num_atoms, m->group))),
                "too few partial slots for this force schedule");
    return NDN_OK;
}

}  // namespace ndn

using namespace ndn;

extern "C" {

int ndn_md_workspace(int dtype, const ndn_potential* pot, int64_t num_atoms,
                     int64_t* partial_slots, int64_t* coefficient_bytes_out)
{
    const int checked = check_potential(pot, num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    NDN_REQUIRE(num_atoms >= 1 && partial_slots && coefficient_bytes_out, "bad arguments");
    // the most blocks any schedule may launch: the widest lane group's tile count
    *partial_slots = std::max<std::int64_t>(xcd_grid(force_tiles(num_atoms, kMaxGroup)), 8);
    *coefficient_bytes_out = coefficient_bytes(pot, dtype == NDN_F32 ? 4 : 8);
    return NDN_OK;
}

int64_t ndn_md_scratch_bytes(int dtype, int64_t num_atoms)
{
    return num_atoms * (dtype == NDN_F32 ? static_cast<int64_t>(sizeof(Packed<float>))
                                         : static_cast<int64_t>(sizeof(Packed<double>)));
}

int ndn_md_rebuild(int dtype, const ndn_md_system* system, const ndn_list_state* list,
                   const ndn_md_list* md_list, int forced, ndn_stream_t stream)
{
    const int checked = check_system(system, md_list);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(list != nullptr, "the list state must not be null");
    NDN_REQUIRE(forced || list->rebuild != nullptr, "an unforced rebuild needs the flag");
    if (dtype == NDN_F32)
    {
        const StepArgs<float> a = step_args<float>(*system, nullptr, list, *md_list);
        return launch_md_rebuild<float>(a, list->radius, forced, stream);
    }
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    const StepArgs<double> a = step_args<double>(*system, nullptr, list, *md_list);
    return launch_md_rebuild<double>(a, list->radius, forced, stream);
}

int ndn_md_step(int dtype, const ndn_md_system* system, const ndn_md_settings* settings,
                const ndn_potential* pot, const ndn_list_state* list, const ndn_md_list* md_list,
                ndn_stream_t stream)
{
    int checked = check_system(system, md_list);
    if (checked != NDN_OK) return checked;
    checked = check_potential(pot, system->num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(settings && list, "settings and list must not be null");
    NDN_REQUIRE(settings->timestep > 0, "the timestep must be positive");
    NDN_REQUIRE(!settings->thermostat || system->chain_masses, "NVT needs chain masses");
    NDN_REQUIRE(list->cutoff == pot->cutoff, "the list's cutoff must be the potential's");
    NDN_REQUIRE(md_list->stride > 0, "the list has no rows yet");
    NDN_REQUIRE(pot->num_types == 1 || md_list->types != nullptr,
                "several atom types need the slot-ordered types");
    if (dtype == NDN_F32)
        return md_step_impl<float>(*system, settings, pot, list, *md_list, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return md_step_impl<double>(*system, settings, pot, list, *md_list, stream);
}

int ndn_md_evaluate(int dtype, const ndn_md_system* system, const ndn_potential* pot,
                    const ndn_md_list* md_list, int virial, int kick, double timestep,
                    ndn_stream_t stream)
{
    int checked = check_system(system, md_list);
    if (checked != NDN_OK) return checked;
    checked = check_potential(pot, system->num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(md_list->stride > 0, "the list has no rows yet");
    NDN_REQUIRE(pot->num_types == 1 || md_list->types != nullptr,
                "several atom types need the slot-ordered types");
    if (dtype == NDN_F32)
        return md_evaluate_impl<float>(*system, pot, *md_list, virial, kick, timestep, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return md_evaluate_impl<double>(*system, pot, *md_list, virial, kick, timestep, stream);
}

int ndn_md_force(int dtype, const ndn_md_system* system, const ndn_potential* pot,
                 const ndn_md_list* md_list, int virial, int kick, double timestep,
                 ndn_stream_t stream)
{
    int checked = check_system(system, md_list);
    if (checked != NDN_OK) return checked;
    checked = check_potential(pot, system->num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(md_list->stride > 0, "the list has no rows yet");
    if (dtype == NDN_F32)
    {
        StepArgs<float> a = step_args<float>(*system, nullptr, nullptr, *md_list);
        a.timestep = timestep;
        int grid = 0;
        return launch_forces<float>(a, pot, system->coefficients, md_list->group, virial != 0,
                                    kick != 0, system->partials, system->partial_slots, stream,
                                    &grid);
    }
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    StepArgs<double> a = step_args<double>(*system, nullptr, nullptr, *md_list);
    a.timestep = timestep;
    int grid = 0;
    return launch_forces<double>(a, pot, system->coefficients, md_list->group, virial != 0,
                                 kick != 0, system->partials, system->partial_slots, stream,
                                 &grid);
}

int ndn_md_settle(int dtype, const ndn_md_system* system, ndn_stream_t stream)
{
    NDN_REQUIRE(system && system->velocities && system->thermostat, "bad arguments");
    const std::int64_t count = 3 * system->num_atoms;
    if (dtype == NDN_F32)
        settle_kernel<float><<<blocks_for(count), kBlock, 0, stream>>>(
            static_cast<float*>(system->velocities), count, system->thermostat);
    else
        settle_kernel<double><<<blocks_for(count), kBlock, 0, stream>>>(
            static_cast<double*>(system->velocities), count, system->thermostat);
    NDN_LAUNCHED("md_settle");
    settle_reset_kernel<<<1, 1, 0, stream>>>(system->thermostat);
    NDN_LAUNCHED("md_settle_reset");
    return NDN_OK;
}

int ndn_md_to_atom_order(int element_bytes, const void* in, const int32_t* ids, int64_t num_atoms,
                         int width, void* out, ndn_stream_t stream)
{
    NDN_REQUIRE(in && ids && out && num_atoms >= 1 && width >= 1, "bad arguments");
    const std::int64_t count = num_atoms * width;
    if (element_bytes == 4)
        to_atom_order_kernel<std::int32_t><<<blocks_for(count), kBlock, 0, stream>>>(
            static_cast<const std::int32_t*>(in), ids, num_atoms, width,
            static_cast<std::int32_t*>(out));
    else
    {
        NDN_REQUIRE(element_bytes == 8, "element_bytes must be 4 or 8");
        to_atom_order_kernel<std::int64_t><<<blocks_for(count), kBlock, 0, stream>>>(
            static_cast<const std::int64_t*>(in), ids, num_atoms, width,
            static_cast<std::int64_t*>(out));
    }
    NDN_LAUNCHED("md_to_atom_order");
    return NDN_OK;
}

}  // extern "C"
