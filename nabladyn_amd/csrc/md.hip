// The molecular-dynamics step on gfx950: velocity Verlet with an optional Nose-Hoover chain of
// three, the neighbor update and the pair forces, replayed as one HIP graph per step.
//
// The resident frame is kept in cell order ("slots"): at every rebuild the atoms are counting-
// sorted into sub-bins at least half the list radius wide, ordered by atom id inside a sub-bin
// (so every run is deterministic), and the whole frame is permuted with them. Every per-atom
// access of a step is then coalesced. The frame is read out in atom order only on demand.
//
// The list works in bricks (boxes of sub-bins): a brick's halo (every sub-bin within 2 of it) is
// listed once per build, and a row names its neighbors by 16-bit indices into its brick's halo.
//
//   prologue     one block: the chain's opening half step (fp64), the reuse rule's preparation
//                (update/update.cuh:27-75), the potential's coefficient rows
//   first_half   per slot: chain scale, half kick, drift, wrap, image count, the float4 copy of
//                the positions, and the squared displacements since the last build and refresh
//                reduced to the rebuild and refresh flags (update/update.cuh:79-116)
//   rebuild      gated on the rebuild flag: sub-bin counting sort, permutation of the frame, the
//                bricks' halos, then one wave per sub-bin tests the candidates of its culled
//                5 x 5 x 5 stencil 64 at a time and ballots the hits into its row of a
//                fixed-stride full list
//   prune        gated on a build, or on an atom having moved half the pruned list's skin since
//                its last refresh: keeps the full rows' pairs within cutoff + that skin, so the
//                force kernel reads about two thirds of the entries
//   force        one block per brick stages the brick's halo in LDS; a lane group per row reads
//                its pruned row and the neighbors' positions from LDS, sums the force in
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
#include <type_traits>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "neighbors.hpp"
#include "pair_forces.hpp"
#include "potentials.hpp"

namespace ndn {

// Thermostat slots (fp64): chain positions [0..2], momenta [3..5], the pending closing factor
// [6], the opening factor of the next (or current) step [7], and the chain after that step's
// opening half step, positions [8..10] and momenta [11..13], which the step's finish commits
// before its closing half step (without a barostat).
constexpr int kChainQ = 0;
constexpr int kChainP = 3;
constexpr int kPending = 6;
constexpr int kOpening = 7;
constexpr int kOpenQ = 8;
constexpr int kOpenP = 11;
constexpr int kStencil = 5;          // sub-bins per axis a stencil spans

// Barostat slots (fp64, NPT only): the box momentum p_eps [0], its chain's positions [1..3] and
// momenta [4..6], that chain's masses [7..9], the master cell [10..12], and this step's drag
// exp(-alpha v_eps dt/2) [13] and growth exp(v_eps dt/2) [14].
constexpr int kBoxP = 0;
constexpr int kBoxChainQ = 1;
constexpr int kBoxChainP = 4;
constexpr int kBoxChainMass = 7;
constexpr int kMasterCell = 10;
constexpr int kDrag = 13;
constexpr int kGrowth = 14;

// The chain's half step against the kinetic energy K it thermostats (_chain.py:29-76), in the
// reference's operation order. Updates q, p and K; returns the velocity factor.
__device__ __forceinline__ double chain_half_step(double* kinetic, double* q, double* p,
                                                  const double* mass, double g, double kT,
                                                  double dt)
{
#pragma clang fp contract(off)
    const double quarter = 0.25 * dt;
    const double eighth = 0.125 * dt;
    // the chain in registers: read once, written once (through the pointers, which the compiler
    // must assume overlap, every read after a write would go back to memory)
    double pc[3], qc[3], mc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        pc[k] = p[k];
        qc[k] = q[k];
        mc[k] = mass[k];
    }
    double damping[2];
    double K = *kinetic;
    auto driving = [&](int index) {
        return index == 0 ? 2.0 * K - g * kT : pc[index - 1] * pc[index - 1] / mc[index - 1] - kT;
    };
#pragma unroll
    for (int index = 2; index >= 0; --index)
    {
        if (index == 2)
            pc[2] = pc[2] + quarter * driving(2);
        else
        {
            const double kick = quarter * driving(index);
            pc[index] = damping[index] * (damping[index] * pc[index] + kick);
        }
        if (index > 0) damping[index - 1] = exp(-eighth * pc[index] / mc[index]);
    }
    const double factor = exp(-0.5 * dt * pc[0] / mc[0]);
    K = K * (factor * factor);
#pragma unroll
    for (int k = 0; k < 3; ++k) qc[k] = qc[k] + 0.5 * dt * pc[k] / mc[k];
#pragma unroll
    for (int index = 0; index <= 2; ++index)
    {
        if (index == 2)
            pc[2] = pc[2] + quarter * driving(2);
        else
        {
            const double kick = quarter * driving(index);
            pc[index] = damping[index] * (damping[index] * pc[index] + kick);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        p[k] = pc[k];
        q[k] = qc[k];
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

// The sentinel that pads the rows of the MD list: entry 0 of every brick's halo holds
// (kFarAway, kFarAway, kFarAway), far beyond any cutoff from every atom, yet its squared
// distances stay finite in fp32.
template <typename T>
constexpr T kFarAway = T(1e18);

// The force kernel's potentials. A pair's two types arrive as the w of the two gathered
// float4s (small integers, exact); bind() runs once per kernel and returns what the pair loop
// evaluates: eval(type_i, type_j, r2, u, fs) with fs = -(du/dr)/r.
//
// LJ with one type: the row's coefficients held in registers for the whole kernel; the hardware
// reciprocal (1 ulp in fp32).
template <typename T>
struct OneTypeLJ
{
    // whether eval reads the pair's types (the w of the staged positions)
    static constexpr bool kTypes = false;

    template <typename E>
    static constexpr int waves = sizeof(E) == 4 ? 8 : 1;

    const LJRow<T>* __restrict__ rows;

    // With p = 1 / r^2: u = p^3 (4 eps sigma^12 p^3 - 4 eps sigma^6) - shift and
    // fs = p^4 (48 eps sigma^12 p^3 - 24 eps sigma^6), the sigma powers folded into the
    // coefficients once per kernel.
    struct Bound
    {
        T e12, e6, f12, f6, shift;

        __device__ __forceinline__ void eval(T, T, T r2, T& u, T& fs) const
        {
            const T p = reciprocal(r2);
            const T p3 = p * p * p;
            u = p3 * (p3 * e12 - e6) - shift;
            fs = p3 * (p3 * f12 - f6) * p;
        }
    };

    __device__ __forceinline__ Bound bind() const
    {
        const LJRow<T> c = rows[0];
        const T s6 = c.sigma_squared * c.sigma_squared * c.sigma_squared;
        const T s12 = s6 * s6;
        return Bound{c.four_eps * s12, c.four_eps * s6, (c.twentyfour_eps + c.twentyfour_eps) * s12,
                     c.twentyfour_eps * s6, c.shift};
    }
};

// LJ with several types: the pair's row read from the small, cache-resident coefficient table.
template <typename T>
struct TypedLJ
{
    // whether eval reads the pair's types (the w of the staged positions)
    static constexpr bool kTypes = true;

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
    // whether eval reads the pair's types (the w of the staged positions)
    static constexpr bool kTypes = true;

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

// A table with one type: each block first writes every segment's cubic into its LDS (after the
// halo and the brick's bookkeeping), as two entries of 4. With t the coordinate within the
// segment, s = max(t, 0) and a, b, c the Catmull-Rom terms of Table::eval:
//   force  (A, B, C, u1): -du/dr = -(0.5 a + b s + 1.5 c s^2) / h, so fs = (A + s (B + s C)) / r
//          with A = -0.5 a / h, B = -b / h, C = -1.5 c / h (below the first knot s = 0 leaves
//          the tangent's slope, as the reference's);
//   energy (0.5 a, 0.5 b, 0.5 c, 0): u = u1 + s (0.5 a + s (0.5 b + s 0.5 c)) + 0.5 a min(t, 0).
// A pair of a kick step (no energy) reads one aligned 16-byte (fp64: 32-byte) entry. r and the
// knot coordinate x = (r - start) / h come from the reference's own operations, each correctly
// rounded (square root, subtraction, division), so every pair's segment and t carry the
// reference's bits; 1/r, which only scales the force, is the hardware reciprocal.
template <typename T>
struct OneTypeTable
{
    // whether eval reads the pair's types (the w of the staged positions)
    static constexpr bool kTypes = false;

    template <typename E>
    static constexpr int waves = sizeof(E) == 4 ? 8 : 1;

    using V4 = typename Vec4<T>::type;

    const T* __restrict__ padded;  // [K + 2] (Table::padded of row 0)
    const T* __restrict__ start;
    const T* __restrict__ spacing;
    int knots;

    // Segments of the table: K - 1, the last starting at knot K - 2; two entries each.
    __host__ __device__ int segments() const { return knots - 1; }

    __host__ __device__ std::size_t lds_bytes() const
    {
        return static_cast<std::size_t>(segments()) * 2 * sizeof(V4);
    }

    __device__ __forceinline__ void stage(unsigned char* lds) const
    {
        V4* out = reinterpret_cast<V4*>(lds);
        const T h = spacing[0];
        for (int s = threadIdx.x; s < segments(); s += blockDim.x)
        {
            const T u0 = padded[s], u1 = padded[s + 1], u2 = padded[s + 2], u3 = padded[s + 3];
            const T a = u2 - u0;
            const T b = T(2) * u0 - T(5) * u1 + T(4) * u2 - u3;
            const T c = T(3) * (u1 - u2) + u3 - u0;
            V4 force, energy;
            force.x = -(T(0.5) * a) / h;
            force.y = -b / h;
            force.z = -(T(1.5) * c) / h;
            force.w = u1;
            energy.x = T(0.5) * a;
            energy.y = T(0.5) * b;
            energy.z = T(0.5) * c;
            energy.w = T(0);
            out[2 * s] = force;
            out[2 * s + 1] = energy;
        }
    }

    struct Bound
    {
        const V4* segments;
        T start, h;
        T last;  // the last segment, K - 2

        __device__ __forceinline__ void eval(T, T, T r2, T& u, T& fs) const
        {
            const T r = sqrt(r2);
            const T x = (r - start) / h;
            // the segment, in range whatever r2 was: the clamps send a NaN to 0, so the
            // conversion always sees a value in [0, last]
            T clamped = x > T(0) ? x : T(0);
            clamped = clamped < last ? clamped : last;
            const int segment = static_cast<int>(clamped);
            const T t = x - static_cast<T>(segment);
            const T s = t > T(0) ? t : T(0);
            const V4 k = segments[2 * segment];
            fs = (k.x + s * (k.y + s * k.z)) * reciprocal(r);
            const V4 e = segments[2 * segment + 1];
            u = k.w + s * (e.x + s * (e.y + s * e.z)) + e.x * (t < T(0) ? t : T(0));
        }
    };

    __device__ __forceinline__ Bound bind(const unsigned char* lds) const
    {
        return Bound{reinterpret_cast<const V4*>(lds), start[0], spacing[0],
                     static_cast<T>(knots - 2)};
    }
};

// One model of a sum of two (potential.sum): LJ or a table, typed, with its own cutoff. The
// kind is the same for every pair of a launch, so the branch does not diverge.
template <typename T>
struct SumPart
{
    int kind;
    TypedLJ<T> lj;
    TypedTable<T> table;
    T cutoff_squared;

    __device__ __forceinline__ void eval(T type_i, T type_j, T r2, T& u, T& fs) const
    {
        if (kind == NDN_LJ)
            lj.eval(type_i, type_j, r2, u, fs);
        else
            table.eval(type_i, type_j, r2, u, fs);
        // beyond this model's cutoff the pair counts nothing for it
        // (potential/_sum/_potential.py:86-99 hands each model the list of its own cutoff)
        const bool inside = r2 < cutoff_squared;
        u = inside ? u : T(0);
        fs = inside ? fs : T(0);
    }
};

// A sum of two models in one walk: the pair loop's cutoff is the larger one, each model masks
// its own; u and fs of the two are added per pair.
template <typename T>
struct PairSum
{
    // whether eval reads the pair's types (the w of the staged positions)
    static constexpr bool kTypes = true;

    template <typename E>
    static constexpr int waves = sizeof(E) == 4 ? 4 : 1;

    SumPart<T> first, second;

    __device__ __forceinline__ const PairSum& bind() const { return *this; }

    __device__ __forceinline__ void eval(T type_i, T type_j, T r2, T& u, T& fs) const
    {
        T u1, f1, u2, f2;
        first.eval(type_i, type_j, r2, u1, f1);
        second.eval(type_i, type_j, r2, u2, f2);
        u = u1 + u2;
        fs = f1 + f2;
    }
};

// LDS a potential's evaluator stages per block (after the halo and the brick's bookkeeping), its
// staging, and its bound evaluator: nothing for the potentials read from registers or memory.
template <class Pot>
__host__ __device__ std::size_t potential_lds(const Pot&)
{
    return 0;
}

template <typename T>
__host__ __device__ std::size_t potential_lds(const OneTypeTable<T>& pot)
{
    return pot.lds_bytes();
}

template <class Pot>
__device__ __forceinline__ void stage_potential(const Pot&, unsigned char*)
{
}

template <typename T>
__device__ __forceinline__ void stage_potential(const OneTypeTable<T>& pot, unsigned char* lds)
{
    pot.stage(lds);
}

template <class Pot>
__device__ __forceinline__ auto bind_potential(const Pot& pot, const unsigned char*)
{
    return pot.bind();
}

template <typename T>
__device__ __forceinline__ auto bind_potential(const OneTypeTable<T>& pot, const unsigned char* lds)
{
    return pot.bind(lds);
}

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
    // barostat (NPT: barostat non-null; the cell then changes every step)
    double* barostat;
    double pressure;
    double box_mass;
    double alpha;
    T* cell_out;
    // reuse rule
    T cutoff;
    T radius;
    T* reference_positions;
    T* reference_cell;
    const index_t* limits;
    T* reuse;
    index_t* rebuild;
    index_t* rebuild_count;
    index_t* batch_rebuild;  // a batch member: the batch's flag of any member rebuilding
    // list
    index_t* status;
    index_t* dims_scratch;
    T* widths_scratch;
    index_t* total_scratch;
    ListHeader* header;
    index_t nx, ny, nz;
    T* grid_widths;  // [3] sub-bin widths, the cell over nx, ny, nz (NPT: rewritten each step)
    std::int64_t bins;
    std::int64_t stride;
    index_t* bin_counts;
    index_t* bin_offsets;
    index_t* bin_of;
    index_t* unordered;
    index_t* order;
    index_t* counts;
    std::uint16_t* index;
    index_t* ids;
    std::int64_t* types;
    T* positions4;
    std::uint8_t* scratch;
    std::uint8_t* scan_temp;
    // the pruned list (inner_index null: none)
    T inner_radius;
    index_t* inner_counts;
    std::uint16_t* inner_index;
    T* prune_positions;
    // bricks and their halos
    index_t brick[3];
    std::int64_t bricks;
    std::int64_t halo_capacity;
    index_t* halo_slots;
    index_t* halo_runs;
    index_t* halo_counts;
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

// The largest halo of the last build (for sizing halo_capacity), after the pruned list's slots.
__device__ __forceinline__ index_t* halo_needed(ListHeader* header)
{
    return reinterpret_cast<index_t*>(reinterpret_cast<std::uint8_t*>(header) + 656);
}

// Whether the gated refresh of the pruned list runs: after a build (the full rows changed) or
// when the check asked for it, and no status bit is set.
template <typename T>
__device__ __forceinline__ bool refreshing(const StepArgs<T>& a)
{
    return a.inner_index != nullptr && a.status[0] == 0 &&
           (a.rebuild[0] != 0 || prune_state(a.header)->refresh != 0);
}

// The block a kernel body runs as and the number of blocks of its system: the hardware's own in
// a launch over one system, the system's share in a launch over a batch of systems (the bodies
// read no blockIdx or gridDim, so a system computes the same bits alone and in a batch).
struct Launch
{
    index_t block;
    index_t blocks;
};

__device__ __forceinline__ Launch own_launch()
{
    return Launch{static_cast<index_t>(blockIdx.x), static_cast<index_t>(gridDim.x)};
}

// ---- the barostat (NPT) ---------------------------------------------------------------------

// G = alpha 2K + trace W - 3 P V (md/_ensembles/_npt.py:232-238): the carried K, the trace of
// the frame's virial in its dtype, the master cell's volume.
template <typename T>
__device__ __forceinline__ double box_force(const StepArgs<T>& a)
{
#pragma clang fp contract(off)
    const double* b = a.barostat;
    const T trace = (a.virial[0] + a.virial[4]) + a.virial[8];
    const double volume = (b[kMasterCell] * b[kMasterCell + 1]) * b[kMasterCell + 2];
    return ((a.alpha * 2.0) * a.kinetic[0] + static_cast<double>(trace)) -
           (3.0 * a.pressure) * volume;
}

// The barostat chain's half step against the box's kinetic energy p_eps^2 / (2M); scales p_eps
// (_npt.py:257-268).
template <typename T>
__device__ __forceinline__ void box_chain_half_step(const StepArgs<T>& a)
{
#pragma clang fp contract(off)
    double* b = a.barostat;
    double kinetic = b[kBoxP] * b[kBoxP] / (2.0 * a.box_mass);
    const double factor = chain_half_step(&kinetic, b + kBoxChainQ, b + kBoxChainP,
                                          b + kBoxChainMass, 1.0, a.kT, a.timestep);
    b[kBoxP] = b[kBoxP] * factor;
}

// The opening half of an NPT step on the box (_npt.py:154-174, after the particles' chain): the
// box momentum's kick, this step's drag and growth, the cell grown twice by the growth, and the
// sub-bin widths of the new cell. A width below what the fixed stencil covers sets
// NDN_STATUS_CELL_SHRANK.
template <typename T>
__device__ __forceinline__ void box_first_half(const StepArgs<T>& a)
{
#pragma clang fp contract(off)
    double* b = a.barostat;
    const double half = 0.5 * a.timestep;
    b[kBoxP] = b[kBoxP] + half * box_force(a);
    const double rate = b[kBoxP] / a.box_mass;
    b[kDrag] = exp(-a.alpha * rate * half);
    const double growth = exp(rate * half);
    b[kGrowth] = growth;
    const index_t dims[3] = {a.nx, a.ny, a.nz};
    const double epsilon = Epsilon<T>::value;
    for (int axis = 0; axis < 3; ++axis)
    {
        b[kMasterCell + axis] = (b[kMasterCell + axis] * growth) * growth;
        const T length = static_cast<T>(b[kMasterCell + axis]);
        a.cell_out[axis] = length;
        const T width = length / static_cast<T>(dims[axis]);
        a.grid_widths[axis] = width;
        const double needed = 0.5 * static_cast<double>(a.radius) * (1.0 + 1e-6) +
                              4.0 * epsilon * static_cast<double>(length);
        if (static_cast<double>(width) < needed && dims[axis] > 1)
            atomicOr(a.status, NDN_STATUS_CELL_SHRANK);
    }
}

// ---- prologue -------------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ void write_coefficients(const ndn_potential& pot, void* coefficients);

// The prologue's `second` when the model is not a sum.
__host__ __device__ inline ndn_potential no_second_model()
{
    ndn_potential none{};
    none.kind = -1;
    return none;
}

// The list check's preparation for a step, by one thread: the reuse rule's (prepare_update) and
// the pruned list's threshold, the outer rule's allowance at the pruned list's radius (the cell
// is fixed between refreshes; none left means a refresh every step).
template <typename T>
__device__ __forceinline__ void prepare_checks(const StepArgs<T>& a, bool reset_moved)
{
    prepare_update(a.cell, a.reference_cell, a.cutoff, a.radius, a.limits, a.dims_scratch,
                   a.widths_scratch, a.reuse, a.rebuild, a.status, a.total_scratch, a.header,
                   reset_moved);
    if (a.inner_index)
    {
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

// The opening of a step without a barostat, by one thread: the checks' preparation and the
// chain's opening half step on a copy of the chain: its velocity factor, and the opened chain in
// the thermostat's opened slots for the step's finish to commit. Nothing in the step reads the
// chain or K before the finish, so the previous step's finish can open the next step (and a
// step that continues a run needs no prologue launch).
template <typename T>
__device__ __forceinline__ void open_step(const StepArgs<T>& a, bool reset_moved)
{
    if (a.nvt)
    {
        double* t = a.thermostat;
        double q[3], p[3];
        double kinetic = a.kinetic[0];
        for (int k = 0; k < 3; ++k)
        {
            q[k] = t[kChainQ + k];
            p[k] = t[kChainP + k];
        }
        t[kOpening] = chain_half_step(&kinetic, q, p, a.chain_masses, a.degrees, a.kT, a.timestep);
        for (int k = 0; k < 3; ++k)
        {
            t[kOpenQ + k] = q[k];
            t[kOpenP + k] = p[k];
        }
    }
    prepare_checks(a, reset_moved);
}

// `second` is the second model of a sum (kind -1 without one), its table at second_table.
template <typename T>
__device__ __forceinline__ void prologue_body(const StepArgs<T>& a, const ndn_potential& pot,
                                              void* coefficients, const ndn_potential& second,
                                              void* second_table)
{
    if (threadIdx.x == 0)
    {
        if (a.barostat)
        {
            // NPT: the barostat's chain, the particles' chain, then the box (_npt.py:154-174)
            box_chain_half_step(a);
            if (a.nvt)
            {
                double* t = a.thermostat;
                t[kOpening] = chain_half_step(a.kinetic, t + kChainQ, t + kChainP,
                                              a.chain_masses, a.degrees, a.kT, a.timestep);
            }
            box_first_half(a);
            prepare_checks(a, true);
        }
        else
            open_step(a, true);
    }
    write_coefficients<T>(pot, coefficients);
    if (second.kind >= 0) write_coefficients<T>(second, second_table);
}

// A model's coefficient table, by the block: LJ rows (sigma^2, 4 eps, 24 eps, the shift at the
// cutoff), or the table's padded knots [P, K + 2] and spacing [P].
template <typename T>
__device__ __forceinline__ void write_coefficients(const ndn_potential& pot, void* coefficients)
{
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
        T* spacing_out = padded + rows * width;
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
        for (int row = threadIdx.x; row < rows; row += blockDim.x) spacing_out[row] = spacing[row];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) prologue_kernel(StepArgs<T> a, const ndn_potential pot,
                                                          void* coefficients,
                                                          const ndn_potential second,
                                                          void* second_table)
{
    prologue_body(a, pot, coefficients, second, second_table);
}

// ---- first half -----------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ void first_half_body(const StepArgs<T>& a, const Launch launch,
                                                T timestep, T half_timestep)
{
#pragma clang fp contract(off)
    __shared__ T block_max[kBlock / kWave];
    const std::int64_t atom = static_cast<std::int64_t>(launch.block) * blockDim.x + threadIdx.x;
    const bool check = a.status[0] == 0;
    // the box and the reuse rule's scalars, read before any write
    const T cell[3] = {a.cell[0], a.cell[1], a.cell[2]};
    const T reuse[4] = {a.reuse[0], a.reuse[1], a.reuse[2], a.reuse[3]};
    bool finite = true;
    T moved = 0, moved_inner = 0;
    if (atom < a.num_atoms)
    {
        T* v = a.velocities + 3 * atom;
        const T* f = a.forces + 3 * atom;
        T* x = a.positions + 3 * atom;
        // every read of the atom first, then the writes: the compiler cannot tell the arrays
        // apart, and a read after a write would wait for it (one memory round trip per atom
        // instead of a chain of them)
        const T inverse_mass = a.inverse_masses[atom];
        T vx = v[0], vy = v[1], vz = v[2];
        const T f_atom[3] = {f[0], f[1], f[2]};
        const T x_atom[3] = {x[0], x[1], x[2]};
        T reference[3] = {0, 0, 0}, at_refresh[3] = {0, 0, 0};
        if (check)
        {
            for (int axis = 0; axis < 3; ++axis)
                reference[axis] = a.reference_positions[3 * atom + axis];
            if (a.inner_index)
                for (int axis = 0; axis < 3; ++axis)
                    at_refresh[axis] = a.prune_positions[3 * atom + axis];
        }
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
        T scale = T(1);
        if (a.barostat)
        {
            // NPT: the drag before the kick, and the box's growth on both sides of the drift
            const T drag = static_cast<T>(a.barostat[kDrag]);
            vx = vx * drag;
            vy = vy * drag;
            vz = vz * drag;
            scale = static_cast<T>(a.barostat[kGrowth]);
        }
        vx = vx + half_timestep * f_atom[0] * inverse_mass;
        vy = vy + half_timestep * f_atom[1] * inverse_mass;
        vz = vz + half_timestep * f_atom[2] * inverse_mass;
        T out[3];
        const T vel[3] = {vx, vy, vz};
        std::int64_t crossings[3];
        bool crossed = false;
        for (int axis = 0; axis < 3; ++axis)
        {
            const T length = cell[axis];
            const T drifted = a.barostat ? (x_atom[axis] * scale + timestep * vel[axis]) * scale
                                         : x_atom[axis] + timestep * vel[axis];
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
                moved = moved_squared(out[0], out[1], out[2], reference, reuse, cell);
                if (a.inner_index)
                {
                    const T dx = minimum_image(out[0], at_refresh[0], cell[0]);
                    const T dy = minimum_image(out[1], at_refresh[1], cell[1]);
                    const T dz = minimum_image(out[2], at_refresh[2], cell[2]);
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
        if (check && wave_moved >= reuse[0] && a.rebuild[0] == 0)
        {
            a.rebuild[0] = 1;
            if (a.batch_rebuild) a.batch_rebuild[0] = 1;
        }
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
        if (m > T(0)) atomicMax(moved_slot(a.header, launch.block), order_bits(m));
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) first_half_kernel(StepArgs<T> a, T timestep,
                                                            T half_timestep)
{
    first_half_body(a, own_launch(), timestep, half_timestep);
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
__device__ __forceinline__ void md_assign_body(const StepArgs<T>& a, const Launch launch,
                                               int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t slot = static_cast<std::int64_t>(launch.block) * blockDim.x + threadIdx.x;
         slot < a.num_atoms; slot += static_cast<std::int64_t>(launch.blocks) * blockDim.x)
    {
        const T* p = a.positions + 3 * slot;
        const T* w = a.grid_widths;
        const index_t bin = static_cast<index_t>(
            sub_bin(p[0], w[0], a.nx) + row_first_bin(sub_bin(p[1], w[1], a.ny),
                                                      sub_bin(p[2], w[2], a.nz), a.nx, a.ny, a.nz));
        a.bin_of[slot] = bin;
        const LaneRun run = equal_runs(bin);
        if (run.head) atomicAdd(a.bin_counts + bin, run.length);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) md_assign_kernel(StepArgs<T> a, int forced)
{
    md_assign_body(a, own_launch(), forced);
}

template <typename T>
__device__ __forceinline__ void md_scatter_body(const StepArgs<T>& a, const Launch launch,
                                                int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t slot = static_cast<std::int64_t>(launch.block) * blockDim.x + threadIdx.x;
         slot < a.num_atoms; slot += static_cast<std::int64_t>(launch.blocks) * blockDim.x)
    {
        const index_t bin = a.bin_of[slot];
        // one atomic per run of a sub-bin's atoms; the run's lanes take the places below the
        // count it returns, one each (the order kernel ranks a sub-bin's atoms by id after)
        const LaneRun run = equal_runs(bin);
        index_t count = 0;
        if (run.head) count = atomicSub(a.bin_counts + bin, run.length);
        count = __shfl(count, run.first, kWave);
        const int lane = lane_id();
        a.unordered[a.bin_offsets[bin] + count - 1 - (lane - run.first)] =
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
// whatever the atomics did, and gathers its whole frame entry into scratch at its new slot (and
// the gather copy positions4 there); md_unpack_body puts the frame home.
template <typename T>
__device__ __forceinline__ void md_order_body(const StepArgs<T>& a, const Launch launch,
                                              int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t i = static_cast<std::int64_t>(launch.block) * blockDim.x + threadIdx.x;
         i < a.num_atoms; i += static_cast<std::int64_t>(launch.blocks) * blockDim.x)
    {
        const std::int64_t old = a.unordered[i];
        const index_t bin = a.bin_of[old];
        const index_t id = a.ids[old];
        // the atom's frame entry is read before the rank walk and any write (the compiler cannot
        // tell the arrays apart, and a read after a write waits for it)
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
        const index_t begin = a.bin_offsets[bin];
        const index_t end = a.bin_offsets[bin + 1];
        index_t rank = 0;
        for (index_t k = begin; k < end; ++k) rank += a.ids[a.unordered[k]] < id;
        const std::int64_t slot = begin + rank;
        a.order[slot] = static_cast<index_t>(old);
        reinterpret_cast<Packed<T>*>(a.scratch)[slot] = p;
        // the gather copy in the new order, so the fill need not wait for the unpack
        typename Vec4<T>::type p4;
        p4.x = p.x[0];
        p4.y = p.x[1];
        p4.z = p.x[2];
        p4.w = static_cast<T>(p.type);
        reinterpret_cast<typename Vec4<T>::type*>(a.positions4)[slot] = p4;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) md_order_kernel(StepArgs<T> a, int forced)
{
    md_order_body(a, own_launch(), forced);
}

template <typename T>
__device__ __forceinline__ void md_unpack_body(const StepArgs<T>& a, const Launch launch,
                                               int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    for (std::int64_t slot = static_cast<std::int64_t>(launch.block) * blockDim.x + threadIdx.x;
         slot < a.num_atoms; slot += static_cast<std::int64_t>(launch.blocks) * blockDim.x)
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
        if (slot == 0)
        {
            for (int k = 0; k < 3; ++k) a.reference_cell[k] = a.cell[k];
            if (a.rebuild_count) ++a.rebuild_count[0];
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

// ---- bricks ---------------------------------------------------------------------------------
//
// The force kernel works brick by brick. A brick is a box of brick_x x brick_y x brick_z
// sub-bins (brick_z divides kLayerBlock, so a brick never straddles a layer block), numbered x
// fastest, then y, then z. Its halo is every sub-bin within 2 of it along each axis (the
// stencil of each of its sub-bins), periodic, the whole axis where that would reach around it.
// The halo's atoms are listed once per build, as runs of x-consecutive sub-bins (contiguous
// slots), and a row of the list names its neighbors by their position in its brick's halo: a
// 16-bit local index, 0 the sentinel (a far point), 1..H the halo. The force kernel stages the
// halo's positions in LDS and reads every neighbor there.

constexpr int kMaxBrickRuns = 64;    // brick_y * brick_z: the brick's own x-runs
constexpr int kMaxHaloRuns = 128;    // (brick_y + 4) * (brick_z + 4) at most
// Threads of a force block (one block per brick): 1024 for fp32 (16 waves; two blocks fill a CU
// with the largest halos), 512 for fp64 (whose kernels need up to 256 registers).
template <typename T>
constexpr int kBrickThreads = sizeof(T) == 4 ? 1024 : 512;

struct BrickBox
{
    index_t lo[3], hi[3];
};

__device__ __forceinline__ BrickBox brick_box(index_t brick, const index_t dims[3],
                                              const index_t size[3])
{
    const index_t across = (dims[0] + size[0] - 1) / size[0];
    const index_t down = (dims[1] + size[1] - 1) / size[1];
    const index_t ix = brick % across;
    const index_t rest = brick / across;
    const index_t coords[3] = {ix, rest % down, rest / down};
    BrickBox box;
    for (int axis = 0; axis < 3; ++axis)
    {
        box.lo[axis] = coords[axis] * size[axis];
        box.hi[axis] = box.lo[axis] + size[axis] < dims[axis] ? box.lo[axis] + size[axis]
                                                             : dims[axis];
    }
    return box;
}

__device__ __forceinline__ index_t brick_of(index_t x, index_t y, index_t z, const index_t dims[3],
                                            const index_t size[3])
{
    const index_t across = (dims[0] + size[0] - 1) / size[0];
    const index_t down = (dims[1] + size[1] - 1) / size[1];
    return (z / size[2] * down + y / size[1]) * across + x / size[0];
}

// The halo along one axis: its first sub-bin (before wrapping; -2 or more) and its length.
__device__ __forceinline__ void halo_axis(index_t lo, index_t hi, index_t n, index_t* start,
                                          index_t* span)
{
    if (hi - lo + 4 >= n)
    {
        *start = 0;
        *span = n;
    }
    else
    {
        *start = lo - 2;
        *span = hi - lo + 4;
    }
}

__device__ __forceinline__ index_t wrap_index(index_t i, index_t n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// The x-range [start, start + span) of a halo run as at most two runs of the row: [a0, a1], then
// [b0, b1] (empty when b1 < b0) past the x faces.
__device__ __forceinline__ void x_parts(index_t start, index_t span, index_t n, index_t* a0,
                                        index_t* a1, index_t* b0, index_t* b1)
{
    *b0 = 0;
    *b1 = -1;
    if (start < 0)
    {
        *a0 = start + n;
        *a1 = n - 1;
        *b1 = start + span - 1;
    }
    else if (start + span > n)
    {
        *a0 = start;
        *a1 = n - 1;
        *b1 = start + span - 1 - n;
    }
    else
    {
        *a0 = start;
        *a1 = start + span - 1;
    }
}

struct Halo
{
    index_t xs, sx, ys, sy, zs, sz;
};

__device__ __forceinline__ Halo brick_halo(const BrickBox& box, const index_t dims[3])
{
    Halo h;
    halo_axis(box.lo[0], box.hi[0], dims[0], &h.xs, &h.sx);
    halo_axis(box.lo[1], box.hi[1], dims[1], &h.ys, &h.sy);
    halo_axis(box.lo[2], box.hi[2], dims[2], &h.zs, &h.sz);
    return h;
}

// One wave per brick: the runs of its halo, their local starts (halo_runs, 1-based), the slots
// in local order (halo_slots) and the count. With halo_capacity 0 it only counts. The largest
// count goes to the header (byte 656) for sizing; a count above the capacity sets
// NDN_STATUS_HALO_OVERFLOW.
template <typename T>
__device__ __forceinline__ void md_halo_body(const StepArgs<T>& a, const Launch launch,
                                             int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    const int lane = lane_id();
    const index_t dims[3] = {a.nx, a.ny, a.nz};
    const std::int64_t waves = static_cast<std::int64_t>(launch.blocks) * (blockDim.x / kWave);
    for (std::int64_t brick =
             static_cast<std::int64_t>(launch.block) * (blockDim.x / kWave) + threadIdx.x / kWave;
         brick < a.bricks; brick += waves)
    {
        const BrickBox box = brick_box(static_cast<index_t>(brick), dims, a.brick);
        const Halo h = brick_halo(box, dims);
        index_t a0, a1, b0, b1;
        x_parts(h.xs, h.sx, a.nx, &a0, &a1, &b0, &b1);
        const index_t runs = h.sy * h.sz;
        index_t* starts = a.halo_runs + brick * kMaxHaloRuns;
        index_t* slots = a.halo_slots + brick * a.halo_capacity;
        index_t carry = 1;
        for (index_t base = 0; base < runs; base += kWave)
        {
            const index_t q = base + lane;
            index_t first_a = 0, length_a = 0, first_b = 0, length_b = 0;
            if (q < runs)
            {
                const index_t y = wrap_index(h.ys + q % h.sy, a.ny);
                const index_t z = wrap_index(h.zs + q / h.sy, a.nz);
                const std::int64_t row = row_first_bin(y, z, a.nx, a.ny, a.nz);
                first_a = a.bin_offsets[row + a0];
                length_a = a.bin_offsets[row + a1 + 1] - first_a;
                if (b1 >= b0)
                {
                    first_b = a.bin_offsets[row + b0];
                    length_b = a.bin_offsets[row + b1 + 1] - first_b;
                }
            }
            const index_t length = length_a + length_b;
            const index_t inclusive = wave_inclusive_sum(length);
            const index_t start = carry + inclusive - length;
            if (q < runs) starts[q] = start;
            const index_t here = runs - base < kWave ? runs - base : kWave;
            for (index_t k = 0; k < here; ++k)
            {
                const index_t s = read_lane(start, k);
                const index_t fa = read_lane(first_a, k);
                const index_t la = read_lane(length_a, k);
                const index_t fb = read_lane(first_b, k);
                const index_t total = la + read_lane(length_b, k);
                for (index_t t = lane; t < total; t += kWave)
                {
                    const index_t place = s - 1 + t;
                    if (place < a.halo_capacity) slots[place] = t < la ? fa + t : fb + (t - la);
                }
            }
            carry += read_lane(inclusive, kWave - 1);
        }
        if (lane == 0)
        {
            const index_t count = carry - 1;
            a.halo_counts[brick] = count;
            atomicMax(halo_needed(a.header), count);
            if (a.halo_capacity > 0 && count > a.halo_capacity)
                atomicOr(a.status, kStatusHaloOverflow);
        }
    }
}

// The build's exclusive scan of the sub-bin counts into bin_offsets (bins + 1 values). A pass
// takes 16 rows of 1024 consecutive values, thread t holding value t of each row (every load and
// store coalesced): each row is scanned within its waves, one wave scans the pass's 16 x 16 wave
// totals in order, and every value adds its row's and wave's prefix; the running total carries
// from pass to pass. Each block of the launch takes its own run of passes and first sums every
// count before its run itself (integer sums: the same bits in any order), so the blocks need
// nothing from each other: one launch where a tiled scan takes three, and the steps without a
// build, nearly all, pay one empty launch.
constexpr int kListScanThreads = 1024;
constexpr int kListScanRows = 16;
constexpr int kListScanWaves = kListScanThreads / kWave;
constexpr std::int64_t kListScanPass = static_cast<std::int64_t>(kListScanThreads) * kListScanRows;

// Blocks of the scan of `bins` sub-bins: one per pass.
inline std::int64_t scan_blocks(std::int64_t bins)
{
    return (bins + 1 + kListScanPass - 1) / kListScanPass;
}

template <typename T>
__device__ __forceinline__ void md_scan_body(const StepArgs<T>& a, const Launch launch)
{
    // [row][wave] totals, scanned in place; then the pass total
    __shared__ index_t totals[kListScanRows * kListScanWaves + 1];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const std::int64_t count = a.bins + 1;
    constexpr std::int64_t kPass = kListScanPass;
    const std::int64_t passes = (count + kPass - 1) / kPass;
    const std::int64_t per_block = (passes + launch.blocks - 1) / launch.blocks;
    const std::int64_t begin = launch.block * per_block * kPass;
    if (begin >= count) return;
    const std::int64_t end = begin + per_block * kPass < count ? begin + per_block * kPass : count;
    // the counts before this block's run: 4 at a time (the run starts at a multiple of 16384)
    index_t before = 0;
    const int4* counts4 = reinterpret_cast<const int4*>(a.bin_counts);
    for (std::int64_t q = threadIdx.x; q < begin / 4; q += kListScanThreads)
    {
        const int4 c = counts4[q];
        before += (c.x + c.y) + (c.z + c.w);
    }
    before = wave_inclusive_sum(before);
    if (lane == kWave - 1) totals[wave] = before;
    __syncthreads();
    index_t carry = 0;
    for (int w = 0; w < kListScanWaves; ++w) carry += totals[w];
    __syncthreads();
    for (std::int64_t base = begin; base < end; base += kPass)
    {
        index_t v[kListScanRows], inclusive[kListScanRows];
#pragma unroll
        for (int j = 0; j < kListScanRows; ++j)
        {
            const std::int64_t at = base + static_cast<std::int64_t>(j) * kListScanThreads + threadIdx.x;
            v[j] = at < end ? a.bin_counts[at] : 0;
        }
#pragma unroll
        for (int j = 0; j < kListScanRows; ++j) inclusive[j] = wave_inclusive_sum(v[j]);
        if (lane == kWave - 1)
#pragma unroll
            for (int j = 0; j < kListScanRows; ++j) totals[j * kListScanWaves + wave] = inclusive[j];
        __syncthreads();
        if (wave == 0)
        {
            // 256 totals in (row, wave) order, 4 per lane
            constexpr int kPer = kListScanRows * kListScanWaves / kWave;
            index_t mine[kPer];
            index_t sum = 0;
#pragma unroll
            for (int e = 0; e < kPer; ++e)
            {
                mine[e] = totals[lane * kPer + e];
                sum += mine[e];
            }
            index_t before = wave_inclusive_sum(sum) - sum;
#pragma unroll
            for (int e = 0; e < kPer; ++e)
            {
                totals[lane * kPer + e] = before;
                before += mine[e];
            }
            if (lane == kWave - 1) totals[kListScanRows * kListScanWaves] = before;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kListScanRows; ++j)
        {
            const std::int64_t at = base + static_cast<std::int64_t>(j) * kListScanThreads + threadIdx.x;
            if (at < end)
                a.bin_offsets[at] = carry + totals[j * kListScanWaves + wave] + inclusive[j] - v[j];
        }
        carry += totals[kListScanRows * kListScanWaves];
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(kListScanThreads) md_scan_kernel(StepArgs<T> a, int forced)
{
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    md_scan_body(a, own_launch());
}

// The scatter and the halo listing, both after the scan, in one launch: the first
// `scatter_blocks` blocks scatter, the rest list the halos.
template <typename T>
__device__ __forceinline__ void md_scatter_halo_body(const StepArgs<T>& a, const Launch launch,
                                                     index_t scatter_blocks, int forced)
{
    if (launch.block < scatter_blocks)
        md_scatter_body(a, Launch{launch.block, scatter_blocks}, forced);
    else
        md_halo_body(a, Launch{launch.block - scatter_blocks, launch.blocks - scatter_blocks},
                     forced);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) md_scatter_halo_kernel(StepArgs<T> a,
                                                                 index_t scatter_blocks, int forced)
{
    md_scatter_halo_body(a, own_launch(), scatter_blocks, forced);
}

// Systems of at most kSmallBuild atoms: the whole build (assign, scan, scatter with the halo
// listing, order, unpack, fill) by one block of 1,024 threads, its phases separated by the
// block's barriers instead of kernel boundaries: one launch on every step where the build takes
// six. The gate is read once per phase by one thread (a phase may set a status bit, and the
// phases after it then do nothing, as the launches of a large build do); every phase is the same
// body the large build runs, so the list is the same bit for bit.
constexpr std::int64_t kSmallBuild = 2048;

template <typename T>
__global__ void __launch_bounds__(kListScanThreads) md_build_small_kernel(StepArgs<T> a, T radius,
                                                                          T cull_squared,
                                                                          int forced)
{
    __shared__ int open;
    const Launch one{0, 1};
    auto gate = [&]() {
        if (threadIdx.x == 0) open = forced || rebuilding(a.rebuild, a.status);
        __syncthreads();
        const bool go = open != 0;
        // before thread 0 reads the gate again
        __syncthreads();
        return go;
    };
    if (!gate()) return;
    md_assign_body(a, one, 1);
    __syncthreads();
    if (!gate()) return;
    md_scan_body(a, Launch{0, 1});
    __syncthreads();
    if (!gate()) return;
    md_scatter_body(a, one, 1);
    md_halo_body(a, one, 1);
    __syncthreads();
    if (!gate()) return;
    md_order_body(a, one, 1);
    __syncthreads();
    if (!gate()) return;
    md_unpack_body(a, one, 1);
    __syncthreads();
    if (!gate()) return;
    md_fill_body(a, one, radius, cull_squared, 1);
}

// The periodic frame of a brick. Along each axis every position p is replaced by its image nearest
// the brick, written so that a plain difference of two images is bit for bit the reference's
// minimum-image displacement (space.py:11-19) of every pair within `reach` of the brick's atoms:
// a brick in the lower half keeps p, or takes p - L for an atom more than L / 2 above its center
// (then p > L / 2, so p - L is exact, and c - (p - L) is the reference's c - (o - L)); a brick
// whose halo passes the upper face moves its whole frame down by L: p - L for every atom but
// those more than L / 2 below its center, which keep p (p - L is exact for all atoms within
// reach, which lie in the upper half, and (c - L) - p is the reference's (c - L) - o). The frame
// exists only when every atom within reach of the brick lies within L / 2 of its center (and, for
// the moved frame, in the upper half); otherwise `exact` is false and the walk takes the
// reference's displacement per pair on unmoved positions.
template <typename T>
struct BrickFrame
{
    T center[3], length[3];
    bool moved[3];
    bool exact;
    // placing is the identity for every atom the brick reads: its halo's sub-bins lie at least
    // one sub-bin inside every face (an atom drifts less than a sub-bin between builds, so none
    // of them crosses a face), far from the half-box distance, and the frame is not moved
    bool plain;

    __device__ __forceinline__ T image(T p, int axis) const
    {
        const T offset = p - center[axis];
        const T half = T(0.5) * length[axis];
        if (moved[axis]) return offset < -half ? p : p - length[axis];
        return offset > half ? p - length[axis] : p;
    }

    template <typename V4>
    __device__ __forceinline__ V4 place(V4 p) const
    {
        if (!exact || plain) return p;
        p.x = image(p.x, 0);
        p.y = image(p.y, 1);
        p.z = image(p.z, 2);
        return p;
    }
};

template <typename T>
__device__ __forceinline__ BrickFrame<T> brick_frame(const StepArgs<T>& a, const BrickBox& box,
                                                     T reach)
{
    BrickFrame<T> f;
    f.exact = true;
    const T widths[3] = {a.grid_widths[0], a.grid_widths[1], a.grid_widths[2]};
    for (int axis = 0; axis < 3; ++axis)
    {
        const T w = widths[axis];
        const T lo = static_cast<T>(box.lo[axis]) * w;
        const T hi = static_cast<T>(box.hi[axis]) * w;
        const T length = a.cell[axis];
        f.center[axis] = T(0.5) * (lo + hi);
        f.length[axis] = length;
        // the brick's atoms may have left its sub-bins by up to half the list's skin since the
        // build; a sub-bin of margin (at least half the list's radius) covers that and the
        // rounding of the sub-bin bounds
        const T extent = T(0.5) * (hi - lo) + reach + w;
        f.moved[axis] = f.center[axis] + extent > length;
        const bool inside = extent < T(0.5) * length;
        const bool upper = !f.moved[axis] || f.center[axis] - extent > T(0.5) * length;
        f.exact = f.exact && inside && upper;
    }
    // the halo's sub-bins [lo - 2, hi + 2) with a sub-bin more inside every face, and the whole
    // span plus that margin within half the box
    const index_t dims[3] = {a.nx, a.ny, a.nz};
    f.plain = f.exact;
    for (int axis = 0; axis < 3; ++axis)
        f.plain = f.plain && !f.moved[axis] && box.lo[axis] >= 3 &&
                  box.hi[axis] + 3 <= dims[axis] && box.hi[axis] - box.lo[axis] + 6 < dims[axis];
    return f;
}

// Centers a fill wave tests together: their coordinates and counts stay in scalar registers.
template <typename T>
constexpr int kFillCenters = sizeof(T) == 4 ? 8 : 4;

// How a fill wave forms a displacement: plain differences of raw positions, of positions placed
// in the sub-bin's frame, or the minimum image per pair.
constexpr int kFramePlain = 0, kFramePlaced = 1, kFrameImage = 2;

// One wave per sub-bin. Its atoms (the centers) share one stencil: lanes 0..49 hold the slot
// ranges of the 25 stencil rows (runs of x-consecutive sub-bins; two parts where a run wraps
// around x), culled by the distance between sub-bins, so every sub-bin a pair within R could
// touch is kept, and the offset from a slot of the range to its local index in the brick's halo
// (each range lies inside one run of the halo). The candidates are packed 64 at a time; each
// chunk is gathered once and tested against a group of centers in turn with the exact
// minimum-image expression, and each center's hits are ballot-compacted into its row, in the
// order met (a sub-bin with more atoms than a group walks its candidates once per group).
// Rows are then padded with the sentinel 0 up to their count rounded up to 32, so the force kernel
// reads whole blocks unmasked. With stride 0 it only counts (the first frame, to size the rows).
template <typename T>
__device__ __forceinline__ void md_fill_body(const StepArgs<T>& a, const Launch launch,
                                             T radius, T cull_squared, int forced)
{
#pragma clang fp contract(off)
    using V4 = typename Vec4<T>::type;
    if (!forced && !rebuilding(a.rebuild, a.status)) return;
    const T radius_squared = radius * radius;
    const int lane = lane_id();
    const V4* pos = reinterpret_cast<const V4*>(a.positions4);
    const T lx = a.cell[0], ly = a.cell[1], lz = a.cell[2];
    const index_t nx = a.nx, ny = a.ny, nz = a.nz;
    const index_t dims[3] = {nx, ny, nz};
    const bool full_x = nx < kStencil, full_y = ny < kStencil, full_z = nz < kStencil;
    const int span_x = full_x ? nx : kStencil;
    const int span_y = full_y ? ny : kStencil;
    const int span_z = full_z ? nz : kStencil;
    const bool writing = a.stride > 0;
    const index_t stride = static_cast<index_t>(a.stride);
    // one wave per sub-bin, the waves of the grid striding over the sub-bins
    const std::int64_t waves = static_cast<std::int64_t>(launch.blocks) * (blockDim.x / kWave);
    for (std::int64_t bin =
             static_cast<std::int64_t>(launch.block) * (blockDim.x / kWave) + threadIdx.x / kWave;
         bin < a.bins; bin += waves)
    {
        // the sub-bin is the wave's: its bounds, and every center's row address below, scalar
        const index_t begin = __builtin_amdgcn_readfirstlane(a.bin_offsets[bin]);
        const index_t end = __builtin_amdgcn_readfirstlane(a.bin_offsets[bin + 1]);
        if (begin == end) continue;
        index_t hx, hy, hz;
        bin_coordinates(bin, nx, ny, nz, &hx, &hy, &hz);
        const index_t brick = brick_of(hx, hy, hz, dims, a.brick);
        // the sub-bin's own frame: the plain difference of placed positions is the reference's
        // minimum image for every pair within R of its atoms
        const BrickBox own{{hx, hy, hz}, {hx + 1, hy + 1, hz + 1}};
        const BrickFrame<T> frame = brick_frame(a, own, radius);
        const Halo h = brick_halo(brick_box(brick, dims, a.brick), dims);
        index_t a0, a1, b0, b1;
        x_parts(h.xs, h.sx, nx, &a0, &a1, &b0, &b1);
        // this lane's range: stencil row lane / 2, part lane % 2 of its run of x sub-bins
        index_t lo = 0, length = 0, to_local = 0;
        const int row = lane / 2, part = lane % 2;
        if (row < span_y * span_z)
        {
            const int oy = row % span_y, oz = row / span_y;
            const T dy = full_y ? T(0) : bin_gap(oy - 2, a.grid_widths[1]);
            const T dz = full_z ? T(0) : bin_gap(oz - 2, a.grid_widths[2]);
            const T dyz = dy * dy + dz * dz;
            if (dyz < cull_squared)
            {
                int first = span_x, last = -1;
                for (int ox = 0; ox < span_x; ++ox)
                {
                    const T dx = full_x ? T(0) : bin_gap(ox - 2, a.grid_widths[0]);
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
                        // the halo run of row (wy, wz) and the part of it this range lies in
                        if (writing)
                        {
                            const index_t q = wrap_index(wz - h.zs, nz) * h.sy +
                                              wrap_index(wy - h.ys, ny);
                            const index_t start = a.halo_runs[brick * kMaxHaloRuns + q];
                            const index_t first_a = a.bin_offsets[base + a0];
                            if (run_lo >= a0 && run_lo <= a1)
                                to_local = start - first_a;
                            else
                                to_local = start + (a.bin_offsets[base + a1 + 1] - first_a) -
                                           a.bin_offsets[base + b0];
                        }
                    }
                }
            }
        }
        const index_t inclusive = wave_inclusive_sum(length);
        const index_t start = inclusive - length;
        // wave-uniform (a scalar read), like everything the walk's masks are built from
        const index_t total = read_lane(inclusive, kWave - 1);
        // The frame kind: kFramePlain where placing is the identity for every candidate (the
        // stencil's sub-bins lie inside the box, far from the half-box distance), kFramePlaced
        // where the sub-bin's frame moves some (plain differences of placed positions),
        // kFrameImage without a frame (a short axis: the minimum image per pair).
        int kind = frame.exact ? kFramePlain : kFrameImage;
        if (kind == kFramePlain)
        {
            const index_t at[3] = {hx, hy, hz};
            for (int axis = 0; axis < 3; ++axis)
                if (frame.moved[axis] || dims[axis] < 8 || at[axis] < 2 ||
                    at[axis] + 2 > dims[axis] - 1)
                    kind = kFramePlaced;
        }
        // candidate t of the stencil: its slot, its local index in the brick's halo (past an
        // overflowing halo's capacity the sentinel; the status reports it), whether t is one;
        // the range holding t is the last lane whose start is <= t
        auto candidate = [&](index_t t, index_t& slot, index_t& local, bool& valid) {
            int r = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
            {
                const index_t s = __shfl(start, r + step, kWave);
                if (s <= t) r += step;
            }
            valid = t < total;
            slot = __shfl(lo, r, kWave) + (t - __shfl(start, r, kWave));
            local = slot + __shfl(to_local, r, kWave);
            local = local <= a.halo_capacity ? local : 0;
        };
        // Every slot is one candidate, and the sub-bin's own slots are consecutive candidates
        // (they lie in one range): center `first` + k is candidate own_first(first) + k.
        auto own_first = [&](index_t first) {
            const std::uint64_t holder =
                ballot(length > 0 && lo <= first && first < lo + length);
            if (holder == 0) return total;
            const int r = __builtin_ctzll(holder);
            return read_lane(start, r) + (first - read_lane(lo, r));
        };
        // Center k's hits in one chunk, `mask` (its own candidate excluded), go to its
        // row at `count` onwards; a row that overflows keeps its first `stride` entries (the
        // status reports it).
        auto write_hits = [&](std::uint64_t mask, index_t count, char* rows, int k, index_t local) {
            const bool hit = __builtin_amdgcn_inverse_ballot_w64(mask);
            const unsigned position =
                static_cast<unsigned>(count) +
                __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
            // `rows` is the group's first row, wave-uniform; the entry is a 32-bit byte offset
            // from it (k < 64 rows on, each shorter than a brick's halo with its margin)
            if (hit && position < static_cast<unsigned>(stride))
                *reinterpret_cast<std::uint16_t*>(
                    rows + ((position + static_cast<unsigned>(k * stride)) << 1)) =
                    static_cast<std::uint16_t>(local);
        };
        // the group's counts (lane k holds center first + k's), the overflow flag, and the
        // sentinel padding of each row
        auto close_rows = [&](index_t first, int centers, index_t mine_count) {
            if (lane < centers)
            {
                a.counts[first + lane] = mine_count;
                if (writing && mine_count > a.stride)
                {
                    atomicOr(a.status, kStatusFullOverflow);
                    atomicMax(a.status + 1, mine_count);
                }
            }
            if (writing)
                for (int k = 0; k < centers; ++k)
                {
                    const index_t count = read_lane(mine_count, k);
                    std::uint16_t* row_out = a.index + static_cast<std::int64_t>(first + k) * a.stride;
                    const index_t padded = (count + 31) & ~31;
                    for (index_t p = count + lane; p < padded && p < stride; p += kWave)
                        row_out[p] = 0;
                }
        };
        if (kind != kFrameImage)
        {
            // kFillCenters<T> centers at a time: their coordinates and counts are wave-uniform
            // (scalar registers) for the whole walk over the candidates, so a test costs the
            // displacement, the compare and, on a hit, the compacted store
            constexpr int kCenters = kFillCenters<T>;
            for (index_t first = begin; first < end; first += kCenters)
            {
                const int centers = end - first < kCenters ? end - first : kCenters;
                const V4 mine = frame.place(pos[lane < centers ? first + lane : first]);
                T cx[kCenters], cy[kCenters], cz[kCenters];
                index_t count[kCenters];
#pragma unroll
                for (int k = 0; k < kCenters; ++k)
                {
                    cx[k] = read_lane(mine.x, k);
                    cy[k] = read_lane(mine.y, k);
                    cz[k] = read_lane(mine.z, k);
                    count[k] = 0;
                }
                const index_t own = own_first(first);
                char* rows = reinterpret_cast<char*>(a.index + static_cast<std::int64_t>(first) * a.stride);
                // each chunk's positions are read while the previous chunk is tested
                index_t next_slot, next_local;
                bool next_valid;
                candidate(lane, next_slot, next_local, next_valid);
                V4 next_raw = pos[next_valid ? next_slot : first];
                // whether the frame moves candidates, and whether rows are written or only
                // counted, are compile-time choices inside the walk
                auto chunks = [&](auto placed, auto writes) {
                    for (index_t chunk = 0; chunk < total; chunk += kWave)
                    {
                        const index_t local = next_local;
                        const std::uint64_t valid = ballot(next_valid);
                        const V4 q = decltype(placed)::value ? frame.place(next_raw) : next_raw;
                        if (chunk + kWave < total)
                        {
                            candidate(chunk + kWave + lane, next_slot, next_local, next_valid);
                            next_raw = pos[next_valid ? next_slot : first];
                        }
                        // the look-ahead read stays here: the scheduler would otherwise sink it
                        __builtin_amdgcn_sched_barrier(0);
                        // this lane's candidate is center k's own where rank == k (held
                        // opaque, so each center's test compares it with the constant k rather
                        // than keeping a rebased copy per center)
                        index_t rank = chunk + lane - own;
                        asm volatile("" : "+v"(rank));
#pragma unroll
                        for (int k = 0; k < kCenters; ++k)
                        {
                            if (k < centers)
                            {
                                const T dx = cx[k] - q.x;
                                const T dy = cy[k] - q.y;
                                const T dz = cz[k] - q.z;
                                const T xx = dx * dx;
                                const T yy = dy * dy;
                                const T zz = dz * dz;
                                // each ballot of one compare is that compare's mask
                                const std::uint64_t mask =
                                    ballot((xx + yy) + zz < radius_squared) & ballot(rank != k) &
                                    valid;
                                if constexpr (decltype(writes)::value)
                                    write_hits(mask, count[k], rows, k, local);
                                count[k] += static_cast<index_t>(__popcll(mask));
                                // the store and the count stay here: sunk below the later
                                // centers' tests, they would keep every center's mask live
                                asm volatile("" : "+s"(count[k]) : : "memory");
                            }
                        }
                    }
                };
                if (writing)
                {
                    if (kind == kFramePlain)
                        chunks(std::false_type{}, std::true_type{});
                    else
                        chunks(std::true_type{}, std::true_type{});
                }
                else
                {
                    if (kind == kFramePlain)
                        chunks(std::false_type{}, std::false_type{});
                    else
                        chunks(std::true_type{}, std::false_type{});
                }
                index_t mine_count = count[0];
#pragma unroll
                for (int k = 1; k < kCenters; ++k) mine_count = lane == k ? count[k] : mine_count;
                close_rows(first, centers, mine_count);
            }
        }
        else
        {
            // the minimum image per pair: 64 centers at a time, lane k holding center first + k
            // and its count
            for (index_t first = begin; first < end; first += kWave)
            {
                const int centers = end - first < kWave ? end - first : kWave;
                const V4 mine = pos[lane < centers ? first + lane : first];
                index_t mine_count = 0;
                const index_t own = own_first(first);
                char* rows = reinterpret_cast<char*>(a.index + static_cast<std::int64_t>(first) * a.stride);
                index_t next_slot, next_local;
                bool next_valid;
                candidate(lane, next_slot, next_local, next_valid);
                V4 next_raw = pos[next_valid ? next_slot : first];
                for (index_t chunk = 0; chunk < total; chunk += kWave)
                {
                    const index_t local = next_local;
                    const std::uint64_t valid = ballot(next_valid);
                    const V4 q = next_raw;
                    if (chunk + kWave < total)
                    {
                        candidate(chunk + kWave + lane, next_slot, next_local, next_valid);
                        next_raw = pos[next_valid ? next_slot : first];
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    const index_t rank = chunk + lane - own;
                    for (int k = 0; k < centers; ++k)
                    {
                        const T dx = minimum_image(read_lane(mine.x, k), q.x, lx);
                        const T dy = minimum_image(read_lane(mine.y, k), q.y, ly);
                        const T dz = minimum_image(read_lane(mine.z, k), q.z, lz);
                        const T xx = dx * dx;
                        const T yy = dy * dy;
                        const T zz = dz * dz;
                        const std::uint64_t mask = ballot((xx + yy) + zz < radius_squared) &
                                                   ballot(rank != k) & valid;
                        const index_t count = read_lane(mine_count, k);
                        if (writing) write_hits(mask, count, rows, k, local);
                        mine_count += lane == k ? static_cast<index_t>(__popcll(mask)) : 0;
                    }
                }
                close_rows(first, centers, mine_count);
            }
        }
    }
}

// The unpack and the fill in one launch: neither reads what the other writes (the fill reads the
// gather copy the order wrote, the halo lists and the sub-bin offsets; the unpack writes the
// frame home), so the first `unpack_blocks` blocks unpack and the rest fill.
template <typename T>
__device__ __forceinline__ void md_unpack_fill_body(const StepArgs<T>& a, const Launch launch,
                                                    index_t unpack_blocks, T radius,
                                                    T cull_squared, int forced)
{
    if (launch.block < unpack_blocks)
        md_unpack_body(a, Launch{launch.block, unpack_blocks}, forced);
    else
        md_fill_body(a, Launch{launch.block - unpack_blocks, launch.blocks - unpack_blocks},
                     radius, cull_squared, forced);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
md_unpack_fill_kernel(StepArgs<T> a,
                                                                index_t unpack_blocks, T radius,
                                                                T cull_squared, int forced)
{
    md_unpack_fill_body(a, own_launch(), unpack_blocks, radius, cull_squared, forced);
}

// ---- forces ---------------------------------------------------------------------------------

inline unsigned xcd_grid(std::int64_t blocks)
{
    return static_cast<unsigned>((blocks + 7) / 8 * 8);
}

// The value of lane (lane ^ 1), (lane ^ 2) within its quad, or of lane 7 - i (i the lane within
// its 8) or 15 - i (within its 16): data-parallel moves, no LDS round trip.
template <int kControl>
__device__ __forceinline__ float dpp_move(float value)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(value), kControl, 0xF, 0xF, false));
}

template <int kControl>
__device__ __forceinline__ double dpp_move(double value)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(value), kControl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(value), kControl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

constexpr int kDppSwapOne = 0xB1;      // quad_perm [1, 0, 3, 2]
constexpr int kDppSwapTwo = 0x4E;      // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;  // lane 7 - i within each 8
constexpr int kDppMirror = 0x140;      // lane 15 - i within each 16

// The sum over each aligned group of G lanes (G = 2, 4, 8 or 16), the same bits on every lane of
// the group (each step adds the same two partial sums, a + b on one lane and b + a on the other).
template <int G, typename V>
__device__ __forceinline__ V group_sum(V value)
{
    static_assert(G == 2 || G == 4 || G == 8 || G == 16, "groups of 2 to 16 lanes");
    value += dpp_move<kDppSwapOne>(value);
    if (G >= 4) value += dpp_move<kDppSwapTwo>(value);
    if (G >= 8) value += dpp_move<kDppHalfMirror>(value);
    if (G >= 16) value += dpp_move<kDppMirror>(value);
    return value;
}

// The largest value over each aligned group of G lanes.
template <int G>
__device__ __forceinline__ index_t group_max(index_t value)
{
    static_assert(G == 2 || G == 4 || G == 8 || G == 16, "groups of 2 to 16 lanes");
    auto take = [](index_t v, index_t other) { return other > v ? other : v; };
    value = take(value, __builtin_amdgcn_update_dpp(0, value, kDppSwapOne, 0xF, 0xF, false));
    if (G >= 4) value = take(value, __builtin_amdgcn_update_dpp(0, value, kDppSwapTwo, 0xF, 0xF, false));
    if (G >= 8) value = take(value, __builtin_amdgcn_update_dpp(0, value, kDppHalfMirror, 0xF, 0xF, false));
    if (G >= 16) value = take(value, __builtin_amdgcn_update_dpp(0, value, kDppMirror, 0xF, 0xF, false));
    return value;
}

// Four entries of a row: one 8-byte load. A plain load, not a non-temporal one: a row's lanes
// read 32 bytes of it per pass, and a line kept in L2 serves the row's next passes (non-temporal
// reads fetched a third more from HBM and ran 7% slower).
typedef unsigned Pair2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ Pair2 stream_entries(const std::uint16_t* entries)
{
    return *reinterpret_cast<const Pair2*>(entries);
}

// The pairs of four entries of a row (one lane's 8-byte read): positions from the halo in LDS,
// the minimum image (or the direct difference on a fast wave), the cutoff mask and the sums.
template <bool kFast, bool kVirial, bool kEnergy, typename T, typename V4, typename Ev>
__device__ __forceinline__ void walk_quad(const Ev& ev, const V4& c, const V4* halo, Pair2 raw,
                                          bool any, T lx, T ly, T lz, T cutoff_squared, T& fx,
                                          T& fy, T& fz, T& energy, T* w)
{
    const unsigned local[4] = {raw.x & 0xffffu, raw.x >> 16, raw.y & 0xffffu, raw.y >> 16};
    const V4 q[4] = {halo[local[0]], halo[local[1]], halo[local[2]], halo[local[3]]};
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
        T pu, fs;
        ev.eval(c.w, q[u].w, r2, pu, fs);
        // a select, not a product: a masked pair's fs may be inf or NaN
        fs = inside ? fs : T(0);
        if (kEnergy) energy += inside ? pu : T(0);
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

// One group's walk over its row: lane s of the row's G lanes takes the 4 entries at 4G m + 4s
// (one 8-byte read) for m = 0, 1, ..., two at a time with both reads issued first (the second's
// latency hides behind the first's arithmetic), up to `limit`, the row's count rounded up to a
// block of 32. A row is padded with the sentinel 0, a point far from every atom, so entries past
// the count fail the cutoff test by themselves and need no mask.
template <int G, bool kFast, bool kVirial, bool kEnergy, typename T, typename V4, typename Ev>
__device__ __forceinline__ void walk_row(const Ev& ev, const V4& c, const V4* halo,
                                         const std::uint16_t* __restrict__ entries, index_t limit,
                                         T lx, T ly, T lz, T cutoff_squared, T& fx, T& fy, T& fz,
                                         T& energy, T* w)
{
    for (index_t first = 4 * (lane_id() % G); first < limit; first += 8 * G)
    {
        const index_t second = first + 4 * G;
        const bool any_second = second < limit;
        const Pair2 raw_first = stream_entries(entries + first);
        const Pair2 raw_second = stream_entries(entries + (any_second ? second : first));
        // both reads stay here: the scheduler would otherwise sink the second to its use
        __builtin_amdgcn_sched_barrier(0);
        walk_quad<kFast, kVirial, kEnergy>(ev, c, halo, raw_first, true, lx, ly, lz,
                                           cutoff_squared, fx, fy, fz, energy, w);
        // past the row's end the second quad is all masked (no branch: the wave walks as far as
        // its longest row anyway)
        walk_quad<kFast, kVirial, kEnergy>(ev, c, halo, raw_second, any_second, lx, ly, lz,
                                           cutoff_squared, fx, fy, fz, energy, w);
    }
}

// The block of brick `b`: blocks are dealt to the 8 XCDs round-robin (block % 8: a batch's systems start
// on multiples of 8), and the
// blocks of one XCD take a contiguous eighth of the bricks, so bricks side by side, whose halos
// overlap, share that XCD's L2. Blocks past the last brick idle.
__device__ __forceinline__ index_t xcd_brick(index_t bricks, const Launch launch)
{
    const index_t per_xcd = (bricks + 7) / 8;
    return (launch.block % 8) * per_xcd + launch.block / 8;
}

// The LDS of a brick's block, one dynamic allocation: the halo (positions and types; entry 0 the
// far point) from offset 0, so a neighbor's address is its local index times 16 (32 in fp64),
// then the brick's runs and the block's sums.
struct BrickShared
{
    index_t run_first[kMaxBrickRuns];
    index_t run_rows[kMaxBrickRuns + 1];
    double sums[16];
};

// Bytes of a block's LDS past the halo.
constexpr std::size_t kBrickSharedBytes = (sizeof(BrickShared) + 15) / 16 * 16;

template <typename T>
__host__ __device__ __forceinline__ std::size_t halo_bytes(std::int64_t halo_capacity)
{
    return static_cast<std::size_t>(halo_capacity + 1) * sizeof(typename Vec4<T>::type);
}

// Dynamic LDS of a brick's block: the halo (with the sentinel) and the block's bookkeeping (a
// potential's staging follows).
template <typename T>
__host__ __device__ __forceinline__ std::size_t brick_block_bytes(std::int64_t halo_capacity)
{
    return halo_bytes<T>(halo_capacity) + kBrickSharedBytes;
}

template <typename T>
__device__ __forceinline__ BrickShared* brick_shared(unsigned char* lds, std::int64_t halo_capacity)
{
    return reinterpret_cast<BrickShared*>(lds + halo_bytes<T>(halo_capacity));
}

__device__ __forceinline__ index_t brick_row_slot(index_t r, index_t runs,
                                                  const index_t* run_first,
                                                  const index_t* run_rows)
{
    int run = 0;
    for (int step = kMaxBrickRuns / 2; step > 0; step >>= 1)
        if (run + step < runs && run_rows[run + step] <= r) run += step;
    return run_first[run] + (r - run_rows[run]);
}

// A brick staged by its block: the halo in LDS, placed in the brick's frame, and the brick's
// x-runs (y outer, z inner: slot order) with their first rows; row r of the walk is slot_of(r).
// (Walking the rows sorted by length instead measured slower: rows of one wave then come from
// all over the brick.)
template <typename T>
struct StagedBrick
{
    index_t runs;
    index_t rows;
    const BrickShared* shared;
    BrickFrame<T> frame;

    __device__ __forceinline__ index_t slot_of(index_t r) const
    {
        return brick_row_slot(r, runs, shared->run_first, shared->run_rows);
    }
};

// A slot's position as the step's kernels see it, from the frame's positions (current after
// every drift), with the atom's type as w when the potential reads types (else 0). The float4
// copy of the frame is written by the build's order kernel for the build's own fill only.
template <bool kTypes, typename T>
__device__ __forceinline__ typename Vec4<T>::type current_position(const StepArgs<T>& a,
                                                                    index_t slot)
{
    const T* x = a.positions + 3LL * slot;
    typename Vec4<T>::type p;
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    p.w = kTypes && a.types != nullptr ? static_cast<T>(a.types[slot]) : T(0);
    return p;
}

template <int kThreads, bool kTypes, typename T, typename V4>
__device__ __forceinline__ StagedBrick<T> stage_brick(const StepArgs<T>& a, index_t brick,
                                                      V4* halo, BrickShared* sh, T reach)
{
    const index_t dims[3] = {a.nx, a.ny, a.nz};
    const BrickBox box = brick_box(brick, dims, a.brick);
    const BrickFrame<T> frame = brick_frame(a, box, reach);
    // an overflowing halo (the status reports it) loads what its slots hold
    const index_t listed = a.halo_counts[brick];
    const index_t halo_size = listed < a.halo_capacity ? listed : a.halo_capacity;
    const index_t* slots = a.halo_slots + static_cast<std::int64_t>(brick) * a.halo_capacity;
    for (index_t k = threadIdx.x; k < halo_size; k += kThreads)
        halo[k + 1] = frame.place(current_position<kTypes>(a, slots[k]));
    if (threadIdx.x == 0)
    {
        V4 far;
        far.x = far.y = far.z = kFarAway<T>;
        far.w = T(0);
        halo[0] = far;
    }
    const index_t layers = box.hi[2] - box.lo[2];
    const index_t runs = (box.hi[1] - box.lo[1]) * layers;
    if (threadIdx.x < kWave)
    {
        const int lane = threadIdx.x;
        index_t length = 0;
        if (lane < runs)
        {
            const std::int64_t row = row_first_bin(box.lo[1] + lane / layers,
                                                   box.lo[2] + lane % layers, a.nx, a.ny, a.nz);
            const index_t first = a.bin_offsets[row + box.lo[0]];
            length = a.bin_offsets[row + box.hi[0]] - first;
            sh->run_first[lane] = first;
        }
        const index_t inclusive = wave_inclusive_sum(length);
        if (lane < runs) sh->run_rows[lane + 1] = inclusive;
        if (lane == 0) sh->run_rows[0] = 0;
    }
    __syncthreads();
    return StagedBrick<T>{runs, sh->run_rows[runs], sh, frame};
}

// Lanes per row of the prune kernel: 16, each reading 4 entries per pass.
constexpr int kPruneGroup = 16;

// Refreshes the pruned rows of a brick staged with images exact out to the pruned radius (the
// halo in LDS), by its block: 16 lanes per row test the row's pairs against the pruned radius
// with the exact minimum-image expression and write those kept from the row's start in pass
// order (ballot ranks, so each store is one contiguous run per row), padded with the sentinel 0
// up to a multiple of 32; the count is the number kept. Since every pair kept is in the full
// row, a pruned row never outgrows the stride. The frame of the refresh is kept for the check.
template <typename T, typename V4>
__device__ __forceinline__ void prune_brick_rows(const StepArgs<T>& a,
                                                 const StagedBrick<T>& staged, const V4* halo)
{
#pragma clang fp contract(off)
    const T lx = a.cell[0], ly = a.cell[1], lz = a.cell[2];
    const T inner_squared = a.inner_radius * a.inner_radius;
    const bool fast = staged.frame.exact;
    const int lane = lane_id();
    const int sub = lane % kPruneGroup;
    const std::uint64_t group = ((1ull << kPruneGroup) - 1ull) << (lane - sub);
    const std::uint64_t below = (1ull << lane) - 1ull;
    const index_t stride = static_cast<index_t>(a.stride);
    constexpr index_t kRowsPerPass = kBrickThreads<T> / kPruneGroup;
    constexpr index_t kPass = 4 * kPruneGroup;
    // the row of this lane's first turn; each turn loads the next turn's count and position
    index_t r = static_cast<index_t>(threadIdx.x) / kPruneGroup;
    bool live = r < staged.rows;
    index_t slot = live ? staged.slot_of(r) : 0;
    index_t listed = live ? a.counts[slot] : 0;
    V4 p = current_position<false>(a, slot);
    for (index_t first_row = 0; first_row < staged.rows; first_row += kRowsPerPass)
    {
        V4 c = staged.frame.place(p);
        if (!live) c.x = c.y = c.z = -kFarAway<T>;
        const index_t count = listed < stride ? listed : stride;
        const index_t limit = (count + 31) & ~31;
        const std::int64_t row_start = static_cast<std::int64_t>(slot) * a.stride;
        const std::uint16_t* entries = a.index + row_start;
        std::uint16_t* out_row = a.inner_index + row_start;
        // the row's first four passes in flight together (a pass past the row reads its first
        // entries, which are unused)
        Pair2 raw[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const index_t base = k * kPass + 4 * sub;
            raw[k] = stream_entries(entries + (base < limit ? base : 0));
        }
        const index_t r_next = r + kRowsPerPass;
        const bool live_next = r_next < staged.rows;
        const index_t slot_next = live_next ? staged.slot_of(r_next) : 0;
        const index_t listed_next = live_next ? a.counts[slot_next] : 0;
        const V4 p_next = current_position<false>(a, slot_next);
        index_t kept = 0;
        // one pass of a row: its 64 entries, kept in (entry of the lane, lane) order
        auto prune_pass = [&](const Pair2 quad, const index_t pass) {
            const bool any = pass + 4 * sub < limit;
            const unsigned local[4] = {quad.x & 0xffffu, quad.x >> 16, quad.y & 0xffffu,
                                       quad.y >> 16};
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                const V4 q = halo[local[u]];
                T dx, dy, dz;
                if (fast)
                {
                    dx = c.x - q.x;
                    dy = c.y - q.y;
                    dz = c.z - q.z;
                }
                else
                {
                    dx = minimum_image(c.x, q.x, lx);
                    dy = minimum_image(c.y, q.y, ly);
                    dz = minimum_image(c.z, q.z, lz);
                }
                const T xx = dx * dx;
                const T yy = dy * dy;
                const T zz = dz * dz;
                const bool keep = any && (xx + yy) + zz < inner_squared;
                const std::uint64_t mask = ballot(keep) & group;
                if (keep)
                    out_row[kept + __popcll(mask & below)] = static_cast<std::uint16_t>(local[u]);
                kept += static_cast<index_t>(__popcll(mask));
            }
        };
        // every lane of a row takes the same passes (the rank carries across them)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k * kPass < limit) prune_pass(raw[k], k * kPass);
        for (index_t pass = 4 * kPass; pass < limit; pass += kPass)
        {
            const index_t base = pass + 4 * sub;
            prune_pass(stream_entries(entries + (base < limit ? base : 0)), pass);
        }
        if (live)
        {
            for (index_t q = kept + sub; q < ((kept + 31) & ~31); q += kPruneGroup) out_row[q] = 0;
            if (sub == 0) a.inner_counts[slot] = kept;
            if (sub < 3) a.prune_positions[3LL * slot + sub] = a.positions[3LL * slot + sub];
        }
        r = r_next;
        live = live_next;
        slot = slot_next;
        listed = listed_next;
        p = p_next;
    }
}

// The refresh of every brick's pruned rows, one block per brick; gated on the refresh flag
// unless `forced`.
template <typename T>
__device__ __forceinline__ void md_prune_body(const StepArgs<T>& a, const Launch launch,
                                              int forced)
{
    using V4 = typename Vec4<T>::type;
    if (!forced && !refreshing(a)) return;
    extern __shared__ __align__(16) unsigned char dynamic_shared[];
    V4* halo = reinterpret_cast<V4*>(dynamic_shared);
    BrickShared* sh = brick_shared<T>(dynamic_shared, a.halo_capacity);
    const index_t brick = xcd_brick(static_cast<index_t>(a.bricks), launch);
    if (brick >= a.bricks) return;
    if (launch.block == 0 && threadIdx.x == 0) ++prune_state(a.header)->refreshes;
    // images exact out to the pruned radius
    const StagedBrick<T> staged =
        stage_brick<kBrickThreads<T>, false>(a, brick, halo, sh, a.inner_radius);
    prune_brick_rows(a, staged, halo);
}

template <typename T>
__global__ void __launch_bounds__(kBrickThreads<T>) md_prune_kernel(StepArgs<T> a, int forced)
{
    md_prune_body(a, own_launch(), forced);
}

// Waves per SIMD the force kernel is compiled for: each potential's `waves` (what its fp32
// kernels without the virial reach without spilling). The virial's six more sums and fp64 need
// more registers: those are left to the compiler (a block of kBrickThreads keeps at least 2
// waves per SIMD).
template <typename T, class Pot, bool kVirial>
constexpr int kForceWaves = kVirial ? 1 : Pot::template waves<T>;

// One block per brick: the brick's halo goes to LDS, then G lanes per row walk its rows,
// kBrickThreads / G rows at a time, in slot order along the brick's x-runs, each row up to its
// own count rounded up to 32; the pruned rows when there are any. Each block sums its E, W and
// K in a fixed order into its partial slot, so every run gives the same bits.
//
// Images come from the brick's frame (BrickFrame): x_i - x_j of the placed positions is the
// reference's minimum-image displacement for every pair within the cutoff, and any other pair is
// farther than the cutoff either way, so the mask drops it exactly as the reference does. A brick
// without a frame (a box short against the cutoff) takes the reference's displacement
// (space.py:11-19) per pair. The kick steps skip E unless they also sum W (nothing in a step
// reads E).
template <typename T, class Pot, int G, bool kVirial, bool kKick>
__device__ __forceinline__ void force_body(const StepArgs<T>& a, const Launch launch,
                                           const Pot& pot, T cutoff, T cutoff_squared,
                                           T half_timestep, double* __restrict__ partials,
                                           int refresh = 0)
{
    using V4 = typename Vec4<T>::type;
    constexpr bool kEnergy = !kKick || kVirial;
    extern __shared__ __align__(16) unsigned char dynamic_shared[];
    V4* halo = reinterpret_cast<V4*>(dynamic_shared);
    BrickShared* sh = brick_shared<T>(dynamic_shared, a.halo_capacity);
    // the block's sums by column: term k of block b at partials[k * blocks + b]
    double* out = partials + launch.block;
    const std::int64_t column = launch.blocks;
    const index_t brick = xcd_brick(static_cast<index_t>(a.bricks), launch);
    if (brick >= a.bricks)
    {
        if (threadIdx.x < 8) out[threadIdx.x * column] = 0.0;
        return;
    }
    // the pruned rows when there are any; a full row that overflowed holds only `stride`
    // entries (the status bit reports it)
    const bool pruned = a.inner_index != nullptr;
    const index_t* counts = pruned ? a.inner_counts : a.counts;
    const std::uint16_t* index = pruned ? a.inner_index : a.index;
    // the potential's own LDS after the brick's (the halo staging's barrier covers it)
    unsigned char* potential_shared = dynamic_shared + brick_block_bytes<T>(a.halo_capacity);
    stage_potential(pot, potential_shared);
    // a step that refreshes the pruned list refreshes this brick's rows first (the rows the walk
    // reads are the brick's own), on the same staged halo. A step that may refresh stages images
    // exact out to the pruned radius, so the staging need not wait for the flag (a pair within
    // the cutoff gets the same displacement in either frame); the others out to the cutoff.
    const bool may_refresh = refresh && pruned;
    const bool refreshes = may_refresh && refreshing(a);
    const StagedBrick<T> staged = stage_brick<kBrickThreads<T>, Pot::kTypes>(
        a, brick, halo, sh, may_refresh ? static_cast<T>(a.inner_radius) : cutoff);
    if (refreshes && launch.block == 0 && threadIdx.x == 0) ++prune_state(a.header)->refreshes;
    if (refreshes)
    {
        prune_brick_rows(a, staged, halo);
        // the rows written, seen by every thread of the block (the barrier orders the block's
        // memory; no other block reads these rows)
        __syncthreads();
    }
    const int sub = lane_id() % G;
    const T lx = a.cell[0], ly = a.cell[1], lz = a.cell[2];
    const bool fast = staged.frame.exact;
    const auto ev = bind_potential(pot, potential_shared);
    double energy_sum = 0.0, kinetic = 0.0;
    double w_sum[6] = {0, 0, 0, 0, 0, 0};
    constexpr index_t kRowsPerPass = kBrickThreads<T> / G;
    for (index_t first_row = 0; first_row < staged.rows; first_row += kRowsPerPass)
    {
        const index_t r = first_row + static_cast<index_t>(threadIdx.x) / G;
        const bool live = r < staged.rows;
        const index_t slot = live ? staged.slot_of(r) : 0;
        // a row past the brick's last walks nothing
        V4 c = staged.frame.place(current_position<Pot::kTypes>(a, slot));
        if (!live) c.x = c.y = c.z = -kFarAway<T>;
        const index_t listed_row = live ? counts[slot] : 0;
        const index_t count = listed_row < a.stride ? listed_row : static_cast<index_t>(a.stride);
        const index_t limit = (count + 31) & ~31;
        const std::uint16_t* entries = index + static_cast<std::int64_t>(slot) * a.stride;
        T fx = 0, fy = 0, fz = 0, energy = 0;
        T w[6] = {0, 0, 0, 0, 0, 0};
        if (fast)
            walk_row<G, true, kVirial, kEnergy>(ev, c, halo, entries, limit, lx, ly, lz,
                                                cutoff_squared, fx, fy, fz, energy, w);
        else
            walk_row<G, false, kVirial, kEnergy>(ev, c, halo, entries, limit, lx, ly, lz,
                                                 cutoff_squared, fx, fy, fz, energy, w);
        fx = group_sum<G>(fx);
        fy = group_sum<G>(fy);
        fz = group_sum<G>(fz);
        if (kEnergy) energy_sum += static_cast<double>(energy);
        if (kVirial)
            for (int k = 0; k < 6; ++k) w_sum[k] += static_cast<double>(w[k]);
        if (live && sub == 0)
        {
            T* f = a.forces + 3LL * slot;
            f[0] = fx;
            f[1] = fy;
            f[2] = fz;
            if (kKick)
            {
#pragma clang fp contract(off)
                T* v = a.velocities + 3LL * slot;
                const T inverse_mass = a.inverse_masses[slot];
                T vx = v[0] + half_timestep * fx * inverse_mass;
                T vy = v[1] + half_timestep * fy * inverse_mass;
                T vz = v[2] + half_timestep * fz * inverse_mass;
                if (a.barostat)
                {
                    // NPT: the drag after the kick (_npt.py:176-186)
                    const T drag = static_cast<T>(a.barostat[kDrag]);
                    vx = vx * drag;
                    vy = vy * drag;
                    vz = vz * drag;
                }
                v[0] = vx;
                v[1] = vy;
                v[2] = vz;
                kinetic += 0.5 * static_cast<double>(a.masses[slot]) *
                           (static_cast<double>(vx) * vx + static_cast<double>(vy) * vy +
                            static_cast<double>(vz) * vz);
            }
        }
    }
    double* shared = sh->sums;
    const double e = kEnergy ? block_sum(energy_sum, shared) : 0.0;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (kVirial)
        for (int k = 0; k < 6; ++k) v[k] = block_sum(w_sum[k], shared);
    const double k_sum = kKick ? block_sum(kinetic, shared) : 0.0;
    if (threadIdx.x == 0)
    {
        out[0] = e;
        for (int k = 0; k < 6; ++k) out[(1 + k) * column] = v[k];
        out[7 * column] = k_sum;
    }
}

template <typename T, class Pot, int G, bool kVirial, bool kKick>
__global__ void __launch_bounds__(kBrickThreads<T>, (kForceWaves<T, Pot, kVirial>))
    force_kernel(StepArgs<T> a, Pot pot, T cutoff, T cutoff_squared, T half_timestep,
                 double* __restrict__ partials, int refresh)
{
    force_body<T, Pot, G, kVirial, kKick>(a, own_launch(), pot, cutoff, cutoff_squared,
                                          half_timestep, partials, refresh);
}

// The rows of the list as slots, for readout: out[slot * stride + k] = the slot's k-th listed
// neighbor (in the row's own order), from its brick's halo; the row's entries past its
// neighbors, up to its count rounded up to 32, get -1. One thread per slot.
__global__ void __launch_bounds__(kBlock) md_entries_kernel(
    const index_t* __restrict__ bin_offsets, std::int64_t bins, index_t nx, index_t ny,
    index_t nz, index_t bx, index_t by, index_t bz, const index_t* __restrict__ counts,
    const std::uint16_t* __restrict__ index, std::int64_t stride,
    const index_t* __restrict__ halo_slots, std::int64_t halo_capacity, index_t num_atoms,
    index_t* __restrict__ out)
{
    const std::int64_t slot = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (slot >= num_atoms) return;
    // the slot's sub-bin: the last whose first slot is at most `slot` (and which is not empty)
    std::int64_t lo = 0, hi = bins - 1;
    while (lo < hi)
    {
        const std::int64_t mid = (lo + hi + 1) / 2;
        if (bin_offsets[mid] <= slot)
            lo = mid;
        else
            hi = mid - 1;
    }
    index_t x, y, z;
    bin_coordinates(lo, nx, ny, nz, &x, &y, &z);
    const index_t dims[3] = {nx, ny, nz};
    const index_t size[3] = {bx, by, bz};
    const std::int64_t brick = brick_of(x, y, z, dims, size);
    const index_t count = counts[slot] < stride ? counts[slot] : static_cast<index_t>(stride);
    const index_t limit = (count + 31) & ~31;
    index_t listed = 0;
    for (index_t p = 0; p < limit; ++p)
    {
        const unsigned local = index[slot * stride + p];
        if (local != 0 && local <= halo_capacity)
            out[slot * stride + listed++] = halo_slots[brick * halo_capacity + local - 1];
    }
    while (listed < limit) out[slot * stride + listed++] = -1;
}

// One block of 1024 threads reduces the force kernel's partials.
constexpr int kFinishThreads = 1024;

// E, W and K from the partials (pairs counted twice, so E and W halve); then the chain's
// closing half step, whose velocity factor waits in the thermostat for the next step. The
// partials are by column (term k of block b at partials[k * blocks + b]), and only the columns
// the force launch summed are read: E and W when it summed them (`with_terms`), K when the
// kinetic energy comes from them; the others are zero.
template <typename T>
__device__ __forceinline__ void finish_step_body(const StepArgs<T>& a,
                                                 const double* __restrict__ partials, int blocks,
                                                 int kinetic_from_partials, int with_terms)
{
    __shared__ double column[kFinishThreads];
    __shared__ double totals[8];
    // without a barostat, thread 0 reads the chain and its masses while the partials are summed
    const bool chain_in_registers =
        threadIdx.x == 0 && kinetic_from_partials && a.nvt && !a.barostat;
    double chain[kOpenP + 3], masses[3];
    if (chain_in_registers)
    {
        for (int k = 0; k < kOpenP + 3; ++k) chain[k] = a.thermostat[k];
        for (int k = 0; k < 3; ++k) masses[k] = a.chain_masses[k];
    }
    // without a barostat, the next step's checks, by another wave meanwhile (nothing here reads
    // what they write); with the chain's opening below, the next step's prologue then only
    // writes the potential's coefficients, which a step continuing a run already has. The
    // record of the largest displacement is kept for reading until the next prologue.
    if (threadIdx.x == kWave && kinetic_from_partials && !a.barostat) prepare_checks(a, false);
    // the columns read, first..last: E (0) and W (1-6) with the terms, K (7) with the kinetic
    // energy; `span` threads (whole waves) per column, thread r of a column summing its blocks
    // r, r + span, ... in that order, 16 reads in flight
    constexpr int kInFlight = 16;
    const int first = with_terms ? 0 : 7;
    const int last = kinetic_from_partials ? 7 : 6;
    const int count = last - first + 1;
    const int span = count > 0 ? kFinishThreads / count / kWave * kWave : kFinishThreads;
    const int group = threadIdx.x / span;
    const int r = threadIdx.x % span;
    double sum = 0.0;
    if (group < count)
    {
        const double* values = partials + static_cast<std::int64_t>(first + group) * blocks;
        for (int b = r; b < blocks; b += kInFlight * span)
        {
            double v[kInFlight];
#pragma unroll
            for (int k = 0; k < kInFlight; ++k)
            {
                const int at = b + k * span;
                v[k] = at < blocks ? values[at] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < kInFlight; ++k)
                if (b + k * span < blocks) sum += v[k];
        }
    }
    column[threadIdx.x] = sum;
    if (threadIdx.x < 8) totals[threadIdx.x] = 0.0;
    __syncthreads();
    // then wave c sums column c's span values: lane l those at l, l + 64, ..., then the lanes
    // by a fixed butterfly
    const int wave = threadIdx.x / kWave;
    if (wave < count)
    {
        const int lane = lane_id();
        double s = 0.0;
        for (int k = lane; k < span; k += kWave) s += column[wave * span + k];
        s = wave_sum(s);
        if (lane == 0) totals[first + wave] = s;
    }
    __syncthreads();
    double sums[8];
    for (int k = 0; k < 8; ++k) sums[k] = totals[k];
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
    // NPT closing (_npt.py:188-191): the box momentum's kick with the new K and W, then the
    // particles' chain, then the barostat's
    if (a.barostat && kinetic_from_partials)
    {
#pragma clang fp contract(off)
        a.barostat[kBoxP] = a.barostat[kBoxP] + (0.5 * a.timestep) * box_force(a);
    }
    if (chain_in_registers)
    {
        // the step's opening half step (it waits in the opened slots), the closing half step,
        // and the next step's opening on a copy, in registers; then every slot written once
        double kinetic = sums[7];
        double q[3], p[3];
        for (int k = 0; k < 3; ++k)
        {
            q[k] = chain[kOpenQ + k];
            p[k] = chain[kOpenP + k];
        }
        const double closing =
            chain_half_step(&kinetic, q, p, masses, a.degrees, a.kT, a.timestep);
        double next_kinetic = kinetic;
        double next_q[3] = {q[0], q[1], q[2]}, next_p[3] = {p[0], p[1], p[2]};
        const double opening =
            chain_half_step(&next_kinetic, next_q, next_p, masses, a.degrees, a.kT, a.timestep);
        double* t = a.thermostat;
        for (int k = 0; k < 3; ++k)
        {
            t[kChainQ + k] = q[k];
            t[kChainP + k] = p[k];
            t[kOpenQ + k] = next_q[k];
            t[kOpenP + k] = next_p[k];
        }
        t[kPending] = closing;
        t[kOpening] = opening;
        a.kinetic[0] = kinetic;
    }
    else if (a.nvt && kinetic_from_partials)
    {
        double* t = a.thermostat;
        t[kPending] = chain_half_step(a.kinetic, t + kChainQ, t + kChainP, a.chain_masses,
                                      a.degrees, a.kT, a.timestep);
    }
    if (a.barostat && kinetic_from_partials) box_chain_half_step(a);
}

template <typename T>
__global__ void __launch_bounds__(kFinishThreads) finish_step_kernel(StepArgs<T> a,
                                                                     const double* __restrict__ partials,
                                                                     int blocks,
                                                                     int kinetic_from_partials,
                                                                     int with_terms)
{
    finish_step_body(a, partials, blocks, kinetic_from_partials, with_terms);
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
        if (s.barostat != nullptr)
        {
            a.barostat = s.barostat;
            a.pressure = settings->pressure;
            a.box_mass = settings->box_mass;
            a.alpha = settings->alpha;
            a.cell_out = static_cast<T*>(const_cast<void*>(s.cell));
        }
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
    a.grid_widths = static_cast<T*>(m.grid_widths);
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
    for (int k = 0; k < 3; ++k) a.brick[k] = m.brick[k];
    a.bricks = m.bricks;
    a.halo_capacity = m.halo_capacity;
    a.halo_slots = m.halo_slots;
    a.halo_runs = m.halo_runs;
    a.halo_counts = m.halo_counts;
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
int launch_md_rebuild(const StepArgs<T>& a, double radius, int forced, hipStream_t stream)
{
    const std::int64_t n = a.num_atoms;
    const T r = static_cast<T>(radius);
    // culling keeps every sub-bin a pair within R could sit in, with room for rounding
    const T cull = static_cast<T>(radius * (1.0 + 1e-4) + 1e-6);
    if (n <= kSmallBuild)
    {
        md_build_small_kernel<T><<<1, kListScanThreads, 0, stream>>>(a, r, cull * cull, forced);
        NDN_LAUNCHED("md_build_small");
        return NDN_OK;
    }
    md_assign_kernel<T><<<gated_grid(n), kBlock, 0, stream>>>(a, forced);
    NDN_LAUNCHED("md_assign");
    md_scan_kernel<T><<<static_cast<unsigned>(scan_blocks(a.bins)), kListScanThreads, 0, stream>>>(
        a, forced);
    NDN_LAUNCHED("md_scan");
    const unsigned scatter = gated_grid(n);
    const unsigned halo = gated_grid(a.bricks, kBlock / kWave);
    md_scatter_halo_kernel<T><<<scatter + halo, kBlock, 0, stream>>>(
        a, static_cast<index_t>(scatter), forced);
    NDN_LAUNCHED("md_scatter_halo");
    md_order_kernel<T><<<gated_grid(n), kBlock, 0, stream>>>(a, forced);
    NDN_LAUNCHED("md_order");
    const unsigned unpack = gated_grid(n);
    const unsigned fill = gated_grid(a.bins, kBlock / kWave);
    md_unpack_fill_kernel<T><<<unpack + fill, kBlock, 0, stream>>>(
        a, static_cast<index_t>(unpack), r, cull * cull, forced);
    NDN_LAUNCHED("md_unpack_fill");
    return NDN_OK;
}


// The LDS a brick block's halo may take: what one block may hold, and no more than half a CU's
// LDS (two blocks, 32 waves of fp32, share a CU), less the block's bookkeeping.
std::size_t max_halo_bytes()
{
    int device = 0, per_block = 65536, per_unit = 0;
    if (hipGetDevice(&device) != hipSuccess) return 65536 - kBrickSharedBytes;
    if (hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
        hipSuccess)
        per_block = 65536;
    if (hipDeviceGetAttribute(&per_unit, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                              device) != hipSuccess)
        per_unit = 0;
    std::size_t bytes = static_cast<std::size_t>(std::max(per_block, 65536));
    if (per_unit > 0) bytes = std::min(bytes, static_cast<std::size_t>(per_unit) / 2);
    return bytes - kBrickSharedBytes;
}

// The LDS one block may take (a staged potential may go past max_halo_bytes' half a CU).
std::size_t max_block_bytes()
{
    int device = 0, per_block = 65536;
    if (hipGetDevice(&device) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
        hipSuccess)
        return 65536;
    return static_cast<std::size_t>(std::max(per_block, 65536));
}

// Kernels taking more than 64 KiB of dynamic LDS say so before their launch (a refusal shows
// as the launch's own error).
template <typename Kernel>
void allow_shared(Kernel kernel, std::size_t bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(bytes));
}

// The force launch of one schedule.
template <typename T, class Pot, int G, bool kVirial, bool kKick>
void launch_force_kernel(const StepArgs<T>& a, const Pot& pot, T cutoff, T cutoff_squared,
                         T half_timestep, double* partials, unsigned grid, hipStream_t stream,
                         int refresh)
{
    const std::size_t lds = brick_block_bytes<T>(a.halo_capacity) + potential_lds(pot);
    allow_shared(force_kernel<T, Pot, G, kVirial, kKick>, lds);
    force_kernel<T, Pot, G, kVirial, kKick><<<grid, kBrickThreads<T>, lds, stream>>>(
        a, pot, cutoff, cutoff_squared, half_timestep, partials, refresh);
}

template <typename T, class Pot, int G>
void launch_force_g(const StepArgs<T>& a, const Pot& pot, T cutoff, T cutoff_squared,
                    T half_timestep, bool virial, bool kick, double* partials, unsigned grid,
                    hipStream_t stream, int refresh)
{
    if (virial && kick)
        launch_force_kernel<T, Pot, G, true, true>(a, pot, cutoff, cutoff_squared, half_timestep,
                                                   partials, grid, stream, refresh);
    else if (kick)
        launch_force_kernel<T, Pot, G, false, true>(a, pot, cutoff, cutoff_squared, half_timestep,
                                                    partials, grid, stream, refresh);
    else if (virial)
        launch_force_kernel<T, Pot, G, true, false>(a, pot, cutoff, cutoff_squared, half_timestep,
                                                    partials, grid, stream, refresh);
    else
        launch_force_kernel<T, Pot, G, false, false>(a, pot, cutoff, cutoff_squared,
                                                     half_timestep, partials, grid, stream,
                                                     refresh);
}

template <typename T, class Pot>
void launch_force(const StepArgs<T>& a, const Pot& pot, int group, T cutoff, T cutoff_squared,
                  T half_timestep, bool virial, bool kick, double* partials, unsigned grid,
                  hipStream_t stream, int refresh)
{
    switch (group)
    {
        case 4:
            launch_force_g<T, Pot, 4>(a, pot, cutoff, cutoff_squared, half_timestep, virial, kick,
                                      partials, grid, stream, refresh);
            break;
        case 8:
            launch_force_g<T, Pot, 8>(a, pot, cutoff, cutoff_squared, half_timestep, virial, kick,
                                      partials, grid, stream, refresh);
            break;
        default:
            launch_force_g<T, Pot, 16>(a, pot, cutoff, cutoff_squared, half_timestep, virial,
                                       kick, partials, grid, stream, refresh);
    }
}

int valid_group(int group)
{
    return group == 4 || group == 8 || group == 16;
}

// The force launches' grid: one block per brick, rounded up to a multiple of 8 (one share per
// XCD); each block owns one partial slot.
inline unsigned force_grid(std::int64_t bricks)
{
    return xcd_grid(std::max<std::int64_t>(bricks, 1));
}

// The larger cutoff of a sum of two, else the model's.
inline double model_cutoff(const ndn_potential* pot)
{
    return pot->next != nullptr ? std::max(pot->cutoff, pot->next->cutoff) : pot->cutoff;
}

// Coefficient bytes of the MD step's model: a sum's second table follows the first's.
inline std::int64_t model_coefficient_bytes(const ndn_potential* pot, int element)
{
    ndn_potential first = *pot;
    first.next = nullptr;
    const std::int64_t bytes = coefficient_bytes(&first, element);
    return pot->next != nullptr ? bytes + coefficient_bytes(pot->next, element) : bytes;
}

// Where a sum's second coefficient table starts in dtype T (null without a second model).
template <typename T>
void* second_coefficients(const ndn_potential* pot, void* coefficients)
{
    if (pot->next == nullptr) return nullptr;
    ndn_potential first = *pot;
    first.next = nullptr;
    return static_cast<unsigned char*>(coefficients) +
           coefficient_bytes(&first, static_cast<int>(sizeof(T)));
}

// One model of a sum over its coefficient table (written by the prologue or prepare_model).
template <typename T>
SumPart<T> sum_part(const ndn_potential& pot, void* coefficients)
{
    SumPart<T> part{};
    part.kind = pot.kind;
    part.cutoff_squared = static_cast<T>(pot.cutoff * pot.cutoff);
    if (pot.kind == NDN_LJ)
    {
        part.lj = TypedLJ<T>{static_cast<const LJRow<T>*>(coefficients), pot.num_types};
    }
    else
    {
        const int rows = pot.num_types * (pot.num_types + 1) / 2;
        T* padded = static_cast<T*>(coefficients);
        part.table = TypedTable<T>{Table<T>{
            padded, static_cast<const T*>(pot.start),
            padded + static_cast<std::int64_t>(rows) * (pot.knots + 2), nullptr, pot.num_types,
            static_cast<int>(pot.knots)}};
    }
    return part;
}

// The coefficient tables of the MD step's model from its parameters, outside a step (the
// prologue writes them in a step).
template <typename T>
int prepare_model(const ndn_potential* pot, void* coefficients, hipStream_t stream)
{
    ndn_potential first = *pot;
    first.next = nullptr;
    PotentialView<T> view;
    int prepared = prepare_potential<T>(&first, coefficients, stream, &view);
    if (prepared != NDN_OK || pot->next == nullptr) return prepared;
    return prepare_potential<T>(pot->next, second_coefficients<T>(pot, coefficients), stream,
                              &view);
}

template <typename T>
OneTypeTable<T> bind_one_type_table(const ndn_potential* pot, void* coefficients)
{
    T* padded = static_cast<T*>(coefficients);
    return OneTypeTable<T>{padded, static_cast<const T*>(pot->start), padded + (pot->knots + 2),
                           static_cast<int>(pot->knots)};
}

// Whether a one-type table's segments are staged in LDS: when they fit beside the halo.
template <typename T>
bool stages_table(const ndn_potential* pot, std::int64_t halo_capacity)
{
    if (pot->kind == NDN_LJ || pot->num_types != 1) return false;
    const std::size_t segments = static_cast<std::size_t>(pot->knots - 1) * 2 *
                                 sizeof(typename Vec4<T>::type);
    return brick_block_bytes<T>(halo_capacity) + segments <= max_block_bytes();
}

template <typename T>
int launch_forces(const StepArgs<T>& a, const ndn_potential* pot, void* coefficients, int group,
                  bool virial, bool kick, double* partials, std::int64_t partial_slots,
                  hipStream_t stream, int* grid_out, int refresh = 0)
{
    NDN_REQUIRE(valid_group(group), "group must be 4, 8 or 16");
    const unsigned grid = force_grid(a.bricks);
    NDN_REQUIRE(static_cast<std::int64_t>(grid) <= partial_slots,
                "the force launch has more blocks than partial slots");
    NDN_REQUIRE(a.stride >= 64 && a.stride % 64 == 0,
                "the list stride must be a positive multiple of 64");
    NDN_REQUIRE(a.halo_capacity >= 1 && halo_bytes<T>(a.halo_capacity) <= max_halo_bytes(),
                "the halo capacity must be at least 1 and its positions fit in the LDS a force "
                "block may take (ndn_md_halo_limit)");
    const double reach = model_cutoff(pot);
    const T cutoff = static_cast<T>(reach);
    const T cutoff_squared = static_cast<T>(reach * reach);
    const T half_timestep = static_cast<T>(0.5 * a.timestep);
    const int rows = pot->num_types * (pot->num_types + 1) / 2;
    if (pot->next != nullptr)
    {
        const PairSum<T> sum{sum_part<T>(*pot, coefficients),
                             sum_part<T>(*pot->next, second_coefficients<T>(pot, coefficients))};
        launch_force<T>(a, sum, group, cutoff, cutoff_squared, half_timestep, virial, kick,
                        partials, grid, stream, refresh);
    }
    else if (pot->kind == NDN_LJ && pot->num_types == 1)
    {
        const OneTypeLJ<T> lj{static_cast<const LJRow<T>*>(coefficients)};
        launch_force<T>(a, lj, group, cutoff, cutoff_squared, half_timestep, virial, kick,
                        partials, grid, stream, refresh);
    }
    else if (pot->kind == NDN_LJ)
    {
        const TypedLJ<T> lj{static_cast<const LJRow<T>*>(coefficients), pot->num_types};
        launch_force<T>(a, lj, group, cutoff, cutoff_squared, half_timestep, virial, kick,
                        partials, grid, stream, refresh);
    }
    else if (stages_table<T>(pot, a.halo_capacity))
    {
        launch_force<T>(a, bind_one_type_table<T>(pot, coefficients), group, cutoff,
                        cutoff_squared, half_timestep, virial, kick, partials, grid, stream,
                        refresh);
    }
    else
    {
        T* padded = static_cast<T*>(coefficients);
        const TypedTable<T> table{Table<T>{
            padded, static_cast<const T*>(pot->start),
            padded + static_cast<std::int64_t>(rows) * (pot->knots + 2), nullptr,
            pot->num_types, static_cast<int>(pot->knots)}};
        launch_force<T>(a, table, group, cutoff, cutoff_squared, half_timestep, virial, kick,
                        partials, grid, stream, refresh);
    }
    NDN_LAUNCHED("md_force");
    *grid_out = static_cast<int>(grid);
    return NDN_OK;
}

// The prune kernel over every brick: gated on the refresh flag, or forced.
template <typename T>
int launch_prune(const StepArgs<T>& a, int forced, hipStream_t stream)
{
    if (a.inner_index == nullptr) return NDN_OK;
    NDN_REQUIRE(a.halo_capacity >= 1 && halo_bytes<T>(a.halo_capacity) <= max_halo_bytes(),
                "the halo capacity must be at least 1 and fit the LDS of a block");
    const std::size_t lds = brick_block_bytes<T>(a.halo_capacity);
    allow_shared(md_prune_kernel<T>, lds);
    md_prune_kernel<T><<<force_grid(a.bricks), kBrickThreads<T>, lds, stream>>>(a, forced);
    NDN_LAUNCHED("md_prune");
    return NDN_OK;
}

template <typename T>
int md_step_impl(const ndn_md_system& s, const ndn_md_settings* settings, const ndn_potential* pot,
                 const ndn_list_state* list, const ndn_md_list& m, hipStream_t stream)
{
    StepArgs<T> a = step_args<T>(s, settings, list, m);
    const T timestep = static_cast<T>(settings->timestep);
    const T half_timestep = static_cast<T>(0.5 * settings->timestep);
    const ndn_potential second = pot->next != nullptr ? *pot->next : no_second_model();
    // a step continuing a run without a barostat was opened by the previous step's finish
    if (!settings->continues || s.barostat != nullptr)
    {
        prologue_kernel<T><<<1, kBlock, 0, stream>>>(a, *pot, s.coefficients, second,
                                                     second_coefficients<T>(pot, s.coefficients));
        NDN_LAUNCHED("md_prologue");
    }
    first_half_kernel<T><<<blocks_for(s.num_atoms), kBlock, 0, stream>>>(a, timestep, half_timestep);
    NDN_LAUNCHED("md_first_half");
    const int rebuilt = launch_md_rebuild<T>(a, list->radius, 0, stream);
    if (rebuilt != NDN_OK) return rebuilt;
    // with a pruned list, each force block refreshes its brick's pruned rows when the step
    // refreshes them (a build, or the check's flag), then walks them
    int blocks = 0;
    const int forces = launch_forces<T>(a, pot, s.coefficients, m.group, settings->virial != 0,
                                        true, s.partials, s.partial_slots, stream, &blocks, 1);
    if (forces != NDN_OK) return forces;
    finish_step_kernel<T><<<1, kFinishThreads, 0, stream>>>(a, s.partials, blocks, 1,
                                                            settings->virial != 0);
    NDN_LAUNCHED("md_finish");
    return NDN_OK;
}

template <typename T>
int md_evaluate_impl(const ndn_md_system& s, const ndn_potential* pot, const ndn_md_list& m,
                     int virial, int kick, double timestep, hipStream_t stream, int refresh = 0)
{
    StepArgs<T> a = step_args<T>(s, nullptr, nullptr, m);
    a.timestep = timestep;
    int prepared = prepare_model<T>(pot, s.coefficients, stream);
    if (prepared != NDN_OK) return prepared;
    if (refresh)
    {
        const int pruned = launch_prune<T>(a, 1, stream);
        if (pruned != NDN_OK) return pruned;
    }
    int blocks = 0;
    const int forces = launch_forces<T>(a, pot, s.coefficients, m.group, virial != 0, kick != 0,
                                        s.partials, s.partial_slots, stream, &blocks);
    if (forces != NDN_OK) return forces;
    finish_step_kernel<T><<<1, kFinishThreads, 0, stream>>>(a, s.partials, blocks, 0,
                                                            virial != 0 || kick == 0);
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
                    m->status && m->prepare_dims && m->prepare_widths && m->prepare_total &&
                    m->grid_widths,
                "a list buffer is null");
    NDN_REQUIRE(m->dims[0] >= 1 && m->dims[1] >= 1 && m->dims[2] >= 1, "bad sub-bin dims");
    const std::int64_t bins = static_cast<std::int64_t>(m->dims[0]) * m->dims[1] * m->dims[2];
    NDN_REQUIRE(bins < INT32_MAX, "too many sub-bins");
    NDN_REQUIRE(m->scan_temp_bytes >= scan_bytes(bins + 1) && m->scan_temp_bytes >= 1024,
                "scan_temp is too small");
    NDN_REQUIRE(m->stride == 0 || m->index != nullptr, "a list with rows needs its index");
    NDN_REQUIRE(m->stride >= 0 && m->stride % 64 == 0 && m->stride * s->num_atoms <= (1LL << 40),
                "the stride must be a nonnegative multiple of 64");
    NDN_REQUIRE(valid_group(m->group), "group must be 4, 8 or 16");
    for (int k = 0; k < 3; ++k) NDN_REQUIRE(m->brick[k] >= 1, "brick sizes must be positive");
    NDN_REQUIRE(kLayerBlock % m->brick[2] == 0, "brick[2] must divide 8");
    NDN_REQUIRE(m->brick[1] * m->brick[2] <= kMaxBrickRuns &&
                    (m->brick[1] + 4) * (m->brick[2] + 4) <= kMaxHaloRuns,
                "bricks too large: brick[1] * brick[2] <= 64 and "
                "(brick[1] + 4) * (brick[2] + 4) <= 128");
    std::int64_t bricks = 1;
    for (int k = 0; k < 3; ++k) bricks *= (m->dims[k] + m->brick[k] - 1) / m->brick[k];
    NDN_REQUIRE(m->bricks == bricks, "bricks must be the product of ceil(dims / brick)");
    NDN_REQUIRE(m->halo_slots && m->halo_runs && m->halo_counts, "a halo buffer is null");
    NDN_REQUIRE(m->halo_capacity >= 0 && m->halo_capacity < 65535 &&
                    m->halo_capacity * bricks <= (1LL << 40),
                "the halo capacity must be below 65535");
    NDN_REQUIRE(s->partial_slots >= static_cast<std::int64_t>(force_grid(bricks)),
                "too few partial slots: one per brick, rounded up to a multiple of 8");
    return NDN_OK;
}

// ---- batches of systems ---------------------------------------------------------------------
//
// A batch steps several systems with one launch per kernel of the step. Each launch's blocks are
// the members' blocks end to end, each member with as many as its own step would launch; a
// block finds its member and its block within the member in the launch's block map (one word
// per block: member << 24 | block) and runs the same body with that member's arguments. Every
// member therefore computes the bits it computes alone. Only the force kernel depends on the
// potential: it is launched once for each kind of potential the members bind, over those
// members. The members, their bound potentials, the members' block counts of every part of the
// step and the launches' block maps live in one device table written once by the preparation.

// The parts of a step, each a body with its own block count per member.
enum BatchPart
{
    kPartPrologue,
    kPartFirstHalf,
    kPartAssign,
    kPartScan,
    kPartScatter,
    kPartHalo,
    kPartOrder,
    kPartUnpack,
    kPartFill,
    kPartPrune,
    kPartForce,
    kPartFinish,
    kBatchParts
};

// The kinds of potential a member binds.
enum BatchPotential
{
    kBatchOneTypeLJ,
    kBatchTypedLJ,
    kBatchTable,
    kBatchOneTypeTable,  // staged in LDS, when it fits beside the batch's largest halo
    kBatchPotentials
};

// The launches of a step: one per part but the scatter and the halo listing, which share one,
// the unpack and the fill, which share one, and the forces, one launch per kind of potential
// (kLaunchForce + kind).
enum BatchLaunch
{
    kLaunchPrologue,
    kLaunchFirstHalf,
    kLaunchAssign,
    kLaunchScan,
    kLaunchScatterHalo,
    kLaunchOrder,
    kLaunchUnpackFill,
    kLaunchPrune,
    kLaunchForce,
    kLaunchFinish = kLaunchForce + kBatchPotentials,
    kBatchLaunches
};
static_assert(kBatchLaunches == NDN_MD_BATCH_KERNELS, "the header's launch count");

// The parts a launch runs, in block order within each member (kBatchParts: none).
constexpr int kLaunchParts[kBatchLaunches][2] = {
    {kPartPrologue, kBatchParts}, {kPartFirstHalf, kBatchParts}, {kPartAssign, kBatchParts},
    {kPartScan, kBatchParts},     {kPartScatter, kPartHalo},     {kPartOrder, kBatchParts},
    {kPartUnpack, kPartFill},     {kPartPrune, kBatchParts},
    {kPartForce, kBatchParts},    {kPartForce, kBatchParts},     {kPartForce, kBatchParts},
    {kPartForce, kBatchParts},    {kPartFinish, kBatchParts}};

constexpr int kBatchMemberBits = 24;
constexpr std::int64_t kBatchMaxMembers = 256;
constexpr std::int64_t kBatchMaxBlocks = 1LL << kBatchMemberBits;

// One member of a batch: its step arguments and the scalars its kernels take.
template <typename T>
struct BatchMember
{
    StepArgs<T> a;
    ndn_potential pot;  // the prologue's coefficient tables
    void* coefficients;
    double* partials;
    T radius;        // the list radius of its builds
    T cull_squared;  // the fill's sub-bin culling
    T timestep;
    T half_timestep;
    T cutoff;
    T cutoff_squared;
    index_t force_blocks;
    index_t kind;        // BatchPotential
    index_t with_terms;  // its force launch sums E and W (the virial setting)
};

// A member's bound potential, in the field of its kind.
template <typename T>
struct MemberPotential
{
    OneTypeLJ<T> one_type_lj;
    TypedLJ<T> typed_lj;
    TypedTable<T> table;
    OneTypeTable<T> one_type_table;
};

template <typename T>
__host__ __device__ const OneTypeLJ<T>& pick(const MemberPotential<T>& p, const OneTypeLJ<T>*)
{
    return p.one_type_lj;
}

template <typename T>
__host__ __device__ const TypedLJ<T>& pick(const MemberPotential<T>& p, const TypedLJ<T>*)
{
    return p.typed_lj;
}

template <typename T>
__host__ __device__ const TypedTable<T>& pick(const MemberPotential<T>& p, const TypedTable<T>*)
{
    return p.table;
}

template <typename T>
__host__ __device__ const OneTypeTable<T>& pick(const MemberPotential<T>& p,
                                                const OneTypeTable<T>*)
{
    return p.one_type_table;
}

// The member and its block of this launch's block, with the member's blocks of `part`.
struct BatchBlock
{
    index_t member;
    Launch launch;
};

__device__ __forceinline__ BatchBlock batch_block(const std::uint32_t* __restrict__ map,
                                                  const index_t* __restrict__ blocks, int part)
{
    const std::uint32_t word = map[blockIdx.x];
    const index_t member = static_cast<index_t>(word >> kBatchMemberBits);
    const index_t block = static_cast<index_t>(word & (kBatchMaxBlocks - 1));
    return BatchBlock{member, Launch{block, blocks[member * kBatchParts + part]}};
}

// The arguments of a batched launch: the table's members and block counts, and this launch's map.
template <typename T>
struct BatchTable
{
    const BatchMember<T>* __restrict__ members;
    const index_t* __restrict__ blocks;  // [member][part]
    const std::uint32_t* __restrict__ map;
    // set when some member rebuilds this step (by the first half), cleared by the finish: the
    // gated build launches' blocks read it before their member's arguments
    index_t* any_rebuild;
};

// Whether the gated build launches have work this step: one read of the batch's flag.
template <typename T>
__device__ __forceinline__ bool batch_builds(const BatchTable<T>& t)
{
    return t.any_rebuild[0] != 0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_prologue_kernel(BatchTable<T> t)
{
    const BatchBlock b = batch_block(t.map, t.blocks, kPartPrologue);
    const BatchMember<T>& m = t.members[b.member];
    const StepArgs<T> a = m.a;
    const ndn_potential pot = m.pot;
    prologue_body(a, pot, m.coefficients, no_second_model(), nullptr);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_first_half_kernel(BatchTable<T> t)
{
    const BatchBlock b = batch_block(t.map, t.blocks, kPartFirstHalf);
    const BatchMember<T>& m = t.members[b.member];
    const StepArgs<T> a = m.a;
    // a rebuild decided before this launch (the prologue's checks) raises the batch's flag
    // here; one the first half decides raises it where it is decided
    if (b.launch.block == 0 && threadIdx.x == 0 && a.rebuild[0] != 0) t.any_rebuild[0] = 1;
    first_half_body(a, b.launch, m.timestep, m.half_timestep);
}

// The gated build kernels taking only the arguments and the block.
#define NDN_BATCH_GATED(name, part)                                                             \
    template <typename T>                                                                        \
    __global__ void __launch_bounds__(kBlock) batch_##name##_kernel(BatchTable<T> t)             \
    {                                                                                            \
        if (!batch_builds(t)) return;                                                            \
        const BatchBlock b = batch_block(t.map, t.blocks, part);                                 \
        const StepArgs<T> a = t.members[b.member].a;                                             \
        name##_body(a, b.launch, 0);                                                             \
    }
NDN_BATCH_GATED(md_assign, kPartAssign)
NDN_BATCH_GATED(md_order, kPartOrder)
#undef NDN_BATCH_GATED

template <typename T>
__global__ void __launch_bounds__(kListScanThreads) batch_scan_kernel(BatchTable<T> t)
{
    if (!batch_builds(t)) return;
    const BatchBlock b = batch_block(t.map, t.blocks, kPartScan);
    const StepArgs<T> a = t.members[b.member].a;
    if (!rebuilding(a.rebuild, a.status)) return;
    md_scan_body(a, b.launch);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_scatter_halo_kernel(BatchTable<T> t)
{
    if (!batch_builds(t)) return;
    const BatchBlock b = batch_block(t.map, t.blocks, kPartScatter);
    const StepArgs<T> a = t.members[b.member].a;
    const index_t scatter = b.launch.blocks;
    const index_t halo = t.blocks[b.member * kBatchParts + kPartHalo];
    md_scatter_halo_body(a, Launch{b.launch.block, scatter + halo}, scatter, 0);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_unpack_fill_kernel(BatchTable<T> t)
{
    if (!batch_builds(t)) return;
    const BatchBlock b = batch_block(t.map, t.blocks, kPartUnpack);
    const BatchMember<T>& m = t.members[b.member];
    const StepArgs<T> a = m.a;
    const index_t unpack = b.launch.blocks;
    const index_t fill = t.blocks[b.member * kBatchParts + kPartFill];
    md_unpack_fill_body(a, Launch{b.launch.block, unpack + fill}, unpack, m.radius,
                        m.cull_squared, 0);
}

template <typename T>
__global__ void __launch_bounds__(kBrickThreads<T>) batch_prune_kernel(BatchTable<T> t)
{
    const BatchBlock b = batch_block(t.map, t.blocks, kPartPrune);
    const StepArgs<T> a = t.members[b.member].a;
    md_prune_body(a, b.launch, 0);
}

// The forces of the members binding `Pot` (the launch's map holds only theirs).
template <typename T, class Pot, int G, bool kVirial>
__global__ void __launch_bounds__(kBrickThreads<T>, (kForceWaves<T, Pot, kVirial>))
    batch_force_kernel(BatchTable<T> t, const MemberPotential<T>* __restrict__ pots)
{
    const BatchBlock b = batch_block(t.map, t.blocks, kPartForce);
    const BatchMember<T>& m = t.members[b.member];
    const StepArgs<T> a = m.a;
    const Pot pot = pick(pots[b.member], static_cast<const Pot*>(nullptr));
    // each member's force blocks refresh its bricks' pruned rows when the member refreshes
    force_body<T, Pot, G, kVirial, true>(a, b.launch, pot, m.cutoff, m.cutoff_squared,
                                         m.half_timestep, m.partials, 1);
}

template <typename T>
__global__ void __launch_bounds__(kFinishThreads) batch_finish_kernel(BatchTable<T> t)
{
    const BatchBlock b = batch_block(t.map, t.blocks, kPartFinish);
    const BatchMember<T>& m = t.members[b.member];
    const StepArgs<T> a = m.a;
    // every gated build launch of this step has run: the flag starts the next step clear
    if (b.member == 0 && threadIdx.x == 0) t.any_rebuild[0] = 0;
    finish_step_body(a, m.partials, m.force_blocks, 1, m.with_terms);
}

int batch_potential(const ndn_potential* pot)
{
    if (pot->kind == NDN_LJ) return pot->num_types == 1 ? kBatchOneTypeLJ : kBatchTypedLJ;
    return pot->num_types == 1 ? kBatchOneTypeTable : kBatchTable;
}

template <typename T>
TypedTable<T> bind_table(const ndn_potential* pot, void* coefficients)
{
    const int rows = pot->num_types * (pot->num_types + 1) / 2;
    T* padded = static_cast<T*>(coefficients);
    return TypedTable<T>{Table<T>{padded, static_cast<const T*>(pot->start),
                                  padded + static_cast<std::int64_t>(rows) * (pot->knots + 2),
                                  nullptr, pot->num_types, static_cast<int>(pot->knots)}};
}

template <typename T>
MemberPotential<T> bind_member(const ndn_potential* pot, void* coefficients, int kind)
{
    MemberPotential<T> p{};
    switch (kind)
    {
        case kBatchOneTypeLJ:
            p.one_type_lj = OneTypeLJ<T>{static_cast<const LJRow<T>*>(coefficients)};
            break;
        case kBatchTypedLJ:
            p.typed_lj = TypedLJ<T>{static_cast<const LJRow<T>*>(coefficients), pot->num_types};
            break;
        case kBatchOneTypeTable:
            p.one_type_table = bind_one_type_table<T>(pot, coefficients);
            break;
        default:
            p.table = bind_table<T>(pot, coefficients);
    }
    return p;
}

// The blocks a member's own step launches for each part (md_step_impl's grids).
template <typename T>
void member_blocks(const StepArgs<T>& a, index_t out[kBatchParts])
{
    const std::int64_t n = a.num_atoms;
    out[kPartPrologue] = 1;
    out[kPartFirstHalf] = static_cast<index_t>(blocks_for(n));
    out[kPartAssign] = static_cast<index_t>(gated_grid(n));
    out[kPartScan] = static_cast<index_t>(scan_blocks(a.bins));
    out[kPartScatter] = static_cast<index_t>(gated_grid(n));
    out[kPartHalo] = static_cast<index_t>(gated_grid(a.bricks, kBlock / kWave));
    out[kPartOrder] = static_cast<index_t>(gated_grid(n));
    out[kPartUnpack] = static_cast<index_t>(gated_grid(n));
    out[kPartFill] = static_cast<index_t>(gated_grid(a.bins, kBlock / kWave));
    out[kPartPrune] = a.inner_index ? static_cast<index_t>(force_grid(a.bricks)) : 0;
    out[kPartForce] = static_cast<index_t>(force_grid(a.bricks));
    out[kPartFinish] = 1;
}

// A member's blocks in a launch (a force launch holds only the members of its kind).
std::int64_t launch_blocks(const index_t* own, int kind, int launch)
{
    if (launch >= kLaunchForce && launch < kLaunchFinish && launch - kLaunchForce != kind) return 0;
    std::int64_t blocks = 0;
    for (const int part : kLaunchParts[launch])
        if (part != kBatchParts) blocks += own[part];
    return blocks;
}

std::int64_t align_table(std::int64_t bytes)
{
    return (bytes + 255) / 256 * 256;
}

// Lays out and writes the table (when `batch->table` is set), and fills the batch's launch
// fields. With no table it only sizes it. `kinds` holds each member's BatchPotential.
template <typename T>
int prepare_batch(ndn_md_batch* batch, const ndn_md_member* members, const int* kinds, int count,
                  std::int64_t staged_bytes, hipStream_t stream)
{
    std::vector<BatchMember<T>> hosts(static_cast<std::size_t>(count));
    std::vector<MemberPotential<T>> pots(static_cast<std::size_t>(count));
    std::vector<index_t> blocks(static_cast<std::size_t>(count) * kBatchParts);
    std::int64_t grids[kBatchLaunches] = {};
    // the largest halo of each launch's members (the brick kernels' LDS)
    std::int64_t halos[kBatchLaunches] = {};
    for (int i = 0; i < count; ++i)
    {
        const ndn_md_member& mb = members[i];
        BatchMember<T>& h = hosts[i];
        h.a = step_args<T>(*mb.system, mb.settings, mb.list, *mb.md_list);
        h.pot = *mb.potential;
        h.coefficients = mb.system->coefficients;
        h.partials = mb.system->partials;
        h.radius = static_cast<T>(mb.list->radius);
        h.cull_squared = static_cast<T>(mb.list->radius * (1.0 + 1e-4) + 1e-6);
        h.cull_squared = h.cull_squared * h.cull_squared;
        h.timestep = static_cast<T>(mb.settings->timestep);
        h.half_timestep = static_cast<T>(0.5 * mb.settings->timestep);
        h.cutoff = static_cast<T>(mb.potential->cutoff);
        h.cutoff_squared = static_cast<T>(mb.potential->cutoff * mb.potential->cutoff);
        h.kind = kinds[i];
        h.with_terms = mb.settings->virial != 0;
        pots[i] = bind_member<T>(mb.potential, mb.system->coefficients, kinds[i]);
        index_t* own = blocks.data() + static_cast<std::size_t>(i) * kBatchParts;
        member_blocks<T>(h.a, own);
        h.force_blocks = own[kPartForce];
        for (int k = 0; k < kBatchLaunches; ++k)
        {
            const std::int64_t launched = launch_blocks(own, kinds[i], k);
            NDN_REQUIRE(launched < kBatchMaxBlocks, "a member launches too many blocks for a batch");
            grids[k] += launched;
            if (launched > 0) halos[k] = std::max<std::int64_t>(halos[k], h.a.halo_capacity);
        }
    }
    std::int64_t offset = align_table(sizeof(BatchMember<T>) * count);
    const std::int64_t pots_at = offset;
    offset += align_table(sizeof(MemberPotential<T>) * count);
    const std::int64_t blocks_at = offset;
    offset += align_table(sizeof(index_t) * count * kBatchParts);
    for (int k = 0; k < kBatchLaunches; ++k)
    {
        NDN_REQUIRE(grids[k] <= INT32_MAX, "a batched launch would exceed 2**31 blocks");
        batch->offsets[k] = offset;
        offset += align_table(4 * std::max<std::int64_t>(grids[k], 1));
    }
    const std::int64_t flag_at = offset;  // batch_flag_offset's answer
    offset += align_table(sizeof(index_t));
    const std::int64_t needed = offset;
    batch->members_offset = 0;
    batch->potentials_offset = pots_at;
    batch->blocks_offset = blocks_at;
    for (int k = 0; k < kBatchLaunches; ++k)
    {
        batch->grids[k] = grids[k];
        batch->lds[k] = 0;
    }
    batch->lds[kLaunchPrune] = static_cast<std::int64_t>(brick_block_bytes<T>(halos[kLaunchPrune]));
    for (int kind = 0; kind < kBatchPotentials; ++kind)
    {
        const int launch = kLaunchForce + kind;
        batch->lds[launch] = static_cast<std::int64_t>(brick_block_bytes<T>(halos[launch])) +
                             (kind == kBatchOneTypeTable ? staged_bytes : 0);
    }
    if (batch->table == nullptr)
    {
        batch->table_bytes = needed;
        return NDN_OK;
    }
    NDN_REQUIRE(batch->table_bytes >= needed, "the batch table is too small");
    for (int i = 0; i < count; ++i)
        hosts[i].a.batch_rebuild =
            reinterpret_cast<index_t*>(static_cast<std::uint8_t*>(batch->table) + flag_at);
    std::vector<std::uint8_t> image(static_cast<std::size_t>(needed), 0);
    std::memcpy(image.data(), hosts.data(), sizeof(BatchMember<T>) * count);
    std::memcpy(image.data() + pots_at, pots.data(), sizeof(MemberPotential<T>) * count);
    std::memcpy(image.data() + blocks_at, blocks.data(), sizeof(index_t) * count * kBatchParts);
    for (int k = 0; k < kBatchLaunches; ++k)
    {
        auto* map = reinterpret_cast<std::uint32_t*>(image.data() + batch->offsets[k]);
        std::int64_t at = 0;
        for (int i = 0; i < count; ++i)
        {
            const std::int64_t launched =
                launch_blocks(blocks.data() + static_cast<std::size_t>(i) * kBatchParts, kinds[i], k);
            for (std::int64_t block = 0; block < launched; ++block)
                map[at++] = (static_cast<std::uint32_t>(i) << kBatchMemberBits) |
                            static_cast<std::uint32_t>(block);
        }
        if (at != grids[k])
        {
            set_error("md_batch_prepare: a block map does not cover its launch");
            return NDN_ERR_ARGUMENT;
        }
    }
    hipError_t error = hipMemcpyAsync(batch->table, image.data(), static_cast<std::size_t>(needed),
                                      hipMemcpyHostToDevice, stream);
    // the host image dies with this call
    if (error == hipSuccess) error = hipStreamSynchronize(stream);
    if (error != hipSuccess)
    {
        set_error(std::string("md_batch_prepare: ") + hipGetErrorString(error));
        return NDN_ERR_HIP;
    }
    return NDN_OK;
}

// The batch's rebuild flag: the table's slot after the last launch's block map (prepare_batch's
// layout).
std::int64_t batch_flag_offset(const ndn_md_batch& b)
{
    const int last = kBatchLaunches - 1;
    return b.offsets[last] + align_table(4 * std::max<std::int64_t>(b.grids[last], 1));
}

template <typename T>
BatchTable<T> batch_table(const ndn_md_batch& b, int launch)
{
    auto* base = static_cast<std::uint8_t*>(b.table);
    return BatchTable<T>{reinterpret_cast<const BatchMember<T>*>(base + b.members_offset),
                         reinterpret_cast<const index_t*>(base + b.blocks_offset),
                         reinterpret_cast<const std::uint32_t*>(base + b.offsets[launch]),
                         reinterpret_cast<index_t*>(base + batch_flag_offset(b))};
}

template <typename T, class Pot, int G, bool kVirial>
void launch_batch_force(const ndn_md_batch& b, int launch, hipStream_t stream)
{
    const std::size_t lds = static_cast<std::size_t>(b.lds[launch]);
    allow_shared(batch_force_kernel<T, Pot, G, kVirial>, lds);
    const auto* pots = reinterpret_cast<const MemberPotential<T>*>(
        static_cast<std::uint8_t*>(b.table) + b.potentials_offset);
    batch_force_kernel<T, Pot, G, kVirial>
        <<<static_cast<unsigned>(b.grids[launch]), kBrickThreads<T>, lds, stream>>>(
            batch_table<T>(b, launch), pots);
}

template <typename T, class Pot>
void launch_batch_force_of(const ndn_md_batch& b, int kind, hipStream_t stream)
{
    const int launch = kLaunchForce + kind;
    const bool virial = b.virial != 0;
    switch (b.group)
    {
        case 4:
            virial ? launch_batch_force<T, Pot, 4, true>(b, launch, stream)
                   : launch_batch_force<T, Pot, 4, false>(b, launch, stream);
            break;
        case 8:
            virial ? launch_batch_force<T, Pot, 8, true>(b, launch, stream)
                   : launch_batch_force<T, Pot, 8, false>(b, launch, stream);
            break;
        default:
            virial ? launch_batch_force<T, Pot, 16, true>(b, launch, stream)
                   : launch_batch_force<T, Pot, 16, false>(b, launch, stream);
    }
}

// One step of every member, in md_step_impl's order.
template <typename T>
int batch_step_impl(const ndn_md_batch& b, hipStream_t stream)
{
    auto grid = [&](int launch) { return static_cast<unsigned>(b.grids[launch]); };
    // a step continuing a run was opened by the previous step's finish, as in md_step_impl
    if (!b.continues)
    {
        batch_prologue_kernel<T><<<grid(kLaunchPrologue), kBlock, 0, stream>>>(
            batch_table<T>(b, kLaunchPrologue));
        NDN_LAUNCHED("md_batch_prologue");
    }
    batch_first_half_kernel<T><<<grid(kLaunchFirstHalf), kBlock, 0, stream>>>(
        batch_table<T>(b, kLaunchFirstHalf));
    NDN_LAUNCHED("md_batch_first_half");
    batch_md_assign_kernel<T><<<grid(kLaunchAssign), kBlock, 0, stream>>>(
        batch_table<T>(b, kLaunchAssign));
    NDN_LAUNCHED("md_batch_assign");
    batch_scan_kernel<T><<<grid(kLaunchScan), kListScanThreads, 0, stream>>>(
        batch_table<T>(b, kLaunchScan));
    NDN_LAUNCHED("md_batch_scan");
    batch_scatter_halo_kernel<T><<<grid(kLaunchScatterHalo), kBlock, 0, stream>>>(
        batch_table<T>(b, kLaunchScatterHalo));
    NDN_LAUNCHED("md_batch_scatter_halo");
    batch_md_order_kernel<T><<<grid(kLaunchOrder), kBlock, 0, stream>>>(
        batch_table<T>(b, kLaunchOrder));
    NDN_LAUNCHED("md_batch_order");
    batch_unpack_fill_kernel<T><<<grid(kLaunchUnpackFill), kBlock, 0, stream>>>(
        batch_table<T>(b, kLaunchUnpackFill));
    NDN_LAUNCHED("md_batch_unpack_fill");
    // the pruned rows are refreshed by the force launches (force_body's refresh)
    for (int kind = 0; kind < kBatchPotentials; ++kind)
    {
        if (b.grids[kLaunchForce + kind] == 0) continue;
        switch (kind)
        {
            case kBatchOneTypeLJ:
                launch_batch_force_of<T, OneTypeLJ<T>>(b, kind, stream);
                break;
            case kBatchTypedLJ:
                launch_batch_force_of<T, TypedLJ<T>>(b, kind, stream);
                break;
            case kBatchOneTypeTable:
                launch_batch_force_of<T, OneTypeTable<T>>(b, kind, stream);
                break;
            default:
                launch_batch_force_of<T, TypedTable<T>>(b, kind, stream);
        }
        NDN_LAUNCHED("md_batch_force");
    }
    batch_finish_kernel<T><<<grid(kLaunchFinish), kFinishThreads, 0, stream>>>(
        batch_table<T>(b, kLaunchFinish));
    NDN_LAUNCHED("md_batch_finish");
    return NDN_OK;
}

// The MD entry points' model: one model, or a sum of two whose parts share the atom types;
// `element` bytes per value, checked against the system's coefficient buffer.
int check_md_potential(const ndn_md_system* system, const ndn_potential* pot, int element)
{
    NDN_REQUIRE(pot != nullptr, "the potential must not be null");
    ndn_potential first = *pot;
    first.next = nullptr;
    int checked = check_potential(&first, system->num_atoms);
    if (checked != NDN_OK) return checked;
    if (pot->next != nullptr)
    {
        checked = check_potential(pot->next, system->num_atoms);
        if (checked != NDN_OK) return checked;
        NDN_REQUIRE(pot->next->num_types == pot->num_types &&
                        pot->next->atom_types == pot->atom_types,
                    "the two models of a sum share num_types and atom_types");
    }
    NDN_REQUIRE(system->coefficient_bytes >= model_coefficient_bytes(pot, element),
                "the coefficient buffer is smaller than ndn_md_workspace asks");
    return NDN_OK;
}

// What ndn_md_step checks of one system.
int check_step(int dtype, const ndn_md_system* system, const ndn_md_settings* settings,
               const ndn_potential* pot, const ndn_list_state* list, const ndn_md_list* md_list)
{
    int checked = check_system(system, md_list);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    checked = check_potential(pot, system->num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(settings && list, "settings and list must not be null");
    NDN_REQUIRE(settings->timestep > 0, "the timestep must be positive");
    NDN_REQUIRE(!settings->thermostat || system->chain_masses, "NVT needs chain masses");
    NDN_REQUIRE(list->cutoff == pot->cutoff, "the list's cutoff must be the potential's");
    NDN_REQUIRE(md_list->stride > 0, "the list has no rows yet");
    NDN_REQUIRE(pot->num_types == 1 || md_list->types != nullptr,
                "several atom types need the slot-ordered types");
    NDN_REQUIRE(list->rebuild != nullptr, "the step needs the rebuild flag");
    const std::size_t halo = dtype == NDN_F32 ? halo_bytes<float>(md_list->halo_capacity)
                                              : halo_bytes<double>(md_list->halo_capacity);
    NDN_REQUIRE(md_list->halo_capacity >= 1 && halo <= max_halo_bytes(),
                "the halo capacity must be at least 1 and fit the LDS of a block");
    NDN_REQUIRE(md_list->stride >= 64, "the list stride must be a positive multiple of 64");
    return NDN_OK;
}

}  // namespace ndn

using namespace ndn;

extern "C" {

int ndn_md_workspace(int dtype, const ndn_potential* pot, int64_t num_atoms,
                     int64_t* partial_slots, int64_t* coefficient_bytes_out)
{
    NDN_REQUIRE(pot != nullptr, "the potential must not be null");
    ndn_potential first = *pot;
    first.next = nullptr;
    int checked = check_potential(&first, num_atoms);
    if (checked != NDN_OK) return checked;
    if (pot->next != nullptr)
    {
        checked = check_potential(pot->next, num_atoms);
        if (checked != NDN_OK) return checked;
    }
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    NDN_REQUIRE(num_atoms >= 1 && partial_slots && coefficient_bytes_out, "bad arguments");
    // at most one brick per atom holds atoms; the list may ask for fewer (one per brick,
    // rounded up to a multiple of 8)
    *partial_slots = std::max<std::int64_t>(xcd_grid(num_atoms), 8);
    *coefficient_bytes_out = model_coefficient_bytes(pot, dtype == NDN_F32 ? 4 : 8);
    return NDN_OK;
}

int64_t ndn_md_halo_limit(int dtype)
{
    const std::size_t point = dtype == NDN_F32 ? sizeof(float4) : sizeof(double4);
    return static_cast<int64_t>(max_halo_bytes() / point) - 1;
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
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    checked = check_md_potential(system, pot, dtype == NDN_F32 ? 4 : 8);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(settings && list, "settings and list must not be null");
    NDN_REQUIRE(settings->timestep > 0, "the timestep must be positive");
    NDN_REQUIRE(!settings->thermostat || system->chain_masses, "NVT needs chain masses");
    NDN_REQUIRE(list->cutoff == model_cutoff(pot),
                "the list's cutoff must be the potential's (a sum's larger one)");
    NDN_REQUIRE(md_list->stride > 0, "the list has no rows yet");
    NDN_REQUIRE(pot->num_types == 1 || md_list->types != nullptr,
                "several atom types need the slot-ordered types");
    if (system->barostat != nullptr)
    {
        NDN_REQUIRE(settings->thermostat && system->chain_masses,
                    "NPT thermostats the particles: it needs the chain and its masses");
        NDN_REQUIRE(settings->virial && system->virial, "NPT needs the virial every step");
        NDN_REQUIRE(md_list->inner_index == nullptr, "NPT keeps no pruned list");
        NDN_REQUIRE(settings->box_mass > 0 && settings->alpha > 0,
                    "NPT needs a positive box mass and alpha");
    }
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
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    checked = check_md_potential(system, pot, dtype == NDN_F32 ? 4 : 8);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(md_list->stride > 0, "the list has no rows yet");
    NDN_REQUIRE(pot->num_types == 1 || md_list->types != nullptr,
                "several atom types need the slot-ordered types");
    if (dtype == NDN_F32)
        return md_evaluate_impl<float>(*system, pot, *md_list, virial, kick, timestep, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return md_evaluate_impl<double>(*system, pot, *md_list, virial, kick, timestep, stream);
}

int ndn_md_refresh(int dtype, const ndn_md_system* system, const ndn_potential* pot,
                   const ndn_md_list* md_list, ndn_stream_t stream)
{
    int checked = check_system(system, md_list);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    checked = check_md_potential(system, pot, dtype == NDN_F32 ? 4 : 8);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(md_list->stride > 0, "the list has no rows yet");
    NDN_REQUIRE(md_list->inner_index && md_list->inner_counts && md_list->prune_positions,
                "the list has no pruned list");
    NDN_REQUIRE(pot->num_types == 1 || md_list->types != nullptr,
                "several atom types need the slot-ordered types");
    if (dtype == NDN_F32)
        return md_evaluate_impl<float>(*system, pot, *md_list, 0, 0, 0.0, stream, 1);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return md_evaluate_impl<double>(*system, pot, *md_list, 0, 0, 0.0, stream, 1);
}

int ndn_md_list_entries(const ndn_md_list* md_list, int64_t num_atoms, int pruned, int32_t* out,
                        ndn_stream_t stream)
{
    NDN_REQUIRE(md_list && out && num_atoms >= 1 && num_atoms <= INT32_MAX, "bad arguments");
    NDN_REQUIRE(md_list->stride > 0 && md_list->halo_capacity > 0, "the list has no rows yet");
    const std::uint16_t* index = pruned ? md_list->inner_index : md_list->index;
    const int32_t* counts = pruned ? md_list->inner_counts : md_list->counts;
    NDN_REQUIRE(index && counts && md_list->bin_offsets && md_list->halo_slots,
                "a list buffer is null");
    const std::int64_t bins =
        static_cast<std::int64_t>(md_list->dims[0]) * md_list->dims[1] * md_list->dims[2];
    md_entries_kernel<<<blocks_for(num_atoms), kBlock, 0, stream>>>(
        md_list->bin_offsets, bins, md_list->dims[0], md_list->dims[1], md_list->dims[2],
        md_list->brick[0], md_list->brick[1], md_list->brick[2], counts, index, md_list->stride,
        md_list->halo_slots, md_list->halo_capacity, static_cast<index_t>(num_atoms), out);
    NDN_LAUNCHED("md_list_entries");
    return NDN_OK;
}

int ndn_md_force(int dtype, const ndn_md_system* system, const ndn_potential* pot,
                 const ndn_md_list* md_list, int virial, int kick, double timestep,
                 ndn_stream_t stream)
{
    int checked = check_system(system, md_list);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    checked = check_md_potential(system, pot, dtype == NDN_F32 ? 4 : 8);
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

int ndn_md_batch_prepare(ndn_md_batch* batch, const ndn_md_member* members, int32_t count,
                         ndn_stream_t stream)
{
    NDN_REQUIRE(batch != nullptr && members != nullptr, "batch and members must not be null");
    NDN_REQUIRE(count >= 1 && count <= kBatchMaxMembers, "a batch holds 1 to 256 members");
    const int dtype = batch->dtype;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    std::int64_t halo = 1, knots = 2;
    for (int i = 0; i < count; ++i)
    {
        const ndn_md_member& m = members[i];
        NDN_REQUIRE(m.system && m.settings && m.potential && m.list && m.md_list,
                    "a member's structure is null");
        NDN_REQUIRE(m.system->barostat == nullptr, "a batch steps NVE and NVT members only");
        const int checked = check_step(dtype, m.system, m.settings, m.potential, m.list, m.md_list);
        if (checked != NDN_OK) return checked;
        NDN_REQUIRE(m.md_list->group == members[0].md_list->group,
                    "the members must share the group");
        NDN_REQUIRE((m.settings->virial != 0) == (members[0].settings->virial != 0),
                    "the members must share the virial setting");
        halo = std::max<std::int64_t>(halo, m.md_list->halo_capacity);
        if (batch_potential(m.potential) == kBatchOneTypeTable)
            knots = std::max<std::int64_t>(knots, m.potential->knots);
    }
    // one-type tables are staged when the largest fits beside the batch's largest halo, else
    // gathered like typed tables
    const std::size_t point = dtype == NDN_F32 ? sizeof(float4) : sizeof(double4);
    const std::size_t staged = static_cast<std::size_t>(knots - 1) * 2 * point;
    const std::size_t brick =
        dtype == NDN_F32 ? brick_block_bytes<float>(halo) : brick_block_bytes<double>(halo);
    const bool stage = brick + staged <= max_block_bytes();
    std::vector<int> kinds(static_cast<std::size_t>(count));
    batch->kinds = 0;
    for (int i = 0; i < count; ++i)
    {
        int kind = batch_potential(members[i].potential);
        if (kind == kBatchOneTypeTable && !stage) kind = kBatchTable;
        kinds[i] = kind;
        batch->kinds |= 1 << kind;
    }
    batch->count = count;
    batch->group = members[0].md_list->group;
    batch->virial = members[0].settings->virial != 0;
    batch->continues = 0;
    const std::int64_t staged_bytes = static_cast<std::int64_t>(staged);
    if (dtype == NDN_F32)
        return prepare_batch<float>(batch, members, kinds.data(), count, staged_bytes, stream);
    return prepare_batch<double>(batch, members, kinds.data(), count, staged_bytes, stream);
}

int ndn_md_batch_step(const ndn_md_batch* batch, ndn_stream_t stream)
{
    NDN_REQUIRE(batch != nullptr && batch->table != nullptr && batch->count >= 1,
                "the batch is not prepared");
    std::int64_t force = 0;
    for (int kind = 0; kind < kBatchPotentials; ++kind) force += batch->grids[kLaunchForce + kind];
    NDN_REQUIRE(force >= 8 && batch->grids[kLaunchFinish] == batch->count,
                "the batch is not prepared");
    if (batch->dtype == NDN_F32) return batch_step_impl<float>(*batch, stream);
    NDN_REQUIRE(batch->dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return batch_step_impl<double>(*batch, stream);
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
