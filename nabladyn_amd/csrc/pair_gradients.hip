// The derivatives a differentiable model needs beyond E, F = -dE/dx and W at zero strain: the
// same quantities at a nonzero strain, and dE/dparameters (LJ sigma and epsilon per type-pair
// row, a table's knots). They back Potential.forward's autograd function
// (potential/_base.py:82-106 and its callers' torch.func use, quantity.py:93-117,
// training/_loss.py:94-102).
//
// Strain follows space.strained (space.py:51-64): the minimum-image vector d of a pair is
// deformed to d' = (I + e) d, each component ((d_a + d_x e_a0) + d_y e_a1) + d_z e_a2 rounded
// term by term, and the pair's energy is u(|d'|). Then, with fs = -u'(r) / r at r = |d'|:
//   force on the first atom   fs (I + e)^T d'        (the second atom takes its negative)
//   W_ab = -dE/de_ab          sum over pairs of fs d'_a d_b      (not symmetric when e != 0)
// At e = 0 these are the unstrained kernels' results.
//
// Parameter derivatives, per pair within the cutoff, summed per parameter row:
//   LJ   u = 4 eps (p^2 - p) - 4 eps (q^2 - q), p = (sigma^2 / r^2)^3, q = (sigma^2 / rc^2)^3
//        du/deps   = 4 (p^2 - p) - 4 (q^2 - q)
//        du/dsigma = 24 eps / sigma ((2 p^2 - p) - (2 q^2 - q))
//   table  u is linear in the knots: the Catmull-Rom weights of the segment's four padded
//        knots, mapped back to the real ones (the virtual first knot 3u0 - 3u1 + u2, the virtual
//        last 3u_(K-1) - 3u_(K-2) + u_(K-3), and the last real knot read as 0, so its
//        derivative is 0; potential/_tabulated/_potential.py:59-91).
// Sums are in fp64: per block in LDS when the rows fit, else straight to memory, with atomics
// (so the last bits depend on the order blocks finish).
//
// The vector-Jacobian product for fitting (training/_loss.py:94-125, whose loss reads E, F and
// W): with upstream gE, gF [N, 3] and gW [3, 3], a pair of vector d = x_i - x_j contributes
//   gE du/dtheta + ((gF_i - gF_j) . d + d^T gW d) dfs/dtheta,
// since F_i = sum_j fs d_ij and W = sum fs d (x) d. dfs/dtheta:
//   LJ   fs = 24 eps (2 p^2 - p) / r^2:  dfs/deps = 24 (2 p^2 - p) / r^2,
//        dfs/dsigma = 24 eps (24 p^2 - 6 p) / (sigma r^2)
//   table fs = -(du/dt) / (h r), du/dt linear in the knots with the Catmull-Rom slope weights.

#include <algorithm>
#include <climits>

#include "common.hpp"
#include "pair_forces.hpp"
#include "potentials.hpp"

namespace ndn {

template <typename T>
struct Strain
{
    T e[9];  // row-major: e[3 a + b]
};

// The strain [3, 3] from device memory (row-major), or zero without one.
template <typename T>
__device__ __forceinline__ Strain<T> read_strain(const T* __restrict__ strain)
{
    Strain<T> s;
    for (int k = 0; k < 9; ++k) s.e[k] = strain != nullptr ? strain[k] : T(0);
    return s;
}

// d' = (I + e) d with the reference's term order.
template <typename T>
__device__ __forceinline__ void strain_vector(const Strain<T>& s, const T d[3], T out[3])
{
#pragma clang fp contract(off)
    for (int a = 0; a < 3; ++a)
    {
        const T first = d[a] + d[0] * s.e[3 * a];
        const T second = first + d[1] * s.e[3 * a + 1];
        out[a] = second + d[2] * s.e[3 * a + 2];
    }
}

// The upstream gradients of a vector-Jacobian product; all null: the plain dE/dtheta.
template <typename T>
struct Upstream
{
    const T* energy;   // [1] or null (0)
    const T* forces;   // [N, 3] or null
    const T* virial;   // [3, 3] or null
    bool plain;        // dE/dtheta: gE = 1, no F or W term

    // frame f of frames stacked one after another: gE [F], gF [F, N, 3], gW [F, 3, 3]
    __device__ __forceinline__ Upstream at(std::int64_t frame, index_t num_atoms) const
    {
        if (plain || frame == 0) return *this;
        return Upstream{energy ? energy + frame : nullptr,
                        forces ? forces + 3 * frame * num_atoms : nullptr,
                        virial ? virial + 9 * frame : nullptr, false};
    }

    // gE and the slope weight (gF_i - gF_j) . d + d^T gW d of a pair
    __device__ __forceinline__ void weights(index_t i, index_t j, const T d[3], double& ge,
                                            double& slope) const
    {
        if (plain)
        {
            ge = 1.0;
            slope = 0.0;
            return;
        }
        ge = energy != nullptr ? static_cast<double>(energy[0]) : 0.0;
        slope = 0.0;
        if (forces != nullptr)
            for (int a = 0; a < 3; ++a)
                slope += static_cast<double>(forces[3LL * i + a] - forces[3LL * j + a]) *
                         static_cast<double>(d[a]);
        if (virial != nullptr)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    slope += static_cast<double>(d[a]) * static_cast<double>(virial[3 * a + b]) *
                             static_cast<double>(d[b]);
    }
};

// The pair's minimum-image vector d and, strained, d' and r^2 = |d'|^2 (exact arithmetic).
template <typename T, bool kStrained>
__device__ __forceinline__ T pair_geometry(const T* positions, const T* cell, index_t i, index_t j,
                                           const Strain<T>& strain, T d[3], T dp[3])
{
    const T* p = positions + 3LL * i;
    const T* q = positions + 3LL * j;
    for (int a = 0; a < 3; ++a) d[a] = displacement(p[a], q[a], cell[a]);
    if (kStrained)
        strain_vector(strain, d, dp);
    else
        for (int a = 0; a < 3; ++a) dp[a] = d[a];
    return exact_squared(dp[0], dp[1], dp[2]);
}

__device__ __forceinline__ bool listed_pair(index_t i, index_t j, index_t num_atoms)
{
    return i != j && static_cast<std::uint32_t>(i) < static_cast<std::uint32_t>(num_atoms) &&
           static_cast<std::uint32_t>(j) < static_cast<std::uint32_t>(num_atoms);
}

constexpr int kStrainedSlot = 16;  // doubles per block: E, W[9], padding

// E, F and W at a nonzero strain over a half list: a lane per column, both atoms' forces added
// with atomics.
template <typename T, class Pot>
__global__ void __launch_bounds__(kBlock) strained_pairs_kernel(
    const T* __restrict__ positions, const T* __restrict__ cell, const index_t* __restrict__ pairs,
    std::int64_t capacity, index_t num_atoms, T cutoff_squared, Pot pot,
    const T* __restrict__ strain_matrix, T* __restrict__ forces, double* __restrict__ partials)
{
    __shared__ double shared[16];
    const Strain<T> strain = read_strain(strain_matrix);
    double energy = 0;
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t column = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         column < capacity; column += stride)
    {
        const index_t i = pairs[column];
        const index_t j = pairs[capacity + column];
        if (!listed_pair(i, j, num_atoms)) continue;
        T d[3], dp[3];
        const T r2 = pair_geometry<T, true>(positions, cell, i, j, strain, d, dp);
        if (!(r2 < cutoff_squared)) continue;
        T u, fs;
        pot.eval(pot.row(i, j), r2, u, fs);
        energy += static_cast<double>(u);
        for (int b = 0; b < 3; ++b)
        {
            const T back = (strain.e[b] * dp[0] + strain.e[3 + b] * dp[1]) + strain.e[6 + b] * dp[2];
            const T g = fs * (dp[b] + back);
            if (forces != nullptr)
            {
                unsafeAtomicAdd(forces + 3LL * i + b, g);
                unsafeAtomicAdd(forces + 3LL * j + b, -g);
            }
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) w[3 * a + b] += static_cast<double>(fs * dp[a] * d[b]);
    }
    const double e = block_sum(energy, shared);
    double v[9];
    for (int k = 0; k < 9; ++k) v[k] = block_sum(w[k], shared);
    if (threadIdx.x == 0)
    {
        double* out = partials + static_cast<std::int64_t>(kStrainedSlot) * blockIdx.x;
        out[0] = e;
        for (int k = 0; k < 9; ++k) out[1 + k] = v[k];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) strained_finish_kernel(const double* __restrict__ partials,
                                                                 int blocks, T* energy, T* virial)
{
    __shared__ double shared[16];
    double sums[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < blocks; b += blockDim.x)
        for (int k = 0; k < 10; ++k) sums[k] += partials[static_cast<std::int64_t>(kStrainedSlot) * b + k];
    for (int k = 0; k < 10; ++k) sums[k] = block_sum(sums[k], shared);
    if (threadIdx.x == 0)
    {
        if (energy) energy[0] = static_cast<T>(sums[0]);
        if (virial)
            for (int k = 0; k < 9; ++k) virial[k] = static_cast<T>(sums[1 + k]);
    }
}

template <typename T>
__global__ void zero_values(T* __restrict__ data, std::int64_t count)
{
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t k = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         k < count; k += stride)
        data[k] = T(0);
}

// Gradient rows up to this many doubles are summed per block in LDS first.
constexpr int kLocalGradient = 4096;

// Adds `value` to gradient entry k: in the block's LDS copy when there is one, else in memory.
__device__ __forceinline__ void add_gradient(double* local, double* global, std::int64_t k,
                                             double value)
{
    if (local != nullptr)
        atomicAdd(local + k, value);
    else
        unsafeAtomicAdd(global + k, value);
}

// dE/dsigma [P] then dE/depsilon [P] into out [2P].
template <typename T, bool kStrained>
__global__ void __launch_bounds__(kBlock) lj_parameter_kernel(
    const T* __restrict__ positions, const T* __restrict__ cell, const index_t* __restrict__ pairs,
    std::int64_t capacity, index_t num_atoms, T cutoff_squared, const T* __restrict__ sigma,
    const T* __restrict__ epsilon, const std::int64_t* __restrict__ types, int num_types,
    const T* __restrict__ strain_matrix, Upstream<T> frames_up, double* __restrict__ out)
{
    __shared__ double local_store[kLocalGradient];
    const Strain<T> strain = read_strain(strain_matrix);
    // frames of one system (blockIdx.y) add into the same out
    const std::int64_t frame = blockIdx.y;
    positions += 3 * frame * num_atoms;
    cell += 3 * frame;
    pairs += 2 * frame * capacity;
    const Upstream<T> up = frames_up.at(frame, num_atoms);
    const int rows = num_types * (num_types + 1) / 2;
    double* local = 2 * rows <= kLocalGradient ? local_store : nullptr;
    if (local != nullptr)
        for (int k = threadIdx.x; k < 2 * rows; k += blockDim.x) local[k] = 0.0;
    __syncthreads();
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t column = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         column < capacity; column += stride)
    {
        const index_t i = pairs[column];
        const index_t j = pairs[capacity + column];
        if (!listed_pair(i, j, num_atoms)) continue;
        T d[3], dp[3];
        const T r2 = pair_geometry<T, kStrained>(positions, cell, i, j, strain, d, dp);
        if (!(r2 < cutoff_squared)) continue;
        const int row = num_types == 1 ? 0 : pair_row(types[i], types[j], num_types);
        const double s = static_cast<double>(sigma[row]);
        const double e = static_cast<double>(epsilon[row]);
        const double s2 = s * s;
        const double p = (s2 / static_cast<double>(r2)) * (s2 / static_cast<double>(r2)) *
                         (s2 / static_cast<double>(r2));
        const double c = (s2 / static_cast<double>(cutoff_squared)) *
                         (s2 / static_cast<double>(cutoff_squared)) *
                         (s2 / static_cast<double>(cutoff_squared));
        double ge, slope;
        up.weights(i, j, dp, ge, slope);
        const double inverse_r2 = 1.0 / static_cast<double>(r2);
        double d_sigma = ge * (24.0 * e / s * ((2.0 * p * p - p) - (2.0 * c * c - c)));
        double d_epsilon = ge * (4.0 * (p * p - p) - 4.0 * (c * c - c));
        if (slope != 0.0)
        {
            d_sigma += slope * (24.0 * e * (24.0 * p * p - 6.0 * p) * inverse_r2 / s);
            d_epsilon += slope * (24.0 * (2.0 * p * p - p) * inverse_r2);
        }
        add_gradient(local, out, row, d_sigma);
        add_gradient(local, out, rows + row, d_epsilon);
    }
    if (local != nullptr)
    {
        __syncthreads();
        for (int k = threadIdx.x; k < 2 * rows; k += blockDim.x)
            if (local[k] != 0.0) unsafeAtomicAdd(out + k, local[k]);
    }
}

// dE/dtable [P, K] into out.
template <typename T, bool kStrained>
__global__ void __launch_bounds__(kBlock) table_parameter_kernel(
    const T* __restrict__ positions, const T* __restrict__ cell, const index_t* __restrict__ pairs,
    std::int64_t capacity, index_t num_atoms, T cutoff_squared, Table<T> table,
    const T* __restrict__ strain_matrix, Upstream<T> frames_up, double* __restrict__ out)
{
    __shared__ double local_store[kLocalGradient];
    const Strain<T> strain = read_strain(strain_matrix);
    const std::int64_t frame = blockIdx.y;
    positions += 3 * frame * num_atoms;
    cell += 3 * frame;
    pairs += 2 * frame * capacity;
    const Upstream<T> up = frames_up.at(frame, num_atoms);
    const int knots = table.knots;
    const int rows = table.num_types * (table.num_types + 1) / 2;
    const std::int64_t entries = static_cast<std::int64_t>(rows) * knots;
    double* local = entries <= kLocalGradient ? local_store : nullptr;
    if (local != nullptr)
        for (int k = threadIdx.x; k < entries; k += blockDim.x) local[k] = 0.0;
    __syncthreads();
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t column = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         column < capacity; column += stride)
    {
        const index_t i = pairs[column];
        const index_t j = pairs[capacity + column];
        if (!listed_pair(i, j, num_atoms)) continue;
        T d[3], dp[3];
        const T r2 = pair_geometry<T, kStrained>(positions, cell, i, j, strain, d, dp);
        if (!(r2 < cutoff_squared)) continue;
        const int row = table.row(i, j);
        // the knot coordinate with the reference's operations (Table::eval)
        const T r = sqrt(r2);
        const T x = (r - table.start[row]) / table.spacing[row];
        T segment = floor(x);
        segment = segment < T(0) ? T(0) : (segment > T(knots - 2) ? T(knots - 2) : segment);
        const double t = static_cast<double>(x - segment);
        const double s = t > 0.0 ? t : 0.0;
        double ge, slope_weight;
        up.weights(i, j, dp, ge, slope_weight);
        // du/du_m (value weights) and d(du/dt)/du_m (slope weights) of the padded knots
        double value[4], slope[4];
        if (t < 0.0)
        {
            value[0] = -0.5 * t;
            value[1] = 1.0;
            value[2] = 0.5 * t;
            value[3] = 0.0;
            slope[0] = -0.5;
            slope[1] = 0.0;
            slope[2] = 0.5;
            slope[3] = 0.0;
        }
        else
        {
            const double s2 = s * s, s3 = s2 * s;
            value[0] = -0.5 * s + s2 - 0.5 * s3;
            value[1] = 1.0 - 2.5 * s2 + 1.5 * s3;
            value[2] = 0.5 * s + 2.0 * s2 - 1.5 * s3;
            value[3] = -0.5 * s2 + 0.5 * s3;
            slope[0] = -0.5 + 2.0 * s - 1.5 * s2;
            slope[1] = -5.0 * s + 4.5 * s2;
            slope[2] = 0.5 + 4.0 * s - 4.5 * s2;
            slope[3] = -s + 1.5 * s2;
        }
        // fs = -(du/dt) / (h r)
        const double fs_scale = slope_weight != 0.0
            ? -slope_weight / (static_cast<double>(table.spacing[row]) * static_cast<double>(r))
            : 0.0;
        double weight[4];
        for (int m = 0; m < 4; ++m) weight[m] = ge * value[m] + fs_scale * slope[m];
        const std::int64_t base = static_cast<std::int64_t>(row) * knots;
        const int first = static_cast<int>(segment);  // padded knot of weight[0]
        for (int m = 0; m < 4; ++m)
        {
            const int padded = first + m;  // 0 .. K + 1
            const double wgt = weight[m];
            if (padded == 0)
            {
                add_gradient(local, out, base + 0, 3.0 * wgt);
                add_gradient(local, out, base + 1, -3.0 * wgt);
                add_gradient(local, out, base + 2, wgt);
            }
            else if (padded == knots + 1)
            {
                add_gradient(local, out, base + knots - 2, -3.0 * wgt);
                add_gradient(local, out, base + knots - 3, wgt);
            }
            else if (padded - 1 < knots - 1)
            {
                add_gradient(local, out, base + padded - 1, wgt);
            }
        }
    }
    if (local != nullptr)
    {
        __syncthreads();
        for (std::int64_t k = threadIdx.x; k < entries; k += blockDim.x)
            if (local[k] != 0.0) unsafeAtomicAdd(out + k, local[k]);
    }
}

int gradient_blocks(std::int64_t capacity)
{
    const std::int64_t blocks = (capacity + kBlock - 1) / kBlock;
    return static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(blocks, 2048)));
}

template <typename T>
int strained_impl(const void* positions, const void* cell, std::int64_t num_atoms,
                  const index_t* pairs, std::int64_t capacity, const ndn_potential* pot,
                  const void* strain_matrix, void* energy, void* forces, void* virial,
                  void* workspace, hipStream_t stream)
{
    PotentialView<T> view;
    const int prepared = prepare_potential<T>(pot, workspace, stream, &view);
    if (prepared != NDN_OK) return prepared;
    const int blocks = gradient_blocks(capacity);
    auto* partials = reinterpret_cast<double*>(static_cast<std::uint8_t*>(workspace) +
                                               coefficient_bytes(pot, sizeof(T)));
    T* f = static_cast<T*>(forces);
    if (f != nullptr)
    {
        zero_values<T><<<std::min<unsigned>(blocks_for(3 * num_atoms), 1024u), kBlock, 0, stream>>>(
            f, 3 * num_atoms);
        NDN_LAUNCHED("zero_forces");
    }
    const T* x = static_cast<const T*>(positions);
    const T* L = static_cast<const T*>(cell);
    const index_t n = static_cast<index_t>(num_atoms);
    const T* strain = static_cast<const T*>(strain_matrix);
    if (view.kind == NDN_LJ)
        strained_pairs_kernel<T><<<blocks, kBlock, 0, stream>>>(
            x, L, pairs, capacity, n, view.cutoff_squared, view.lj, strain, f, partials);
    else
        strained_pairs_kernel<T><<<blocks, kBlock, 0, stream>>>(
            x, L, pairs, capacity, n, view.cutoff_squared, view.table, strain, f, partials);
    NDN_LAUNCHED("strained_pairs");
    if (energy != nullptr || virial != nullptr)
    {
        strained_finish_kernel<T><<<1, kBlock, 0, stream>>>(partials, blocks, static_cast<T*>(energy),
                                                            static_cast<T*>(virial));
        NDN_LAUNCHED("strained_finish");
    }
    return NDN_OK;
}

template <typename T>
int parameter_impl(const void* positions, const void* cell, std::int64_t num_atoms,
                   const index_t* pairs, std::int64_t capacity, const ndn_potential* pot,
                   const void* strain, const Upstream<T>& up, double* out,
                   std::int64_t out_count, void* workspace, hipStream_t stream,
                   std::int64_t frames = 1)
{
    PotentialView<T> view;
    const int prepared = prepare_potential<T>(pot, workspace, stream, &view);
    if (prepared != NDN_OK) return prepared;
    zero_values<double><<<std::min<unsigned>(blocks_for(out_count), 256u), kBlock, 0, stream>>>(
        out, out_count);
    NDN_LAUNCHED("zero_gradient");
    // each block adds its rows into `out` once at its end: over many frames, fewer blocks per
    // frame (each striding over more columns) keep those atomics on the same rows few
    const std::int64_t per_frame = std::max<std::int64_t>(
        1, std::min<std::int64_t>(gradient_blocks(capacity), 2048 / frames));
    const dim3 blocks(static_cast<unsigned>(per_frame), static_cast<unsigned>(frames));
    const T* x = static_cast<const T*>(positions);
    const T* L = static_cast<const T*>(cell);
    const index_t n = static_cast<index_t>(num_atoms);
    const T* e = static_cast<const T*>(strain);
    const bool strained = strain != nullptr;
    if (view.kind == NDN_LJ)
    {
        const T* sigma = static_cast<const T*>(pot->sigma);
        const T* epsilon = static_cast<const T*>(pot->epsilon);
        if (strained)
            lj_parameter_kernel<T, true><<<blocks, kBlock, 0, stream>>>(
                x, L, pairs, capacity, n, view.cutoff_squared, sigma, epsilon, pot->atom_types,
                pot->num_types, e, up, out);
        else
            lj_parameter_kernel<T, false><<<blocks, kBlock, 0, stream>>>(
                x, L, pairs, capacity, n, view.cutoff_squared, sigma, epsilon, pot->atom_types,
                pot->num_types, e, up, out);
    }
    else if (strained)
        table_parameter_kernel<T, true><<<blocks, kBlock, 0, stream>>>(
            x, L, pairs, capacity, n, view.cutoff_squared, view.table, e, up, out);
    else
        table_parameter_kernel<T, false><<<blocks, kBlock, 0, stream>>>(
            x, L, pairs, capacity, n, view.cutoff_squared, view.table, e, up, out);
    NDN_LAUNCHED("parameter_gradient");
    return NDN_OK;
}

int check_pair_call(int dtype, const ndn_potential* pot, const void* positions, const void* cell,
                    std::int64_t num_atoms, const int32_t* pairs, std::int64_t capacity,
                    void* workspace, std::int64_t workspace_bytes)
{
    const int checked = check_potential(pot, num_atoms);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(dtype == NDN_F32 || dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    NDN_REQUIRE(positions && cell && pairs && workspace, "a buffer is null");
    NDN_REQUIRE(num_atoms >= 1 && num_atoms <= INT32_MAX, "N must satisfy 1 <= N < 2**31");
    NDN_REQUIRE(capacity >= 1 && capacity <= INT32_MAX, "pairs must be [2, C] with int32 C");
    std::int64_t needed = 0;
    const int sized = ndn_pair_workspace_bytes(dtype, pot, capacity, &needed);
    if (sized != NDN_OK) return sized;
    NDN_REQUIRE(workspace_bytes >= needed, "the workspace is too small");
    return NDN_OK;
}

// The VJP entry points' shared tail: out's size, then the launch over `frames` frames.
int vjp_call(int dtype, const void* positions, const void* cells, std::int64_t frames,
             std::int64_t num_atoms, const int32_t* pairs, std::int64_t capacity,
             const ndn_potential* pot, const void* grad_energy, const void* grad_forces,
             const void* grad_virial, double* out, std::int64_t out_count, void* workspace,
             hipStream_t stream)
{
    const std::int64_t rows = static_cast<std::int64_t>(pot->num_types) * (pot->num_types + 1) / 2;
    const std::int64_t expected = pot->kind == NDN_LJ ? 2 * rows : rows * pot->knots;
    NDN_REQUIRE(out != nullptr && out_count == expected,
                "out must hold 2P doubles for LJ (dsigma, depsilon) or P * K for a table");
    if (dtype == NDN_F32)
        return parameter_impl<float>(
            positions, cells, num_atoms, pairs, capacity, pot, nullptr,
            Upstream<float>{static_cast<const float*>(grad_energy),
                            static_cast<const float*>(grad_forces),
                            static_cast<const float*>(grad_virial), false},
            out, out_count, workspace, stream, frames);
    return parameter_impl<double>(
        positions, cells, num_atoms, pairs, capacity, pot, nullptr,
        Upstream<double>{static_cast<const double*>(grad_energy),
                         static_cast<const double*>(grad_forces),
                         static_cast<const double*>(grad_virial), false},
        out, out_count, workspace, stream, frames);
}

}  // namespace ndn

using namespace ndn;

extern "C" {

int ndn_pair_efw_strained(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                          const int32_t* pairs, int64_t capacity, const ndn_potential* pot,
                          const void* strain, void* energy, void* forces, void* virial,
                          void* workspace, int64_t workspace_bytes, ndn_stream_t stream)
{
    const int checked = check_pair_call(dtype, pot, positions, cell, num_atoms, pairs, capacity,
                                        workspace, workspace_bytes);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(strain != nullptr, "strain must not be null (ndn_pair_efw takes none)");
    if (dtype == NDN_F32)
        return strained_impl<float>(positions, cell, num_atoms, pairs, capacity, pot, strain,
                                    energy, forces, virial, workspace, stream);
    return strained_impl<double>(positions, cell, num_atoms, pairs, capacity, pot, strain, energy,
                                 forces, virial, workspace, stream);
}

int ndn_pair_parameter_gradient(int dtype, const void* positions, const void* cell,
                                int64_t num_atoms, const int32_t* pairs, int64_t capacity,
                                const ndn_potential* pot, const void* strain, double* out,
                                int64_t out_count, void* workspace, int64_t workspace_bytes,
                                ndn_stream_t stream)
{
    const int checked = check_pair_call(dtype, pot, positions, cell, num_atoms, pairs, capacity,
                                        workspace, workspace_bytes);
    if (checked != NDN_OK) return checked;
    const std::int64_t rows = static_cast<std::int64_t>(pot->num_types) * (pot->num_types + 1) / 2;
    const std::int64_t expected = pot->kind == NDN_LJ ? 2 * rows : rows * pot->knots;
    NDN_REQUIRE(out != nullptr && out_count == expected,
                "out must hold 2P doubles for LJ (dsigma, depsilon) or P * K for a table");
    if (dtype == NDN_F32)
        return parameter_impl<float>(positions, cell, num_atoms, pairs, capacity, pot, strain,
                                     Upstream<float>{nullptr, nullptr, nullptr, true}, out,
                                     out_count, workspace, stream);
    return parameter_impl<double>(positions, cell, num_atoms, pairs, capacity, pot, strain,
                                  Upstream<double>{nullptr, nullptr, nullptr, true}, out,
                                  out_count, workspace, stream);
}

int ndn_pair_parameter_vjp(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                           const int32_t* pairs, int64_t capacity, const ndn_potential* pot,
                           const void* grad_energy, const void* grad_forces,
                           const void* grad_virial, double* out, int64_t out_count,
                           void* workspace, int64_t workspace_bytes, ndn_stream_t stream)
{
    const int checked = check_pair_call(dtype, pot, positions, cell, num_atoms, pairs, capacity,
                                        workspace, workspace_bytes);
    if (checked != NDN_OK) return checked;
    return vjp_call(dtype, positions, cell, 1, num_atoms, pairs, capacity, pot, grad_energy,
                    grad_forces, grad_virial, out, out_count, workspace, stream);
}

int ndn_pair_parameter_vjp_frames(int dtype, const void* positions, const void* cells,
                                  int64_t frames, int64_t num_atoms, const int32_t* stacks,
                                  int64_t capacity, const ndn_potential* pot,
                                  const void* grad_energy, const void* grad_forces,
                                  const void* grad_virial, double* out, int64_t out_count,
                                  void* workspace, int64_t workspace_bytes, ndn_stream_t stream)
{
    const int checked = check_pair_call(dtype, pot, positions, cells, num_atoms, stacks, capacity,
                                        workspace, workspace_bytes);
    if (checked != NDN_OK) return checked;
    NDN_REQUIRE(frames >= 1 && frames <= 65535, "frames must lie in [1, 65535]");
    return vjp_call(dtype, positions, cells, frames, num_atoms, stacks, capacity, pot,
                    grad_energy, grad_forces, grad_virial, out, out_count, workspace, stream);
}

}  // extern "C"
