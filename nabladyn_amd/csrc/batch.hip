// Many independent systems of uneven size in one set of launches: energy, forces and virial of
// every system, with no per-system launch and no host round trip between systems.
//
// All systems share one global bin space (each owns a contiguous range of it), so one counting
// sort bins every atom of the batch, one scan covers every bin, and one force launch covers every
// atom. The force launch builds no list: each atom's candidates come from the culled stencil of
// its own system's sub-bins, and the pair terms are summed as they are found. Two schedules:
//   - wave per atom: the atom's candidates 64 at a time across the lanes, reduced over the wave;
//   - thread per atom: a lane per atom walking its own stencil.
// and two sub-bin sizes (at least the cutoff, or half of it). A host-side autotuner times the
// four on the batch and keeps the fastest.
//
// Every sum is per atom and then per system in a fixed order, so a system's results do not
// depend on the other systems of the batch: alone or in the mix, the same schedule gives the
// same bits. Geometry and mask follow the reference (space.py:11-19,
// _lennard_jones/_potential.py:218-234, _tabulated/_potential.py:146-162).

#include <algorithm>
#include <climits>

#include "common.hpp"
#include "neighbors.hpp"
#include "pair_forces.hpp"
#include "potentials.hpp"
#include "scan.hpp"

namespace ndn {

template <typename T>
struct BatchArgs
{
    const T* positions;       // [N, 3] input, anywhere
    index_t num_atoms;
    std::int64_t num_systems;
    const index_t* system_of;      // [N]
    const std::int64_t* atom_start;  // [S + 1]
    const index_t* bin_start;      // [S + 1] first global bin of each system
    const index_t* dims;           // [S, 3]
    const T* widths;               // [S, 3]
    const T* cells;                // [S, 3]
    const T* cutoffs;              // [S]
    const T* cutoffs_squared;      // [S]
    const index_t* kinds;          // [S]
    const LJRow<T>* lj;            // [S]
    const index_t* table_of;       // [S] row of the padded tables, or -1
    const T* tables;               // [R, K + 2]
    const T* table_start;          // [R]
    const T* table_spacing;  // [R]
    int knots;
    std::int64_t total_bins;
    int span;                      // sub-bins a stencil reaches on each side: 1 or 2
    // scratch
    index_t* bin_of;
    index_t* bin_counts;
    index_t* bin_offsets;
    index_t* unordered;
    index_t* order;                // slot -> atom
    T* wrapped4;                   // [N, 4] wrapped positions in atom order
    T* sorted4;                    // [N, 4] the same in slot order
    std::uint8_t* scan_temp;
    // outputs
    T* forces;                     // [N, 3]
    T* atom_terms;                 // [N, 7]: half the pair energy and virial of each atom
    T* energy;                     // [S]
    T* virial;                     // [S, 9]
};

template <typename T>
__device__ __forceinline__ index_t batch_bin(T x, T width, index_t count)
{
    const index_t bin = static_cast<index_t>(floor(x / width));
    return bin < 0 ? 0 : (bin >= count ? count - 1 : bin);
}

// Wraps each atom into its system's cell (the reference's arithmetic) and counts it in its bin.
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_assign_kernel(BatchArgs<T> a)
{
#pragma clang fp contract(off)
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    const std::int64_t atom = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (atom >= a.num_atoms) return;
    const index_t s = a.system_of[atom];
    T w[3];
    index_t b[3];
    for (int axis = 0; axis < 3; ++axis)
    {
        const T length = a.cells[3 * s + axis];
        const T v = a.positions[3 * atom + axis];
        const T k = floor(v / length);
        T x = v - k * length;
        if (x < T(0)) x = x + length;
        if (x >= length) x = x - length;
        w[axis] = x;
        b[axis] = batch_bin(x, a.widths[3 * s + axis], a.dims[3 * s + axis]);
    }
    V4 p;
    p.x = w[0];
    p.y = w[1];
    p.z = w[2];
    p.w = T(0);
    reinterpret_cast<V4*>(a.wrapped4)[atom] = p;
    const index_t nx = a.dims[3 * s], ny = a.dims[3 * s + 1];
    const index_t bin = a.bin_start[s] + b[0] + nx * (b[1] + ny * b[2]);
    a.bin_of[atom] = bin;
    // one atomic per run of a bin's atoms in neighbouring lanes
    const LaneRun run = equal_runs(bin);
    if (run.head) atomicAdd(a.bin_counts + bin, run.length);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_scatter_kernel(BatchArgs<T> a)
{
    const std::int64_t atom = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (atom >= a.num_atoms) return;
    const index_t bin = a.bin_of[atom];
    // one atomic per run: the run's lanes take the places below the count it returns, one each
    // (batch_order_kernel orders a bin's atoms by index after)
    const LaneRun run = equal_runs(bin);
    index_t count = 0;
    if (run.head) count = atomicSub(a.bin_counts + bin, run.length);
    count = __shfl(count, run.first, kWave);
    a.unordered[a.bin_offsets[bin] + count - 1 - (lane_id() - run.first)] =
        static_cast<index_t>(atom);
}

// Thread per bin: orders its atoms by index (deterministic), writes slot -> atom and the sorted
// positions.
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_order_kernel(BatchArgs<T> a)
{
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    const std::int64_t bin = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (bin >= a.total_bins) return;
    const index_t begin = a.bin_offsets[bin];
    const index_t end = a.bin_offsets[bin + 1];
    for (index_t i = begin; i < end; ++i)
    {
        const index_t atom = a.unordered[i];
        index_t rank = 0;
        for (index_t k = begin; k < end; ++k) rank += a.unordered[k] < atom;
        a.order[begin + rank] = atom;
        reinterpret_cast<V4*>(a.sorted4)[begin + rank] = reinterpret_cast<const V4*>(a.wrapped4)[atom];
    }
}

template <typename T>
__device__ __forceinline__ T batch_slab(T c, T lo, T w)
{
    const T below = lo - c;
    const T above = c - (lo + w);
    return below > T(0) ? below : (above > T(0) ? above : T(0));
}

template <typename T>
struct SystemView
{
    index_t nx, ny, nz;
    T wx, wy, wz;
    T lx, ly, lz;
    T cutoff, cutoff_squared;
    index_t bin_start;
    int kind;
    LJRow<T> lj;
    const T* table;
    T table_start, table_spacing;
};

template <typename T>
__device__ __forceinline__ SystemView<T> system_view(const BatchArgs<T>& a, index_t s)
{
    SystemView<T> v;
    v.nx = a.dims[3 * s];
    v.ny = a.dims[3 * s + 1];
    v.nz = a.dims[3 * s + 2];
    v.wx = a.widths[3 * s];
    v.wy = a.widths[3 * s + 1];
    v.wz = a.widths[3 * s + 2];
    v.lx = a.cells[3 * s];
    v.ly = a.cells[3 * s + 1];
    v.lz = a.cells[3 * s + 2];
    v.cutoff = a.cutoffs[s];
    v.cutoff_squared = a.cutoffs_squared[s];
    v.bin_start = a.bin_start[s];
    v.kind = a.kinds[s];
    v.lj = a.lj[s];
    const index_t row = a.table_of[s];
    v.table = row >= 0 ? a.tables + static_cast<std::int64_t>(row) * (a.knots + 2) : nullptr;
    v.table_start = row >= 0 ? a.table_start[row] : T(0);
    v.table_spacing = row >= 0 ? a.table_spacing[row] : T(1);
    return v;
}

// u and -(du/dr)/r of a pair of the system, r^2 below its cutoff.
template <typename T>
__device__ __forceinline__ void pair_terms(const SystemView<T>& v, int knots, T r2, T& u, T& fs)
{
    if (v.kind == NDN_LJ)
    {
        const T inverse = T(1) / r2;
        const T s = v.lj.sigma_squared * inverse;
        const T p6 = s * s * s;
        u = v.lj.four_eps * (p6 * p6 - p6) - v.lj.shift;
        fs = v.lj.twentyfour_eps * (T(2) * p6 * p6 - p6) * inverse;
        return;
    }
    // r and the knot coordinate with the reference's operations, each correctly rounded
    const T r = sqrt(r2);
    const T x = (r - v.table_start) / v.table_spacing;
    T segment = floor(x);
    segment = segment < T(0) ? T(0) : (segment > T(knots - 2) ? T(knots - 2) : segment);
    const T t = x - segment;
    const T* k = v.table + static_cast<int>(segment);
    const T u0 = k[0], u1 = k[1], u2 = k[2], u3 = k[3];
    const T a1 = u2 - u0;
    const T b1 = T(2) * u0 - T(5) * u1 + T(4) * u2 - u3;
    const T c1 = T(3) * (u1 - u2) + u3 - u0;
    const T s = t > T(0) ? t : T(0);
    const bool below = t < T(0);
    u = u1 + T(0.5) * s * (a1 + s * (b1 + s * c1)) + (below ? T(0.5) * a1 * t : T(0));
    const T du_dt = below ? T(0.5) * a1 : T(0.5) * (a1 + s * (T(2) * b1 + T(3) * s * c1));
    fs = -(du_dt / v.table_spacing) / r;
}

// The stencil of a center: one lane (of 64) per candidate run of x-consecutive sub-bins in a
// (y, z) row, culled by distance. Returns this lane's [lo, lo + length) of slots.
template <typename T>
__device__ __forceinline__ void stencil_run(const BatchArgs<T>& a, const SystemView<T>& v, T cx,
                                            T cy, T cz, int lane, T cull_squared, index_t& lo,
                                            index_t& length, bool& wraps)
{
    const int reach = 2 * a.span + 1;
    const index_t hx = batch_bin(cx, v.wx, v.nx), hy = batch_bin(cy, v.wy, v.ny),
                  hz = batch_bin(cz, v.wz, v.nz);
    const bool full_x = v.nx < reach, full_y = v.ny < reach, full_z = v.nz < reach;
    const int span_x = full_x ? v.nx : reach;
    const int span_y = full_y ? v.ny : reach;
    const int span_z = full_z ? v.nz : reach;
    wraps = full_x || full_y || full_z || hx < a.span || hx + a.span >= v.nx || hy < a.span ||
            hy + a.span >= v.ny || hz < a.span || hz + a.span >= v.nz;
    lo = 0;
    length = 0;
    const int row = lane / 2, part = lane % 2;
    if (row >= span_y * span_z) return;
    const int oy = row % span_y, oz = row / span_y;
    const index_t by = full_y ? oy : hy + oy - a.span;
    const index_t bz = full_z ? oz : hz + oz - a.span;
    const T dy = full_y ? T(0) : batch_slab(cy, by * v.wy, v.wy);
    const T dz = full_z ? T(0) : batch_slab(cz, bz * v.wz, v.wz);
    const T dyz = dy * dy + dz * dz;
    if (!(dyz < cull_squared)) return;
    int first = span_x, last = -1;
    for (int ox = 0; ox < span_x; ++ox)
    {
        const index_t bx = full_x ? ox : hx + ox - a.span;
        const T dx = full_x ? T(0) : batch_slab(cx, bx * v.wx, v.wx);
        if (dx * dx + dyz < cull_squared)
        {
            first = ox < first ? ox : first;
            last = ox;
        }
    }
    if (last < first) return;
    const index_t wy = by < 0 ? by + v.ny : (by >= v.ny ? by - v.ny : by);
    const index_t wz = bz < 0 ? bz + v.nz : (bz >= v.nz ? bz - v.nz : bz);
    const std::int64_t base = v.bin_start + static_cast<std::int64_t>(v.nx) *
                                                (wy + static_cast<std::int64_t>(v.ny) * wz);
    const index_t x0 = full_x ? first : hx + first - a.span;
    const index_t x1 = full_x ? last : hx + last - a.span;
    index_t run_lo = 0, run_hi = -1;
    if (x0 < 0)
    {
        run_lo = part == 0 ? x0 + v.nx : 0;
        run_hi = part == 0 ? v.nx - 1 : x1;
    }
    else if (x1 >= v.nx)
    {
        run_lo = part == 0 ? x0 : 0;
        run_hi = part == 0 ? v.nx - 1 : x1 - v.nx;
    }
    else if (part == 0)
    {
        run_lo = x0;
        run_hi = x1;
    }
    if (run_hi < run_lo) return;
    lo = a.bin_offsets[base + run_lo];
    length = a.bin_offsets[base + run_hi + 1] - lo;
}

template <typename T>
__device__ __forceinline__ void pair_vector(bool wraps, T cx, T cy, T cz, T qx, T qy, T qz,
                                            const SystemView<T>& v, T& dx, T& dy, T& dz)
{
    if (wraps)
    {
        dx = displacement(cx, qx, v.lx);
        dy = displacement(cy, qy, v.ly);
        dz = displacement(cz, qz, v.lz);
    }
    else
    {
        dx = cx - qx;
        dy = cy - qy;
        dz = cz - qz;
    }
}

// Wave per atom: the center's candidates 64 at a time, the pair terms of the hits summed over
// the wave in a fixed order.
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_wave_kernel(BatchArgs<T> a, T cull_factor)
{
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    const std::int64_t slot =
        static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
    if (slot >= a.num_atoms) return;
    const int lane = lane_id();
    const V4* sorted = reinterpret_cast<const V4*>(a.sorted4);
    const index_t atom = a.order[slot];
    const index_t s = a.system_of[atom];
    const SystemView<T> v = system_view(a, s);
    const V4 c = sorted[slot];
    const T cull = v.cutoff * cull_factor;
    index_t lo, length;
    bool wraps;
    stencil_run(a, v, c.x, c.y, c.z, lane, cull * cull, lo, length, wraps);
    const index_t inclusive = wave_inclusive_sum(length);
    const index_t start = inclusive - length;
    const index_t total = __shfl(inclusive, kWave - 1, kWave);
    T fx = 0, fy = 0, fz = 0, e = 0;
    T w[6] = {0, 0, 0, 0, 0, 0};
    for (index_t chunk = 0; chunk < total; chunk += kWave)
    {
        const index_t t = chunk + lane;
        int r = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
        {
            const index_t st = __shfl(start, r + step, kWave);
            if (st <= t) r += step;
        }
        const index_t other = __shfl(lo, r, kWave) + (t - __shfl(start, r, kWave));
        if (t < total && other != slot)
        {
            const V4 q = sorted[other];
            T dx, dy, dz;
            pair_vector(wraps, c.x, c.y, c.z, q.x, q.y, q.z, v, dx, dy, dz);
            const T r2 = exact_squared(dx, dy, dz);
            if (r2 < v.cutoff_squared)
            {
                T u, fs;
                pair_terms(v, a.knots, r2, u, fs);
                e += u;
                fx += fs * dx;
                fy += fs * dy;
                fz += fs * dz;
                w[0] += dx * (fs * dx);
                w[1] += dy * (fs * dy);
                w[2] += dz * (fs * dz);
                w[3] += dx * (fs * dy);
                w[4] += dx * (fs * dz);
                w[5] += dy * (fs * dz);
            }
        }
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    e = wave_sum(e);
    for (int k = 0; k < 6; ++k) w[k] = wave_sum(w[k]);
    if (lane == 0)
    {
        T* f = a.forces + 3LL * atom;
        f[0] = fx;
        f[1] = fy;
        f[2] = fz;
        T* terms = a.atom_terms + 7LL * atom;
        terms[0] = T(0.5) * e;
        for (int k = 0; k < 6; ++k) terms[1 + k] = T(0.5) * w[k];
    }
}

// Thread per atom: each lane walks its own stencil runs, one candidate at a time.
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_thread_kernel(BatchArgs<T> a, T cull_factor)
{
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    const std::int64_t slot = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (slot >= a.num_atoms) return;
    const V4* sorted = reinterpret_cast<const V4*>(a.sorted4);
    const index_t atom = a.order[slot];
    const index_t s = a.system_of[atom];
    const SystemView<T> v = system_view(a, s);
    const V4 c = sorted[slot];
    const T cull = v.cutoff * cull_factor;
    T fx = 0, fy = 0, fz = 0, e = 0;
    T w[6] = {0, 0, 0, 0, 0, 0};
    const int reach = 2 * a.span + 1;
    const int rows = (v.ny < reach ? v.ny : reach) * (v.nz < reach ? v.nz : reach);
    for (int range = 0; range < 2 * rows; ++range)
    {
        index_t lo, length;
        bool wraps;
        stencil_run(a, v, c.x, c.y, c.z, range, cull * cull, lo, length, wraps);
        for (index_t other = lo; other < lo + length; ++other)
        {
            if (other == slot) continue;
            const V4 q = sorted[other];
            T dx, dy, dz;
            pair_vector(wraps, c.x, c.y, c.z, q.x, q.y, q.z, v, dx, dy, dz);
            const T r2 = exact_squared(dx, dy, dz);
            if (!(r2 < v.cutoff_squared)) continue;
            T u, fs;
            pair_terms(v, a.knots, r2, u, fs);
            e += u;
            fx += fs * dx;
            fy += fs * dy;
            fz += fs * dz;
            w[0] += dx * (fs * dx);
            w[1] += dy * (fs * dy);
            w[2] += dz * (fs * dz);
            w[3] += dx * (fs * dy);
            w[4] += dx * (fs * dz);
            w[5] += dy * (fs * dz);
        }
    }
    T* f = a.forces + 3LL * atom;
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
    T* terms = a.atom_terms + 7LL * atom;
    terms[0] = T(0.5) * e;
    for (int k = 0; k < 6; ++k) terms[1 + k] = T(0.5) * w[k];
}

// The table's gradient (the vector-Jacobian product of E, F and W with upstream gE [S],
// gF [N, 3], gW [S, 9]): wave per atom over the same stencil; each hit pair adds, to the four
// padded knots of its segment, gE du/dk - ((gF_i - gF_j) . r_ij + r_ij' gW r_ij) / r d(du/dr)/dk,
// halved because every pair is met from both of its atoms. E, F and W are linear in the knots,
// so these weights are the exact derivatives. Per-block sums in LDS (fp64), written per block.
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_table_vjp_kernel(BatchArgs<T> a, T cull_factor,
                                                                 const T* __restrict__ grad_energy,
                                                                 const T* __restrict__ grad_forces,
                                                                 const T* __restrict__ grad_virial,
                                                                 double* __restrict__ partials)
{
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    extern __shared__ double knot_sums[];
    const int width = a.knots + 2;
    for (int k = threadIdx.x; k < width; k += blockDim.x) knot_sums[k] = 0.0;
    __syncthreads();
    const std::int64_t slot =
        static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
    if (slot < a.num_atoms)
    {
        const int lane = lane_id();
        const V4* sorted = reinterpret_cast<const V4*>(a.sorted4);
        const index_t atom = a.order[slot];
        const index_t s = a.system_of[atom];
        const SystemView<T> v = system_view(a, s);
        if (v.kind == NDN_TABLE)
        {
            const V4 c = sorted[slot];
            const T cull = v.cutoff * cull_factor;
            index_t lo, length;
            bool wraps;
            stencil_run(a, v, c.x, c.y, c.z, lane, cull * cull, lo, length, wraps);
            const index_t inclusive = wave_inclusive_sum(length);
            const index_t start = inclusive - length;
            const index_t total = __shfl(inclusive, kWave - 1, kWave);
            const T ge = grad_energy[s];
            const T* gw = grad_virial + 9LL * s;
            const T gix = grad_forces[3LL * atom], giy = grad_forces[3LL * atom + 1],
                    giz = grad_forces[3LL * atom + 2];
            const std::int64_t row_base = static_cast<std::int64_t>(a.table_of[s]) * width;
            (void)row_base;
            for (index_t chunk = 0; chunk < total; chunk += kWave)
            {
                const index_t t = chunk + lane;
                int r = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1)
                {
                    const index_t st = __shfl(start, r + step, kWave);
                    if (st <= t) r += step;
                }
                const index_t other = __shfl(lo, r, kWave) + (t - __shfl(start, r, kWave));
                if (t < total && other != slot)
                {
                    const V4 q = sorted[other];
                    T dx, dy, dz;
                    pair_vector(wraps, c.x, c.y, c.z, q.x, q.y, q.z, v, dx, dy, dz);
                    const T r2 = exact_squared(dx, dy, dz);
                    if (r2 < v.cutoff_squared)
                    {
                        const index_t j = a.order[other];
                        const T dist = sqrt(r2);
                        const T x = (dist - v.table_start) / v.table_spacing;
                        T segment = floor(x);
                        segment = segment < T(0) ? T(0)
                                : (segment > T(a.knots - 2) ? T(a.knots - 2) : segment);
                        const T tt = x - segment;
                        const bool below = tt < T(0);
                        const T sv = below ? T(0) : tt;
                        T du[4], dslope[4];
                        if (below)
                        {
                            du[0] = T(-0.5) * tt;
                            du[1] = T(1);
                            du[2] = T(0.5) * tt;
                            du[3] = T(0);
                            dslope[0] = T(-0.5);
                            dslope[1] = T(0);
                            dslope[2] = T(0.5);
                            dslope[3] = T(0);
                        }
                        else
                        {
                            const T s2 = sv * sv, s3 = s2 * sv;
                            du[0] = T(-0.5) * sv + s2 - T(0.5) * s3;
                            du[1] = T(1) - T(2.5) * s2 + T(1.5) * s3;
                            du[2] = T(0.5) * sv + T(2) * s2 - T(1.5) * s3;
                            du[3] = T(-0.5) * s2 + T(0.5) * s3;
                            dslope[0] = T(-0.5) + T(2) * sv - T(1.5) * s2;
                            dslope[1] = T(-5) * sv + T(4.5) * s2;
                            dslope[2] = T(0.5) + T(4) * sv - T(4.5) * s2;
                            dslope[3] = -sv + T(1.5) * s2;
                        }
                        const T gjx = grad_forces[3LL * j], gjy = grad_forces[3LL * j + 1],
                                gjz = grad_forces[3LL * j + 2];
                        const T projection = (gix - gjx) * dx + (giy - gjy) * dy + (giz - gjz) * dz;
                        const T quadratic =
                            dx * (gw[0] * dx + gw[1] * dy + gw[2] * dz) +
                            dy * (gw[3] * dx + gw[4] * dy + gw[5] * dz) +
                            dz * (gw[6] * dx + gw[7] * dy + gw[8] * dz);
                        // dF and dW weights: d(-(du/dr)/r) = -(d(du/dt) / h) / r
                        const T slope_weight =
                            -((projection + quadratic) / v.table_spacing) / dist;
                        const int base = static_cast<int>(segment);
                        for (int m = 0; m < 4; ++m)
                        {
                            const double g = 0.5 * (static_cast<double>(ge) * du[m] +
                                                    static_cast<double>(slope_weight) * dslope[m]);
                            atomicAdd(knot_sums + base + m, g);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < width; k += blockDim.x)
        partials[static_cast<std::int64_t>(blockIdx.x) * width + k] = knot_sums[k];
}

// Folds the block partials [blocks, K + 2] onto their first `folds` rows: fold f sums rows f,
// f + folds, f + 2 folds, ... in that order (coalesced rows, a thread per padded knot) and writes
// the sum over row f, which only fold f reads.
__global__ void __launch_bounds__(kBlock) table_vjp_fold_kernel(double* __restrict__ partials,
                                                                int blocks, int width)
{
    const int folds = gridDim.x;
    const int f = blockIdx.x;
    for (int k = threadIdx.x; k < width; k += blockDim.x)
    {
        double sum = 0.0;
        for (int b = f; b < blocks; b += folds) sum += partials[static_cast<std::int64_t>(b) * width + k];
        partials[static_cast<std::int64_t>(f) * width + k] = sum;
    }
}

// Sums the block partials per padded knot in a fixed order, then maps the padded knots back to
// the table: the virtual first knot is 3 v0 - 3 v1 + v2, the virtual last 3 v_{K-1} - 3 v_{K-2}
// + v_{K-3} with v_{K-1} held at 0, so the last knot's gradient is 0.
template <typename T>
__global__ void __launch_bounds__(kBlock) table_vjp_finish_kernel(const double* __restrict__ partials,
                                                                  int blocks, int knots,
                                                                  T* __restrict__ grad_table)
{
    __shared__ double padded[4096];
    const int width = knots + 2;
    for (int k = threadIdx.x; k < width; k += blockDim.x)
    {
        double sum = 0.0;
        for (int b = 0; b < blocks; ++b) sum += partials[static_cast<std::int64_t>(b) * width + k];
        padded[k] = sum;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < knots; k += blockDim.x)
    {
        double g = padded[k + 1];
        if (k == 0) g += 3.0 * padded[0];
        if (k == 1) g -= 3.0 * padded[0];
        if (k == 2) g += padded[0];
        if (k == knots - 2) g -= 3.0 * padded[width - 1];
        if (k == knots - 3) g += padded[width - 1];
        if (k == knots - 1) g = 0.0;
        grad_table[k] = static_cast<T>(g);
    }
}

// Block per system: E and W from the atoms' terms, in a fixed order.
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_reduce_kernel(BatchArgs<T> a)
{
    __shared__ double shared[16];
    const std::int64_t s = blockIdx.x;
    if (s >= a.num_systems) return;
    double sums[7] = {0, 0, 0, 0, 0, 0, 0};
    for (std::int64_t atom = a.atom_start[s] + threadIdx.x; atom < a.atom_start[s + 1];
         atom += blockDim.x)
        for (int k = 0; k < 7; ++k) sums[k] += static_cast<double>(a.atom_terms[7 * atom + k]);
    for (int k = 0; k < 7; ++k) sums[k] = block_sum(sums[k], shared);
    if (threadIdx.x == 0)
    {
        a.energy[s] = static_cast<T>(sums[0]);
        const double xx = sums[1], yy = sums[2], zz = sums[3], xy = sums[4], xz = sums[5],
                     yz = sums[6];
        const double out[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
        for (int k = 0; k < 9; ++k) a.virial[9 * s + k] = static_cast<T>(out[k]);
    }
}

template <typename T>
BatchArgs<T> batch_args(const ndn_batch& b);

template <typename T>
int batch_impl(const ndn_batch& b, int schedule, hipStream_t stream)
{
    const BatchArgs<T> a = batch_args<T>(b);
    const std::int64_t n = b.num_atoms;
    batch_assign_kernel<T><<<blocks_for(n), kBlock, 0, stream>>>(a);
    NDN_LAUNCHED("batch_assign");
    const int scanned = exclusive_scan(a.bin_counts, a.bin_offsets, a.total_bins + 1, a.scan_temp,
                                       nullptr, b.status, stream);
    if (scanned != NDN_OK) return scanned;
    batch_scatter_kernel<T><<<blocks_for(n), kBlock, 0, stream>>>(a);
    NDN_LAUNCHED("batch_scatter");
    batch_order_kernel<T><<<blocks_for(a.total_bins), kBlock, 0, stream>>>(a);
    NDN_LAUNCHED("batch_order");
    // sub-bins of the cutoff / span: culling keeps every one a pair within it could sit in
    const T cull_factor = static_cast<T>(1.0 + 1e-4);
    if (schedule == 0)
        batch_thread_kernel<T><<<blocks_for(n), kBlock, 0, stream>>>(a, cull_factor);
    else
        batch_wave_kernel<T><<<blocks_for(n, kBlock / kWave), kBlock, 0, stream>>>(a, cull_factor);
    NDN_LAUNCHED("batch_forces");
    batch_reduce_kernel<T><<<static_cast<unsigned>(b.num_systems), kBlock, 0, stream>>>(a);
    NDN_LAUNCHED("batch_reduce");
    return NDN_OK;
}

template <typename T>
BatchArgs<T> batch_args(const ndn_batch& b)
{
    BatchArgs<T> a{};
    a.positions = static_cast<const T*>(b.positions);
    a.num_atoms = static_cast<index_t>(b.num_atoms);
    a.num_systems = b.num_systems;
    a.system_of = b.system_of;
    a.atom_start = b.atom_start;
    a.bin_start = b.bin_start;
    a.dims = b.dims;
    a.widths = static_cast<const T*>(b.widths);
    a.cells = static_cast<const T*>(b.cells);
    a.cutoffs = static_cast<const T*>(b.cutoffs);
    a.cutoffs_squared = static_cast<const T*>(b.cutoffs_squared);
    a.kinds = b.kinds;
    a.lj = static_cast<const LJRow<T>*>(b.lj_rows);
    a.table_of = b.table_of;
    a.tables = static_cast<const T*>(b.tables);
    a.table_start = static_cast<const T*>(b.table_start);
    a.table_spacing = static_cast<const T*>(b.table_spacing);
    a.knots = static_cast<int>(b.knots);
    a.total_bins = b.total_bins;
    a.span = b.span;
    a.bin_of = b.bin_of;
    a.bin_counts = b.bin_counts;
    a.bin_offsets = b.bin_offsets;
    a.unordered = b.unordered;
    a.order = b.order;
    a.wrapped4 = static_cast<T*>(b.wrapped4);
    a.sorted4 = static_cast<T*>(b.sorted4);
    a.scan_temp = b.scan_temp;
    a.forces = static_cast<T*>(b.forces);
    a.atom_terms = static_cast<T*>(b.atom_terms);
    a.energy = static_cast<T*>(b.energy);
    a.virial = static_cast<T*>(b.virial);
    return a;
}

template <typename T>
int table_vjp_impl(const ndn_batch& b, const void* grad_energy, const void* grad_forces,
                   const void* grad_virial, double* partials, void* grad_table, hipStream_t stream)
{
    const BatchArgs<T> a = batch_args<T>(b);
    const std::int64_t n = b.num_atoms;
    const unsigned blocks = blocks_for(n, kBlock / kWave);
    const T cull_factor = static_cast<T>(1.0 + 1e-4);
    const std::size_t lds = static_cast<std::size_t>(b.knots + 2) * sizeof(double);
    batch_table_vjp_kernel<T><<<blocks, kBlock, lds, stream>>>(
        a, cull_factor, static_cast<const T*>(grad_energy), static_cast<const T*>(grad_forces),
        static_cast<const T*>(grad_virial), partials);
    NDN_LAUNCHED("batch_table_vjp");
    // two fixed-order passes: the block rows folded onto at most 512 rows, then summed
    const int width = static_cast<int>(b.knots) + 2;
    const int folds = static_cast<int>(std::min<std::int64_t>(blocks, 512));
    table_vjp_fold_kernel<<<folds, kBlock, 0, stream>>>(partials, static_cast<int>(blocks), width);
    NDN_LAUNCHED("table_vjp_fold");
    table_vjp_finish_kernel<T><<<1, kBlock, 0, stream>>>(partials, folds,
                                                         static_cast<int>(b.knots),
                                                         static_cast<T*>(grad_table));
    NDN_LAUNCHED("table_vjp_finish");
    return NDN_OK;
}

}  // namespace ndn

using namespace ndn;

extern "C" {

int64_t ndn_batch_scan_bytes(int64_t total_bins)
{
    return scan_bytes(total_bins + 2);
}

int64_t ndn_batch_vjp_bytes(int64_t num_atoms, int64_t knots)
{
    return ((num_atoms + (kBlock / kWave) - 1) / (kBlock / kWave)) * (knots + 2) *
           static_cast<int64_t>(sizeof(double));
}

int ndn_batch_table_vjp(int dtype, const ndn_batch* batch, const void* grad_energy,
                        const void* grad_forces, const void* grad_virial, void* workspace,
                        int64_t workspace_bytes, void* grad_table, ndn_stream_t stream)
{
    NDN_REQUIRE(batch != nullptr && grad_energy && grad_forces && grad_virial && workspace &&
                    grad_table,
                "a buffer is null");
    NDN_REQUIRE(batch->tables != nullptr && batch->knots >= 4 && batch->knots + 2 <= 4096,
                "the table gradient needs tables of 4 to 4094 knots");
    NDN_REQUIRE(workspace_bytes >= ndn_batch_vjp_bytes(batch->num_atoms, batch->knots),
                "the workspace is too small");
    NDN_REQUIRE(batch->span == 1 || batch->span == 2, "span must be 1 or 2");
    if (dtype == NDN_F32)
        return table_vjp_impl<float>(*batch, grad_energy, grad_forces, grad_virial,
                                     static_cast<double*>(workspace), grad_table, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return table_vjp_impl<double>(*batch, grad_energy, grad_forces, grad_virial,
                                  static_cast<double*>(workspace), grad_table, stream);
}

int ndn_batch_efw(int dtype, const ndn_batch* batch, int schedule, ndn_stream_t stream)
{
    NDN_REQUIRE(batch != nullptr, "the batch must not be null");
    const ndn_batch& b = *batch;
    NDN_REQUIRE(b.num_systems >= 1 && b.num_atoms >= 1 && b.num_atoms <= INT32_MAX,
                "a batch needs systems and at most 2**31 - 1 atoms");
    NDN_REQUIRE(b.positions && b.system_of && b.atom_start && b.bin_start && b.dims && b.widths &&
                    b.cells && b.cutoffs && b.cutoffs_squared && b.kinds && b.lj_rows && b.table_of && b.bin_of &&
                    b.bin_counts && b.bin_offsets && b.unordered && b.order && b.wrapped4 &&
                    b.sorted4 && b.scan_temp && b.forces && b.atom_terms && b.energy &&
                    b.virial && b.status,
                "a batch buffer is null");
    NDN_REQUIRE(b.total_bins >= 1 && b.total_bins < INT32_MAX, "bad bin count");
    NDN_REQUIRE(b.scan_temp_bytes >= scan_bytes(b.total_bins + 2), "scan_temp is too small");
    NDN_REQUIRE(b.span == 1 || b.span == 2, "span must be 1 or 2");
    NDN_REQUIRE(schedule == 0 || schedule == 1, "schedule must be 0 (thread) or 1 (wave)");
    NDN_REQUIRE(b.tables == nullptr || b.knots >= 4, "tables need at least 4 knots");
    if (dtype == NDN_F32) return batch_impl<float>(b, schedule, stream);
    NDN_REQUIRE(dtype == NDN_F64, "dtype must be NDN_F32 or NDN_F64");
    return batch_impl<double>(b, schedule, stream);
}

}  // extern "C"
