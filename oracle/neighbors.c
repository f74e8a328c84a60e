/*
 * ORACLE — test infrastructure only. Nothing in the product path links, loads or calls this
 * file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * checker.
 *
 * A scalar C restatement of NablaDyn's neighbor-list operator, with the arithmetic written
 * exactly as the reference's source text states it, evaluated in the list's dtype T (float or
 * double), with no contraction (this file is compiled with -ffp-contract=off) and left-to-right
 * sums. Column order follows the reference too, so a build here can be compared with a GPU
 * build column by column, not only as a set.
 *
 * Pinned by (tests/test_oracle_*.py): a numpy brute-force pair search (independent code, fp64
 * and exact-T), the counts printed by the reference for examples/data/argon_512.npz
 * (README.md:166-172: 29,014 pairs within 1.1 nm, 21,398 within 1.0 nm, capacity 36,268),
 * invariants (each unordered pair once, padding (k mod N, k mod N), i < j on the all-pairs
 * path) and closed forms of the reuse threshold.
 *
 * Reference files followed (paths relative to src/nabladyn/neighbors/):
 *   csrc/common/traversal.cuh:12-38  AxisImage, minimum_image
 *   csrc/common/traversal.cuh:41-147 neighbor_bin, neighbor_shift, visit_neighbors
 *   csrc/grid/grid.cuh:10-118        axis_bin, flat_bin, binning, stable in-bin order by id
 *   csrc/pairs/pairs.cuh:12-157      count, fill, total, pad
 *   csrc/pairs/all_pairs.cuh:29-114  all-pairs test and fill
 *   csrc/update/update.cuh:27-145    prepare_update, check_reuse, commit
 *   _neighbor_list.py:115-140        grid dims and the all-pairs choice (host side, in Python)
 */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define STATUS_INVALID_CELL 1
#define STATUS_NON_FINITE 2
#define STATUS_OVERFLOW 8

/* One body per dtype, generated from a macro so the float and double versions cannot drift. */
#define DEFINE_ORACLE(T, SUF)                                                                     \
                                                                                                  \
    /* traversal.cuh:31-38 */                                                                     \
    static T minimum_image_##SUF(T center, T other, T length)                                     \
    {                                                                                             \
        const T raw = center - other;                                                             \
        const T half = (T)0.5 * length;                                                           \
        const T neg_half = (T)-0.5 * length;                                                      \
        if (raw > half) return (center - length) - other;                                         \
        if (raw < neg_half) return center - (other - length);                                     \
        return raw;                                                                               \
    }                                                                                             \
                                                                                                  \
    /* space.py:22-38, evaluated in T; crossings counted as the reference's int64 */              \
    void orc_wrap_##SUF(const T* x, int64_t n, const T* cell, T* out, int64_t* crossings)         \
    {                                                                                             \
        for (int64_t i = 0; i < n; ++i)                                                           \
        {                                                                                         \
            for (int a = 0; a < 3; ++a)                                                           \
            {                                                                                     \
                const T L = cell[a];                                                              \
                const T v = x[3 * i + a];                                                         \
                T k = floor(v / L);                                                               \
                T w = v - k * L;                                                                  \
                if (w < 0) { w = w + L; k = k - 1; }                                              \
                if (w >= L) { w = w - L; k = k + 1; }                                             \
                out[3 * i + a] = w;                                                               \
                if (crossings) crossings[3 * i + a] = (int64_t)k;                                 \
            }                                                                                     \
        }                                                                                         \
    }                                                                                             \
                                                                                                  \
    /* grid.cuh:10-16 */                                                                          \
    static int32_t axis_bin_##SUF(T x, T width, int32_t count)                                    \
    {                                                                                             \
        const int32_t bin = (int32_t)floor(x / width);                                            \
        return bin < 0 ? 0 : (bin >= count ? count - 1 : bin);                                    \
    }                                                                                             \
                                                                                                  \
    /* Bins the atoms and orders them by (bin, atom id): grid.cuh:23-118. Writes atom_bin [N],  \
     * sorted_ids [N], bin_offsets [bins + 1] and sorted_positions [N, 3]. */                     \
    void orc_grid_##SUF(const T* x, int32_t n, const int32_t* dims, const T* widths,              \
                        int32_t* atom_bin, int32_t* sorted_ids, int32_t* bin_offsets,             \
                        T* sorted_positions)                                                      \
    {                                                                                             \
        const int32_t bins = dims[0] * dims[1] * dims[2];                                         \
        memset(bin_offsets, 0, sizeof(int32_t) * (size_t)(bins + 1));                             \
        for (int32_t i = 0; i < n; ++i)                                                           \
        {                                                                                         \
            const int32_t bx = axis_bin_##SUF(x[3 * (int64_t)i], widths[0], dims[0]);             \
            const int32_t by = axis_bin_##SUF(x[3 * (int64_t)i + 1], widths[1], dims[1]);         \
            const int32_t bz = axis_bin_##SUF(x[3 * (int64_t)i + 2], widths[2], dims[2]);         \
            atom_bin[i] = bx + dims[0] * (by + dims[1] * bz);                                     \
            bin_offsets[atom_bin[i] + 1] += 1;                                                    \
        }                                                                                         \
        for (int32_t b = 0; b < bins; ++b) bin_offsets[b + 1] += bin_offsets[b];                  \
        int32_t* cursor = (int32_t*)malloc(sizeof(int32_t) * (size_t)(bins + 1));                 \
        memcpy(cursor, bin_offsets, sizeof(int32_t) * (size_t)(bins + 1));                        \
        /* ascending ids within a bin: a stable counting sort over ids in order */                \
        for (int32_t i = 0; i < n; ++i) sorted_ids[cursor[atom_bin[i]]++] = i;                    \
        free(cursor);                                                                             \
        for (int32_t s = 0; s < n; ++s)                                                           \
            for (int a = 0; a < 3; ++a)                                                           \
                sorted_positions[3 * (int64_t)s + a] = x[3 * (int64_t)sorted_ids[s] + a];         \
    }                                                                                             \
                                                                                                  \
    static int32_t neighbor_bin_##SUF(int32_t home, int step, int32_t count)                      \
    {                                                                                             \
        const int32_t bin = home + step;                                                          \
        return bin < 0 ? bin + count : (bin >= count ? bin - count : bin);                        \
    }                                                                                             \
                                                                                                  \
    static T neighbor_shift_##SUF(int32_t home, int step, int32_t count, T length)                \
    {                                                                                             \
        const int32_t bin = home + step;                                                          \
        return bin < 0 ? -length : (bin >= count ? length : (T)0);                                \
    }                                                                                             \
                                                                                                  \
    /* The reference's visit order for one sorted center (traversal.cuh:63-147). With out ==    \
     * NULL it only counts; otherwise it writes (id[center], id[neighbor]) from column `column` \
     * on, skipping columns at or past the capacity (pairs.cuh:23-40). Returns the count. */      \
    static int32_t visit_##SUF(const T* sp, const int32_t* ids, const int32_t* atom_bin,          \
                               const int32_t* bin_offsets, const int32_t* dims, const T* cell,    \
                               T r2, int32_t center, int64_t column, int64_t capacity,            \
                               int32_t* pairs)                                                    \
    {                                                                                             \
        const T cx = sp[3 * (int64_t)center], cy = sp[3 * (int64_t)center + 1],                   \
                cz = sp[3 * (int64_t)center + 2];                                                 \
        const int32_t nx = dims[0], ny = dims[1], nz = dims[2];                                   \
        const int32_t home = atom_bin[ids[center]];                                               \
        const int32_t hx = home % nx, hy = (home / nx) % ny, hz = home / (nx * ny);               \
        int32_t count = 0;                                                                        \
        if (nx >= 3 && ny >= 3 && nz >= 3)                                                        \
        {                                                                                         \
            for (int sz = -1; sz <= 1; ++sz)                                                      \
            {                                                                                     \
                const T shz = neighbor_shift_##SUF(hz, sz, nz, cell[2]);                          \
                const T az = shz > 0 ? cz - shz : cz, bhz = shz < 0 ? shz : (T)0;                 \
                const int32_t bz = neighbor_bin_##SUF(hz, sz, nz);                                \
                for (int sy = -1; sy <= 1; ++sy)                                                  \
                {                                                                                 \
                    const T shy = neighbor_shift_##SUF(hy, sy, ny, cell[1]);                      \
                    const T ay = shy > 0 ? cy - shy : cy, bhy = shy < 0 ? shy : (T)0;             \
                    const int32_t by = neighbor_bin_##SUF(hy, sy, ny);                            \
                    for (int sx = -1; sx <= 1; ++sx)                                              \
                    {                                                                             \
                        const T shx = neighbor_shift_##SUF(hx, sx, nx, cell[0]);                  \
                        const T ax = shx > 0 ? cx - shx : cx, bhx = shx < 0 ? shx : (T)0;         \
                        const int32_t bin = neighbor_bin_##SUF(hx, sx, nx) + nx * (by + ny * bz); \
                        const int32_t first = bin_offsets[bin], end = bin_offsets[bin + 1];       \
                        for (int32_t nb = first > center ? first : center + 1; nb < end; ++nb)    \
                        {                                                                         \
                            const T* p = sp + 3 * (int64_t)nb;                                    \
                            const T dx = ax - (p[0] + bhx);                                       \
                            const T dy = ay - (p[1] + bhy);                                       \
                            const T dz = az - (p[2] + bhz);                                       \
                            const T d2 = dx * dx + dy * dy + dz * dz;                             \
                            if (d2 < r2)                                                          \
                            {                                                                     \
                                if (pairs && column < capacity)                                   \
                                {                                                                 \
                                    pairs[column] = ids[center];                                  \
                                    pairs[capacity + column] = ids[nb];                           \
                                }                                                                 \
                                ++column;                                                         \
                                ++count;                                                          \
                            }                                                                     \
                        }                                                                         \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
            return count;                                                                         \
        }                                                                                         \
        const int span_x = nx >= 3 ? 3 : nx, span_y = ny >= 3 ? 3 : ny, span_z = nz >= 3 ? 3 : nz;\
        for (int iz = 0; iz < span_z; ++iz)                                                       \
        {                                                                                         \
            const int32_t bz = nz >= 3 ? neighbor_bin_##SUF(hz, iz - 1, nz) : iz;                 \
            for (int iy = 0; iy < span_y; ++iy)                                                   \
            {                                                                                     \
                const int32_t by = ny >= 3 ? neighbor_bin_##SUF(hy, iy - 1, ny) : iy;             \
                for (int ix = 0; ix < span_x; ++ix)                                               \
                {                                                                                 \
                    const int32_t bin =                                                           \
                        (nx >= 3 ? neighbor_bin_##SUF(hx, ix - 1, nx) : ix) + nx * (by + ny * bz);\
                    const int32_t first = bin_offsets[bin], end = bin_offsets[bin + 1];           \
                    for (int32_t nb = first > center ? first : center + 1; nb < end; ++nb)        \
                    {                                                                             \
                        const T* p = sp + 3 * (int64_t)nb;                                        \
                        const T dx = minimum_image_##SUF(cx, p[0], cell[0]);                      \
                        const T dy = minimum_image_##SUF(cy, p[1], cell[1]);                      \
                        const T dz = minimum_image_##SUF(cz, p[2], cell[2]);                      \
                        const T d2 = dx * dx + dy * dy + dz * dz;                                 \
                        if (d2 < r2)                                                              \
                        {                                                                         \
                            if (pairs && column < capacity)                                       \
                            {                                                                     \
                                pairs[column] = ids[center];                                      \
                                pairs[capacity + column] = ids[nb];                               \
                            }                                                                     \
                            ++column;                                                             \
                            ++count;                                                              \
                        }                                                                         \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
        }                                                                                         \
        return count;                                                                             \
    }                                                                                             \
                                                                                                  \
    /* Counts per sorted center on the grid path (pairs.cuh:43-76). */                            \
    void orc_count_grid_##SUF(const T* sp, const int32_t* ids, const int32_t* atom_bin,           \
                              const int32_t* bin_offsets, const int32_t* dims, const T* cell,     \
                              double radius, int32_t n, int32_t* counts)                          \
    {                                                                                             \
        const T r = (T)radius;                                                                    \
        const T r2 = r * r;                                                                       \
        for (int32_t c = 0; c < n; ++c)                                                           \
            counts[c] = visit_##SUF(sp, ids, atom_bin, bin_offsets, dims, cell, r2, c, 0, 0, 0);  \
    }                                                                                             \
                                                                                                  \
    /* Fills from each center's first column (pairs.cuh:78-112). */                               \
    void orc_fill_grid_##SUF(const T* sp, const int32_t* ids, const int32_t* atom_bin,            \
                             const int32_t* bin_offsets, const int32_t* dims, const T* cell,      \
                             double radius, int32_t n, const int32_t* offsets, int64_t capacity,  \
                             int32_t* pairs)                                                      \
    {                                                                                             \
        const T r = (T)radius;                                                                    \
        const T r2 = r * r;                                                                       \
        for (int32_t c = 0; c < n; ++c)                                                           \
            visit_##SUF(sp, ids, atom_bin, bin_offsets, dims, cell, r2, c, offsets[c], capacity,  \
                        pairs);                                                                   \
    }                                                                                             \
                                                                                                  \
    /* The all-pairs path (all_pairs.cuh:29-114): counts, and with pairs != NULL fills. */        \
    void orc_all_pairs_##SUF(const T* x, const T* cell, double radius, int32_t n,                 \
                             int32_t* counts, const int32_t* offsets, int64_t capacity,           \
                             int32_t* pairs)                                                      \
    {                                                                                             \
        const T r = (T)radius;                                                                    \
        const T r2 = r * r;                                                                       \
        for (int32_t i = 0; i < n; ++i)                                                           \
        {                                                                                         \
            int64_t column = offsets ? offsets[i] : 0;                                            \
            int32_t count = 0;                                                                    \
            for (int32_t j = i + 1; j < n; ++j)                                                   \
            {                                                                                     \
                const T dx = minimum_image_##SUF(x[3 * (int64_t)i], x[3 * (int64_t)j], cell[0]);  \
                const T dy =                                                                      \
                    minimum_image_##SUF(x[3 * (int64_t)i + 1], x[3 * (int64_t)j + 1], cell[1]);   \
                const T dz =                                                                      \
                    minimum_image_##SUF(x[3 * (int64_t)i + 2], x[3 * (int64_t)j + 2], cell[2]);   \
                if (dx * dx + dy * dy + dz * dz < r2)                                             \
                {                                                                                 \
                    if (pairs && column < capacity)                                               \
                    {                                                                             \
                        pairs[column] = i;                                                        \
                        pairs[capacity + column] = j;                                             \
                    }                                                                             \
                    ++column;                                                                     \
                    ++count;                                                                      \
                }                                                                                 \
            }                                                                                     \
            if (counts) counts[i] = count;                                                        \
        }                                                                                         \
    }                                                                                             \
                                                                                                  \
    /* prepare_update_kernel (update.cuh:27-75). Returns nothing; writes rebuild, status, dims, \
     * widths and reuse exactly as the kernel does, in T. */                                      \
    void orc_prepare_update_##SUF(const T* cell, const T* reference_cell, double cutoff_d,        \
                                  double radius_d, const int32_t* limits, int32_t* dims,          \
                                  T* widths, T* reuse, int32_t* rebuild, int32_t* status)         \
    {                                                                                             \
        const T cutoff = (T)cutoff_d, radius = (T)radius_d;                                       \
        const T epsilon = sizeof(T) == 4 ? (T)1.1920928955078125e-07 : (T)2.220446049250313e-16;  \
        rebuild[0] = 0;                                                                           \
        if (status[0] != 0) return;                                                               \
        T smallest_scale = cell[0] / reference_cell[0];                                           \
        T longest = cell[0];                                                                      \
        for (int axis = 0; axis < 3; ++axis)                                                      \
        {                                                                                         \
            const T length = cell[axis];                                                          \
            if (!(isfinite(length) && length > 2 * cutoff))                                       \
            {                                                                                     \
                status[0] |= STATUS_INVALID_CELL;                                                 \
                return;                                                                           \
            }                                                                                     \
            const T scale = length / reference_cell[axis];                                        \
            reuse[1 + axis] = scale;                                                              \
            smallest_scale = scale < smallest_scale ? scale : smallest_scale;                     \
            longest = length > longest ? length : longest;                                        \
            const T fit = floor(length / (radius + 4 * epsilon * length));                        \
            const int32_t count =                                                                 \
                fit < 1 ? 1 : (fit > limits[axis] ? limits[axis] : (int32_t)fit);                 \
            dims[axis] = count;                                                                   \
            widths[axis] = length / count;                                                        \
        }                                                                                         \
        const T allowance = radius * smallest_scale - cutoff - 32 * epsilon * longest;            \
        if (allowance <= 0)                                                                       \
        {                                                                                         \
            rebuild[0] = 1;                                                                       \
            return;                                                                               \
        }                                                                                         \
        reuse[0] = allowance * allowance / 4;                                                     \
    }                                                                                             \
                                                                                                  \
    /* check_reuse_kernel (update.cuh:79-116). */                                                 \
    void orc_check_reuse_##SUF(const T* x, const T* x0, const T* cell, const T* reuse,            \
                               int32_t n, int32_t* rebuild, int32_t* status)                      \
    {                                                                                             \
        if (status[0] != 0) return;                                                               \
        const int32_t asked = rebuild[0];                                                         \
        for (int32_t i = 0; i < n; ++i)                                                           \
        {                                                                                         \
            const T* p = x + 3 * (int64_t)i;                                                      \
            const T* q = x0 + 3 * (int64_t)i;                                                     \
            if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2])))                            \
            {                                                                                     \
                status[0] |= STATUS_NON_FINITE;                                                   \
                continue;                                                                         \
            }                                                                                     \
            if (asked != 0) continue;                                                             \
            T moved = 0;                                                                          \
            for (int axis = 0; axis < 3; ++axis)                                                  \
            {                                                                                     \
                const T d = minimum_image_##SUF(p[axis], reuse[1 + axis] * q[axis], cell[axis]);  \
                moved += d * d;                                                                   \
            }                                                                                     \
            if (moved >= reuse[0]) rebuild[0] = 1;                                                \
        }                                                                                         \
    }

DEFINE_ORACLE(float, f32)
DEFINE_ORACLE(double, f64)

/* Exclusive scan of int32 counts into int32 offsets (the reference's CUB scan). */
void orc_exclusive_scan(const int32_t* in, int64_t n, int32_t* out)
{
    int32_t running = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        out[i] = running;
        running += in[i];
    }
}
