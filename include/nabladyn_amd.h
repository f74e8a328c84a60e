/*
 * nabladyn_amd — C ABI of the MI355X (gfx950) neighbor-list-and-pair-force library.
 *
 * Every entry point takes plain device pointers, element counts and a hipStream_t, launches
 * its kernels on that stream, never synchronises and never allocates (callers own every
 * buffer; sizes come from the *_bytes queries). Argument errors return NDN_ERR_ARGUMENT (or
 * NDN_ERR_HIP for a launch failure) and leave a message for ndn_last_error(); errors in the
 * data (a bad cell, a non-finite coordinate, too many pairs) never fail a call: they set
 * latched bits in the list's device status word, as the reference does.
 *
 * dtype tags: NDN_F32 = float, NDN_F64 = double, for every "T" pointer below.
 */
#ifndef NABLADYN_AMD_H
#define NABLADYN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* ndn_stream_t; /* a hipStream_t */

enum { NDN_OK = 0, NDN_ERR_ARGUMENT = 1, NDN_ERR_HIP = 2 };
enum { NDN_F32 = 0, NDN_F64 = 1 };

/* Status bits, latched until allocate (reference: neighbors/csrc/common/common.cuh:13-16). */
enum {
    NDN_STATUS_INVALID_CELL = 1, /* a cell length not finite or at most twice the cutoff */
    NDN_STATUS_NON_FINITE = 2,   /* a coordinate is not finite */
    NDN_STATUS_OVERFLOW = 8,     /* the pairs exceed the capacity; status[1] = pairs needed */
    NDN_STATUS_FULL_OVERFLOW = 16 /* the full (both-orientation) list exceeds its capacity;
                                     status[1] is unchanged; only lists with a full list set it */
};

/* The device buffers of one neighbor list, in the reference op's argument order
 * (neighbors/csrc/ops.cpp:13-17), plus their element counts. `bits` is NULL (bits_words 0)
 * on the grid path. The full_* fields describe the optional full list (both orientations of
 * every pair, sliced-ELL by row groups of 64) that the fused force kernels read; leave them
 * NULL / 0 for the reference's half list alone. */
typedef struct ndn_neighbor_buffers {
    int32_t* dims;           /* [3] */
    void* widths;            /* [3] T */
    int32_t* atom_bin;       /* [N] */
    int32_t* bin_counts;     /* [bin_slots] */
    int32_t* bin_offsets;    /* [bin_slots] */
    int32_t* unordered_ids;  /* [N] */
    int32_t* sorted_ids;     /* [N] */
    void* sorted_positions;  /* [N, 3] T, the positions in (bin, atom id) order */
    uint8_t* scan_temp;      /* [scan_temp_bytes] */
    int32_t* pair_counts;    /* [N] */
    int32_t* pair_offsets;   /* [N] */
    int32_t* pair_total;     /* [1] */
    int32_t* status;         /* [2] */
    int32_t* pairs;          /* [2, capacity] */
    int32_t* bits;           /* [bits_words] = N * ceil(N / 32), or NULL */
    int64_t num_atoms;
    int64_t bin_slots;       /* bins of the first grid + 1 */
    int64_t scan_temp_bytes;
    int64_t capacity;
    int64_t bits_words;
    /* optional full list */
    int32_t* full_counts;    /* [N] neighbors of each row, both orientations */
    int32_t* full_slices;    /* [ceil(N / 64) + 1] exclusive scan of 64 * slice width */
    int32_t* full_index;     /* [full_capacity] neighbor atom ids, column-major per slice */
    int64_t full_capacity;
} ndn_neighbor_buffers;

/* Message of the last failed call on this thread ("" if none). */
const char* ndn_last_error(void);

/* Scratch bytes for a build scanning up to `items` int32 values; write it to *bytes.
 * Replaces torch.ops.nabladyn_neighbors.scan_temp_bytes (neighbors/csrc/ops.h:13-17,
 * runtime/entry_points.cu:284-301). */
int ndn_scan_temp_bytes(int64_t items, int64_t* bytes);

/* allocate, part one: grid and per-atom pair counts of wrapped positions [N, 3] in cell [3].
 * Replaces nabladyn_neighbors.count_first_frame (ops.h:19-38, entry_points.cu:303-352). */
int ndn_count_first_frame(int dtype, const void* positions, const void* cell, double radius,
                          const ndn_neighbor_buffers* buffers, ndn_stream_t stream);

/* allocate, part two: the pairs into the capacity the host chose, the total and padding.
 * Replaces nabladyn_neighbors.fill_first_frame (ops.h:40-59, entry_points.cu:354-399). */
int ndn_fill_first_frame(int dtype, const void* positions, const void* cell, double radius,
                         const ndn_neighbor_buffers* buffers, ndn_stream_t stream);

/* The in-step update: reuse check, then a rebuild gated by a device flag, then the new
 * reference frame. Graph-capturable: no sync, no allocation.
 * Replaces nabladyn_neighbors.update (ops.h:61-90, entry_points.cu:401-508). */
int ndn_update(int dtype, const void* positions, const void* cell, double cutoff, double radius,
               void* reference_positions, void* reference_cell, const int32_t* limits,
               void* reuse, int32_t* rebuild, int32_t* rebuild_count,
               const ndn_neighbor_buffers* buffers, ndn_stream_t stream);

/* Wrap positions [N, 3] into [0, L) per axis, adding lattice crossings to images [N, 3] int64
 * (images may be NULL). Replaces space.wrap (space.py:22-38); out may alias positions. */
int ndn_wrap(int dtype, const void* positions, const void* cell, int64_t num_atoms, void* out,
             int64_t* images, ndn_stream_t stream);

/* ---- pair potentials -------------------------------------------------------------------- */

enum { NDN_LJ = 0, NDN_TABLE = 1 };

/* One pair potential of num_types atom types, P = T (T + 1) / 2 rows in pair_index order
 * (potential/_lennard_jones/_potential.py:143-151). LJ reads sigma, epsilon [P] T
 * (potential/_lennard_jones/_potential.py:113-120); a table reads table [P, K] T with its last
 * column 0, start [P] T and spacing [P] T (potential/_tabulated/_potential.py:59-91). The
 * cutoff is in nm; atom_types [N] int64 may be NULL with one type. */
typedef struct ndn_potential {
    int kind;
    int num_types;
    const int64_t* atom_types;
    const void* sigma;
    const void* epsilon;
    const void* table;
    const void* start;
    const void* spacing;
    int64_t knots;
    double cutoff;
} ndn_potential;

/* Workspace bytes of ndn_pair_efw for a list of `columns` columns. */
int ndn_pair_workspace_bytes(int dtype, const ndn_potential* pot, int64_t columns, int64_t* bytes);

/* Energy [1], forces [N, 3] and virial [3, 3] of wrapped positions [N, 3] in cell [3] over the
 * half list pairs [2, C] (self-pairs and pairs past the cutoff count nothing). Any of energy,
 * forces, virial may be NULL. Replaces quantity.energy_force_virial(model._energy_wrapped)
 * (quantity.py:93-117) for LJ and tabulated models: E, F = -dE/dx, W = -dE/dstrain at zero
 * strain = sum over pairs of r_ij (x) F_ij. */
int ndn_pair_efw(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                 const int32_t* pairs, int64_t capacity, const ndn_potential* pot, void* energy,
                 void* forces, void* virial, void* workspace, int64_t workspace_bytes,
                 ndn_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NABLADYN_AMD_H */
