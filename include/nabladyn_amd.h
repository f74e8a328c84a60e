/*
 * nabladyn_amd — C ABI of the MI355X (gfx950) neighbor-list-and-pair-force library.
 *
 * Every entry point takes plain device pointers, element counts and a hipStream_t, launches
 * its kernels on that stream and never allocates (callers own every buffer; sizes come from the
 * *_bytes queries and sizing calls); none synchronises but ndn_md_batch_prepare, which copies a
 * host-built table. Argument errors return NDN_ERR_ARGUMENT (or
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
    NDN_STATUS_FULL_OVERFLOW = 16, /* the full (both-orientation) list exceeds its capacity;
                                      status[1] is unchanged; only lists with a full list set it */
    NDN_STATUS_HALO_OVERFLOW = 32, /* (MD list) a brick's halo exceeds halo_capacity; the
                                      largest halo is at byte 656 of scan_temp */
    NDN_STATUS_CELL_SHRANK = 64    /* (MD list, NPT) the cell shrank below what the fixed
                                      sub-bin grid covers (widths under half the radius) */
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
 * cutoff is in nm; atom_types [N] int64 may be NULL with one type.
 *
 * next: NULL, or the second model of a sum of two (potential.sum, potential/_sum/_potential.py:
 * 86-99), for the ndn_md_* entry points only; every other entry point refuses it. The two share
 * num_types and atom_types, and the MD step evaluates both in one walk of its list, each pair
 * masked at each model's own cutoff, their energies and forces added; the list's cutoff is the
 * larger of the two. */
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
    const struct ndn_potential* next;
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

/* ndn_pair_efw over a full list (every pair within the list radius in both orientations, as
 * ndn_fill_first_frame / ndn_update write it with full_* buffers): row r is atom rows[r] (rows
 * NULL: atom r), its k-th neighbor the atom id index[slices[r / 64] + r % 64 + 64 k] for
 * k < counts[r]; `capacity` entries of index. A lane per row sums the row's force in registers
 * and writes it once (no atomics); E and W take half of each of a pair's two meetings. Same
 * results as ndn_pair_efw on the list's half pairs, to rounding. Replaces
 * quantity.energy_force_virial (quantity.py:93-117) for LJ and tabulated models when the list
 * keeps a full list. The workspace is ndn_pair_workspace_bytes. */
int ndn_pair_efw_full(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                      const int32_t* rows, const int32_t* counts, const int32_t* slices,
                      const int32_t* index, int64_t capacity, const ndn_potential* pot,
                      void* energy, void* forces, void* virial, void* workspace,
                      int64_t workspace_bytes, ndn_stream_t stream);

/* ndn_pair_efw at a strain [3, 3] (T, row-major, device memory): each pair's minimum-image
 * vector d is deformed to (I + strain) d as space.strained does (space.py:51-64), and
 * energy = E(strain), forces = -dE/dx, virial = -dE/dstrain at that strain (not symmetric in
 * general; at zero strain it is ndn_pair_efw's). Replaces the strained call of
 * Potential.forward (potential/_base.py:82-106) differentiated by torch.func, as quantity.py:21-32
 * sets it up. The workspace is ndn_pair_workspace_bytes. */
int ndn_pair_efw_strained(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                          const int32_t* pairs, int64_t capacity, const ndn_potential* pot,
                          const void* strain, void* energy, void* forces, void* virial,
                          void* workspace, int64_t workspace_bytes, ndn_stream_t stream);

/* dE/dparameters over the half list in fp64, at `strain` [3, 3] (device, may be NULL for
 * none): LJ out [2P] = dE/dsigma then dE/depsilon per type-pair row; a table out [P, K] =
 * dE/dtable (the last knot's is 0). Replaces autograd of LennardJones._energy_wrapped
 * (potential/_lennard_jones/_potential.py:209-234) and Tabulated._energy_wrapped
 * (potential/_tabulated/_potential.py:146-162) with respect to their parameters. The workspace
 * is ndn_pair_workspace_bytes. */
int ndn_pair_parameter_gradient(int dtype, const void* positions, const void* cell,
                                int64_t num_atoms, const int32_t* pairs, int64_t capacity,
                                const ndn_potential* pot, const void* strain, double* out,
                                int64_t out_count, void* workspace, int64_t workspace_bytes,
                                ndn_stream_t stream);

/* The vector-Jacobian product of (E, F, W) = ndn_pair_efw(...) with respect to the parameters,
 * in fp64: out = gE dE/dtheta + sum gF . dF/dtheta + sum gW : dW/dtheta, given the upstream
 * grad_energy [1], grad_forces [N, 3] and grad_virial [3, 3] (T, device; any may be NULL for
 * zero). Layout of out as for ndn_pair_parameter_gradient. It is the exact mixed second
 * derivative of force and virial matching, per pair (F_i = sum_j fs d_ij, W = sum fs d (x) d):
 * gE du/dtheta + ((gF_i - gF_j) . d + d^T gW d) dfs/dtheta. Replaces the nested
 * torch.func.grad of the reference's fit (training/_loss.py:71-125, training/_fit.py:245-271)
 * for LJ and tabulated models. The workspace is ndn_pair_workspace_bytes. */
int ndn_pair_parameter_vjp(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                           const int32_t* pairs, int64_t capacity, const ndn_potential* pot,
                           const void* grad_energy, const void* grad_forces,
                           const void* grad_virial, double* out, int64_t out_count,
                           void* workspace, int64_t workspace_bytes, ndn_stream_t stream);

/* Workspace bytes of ndn_pair_efw_frames (and enough for ndn_pair_parameter_vjp_frames) for F
 * frames (1 <= F <= 65535) of `columns` columns each. */
int ndn_pair_frames_workspace_bytes(int dtype, const ndn_potential* pot, int64_t frames,
                                    int64_t columns, int64_t* bytes);

/* ndn_pair_efw over F frames of one system in one launch: wrapped positions [F, N, 3], cells
 * [F, 3], pair stacks [F, 2, C] (each row a half list padded with self-pairs, as the fit's
 * pair_stack writes it); energy [F], forces [F, N, 3], virial [F, 3, 3], any may be NULL. Each
 * frame's results equal ndn_pair_efw's on that frame. Replaces torch.func.vmap of the per-frame
 * channels in the reference's fit (training/_loss.py:71-102). The workspace is
 * ndn_pair_frames_workspace_bytes. */
int ndn_pair_efw_frames(int dtype, const void* positions, const void* cells, int64_t frames,
                        int64_t num_atoms, const int32_t* stacks, int64_t capacity,
                        const ndn_potential* pot, void* energy, void* forces, void* virial,
                        void* workspace, int64_t workspace_bytes, ndn_stream_t stream);

/* ndn_pair_efw_frames over full rows instead of the stacks: frame f's rows are its atoms in
 * order, row r's k-th neighbor index[f * capacity + slices[f * (S + 1) + r / 64] + r % 64 + 64 k]
 * for k < counts[f * N + r] (S = ceil(N / 64); the layout of ndn_pair_efw_full per frame, as the
 * fit builds it once from the frames' stacks). A lane per row sums its force in registers: no
 * atomics. Same results as ndn_pair_efw_frames to rounding. The workspace is
 * ndn_pair_frames_workspace_bytes with columns = N. */
int ndn_pair_efw_full_frames(int dtype, const void* positions, const void* cells, int64_t frames,
                             int64_t num_atoms, const int32_t* counts, const int32_t* slices,
                             const int32_t* index, int64_t capacity, const ndn_potential* pot,
                             void* energy, void* forces, void* virial, void* workspace,
                             int64_t workspace_bytes, ndn_stream_t stream);

/* ndn_pair_parameter_vjp over the F frames of ndn_pair_efw_frames, summed into one out: upstream
 * grad_energy [F], grad_forces [F, N, 3], grad_virial [F, 3, 3] (any may be NULL). Replaces the
 * vmapped nested grad of training/_loss.py:71-125 with training/_fit.py:245-271. The workspace
 * is ndn_pair_workspace_bytes of C columns (or the frames' size). */
int ndn_pair_parameter_vjp_frames(int dtype, const void* positions, const void* cells,
                                  int64_t frames, int64_t num_atoms, const int32_t* stacks,
                                  int64_t capacity, const ndn_potential* pot,
                                  const void* grad_energy, const void* grad_forces,
                                  const void* grad_virial, double* out, int64_t out_count,
                                  void* workspace, int64_t workspace_bytes, ndn_stream_t stream);

/* ndn_pair_efw without the workspace argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Its scratch (the potential's coefficients, then 64 bytes per
 * block of the pair launch, at most 2,048 blocks) is a static device array of
 * NDN_PAIR_EFW_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a call whose
 * potential and list need more is refused with NDN_ERR_ARGUMENT. It calls no allocator.
 * Nothing in the scratch needs clearing between calls (each call writes what it reads), but
 * every call on one GPU shares that GPU's array: calls that overlap on one GPU (two streams,
 * two threads) race and must be serialised by the caller. */
#define NDN_PAIR_EFW_NOALLOC_SCRATCH_BYTES (1 << 20)
int ndn_pair_efw_noalloc(int dtype, const void* positions, const void* cell, int64_t num_atoms,
                         const int32_t* pairs, int64_t capacity, const ndn_potential* pot,
                         void* energy, void* forces, void* virial, ndn_stream_t stream);

/* ---- molecular dynamics ----------------------------------------------------------------- */

/* The resident frame of a simulation (md/_state.py:8-21), kept in cell order ("slots"): every
 * per-atom array below is permuted at each rebuild, and ndn_md_list.ids maps a slot to its atom.
 * thermostat [14] f64: chain positions [3], momenta [3], the closing velocity factor still to be
 * applied (1 when settled), the opening factor of the next step, and the chain after that step's
 * opening half step (positions [3], momenta [3]; a step's finish commits them before its closing
 * half step, then opens the next step). kinetic [1] f64 is the K the chain
 * reads (md/_ensembles/_nvt.py:47-49). partials [partial_slots * 8] f64 and coefficients
 * [coefficient_bytes] are scratch sized by ndn_md_workspace. virial may be NULL. */
typedef struct ndn_md_system {
    int64_t num_atoms;
    void* positions;            /* [N, 3] T, wrapped */
    void* velocities;           /* [N, 3] T */
    void* forces;               /* [N, 3] T */
    int64_t* images;            /* [N, 3] */
    void* masses;               /* [N] T */
    void* inverse_masses;       /* [N] T, 1 / m */
    const void* cell;           /* [3] T */
    void* energy;               /* [1] T */
    void* virial;               /* [3, 3] T or NULL */
    double* thermostat;         /* [14] */
    const double* chain_masses; /* [3]: g kT tau^2, kT tau^2, kT tau^2 */
    double* kinetic;            /* [1] */
    double* partials;
    int64_t partial_slots;
    void* coefficients;
    int64_t coefficient_bytes;
    double* barostat;           /* NPT: [16] (box momentum, its chain's q [3], p [3] and masses
                                   [3], the master cell [3], this step's drag and growth);
                                   NULL otherwise. The cell is then rewritten every step. */
} ndn_md_system;

typedef struct ndn_md_settings {
    double timestep;              /* ps */
    double kT;                    /* kJ/mol (NVT) */
    int64_t degrees_of_freedom;   /* 3N - 3 */
    int thermostat;               /* 0: NVE, 1: NVT with a Nose-Hoover chain of 3 */
    int virial;                   /* also compute W each step (NPT: required) */
    /* NPT, isotropic MTK (md/_ensembles/_npt.py:66-266), used when system.barostat is set */
    double pressure;              /* kJ/mol/nm^3 */
    double box_mass;              /* 3 (N + 1) kT tau_b^2 */
    double alpha;                 /* 1 + 1/N */
    /* 1: this step follows a step of the same run on the stream with nothing in between (a
     * captured run's later steps), so it skips the prologue launch: the previous step's finish
     * opened it. Without a barostat only; 0 for a first step. */
    int continues;
} ndn_md_settings;

/* The reuse state of a neighbor list (neighbors/_neighbor_list.py:42-56). */
typedef struct ndn_list_state {
    double cutoff;
    double radius;
    void* reference_positions;    /* [N, 3] T, slot order */
    void* reference_cell;         /* [3] T */
    const int32_t* limits;        /* [3] */
    void* reuse;                  /* [4] T */
    int32_t* rebuild;             /* [1] */
    int32_t* rebuild_count;       /* [1] */
} ndn_list_state;

/* The MD step's own neighbor structure: sub-bins at least radius / 2 wide (dims fixed for the
 * run), and a full list (both orientations of every pair within the radius) in fixed-stride rows,
 * one per slot. The slots fall into bricks of brick[0] x brick[1] x brick[2] sub-bins (brick[2]
 * divides 8; brick[1] * brick[2] <= 64; (brick[1] + 4) * (brick[2] + 4) <= 128); a brick's halo
 * is every sub-bin within 2 of it (periodic), listed once per build in halo_slots, and a row
 * holds 16-bit local indices into its brick's halo: 0 the sentinel (a point far from every
 * atom) that pads the row up to its count rounded up to 32, k >= 1 the slot
 * halo_slots[brick * halo_capacity + k - 1], in no set order. stride is a multiple of 64. `group` lanes of a wave read one row in the force
 * kernel (4, 8 or 16). status [2] latches errors as the neighbor list's does, with
 * NDN_STATUS_FULL_OVERFLOW when a row needs more than `stride` entries (status[1] = the width
 * needed) and NDN_STATUS_HALO_OVERFLOW when a halo needs more than halo_capacity. scan_temp
 * holds a 1024-byte header and the scan tiles. */
typedef struct ndn_md_list {
    int32_t dims[3];
    int32_t group;
    double widths[3];
    int64_t stride;
    int32_t* bin_counts;        /* [bins + 1], zero between builds */
    int32_t* bin_offsets;       /* [bins + 1] */
    int32_t* bin_of;            /* [N] */
    int32_t* unordered;         /* [N] */
    int32_t* order;             /* [N] new slot -> previous slot */
    int32_t* counts;            /* [N] neighbors of each slot */
    uint16_t* index;            /* [N * stride] local indices, row-major, each row padded with 0
                                   to its count rounded up to 32; NULL while stride is 0 */
    int32_t* ids;               /* [N] atom id of each slot */
    int64_t* types;             /* [N] atom type of each slot, or NULL with one type */
    void* positions4;           /* [N + 1, 4] T: x, y, z, type of each slot at the last build,
                                   the build's own copy (row N spare) */
    void* scratch;              /* [ndn_md_scratch_bytes] */
    uint8_t* scan_temp;
    int64_t scan_temp_bytes;
    int32_t* status;            /* [2] */
    int32_t* prepare_dims;      /* [3] scratch of the reuse rule */
    void* prepare_widths;       /* [3] T scratch */
    int32_t* prepare_total;     /* [1] scratch */
    /* The pruned (inner) list: the pairs of the full list within inner_radius = cutoff + a
     * smaller skin, refreshed from the full list after every build and whenever an atom has
     * moved half that skin since the last refresh; the force kernel reads it instead of the full
     * rows. inner_index NULL (inner_radius 0) reads the full rows. */
    double inner_radius;
    int32_t* inner_counts;      /* [N] */
    uint16_t* inner_index;      /* [N * stride] local indices as in `index`, padded likewise */
    void* prune_positions;      /* [N, 3] T: the frame at the last refresh */
    /* bricks and their halos */
    int32_t brick[3];           /* sub-bins per brick along x, y, z */
    int64_t bricks;             /* ceil(dims[k] / brick[k]) multiplied over k */
    int64_t halo_capacity;      /* slots per brick in halo_slots; 0 while sizing */
    int32_t* halo_slots;        /* [bricks * halo_capacity] */
    int32_t* halo_runs;         /* [bricks * 128] local start of each run of the halo */
    int32_t* halo_counts;       /* [bricks] */
    void* grid_widths;          /* [3] T device: the sub-bin widths cell / dims (NPT: the
                                   prologue rewrites them each step) */
} ndn_md_list;

/* Scratch sizes of a system: partial slots and potential-coefficient bytes. */
int ndn_md_workspace(int dtype, const ndn_potential* pot, int64_t num_atoms,
                     int64_t* partial_slots, int64_t* coefficient_bytes);

/* The largest halo_capacity the current device's force blocks hold (the halo's positions in
 * LDS: what one block may take, at most half a CU's). No reference counterpart. */
int64_t ndn_md_halo_limit(int dtype);

/* Bytes of ndn_md_list.scratch for N atoms. */
int64_t ndn_md_scratch_bytes(int dtype, int64_t num_atoms);

/* The list build of the MD path: sort into sub-bins, permute the frame, list the bricks' halos
 * and fill the rows (count only while stride or halo_capacity is 0), store the reference frame.
 * With `forced` it always runs (then the pruned rows wait for ndn_md_refresh); otherwise it is
 * gated on list->rebuild as inside a step. */
int ndn_md_rebuild(int dtype, const ndn_md_system* system, const ndn_list_state* list,
                   const ndn_md_list* md_list, int forced, ndn_stream_t stream);

/* One MD step (NVE or NVT), launched on the stream with no sync or allocation, so a HIP graph
 * can capture it: chain half step, first half kick + drift + wrap, the reuse check and the gated
 * rebuild, the gated refresh of the pruned list, forces over the pruned list with the second
 * half kick and K, then the chain's closing half step (E is not summed: `energy` keeps the last
 * evaluation's). Replaces one replay of the reference's captured step
 * (md/_simulation.py:316-364: first_half, NeighborList._update_wrapped, second_half, closing). */
int ndn_md_step(int dtype, const ndn_md_system* system, const ndn_md_settings* settings,
                const ndn_potential* pot, const ndn_list_state* list, const ndn_md_list* md_list,
                ndn_stream_t stream);

/* E (and W when `virial`) and F of the current frame over the full list; with `kick` also the
 * second half kick of `timestep`. The first frame, the sampled panel (md/_simulation.py:177-180,
 * 366-402), and the force launch alone for timing. */
int ndn_md_evaluate(int dtype, const ndn_md_system* system, const ndn_potential* pot,
                    const ndn_md_list* md_list, int virial, int kick, double timestep,
                    ndn_stream_t stream);

/* The force launch alone (the step's dominant kernel), on potential coefficients a step or an
 * evaluate already prepared, so events can time it; with `kick` it also advances the
 * velocities by the second half kick. No reference counterpart. */
int ndn_md_force(int dtype, const ndn_md_system* system, const ndn_potential* pot,
                 const ndn_md_list* md_list, int virial, int kick, double timestep,
                 ndn_stream_t stream);

/* E and F of the current frame over the full rows, rewriting the pruned rows from them: the
 * first evaluation after a forced ndn_md_rebuild when the list has a pruned list (a step
 * refreshes by itself). The refresh of md/_simulation.py's first frame (md/_simulation.py:177-180);
 * the reference has no pruned list. */
int ndn_md_refresh(int dtype, const ndn_md_system* system, const ndn_potential* pot,
                   const ndn_md_list* md_list, ndn_stream_t stream);

/* The list's rows as slots, for readout and tests: out[slot * stride + k] is the k-th neighbor
 * of `slot` for k < its count (the pruned rows with `pruned`); the rest is left as it is. No
 * reference counterpart (the reference's list holds atom indices already). */
int ndn_md_list_entries(const ndn_md_list* md_list, int64_t num_atoms, int pruned, int32_t* out,
                        ndn_stream_t stream);

/* ---- batches of MD systems ------------------------------------------------------------------
 * Several systems stepped together: each kernel of ndn_md_step is launched once for the whole
 * batch (the force kernel once for each kind of potential the members bind: LJ or a table, of
 * one type or several), each member getting the blocks its own step would launch, so a member
 * computes the same bits as ndn_md_step gives it alone. No reference counterpart: the reference
 * steps one system per compiled step (md/_simulation.py:316-364); a batch is that step for
 * every member. */
#define NDN_MD_BATCH_KERNELS 13

/* One member: the same structures ndn_md_step takes. They are read at preparation only. */
typedef struct ndn_md_member {
    const ndn_md_system* system;
    const ndn_md_settings* settings;
    const ndn_potential* potential;
    const ndn_list_state* list;
    const ndn_md_list* md_list;
} ndn_md_member;

typedef struct ndn_md_batch {
    void* table;          /* device memory of table_bytes bytes, or NULL to size it */
    int64_t table_bytes;  /* set by a sizing call */
    int32_t dtype;        /* NDN_F32 or NDN_F64: every member's */
    /* set by ndn_md_batch_prepare */
    int32_t count;
    int32_t kinds;        /* bit k: some member binds potential kind k (a force launch) */
    int32_t group;        /* every member's lanes per row */
    int32_t virial;       /* every member's ndn_md_settings.virial */
    int32_t continues;    /* set by the caller: 1 when every member's previous step was a
                             batch step of this batch (its finish opened this step), so the
                             prologue launch is skipped; 0 (prepare's value) runs it */
    int64_t members_offset;
    int64_t potentials_offset;
    int64_t blocks_offset;
    int64_t offsets[NDN_MD_BATCH_KERNELS];  /* each launch's block map in the table */
    int64_t grids[NDN_MD_BATCH_KERNELS];    /* each launch's blocks over all members (0: skipped) */
    int64_t lds[NDN_MD_BATCH_KERNELS];      /* each launch's dynamic LDS: its members' largest
                                               halo (and a one-type table's staged segments) */
} ndn_md_batch;

/* Checks every member as ndn_md_step does (and that all share the dtype, the group and the
 * virial setting; 1 <= count <= 256), lays out the batch's table and, when batch->table is set,
 * writes it (synchronizing the stream). With batch->table NULL it sets table_bytes only. Prepare
 * again after any member's list is resized. */
int ndn_md_batch_prepare(ndn_md_batch* batch, const ndn_md_member* members, int32_t count,
                         ndn_stream_t stream);

/* One step of every member (ndn_md_step's kernels, one launch each for the batch and one force
 * launch per kind of potential), with no sync or allocation. */
int ndn_md_batch_step(const ndn_md_batch* batch, ndn_stream_t stream);

/* Applies the pending closing velocity factor, leaving the velocities as the reference's frame
 * holds them after a step. */
int ndn_md_settle(int dtype, const ndn_md_system* system, ndn_stream_t stream);

/* Slot order -> atom order: out[ids[s] * width + k] = in[s * width + k], elements of 4 or 8
 * bytes. */
int ndn_md_to_atom_order(int element_bytes, const void* in, const int32_t* ids, int64_t num_atoms,
                         int width, void* out, ndn_stream_t stream);

/* ---- many systems in one launch ---------------------------------------------------------- */

/* A batch of S independent systems (one atom type each), N atoms in all. Every system owns a
 * contiguous range of atoms (atom_start) and of one global bin space (bin_start, its sub-bins
 * dims [3] of widths [3], each at least cutoff / span wide). Its potential is LJ (kinds = NDN_LJ,
 * lj_rows [4] = sigma^2, 4 eps, 24 eps, the energy at the cutoff) or a padded table row
 * (table_of >= 0 into tables [R, K + 2], table_start [R], table_spacing [R]). The rest
 * is scratch (bin_counts zero between calls) and outputs: forces [N, 3], atom_terms [N, 7],
 * energy [S], virial [S, 3, 3]. Positions may lie outside their cells. No reference
 * counterpart: the reference evaluates one system per call. */
typedef struct ndn_batch {
    int64_t num_systems;
    int64_t num_atoms;
    const void* positions;
    const int32_t* system_of;
    const int64_t* atom_start;
    const int32_t* bin_start;
    const int32_t* dims;
    const void* widths;
    const void* cells;
    const void* cutoffs;
    const void* cutoffs_squared; /* [S] each cutoff squared in double, then rounded to T (the
                                    reference's `squared < cutoff**2` test) */
    const int32_t* kinds;
    const void* lj_rows;
    const int32_t* table_of;
    const void* tables;
    const void* table_start;
    const void* table_spacing;
    int64_t knots;
    int64_t total_bins;
    int32_t span;
    int32_t reserved;
    int32_t* bin_of;
    int32_t* bin_counts;
    int32_t* bin_offsets;
    int32_t* unordered;
    int32_t* order;
    void* wrapped4;
    void* sorted4;
    uint8_t* scan_temp;
    int64_t scan_temp_bytes;
    int32_t* status;
    void* forces;
    void* atom_terms;
    void* energy;
    void* virial;
} ndn_batch;

/* scan_temp bytes of a batch of `total_bins` bins. */
int64_t ndn_batch_scan_bytes(int64_t total_bins);

/* E, F, W of every system of the batch: bin, sort, and one force launch over all atoms with no
 * stored list; schedule 0 = a lane per atom, 1 = a wave per atom. */
int ndn_batch_efw(int dtype, const ndn_batch* batch, int schedule, ndn_stream_t stream);

/* Workspace bytes of ndn_batch_table_vjp. */
int64_t ndn_batch_vjp_bytes(int64_t num_atoms, int64_t knots);

/* The hand-written backward of a batch of tabulated systems sharing one table (K knots): with the
 * batch binned by a preceding ndn_batch_efw, and upstream gradients gE [S], gF [N, 3], gW [S, 3, 3]
 * of its E, F and W, writes grad_table [K] = sum over systems of gE dE/dtable + gF . dF/dtable
 * + gW : dW/dtable (exact: E, F and W are linear in the knots; the last knot gets 0). Replaces the
 * reference's torch.func double backward over vmapped frames (training/_loss.py:71-125,
 * training/_fit.py:245-271) for the table's gradient. */
int ndn_batch_table_vjp(int dtype, const ndn_batch* batch, const void* grad_energy,
                        const void* grad_forces, const void* grad_virial, void* workspace,
                        int64_t workspace_bytes, void* grad_table, ndn_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NABLADYN_AMD_H */
