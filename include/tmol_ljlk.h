/* C ABI of the MI355X LJ+LK pose-scoring library (libtmol_ljlk_hip.so).
 *
 * Replaces, for one path, the reference's op
 *   tmol_ljlk::ljlk_pose_scores   tmol/score/ljlk/potentials/compiled.ops.cpp:533-597
 * and its dispatch
 *   LJLKPoseScoreDispatch<D,Real,Int>::forward   ljlk_pose_score.impl.hh:633-1158
 *   LJLKPoseScoreOp::backward (whole pose)       compiled.ops.cpp:157-174
 *   LJLKPoseScoreDispatch::backward (block pair) ljlk_pose_score.impl.hh:1160-1477
 *
 * Plain C: device pointers, sizes, an opaque hipStream_t (void*), int status.
 * All pointers named in the structs are device pointers unless noted.
 * No entry point allocates device memory: the caller passes a workspace of
 * tmol_ljlk_workspace_bytes() bytes.  Everything runs on the given stream;
 * nothing synchronises except tmol_ljlk_read_counters().
 */
#ifndef TMOL_LJLK_H
#define TMOL_LJLK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TMOL_LJLK_OK 0
#define TMOL_LJLK_EINVAL 1    /* bad sizes / pointers / flags */
#define TMOL_LJLK_EOVERFLOW 2 /* work list capacity exceeded (see read_counters) */
#define TMOL_LJLK_EHIP 3      /* HIP runtime error (launch) */
#define TMOL_LJLK_ENOTSUP 4   /* unsupported configuration */
#define TMOL_LJLK_ESCRATCH 5  /* tmol_ljlk_score_noalloc: the stack needs more scratch than
                                 TMOL_LJLK_NOALLOC_SCRATCH_BYTES */

/* coordinate / parameter dtypes */
#define TMOL_LJLK_F32 0
#define TMOL_LJLK_F64 1

/* stack layouts */
#define TMOL_LJLK_PADDED 0 /* tmol PoseStack: P poses x B blocks, dense count-pair tables */
#define TMOL_LJLK_RAGGED 1 /* CSR over poses, count pair from a sparse bonded table */

/* score flags */
#define TMOL_LJLK_DVDX 2       /* per-term dV/dx [3, NA, 3] (reference whole-pose forward) */
#define TMOL_LJLK_WGRAD 4      /* weighted gradient sum_t w[t] dV_t/dx -> [NA, 3] */
#define TMOL_LJLK_BLOCK_PAIR 8 /* energies (and weights) per block pair: padded [3, P, B, B];
                                  ragged [3, G, row_capacity] at (g1, position in row g1) */
#define TMOL_LJLK_NO_ENERGY 16 /* skip the energy output (backward re-walk) */
#define TMOL_LJLK_ACCUMULATE 32 /* add into dvdx / grad instead of overwriting */
#define TMOL_LJLK_GENERAL_KERNEL 128 /* force the general (any block size) kernel */
#define TMOL_LJLK_ROW_KERNEL 4096 /* energies / WGRAD: score block pair by block pair (the
                                     row kernel) instead of streaming each row's atoms */
#define TMOL_LJLK_DETERMINISTIC 1024 /* float32 DVDX / WGRAD (whole pose): gradients and
                                        per-pose energies summed in 64-bit fixed point in the
                                        workspace, then converted once: bit-identical results
                                        run to run (the workspace must come from
                                        tmol_ljlk_workspace_bytes_flags with this flag) */
/* launch variants of the scoring kernel, chosen by an autotuner (float32
 * whole-pose energy and weighted-gradient modes; other modes keep their
 * defaults): waves per SIMD the kernel is compiled for (3, 4 or 5; 0 = the
 * mode's default, 5 energy-only, 4 with gradients) and rows of the work list
 * taken per queue grab (1..8; 0 = 1) */
#define TMOL_LJLK_WAVES_PER_SIMD(n) ((n) << 16)
#define TMOL_LJLK_ROW_GRAB(n) ((n) << 20)

/* Per block-type tables: the PackedBlockTypes annotations the reference op
 * takes (ljlk_energy_term.py:163-178).  Maximum 32 atoms per tile. */
typedef struct {
  int32_t n_block_types;          /* T */
  int32_t max_n_atoms;            /* Amax */
  int32_t max_n_tiles;            /* nt = (Amax - 1) / 32 + 1 */
  int32_t max_n_conn;             /* C <= 4 */
  int32_t n_atom_types;           /* rows of type_params */
  int32_t dtype;                  /* of type_params / global_params */
  const int32_t* n_atoms;         /* [T] */
  const int32_t* n_heavy_in_tile; /* [T, nt] */
  const int32_t* heavy_in_tile;   /* [T, 32 * nt], -1 padded */
  const int32_t* atom_types;      /* [T, Amax], -1 padded */
  const int32_t* n_conn;          /* [T] */
  const int32_t* conn_atoms;      /* [T, C] */
  const int32_t* path_distance;   /* [T, Amax, Amax] (ljlk_bond_separation) */
  const int32_t* is_ligand;       /* [T] */
  const void* type_params;        /* [n_atom_types, 10] */
  const void* global_params;      /* [5] */
} tmol_ljlk_block_types;

/* The systems.  Blocks are addressed by a global index g.
 *   padded: g = pose * B + block (the reference's "rot" index); G = P * B
 *   ragged: g in [pose_block_start[p], pose_block_start[p + 1]) */
typedef struct {
  int32_t layout;                     /* TMOL_LJLK_PADDED / _RAGGED */
  int32_t n_poses;                    /* P */
  int32_t max_n_blocks;               /* padded: B; ragged: max blocks per pose */
  int32_t n_blocks_total;             /* G */
  int64_t n_atoms_total;              /* NA (rows of coords) */
  const int32_t* block_atom_offset;   /* [G] rot_coord_offset */
  const int32_t* block_type;          /* [G] block type, -1 = padding */
  const int32_t* block_pose;          /* ragged: [G]; padded: may be NULL */
  const int32_t* pose_block_start;    /* ragged: [P + 1]; padded: NULL */
  const int32_t* atom_pose;           /* [NA] pose_ind_for_atom */
  const int32_t* min_bondsep;         /* padded: [P, B, B] */
  const int32_t* inter_block_bondsep; /* padded: [P, B, B, C, C] */
  const int32_t* bonded_start;        /* ragged: [G + 1] CSR offsets into bonded */
  const int32_t* bonded;              /* ragged: [n, 8] forward bonded table, see below */
} tmol_ljlk_stack;

/* Ragged count-pair connectivity, for any bond topology (the sparse form of
 * the padded inter_block_bondsep, pose_stack_builder.py:1271-1349): rows
 * bonded_start[g] .. bonded_start[g + 1] - 1 of `bonded` list the blocks
 * g2 > g of g's pose whose minimum connection-point separation is at most
 * 5, sorted by g2.  Row: {g2, min separation, 0, 0, sep[0], sep[1], sep[2],
 * sep[3]}, where byte c2 of sep[c1] is the bond separation (capped at 6)
 * between connection c1 of g and connection c2 of g2.  Absent pairs are 6
 * bonds or more apart (count-pair weight 1, reported separation 6). */
#define TMOL_LJLK_BONDED_WIDTH 8

/* Neighbor-pass tiles: 64 x 64 block tiles of each pose's upper triangle.
 * Each tile is int32[4] = {row_start_g, col_start_g, pose_end_g, pose}. */
int64_t tmol_ljlk_count_neighbor_tiles(const int32_t* pose_block_start_host,
                                       int32_t n_poses);
int tmol_ljlk_fill_neighbor_tiles(const int32_t* pose_block_start_host,
                                  int32_t n_poses, int32_t* tiles_host);

/* Workspace bytes for a stack, block-type set and work-list row capacity
 * (the most neighbor blocks g2 >= g1 any block g1 may have; a padded stack
 * never needs more than B). */
int64_t tmol_ljlk_workspace_bytes(const tmol_ljlk_stack* stack,
                                  const tmol_ljlk_block_types* bt,
                                  int64_t row_capacity, int32_t dtype);

/* Workspace bytes for score flags: with TMOL_LJLK_DETERMINISTIC, room for the
 * fixed-point gradient planes (after the regions tmol_ljlk_workspace_bytes
 * lays out, which keep their offsets); otherwise as tmol_ljlk_workspace_bytes. */
int64_t tmol_ljlk_workspace_bytes_flags(const tmol_ljlk_stack* stack,
                                        const tmol_ljlk_block_types* bt,
                                        int64_t row_capacity, int32_t dtype, int32_t flags);

/* Device address of the block spheres [G, 4] (dtype) inside the workspace,
 * written by tmol_ljlk_neighbors (centre x, y, z, radius). */
int tmol_ljlk_spheres(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                      void* workspace, int64_t row_capacity, int32_t dtype,
                      void** spheres_out);

/* Neighbor pass (replaces compute_block_spheres + detect_block_neighbors,
 * sphere_overlap.impl.hh:109-196, 287-351): bounding sphere per block,
 * then every b1 <= b2 pair of a pose with d^2 < ((r1 + r2) + max_dis)^2 is
 * appended to row b1 of the work list in the workspace.  Also prepares the
 * per-block-type atom records.  dense_neighbors (padded layout only, may be
 * NULL) receives the reference's [P, B, B] int32 flags. */
int tmol_ljlk_neighbors(const tmol_ljlk_stack* stack,
                        const tmol_ljlk_block_types* bt, const void* coords,
                        int32_t dtype, double max_dis, const int32_t* tiles,
                        int64_t n_tiles, void* workspace, int64_t ws_bytes,
                        int64_t row_capacity, int32_t* dense_neighbors,
                        void* stream);

/* Score the work list (replaces eval_energies / eval_energies_by_block /
 * eval_derivs, ljlk_pose_score.impl.hh:803-1119, 1300-1474).
 *   energies: [3, P] (or [3, P, B, B] with BLOCK_PAIR), dtype; may be NULL
 *             with NO_ENERGY
 *   dvdx:     [3, NA, 3] with DVDX
 *   weights:  [3, P] (or [3, P, B, B]) with WGRAD
 *   grad:     [NA, 3] with WGRAD
 * Outputs are overwritten (zero-filled by kernels) unless ACCUMULATE. */
int tmol_ljlk_score(const tmol_ljlk_stack* stack,
                    const tmol_ljlk_block_types* bt, const void* coords,
                    int32_t dtype, void* workspace, int64_t ws_bytes,
                    int64_t row_capacity, int32_t flags,
                    const void* weights, void* energies, void* dvdx,
                    void* grad, int32_t n_wave_groups, void* stream);

/* No-allocation variant of tmol_ljlk_neighbors + tmol_ljlk_score, in one
 * call.  It has no counterpart in the reference: it is for callers that
 * cannot allocate device memory at all (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc), which therefore cannot
 * hand the other entry points a workspace.  Instead of a workspace argument it
 * uses a device array of TMOL_LJLK_NOALLOC_SCRATCH_BYTES that the library
 * holds on each GPU it runs on until the process ends (its address is looked
 * up on the first call on a device); it calls no allocating routine.
 *   - The stack's workspace (tmol_ljlk_workspace_bytes for row_capacity) must
 *     fit the array, else TMOL_LJLK_ESCRATCH and nothing is launched.
 *   - row_capacity must be at least stack->max_n_blocks, so the work list
 *     cannot overflow (no host read-back).
 *   - flags as for tmol_ljlk_score, without GENERAL_KERNEL and the launch
 *     variant fields; dense_neighbors as for tmol_ljlk_neighbors.
 *   - Every call on a GPU uses that GPU's one array and rewrites all of it
 *     that it reads: calls that overlap on one GPU (other streams, other
 *     threads) race and must be serialised by the caller. */
#define TMOL_LJLK_NOALLOC_SCRATCH_BYTES (64ll << 20)
int64_t tmol_ljlk_noalloc_scratch_bytes(void);
int tmol_ljlk_score_noalloc(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                            const void* coords, int32_t dtype, double max_dis,
                            const int32_t* tiles, int64_t n_tiles, int64_t row_capacity,
                            int32_t flags, const void* weights, void* energies, void* dvdx,
                            void* grad, int32_t* dense_neighbors, void* stream);

/* Branch census of the work list (float32 only): counts of enumerated atom
 * pairs per potential region, for algorithmic FLOP accounting.  census[16]
 * (device uint64): LJ pairs, LJ weight-0, LJ beyond max_dis, LJ linear,
 * LJ 12-6, LJ spline, LK pairs, LK weight-0, LK beyond, LK constant,
 * LK close spline, LK f_desolv, LK far spline, items, intra-block items. */
int tmol_ljlk_census(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                     const void* coords, int32_t dtype, void* workspace,
                     int64_t ws_bytes, int64_t row_capacity, uint64_t* census,
                     void* stream);

/* Whole-pose backward (compiled.ops.cpp:157-174):
 *   grad[a, k] = sum_t grad_scores[t, atom_pose[a]] * dvdx[t, a, k] */
int tmol_ljlk_backward_combine(const void* dvdx, const void* grad_scores,
                               const int32_t* atom_pose, int64_t n_atoms,
                               int32_t n_poses, int32_t dtype, void* grad,
                               void* stream);

/* Copy the work-list counters to the host (synchronises the stream):
 *   counters_host[0] = block pairs found (including any beyond row capacity)
 *   counters_host[1] = row capacity overflow flag */
int tmol_ljlk_read_counters(const void* workspace, int64_t* counters_host,
                            void* stream);

/* Device pointers to the work list inside a workspace: row_count[G]
 * (uint32) and row_items[G, row_capacity] (int32 g2 of row g1). */
int tmol_ljlk_work_list(const tmol_ljlk_stack* stack,
                        const tmol_ljlk_block_types* bt, void* workspace,
                        int64_t row_capacity, int32_t dtype,
                        uint32_t** row_count_out, int32_t** row_items_out);

/* Debug: count-pair separation for every atom pair of each listed block pair,
 * computed by the scoring kernel's own count-pair code (needs the workspace
 * prepared by tmol_ljlk_neighbors).  out[item, 32 * a1 + a2] (uint8, first
 * tile of each block; 255 = not a real atom pair).  items: [n_items, 2]
 * global block indices {g1, g2}, g1 <= g2, same pose. */
int tmol_ljlk_debug_pair_separations(const tmol_ljlk_stack* stack,
                                     const tmol_ljlk_block_types* bt,
                                     const void* workspace, int64_t row_capacity,
                                     int32_t dtype, const int32_t* items,
                                     int64_t n_items, uint8_t* out, void* stream);

/* Number of compute units of the current device (autotuner input). */
int tmol_ljlk_device_cu_count(int32_t* cu_count);

/* ---------------------------------------------------------------------
 * Learned block-pair pruner (no counterpart in the reference: its kernels
 * train and run the MLP that predicts which neighbor block pairs carry
 * |E| < 1e-3 kcal/mol).  All arrays are device pointers; bf16 values are
 * passed as uint16_t bit patterns. */

/* Per-block inputs of the pruner MLP (all device, float32 / int32). */
typedef struct {
  const float* sphere;    /* [G, 4] centre, radius */
  const float* sincos;    /* [G, 10] sin, cos of phi, psi, omega, chi1, chi2 */
  const int32_t* emb_row; /* [G] embedding row of the block's residue type */
  const int32_t* seqpos;  /* [G] position of the block within its pose */
  const float* emb;       /* [n_emb, 8] residue-type embedding */
} tmol_ljlk_prune_blocks;

/* sin, cos of five torsions per block from global atom indices
 * torsion_atoms [G, 5, 4] (-1: torsion absent -> 0, 0). */
int tmol_ljlk_prune_block_torsions(int32_t n_blocks, const int32_t* torsion_atoms,
                                   const void* coords, int32_t dtype, float* sincos,
                                   void* stream);

/* Feature rows [n, 48] (bf16; columns 40..47 zero) of block pairs [n, 2]. */
int tmol_ljlk_prune_pair_features(int64_t n_pairs, const int32_t* pairs,
                                  const tmol_ljlk_prune_blocks* blocks, uint16_t* features,
                                  void* stream);

/* C = A B on the matrix cores, bf16 operands, fp32 accumulation.
 * a_km: A stored [K][M] (else [M][K]); b_kn: B stored [K][N] (else [N][K]).
 * epilogue 0: C fp32 [M, N] (split_k > 1: atomically added, C pre-zeroed);
 *          1: Z = C + bias[n] (fp32), H = SiLU(Z) (bf16);
 *          2: D = C * SiLU'(Z) (fp32, Z read), H = D (bf16). */
int tmol_ljlk_prune_gemm(int32_t epilogue, int32_t a_km, int32_t b_kn, const uint16_t* A,
                         const uint16_t* B, int32_t M, int32_t N, int32_t K, int32_t lda,
                         int32_t ldb, int32_t ldc, int32_t split_k, float* C, const float* bias,
                         float* Z, uint16_t* H, float* D, void* stream);

/* Output layer (128 -> 1) and, with train, the weighted logistic loss and its
 * backward: loss (+=), dZ3 = dlogit w4 SiLU'(Z3) (fp32 and bf16), dw4 (+=),
 * db4 (+=).  labels: 1 = negligible pair; sample_weights [n] may be NULL. */
int tmol_ljlk_prune_head(int32_t n, const uint16_t* H3, const float* Z3, const float* w4,
                         const float* b4, const float* labels, const float* sample_weights,
                         float pos_weight, int32_t train, float* logits, float* loss, float* dZ3,
                         uint16_t* dZ3_bf16, float* dw4, float* db4, void* stream);

/* out[c] += sum over rows of D[r, c] (bias gradients). */
int tmol_ljlk_prune_colsum(int32_t n, int32_t ncol, int32_t ld, const float* D, float* out,
                           void* stream);

/* Embedding gradient: demb[emb_row[g]] += dX[pair, 3..19] of each side. */
int tmol_ljlk_prune_emb_grad(int64_t n_pairs, const int32_t* pairs, const int32_t* emb_row,
                             const float* dX, float* demb, void* stream);

/* Bytes of device scratch tmol_ljlk_prune_flatten needs for n_blocks rows. */
int64_t tmol_ljlk_prune_scratch_bytes(int32_t n_blocks);

/* Flatten the work list: row_offset [G + 1] (exclusive offsets of each row's
 * pairs), pairs [n, 2] (g1, g2) in row order.  The total n is left on the
 * device at scratch[(G + 1023) / 1024] (int64) for tmol_ljlk_prune_infer.
 * pairs must hold G * row_capacity entries. */
int tmol_ljlk_prune_flatten(int32_t n_blocks, const uint32_t* row_count,
                            const int32_t* row_items, int64_t row_capacity, int64_t* scratch,
                            int64_t* row_offset, int32_t* pairs, void* stream);

/* Fused inference on the matrix cores: features, three BF16 layers, head.
 * keep[i] = 0 for a pair predicted negligible (logit > logit_cut) unless it
 * is a block with itself or its sphere gap |c1 - c2| - r1 - r2 (a lower
 * bound on its atom-pair distances) is below min_gap; logits [n] may be NULL.
 * n is read from *n_pairs (device).  n_workgroups <= 0: one per resident
 * slot. */
int tmol_ljlk_prune_infer(const int64_t* n_pairs, const int32_t* pairs,
                          const tmol_ljlk_prune_blocks* blocks, const uint16_t* W1,
                          const uint16_t* W2, const uint16_t* W3, const float* b1,
                          const float* b2, const float* b3, const float* w4, const float* b4,
                          float logit_cut, float min_gap, uint8_t* keep, float* logits,
                          int32_t n_workgroups, void* stream);

/* Remove the pairs with keep == 0 from the work list, in place, keeping the
 * order of each row; row_count is updated. */
int tmol_ljlk_prune_compact(int32_t n_blocks, uint32_t* row_count, int32_t* row_items,
                            int64_t row_capacity, const int64_t* row_offset,
                            const uint8_t* keep, void* stream);

/* ---------------------------------------------------------------------
 * Rotamer-pair scoring for packing.  Replaces the reference's op
 *   tmol_ljlk::ljlk_rotamer_scores   compiled.ops.cpp:282-530, 597-663
 * and its dispatch
 *   LJLKRotamerScoreDispatch::forward / ::backward
 *                                    ljlk_pose_score.impl.hh:1479-2126
 *   compute_rot_spheres, compute_block_spheres_from_rot_spheres,
 *   detect_block_neighbors, rot_neighbor_indices_from_block_neighbors
 *                                    sphere_overlap.impl.hh:23-107, 287-351, 455-610
 * Rotamers are numbered globally, grouped by pose and block; the pose
 * stack's bond-separation tables give the count pair between blocks.
 * Block types of at most 64 atoms (ENOTSUP beyond). */
typedef struct {
  int32_t n_poses;                       /* P */
  int32_t max_n_blocks;                  /* B */
  int32_t n_rots;                        /* R */
  int32_t reserved;                      /* 0 */
  int64_t n_atoms_total;                 /* rows of the rotamer coordinates */
  const int32_t* rot_coord_offset;       /* [R] first atom of each rotamer */
  const int32_t* block_ind_for_rot;      /* [R] */
  const int32_t* block_type_ind_for_rot; /* [R] */
  const int32_t* n_rots_for_block;       /* [P, B] */
  const int32_t* rot_offset_for_block;   /* [P, B] first rotamer of each block */
  const int32_t* first_rot_block_type;   /* [P, B] -1: block without rotamers */
  const int32_t* min_bondsep;            /* [P, B, B] */
  const int32_t* inter_block_bondsep;    /* [P, B, B, C, C] */
} tmol_ljlk_rotamer_set;

/* bytes of one work unit (tmol_ljlk_rotamer_dispatch's `units`) */
#define TMOL_LJLK_ROT_UNIT_BYTES 32

/* Workspace bytes for a rotamer set (spheres, rows, scans, atom records). */
int64_t tmol_ljlk_rotamer_workspace_bytes(const tmol_ljlk_rotamer_set* rs,
                                          const tmol_ljlk_block_types* bt, int32_t dtype);

/* Rotamer spheres, block spheres enclosing each block's rotamers, the block
 * neighbor test (d^2 < ((r1 + r2) + max_dis)^2, bit-identical to the
 * reference), and each row's counts of rotamer pairs and work units, scanned. */
int tmol_ljlk_rotamer_neighbors(const tmol_ljlk_rotamer_set* rs,
                                const tmol_ljlk_block_types* bt, const void* coords,
                                int32_t dtype, double max_dis, void* workspace,
                                int64_t ws_bytes, void* stream);

/* totals_host[0] = rotamer pairs (dispatch entries), [1] = work units.
 * Synchronises the stream (the caller sizes its outputs from them). */
int tmol_ljlk_rotamer_totals(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                             int32_t dtype, const void* workspace, int64_t* totals_host,
                             void* stream);

/* Dispatch indices [3, n_pairs] int32 = (pose, rot1, rot2), in the
 * reference's order (neighbor cells (b1 <= b2) row-major; all n1 x n2 pairs
 * rot1-major off the diagonal, self pairs on it), and n_units work units of
 * TMOL_LJLK_ROT_UNIT_BYTES each.  max_dis as given to _neighbors. */
int tmol_ljlk_rotamer_dispatch(const tmol_ljlk_rotamer_set* rs,
                               const tmol_ljlk_block_types* bt, int32_t dtype, double max_dis,
                               void* workspace, int64_t ws_bytes, int32_t* indices,
                               int64_t n_pairs, void* units, int64_t n_units, void* stream);

/* Score the dispatch list.  flags 0: energies [3, n_pairs] (atr, rep, lk of
 * each rotamer pair, unweighted).  flags WGRAD | NO_ENERGY: the backward,
 * grad [NA, 3] = sum_i sum_t dTdV[t, i] dV_t(pair i)/dx (zeroed first unless
 * ACCUMULATE).  n_workgroups <= 0: one per resident slot. */
int tmol_ljlk_rotamer_score(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                            const void* coords, int32_t dtype, void* workspace,
                            int64_t ws_bytes, const void* units, int64_t n_units,
                            int64_t n_pairs, int32_t flags, const void* dTdV, void* energies,
                            void* grad, int32_t n_workgroups, void* stream);

/* Branch census of the dispatch list (float32 only): counts of enumerated
 * atom pairs per potential region, in tmol_ljlk_census's layout (items =
 * rotamer pairs, intra-block items = self pairs). */
int tmol_ljlk_rotamer_census(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                             const void* coords, int32_t dtype, void* workspace,
                             int64_t ws_bytes, const void* units, int64_t n_units,
                             int64_t n_pairs, uint64_t* census, void* stream);

/* Debug: device addresses of the rotamer spheres [R, 4] and block spheres
 * [P, B, 4] (dtype) inside the workspace, written by _neighbors. */
int tmol_ljlk_rotamer_spheres(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                              int32_t dtype, void* workspace, void** rot_spheres,
                              void** block_spheres);

/* ---------------------------------------------------------------------
 * Pose topology on the device.  inter_block_bondsep [P, B, B, C, C] (int32;
 * 6 = six bonds or more, and for connection pairs that do not exist) and,
 * if non-NULL, min_bondsep [P, B, B] (minimum over the [C, C] entries), from
 * block_type [P, B] (-1 = no block) and inter_residue_connections
 * [P, B, C, 2] ((block, connection) partner, -1 = none; both directions
 * stored): shortest bond paths between connection atoms, capped at 6.
 * C = bt->max_n_conn.  Replaces the reference's connection-point graph and
 * all-pairs shortest paths (pose_stack_builder.py:459-599, 1230-1349;
 * compiled/apsp.cpu.cpp) and the min over connection pairs
 * (bond_dependent_term.py:91-100).  overflow (device, may be NULL) is set
 * when a connection point reaches more than 48 points within 5 bonds. */
int tmol_ljlk_inter_block_bondsep(const tmol_ljlk_block_types* bt, int32_t n_poses,
                                  int32_t max_n_blocks, const int32_t* block_type,
                                  const int32_t* inter_residue_connections,
                                  int32_t* inter_block_bondsep, int32_t* min_bondsep,
                                  uint32_t* overflow, void* stream);

/* ---------------------------------------------------------------------
 * Elec, a sibling two-body term on the same scheduler.  Replaces the
 * reference's op
 *   tmol_elec::elec_pose_scores   tmol/score/elec/potentials/compiled.ops.cpp:526-586
 * and its dispatch
 *   ElecPoseScoreDispatch::forward    elec_pose_score.impl.hh:383-856
 *   ElecPoseScoreOp::backward (whole pose)  compiled.ops.cpp:150-168
 *   ElecPoseScoreDispatch::backward (block pair)  elec_pose_score.impl.hh:858-1125
 * The stack and block-type structs are the LJ+LK ones; elec reads n_atoms,
 * n_conn, conn_atoms and is_ligand (elec_is_ligand_fragment) of the block
 * types, which must be at most 32 atoms (ENOTSUP beyond).  The workspace
 * has tmol_ljlk_workspace_bytes() bytes and holds elec's own work list. */
typedef struct {
  const void* partial_charge;              /* [T, Amax] (block-type dtype) */
  const int32_t* inter_repr_path_distance; /* [T, Amax, Amax]: [i][j] = path_dist[i][rep(j)] */
  const int32_t* intra_repr_path_distance; /* [T, Amax, Amax]: path_dist[rep(i)][rep(j)] */
  float global_params[5];                  /* host values: D, D0, S, min_dis, max_dis */
} tmol_elec_tables;

/* Neighbor pass with elec's reach (max_dis, the reference's host scalar):
 * spheres, the row-major work list, the packed count-pair tables and the
 * intra-block pair table; dense_neighbors as for tmol_ljlk_neighbors. */
int tmol_elec_neighbors(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                        const tmol_elec_tables* elec, const void* coords, int32_t dtype,
                        double max_dis, const int32_t* tiles, int64_t n_tiles,
                        void* workspace, int64_t ws_bytes, int64_t row_capacity,
                        int32_t* dense_neighbors, void* stream);

/* Score the work list.  flags as for tmol_ljlk_score with one term:
 *   energies: [P] (or [P, B, B] / ragged [G, row_capacity] with BLOCK_PAIR)
 *   dvdx:     [NA, 3] with DVDX (the reference's saved dV/dx)
 *   weights:  [P] (or block-pair shaped) with WGRAD; grad [NA, 3]
 * Outputs are overwritten unless ACCUMULATE. */
int tmol_elec_score(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                    const tmol_elec_tables* elec, const void* coords, int32_t dtype,
                    void* workspace, int64_t ws_bytes, int64_t row_capacity, int32_t flags,
                    const void* weights, void* energies, void* dvdx, void* grad,
                    int32_t n_wave_groups, void* stream);

/* LJ+LK and elec in one pass over the LJ+LK work list (tmol_ljlk_neighbors;
 * elec's 5.5 A reach lies inside LJ+LK's 6 A, so the list holds every elec
 * block pair).  No reference counterpart: the reference scores each term in
 * its own launch.  Blocks share staging, cull and the pair tests; elec
 * entries go to a third FIFO.
 *   tmol_ljlk_elec_prepare: elec's packed path distances and intra-block
 *     table into the LJ+LK workspace (once per block-type set).
 *   tmol_ljlk_elec_score: flags 0, WGRAD, WGRAD | NO_ENERGY (ACCUMULATE
 *     allowed); energies [4, P] = atr, rep, lk, elec; weights [4, P]; grad
 *     [NA, 3] = sum_t w[t, p] dE_t/dx.  Block types of at most 32 atoms. */
int tmol_ljlk_elec_prepare(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                           const tmol_elec_tables* elec, void* workspace, int64_t ws_bytes,
                           int64_t row_capacity, int32_t dtype, void* stream);
int tmol_ljlk_elec_score(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                         const tmol_elec_tables* elec, const void* coords, int32_t dtype,
                         void* workspace, int64_t ws_bytes, int64_t row_capacity, int32_t flags,
                         const void* weights, void* energies, void* grad, int32_t n_wave_groups,
                         void* stream);

/* Census of the elec work list (float32 only), census[16] (device uint64):
 * block-pair items, intra-block items, enumerated atom pairs, pairs of two
 * charged atoms, charged pairs of count-pair weight 0, and evaluated pairs by
 * region: flat (d < min_dis - 0.25), low fade, Coulomb, high fade, beyond
 * max_dis.  For algorithmic FLOP accounting. */
int tmol_elec_census(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                     const tmol_elec_tables* elec, const void* coords, int32_t dtype,
                     void* workspace, int64_t ws_bytes, int64_t row_capacity, uint64_t* census,
                     void* stream);

/* Elec rotamer-pair scoring for packing.  Replaces the reference's op
 *   tmol_elec::elec_rotamer_scores   tmol/score/elec/potentials/compiled.ops.cpp:588-650
 *   ElecRotamerScoreDispatch::forward / ::backward
 *                                    elec_pose_score.impl.hh:1132-1700
 * on the LJ+LK rotamer dispatch: size the workspace with
 * tmol_ljlk_rotamer_workspace_bytes, run tmol_elec_rotamer_neighbors (in
 * place of tmol_ljlk_rotamer_neighbors: elec's tables instead of the LJ+LK
 * atom records, max_dis = elec's), then tmol_ljlk_rotamer_totals and
 * tmol_ljlk_rotamer_dispatch with the same max_dis, then
 * tmol_elec_rotamer_score.  The block-type struct is elec's (as for
 * tmol_elec_neighbors); block types of at most 32 atoms (ENOTSUP beyond). */
int tmol_elec_rotamer_neighbors(const tmol_ljlk_rotamer_set* rs,
                                const tmol_ljlk_block_types* bt, const tmol_elec_tables* elec,
                                const void* coords, int32_t dtype, double max_dis,
                                void* workspace, int64_t ws_bytes, void* stream);

/* flags 0: energies [n_pairs] (the elec energy of each rotamer pair; the
 * reference's [1, n] output).  flags WGRAD | NO_ENERGY: the backward,
 * grad [NA, 3] = sum_i dTdV[i] dV(pair i)/dx (zeroed first unless
 * ACCUMULATE).  n_workgroups <= 0: one per resident slot. */
int tmol_elec_rotamer_score(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                            const tmol_elec_tables* elec, const void* coords, int32_t dtype,
                            void* workspace, int64_t ws_bytes, const void* units,
                            int64_t n_units, int64_t n_pairs, int32_t flags, const void* dTdV,
                            void* energies, void* grad, int32_t n_workgroups, void* stream);

/* ---------------------------------------------------------------------
 * lk_ball on the same scheduler.  Replaces, for its pose path, the
 * reference's ops
 *   tmol_lk_ball::gen_pose_waters     tmol/score/lk_ball/potentials/compiled.ops.cpp:22-353
 *                                     (PoseWaterGen forward/backward, pose_watergen_op:283)
 *   tmol_lk_ball::lk_ball_pose_score  compiled.ops.cpp:356-907 (LKBallPoseScoreOp
 *                                     forward/backward, lkball_pose_score:840)
 * and their dispatch
 *   GeneratePoseWaters::forward / backward  gen_pose_waters.impl.hh:28-432
 *   LKBallPoseScoreDispatch::forward / backward  lk_ball_pose_score.impl.hh:358-1000
 * The stack and block-type structs are the LJ+LK ones (block types: n_atoms,
 * n_conn, conn_atoms).  lk_ball's own tables (all device pointers, dtype of
 * the coordinates for the float ones):
 *   occluders: per block type the heavy atoms, those that carry waters
 *     ("polar": donor heavy atoms and acceptors) first, at most 32 per type
 *     of which at most 16 polar (ENOTSUP beyond), with the reference's 9
 *     per-atom parameters
 *     (lj_radius, lk_dgfree, lk_lambda, lk_volume, is_donor, is_hydroxyl,
 *     is_polarh, is_acceptor, is_carbon_lk);
 *   recipes: one row per water slot, TMOL_LKBALL_RECIPE_WIDTH int32:
 *     {atom, slot (0..3), kind, a1, a2, a3, torsion index, 0}; kind 0 donor
 *     (a1 = heavy atom D, a2 = its polar H), 1 sp2 / 2 sp3 / 3 ring acceptor
 *     (a1 = A, a2 = B, a3 = B0; ring: built from (B + B0) / 2 and B0), all
 *     atoms of the same block;
 *   path_distance: [T, Amax, Amax] intra-block bond separations (no ligand
 *     rule: lk_ball counts pairs by plain path distance). */
#define TMOL_LKBALL_RECIPE_WIDTH 8
typedef struct {
  int32_t max_n_occ;             /* Omax (<= 32) */
  int32_t max_n_recipes;         /* Rmax */
  int32_t max_n_polar;           /* most polar atoms of a block type (<= 16) */
  const int32_t* n_occ;          /* [T] */
  const int32_t* n_polar;        /* [T] */
  const int32_t* occ_atoms;      /* [T, Omax] block atom index, polar first */
  const void* occ_params;        /* [T, Omax, 9] */
  const int32_t* n_recipes;      /* [T] */
  const int32_t* recipes;        /* [T, Rmax, 8] */
  const int32_t* path_distance;  /* [T, Amax, Amax] */
  const void* global_params;     /* [5] lj_hbond_dis, lj_hbond_OH_donor_dis, lj_hbond_hdis,
                                    lkb_water_dist, max_dis */
  const void* watergen_params;   /* [4] lkb_water_dist, lkb_water_angle_sp2, _sp3, _ring (rad) */
  const void* water_tors;        /* [3, 4] torsions (rad) of sp2, sp3, ring waters */
} tmol_lkball_tables;

/* Workspace bytes of tmol_lkball_neighbors + tmol_lkball_score (the LJ+LK
 * layout for row_capacity, then per-launch block records). */
int64_t tmol_lkball_workspace_bytes(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                                    const tmol_lkball_tables* lkb, int64_t row_capacity,
                                    int32_t dtype);

/* Waters [NA, 4, 3] (NaN where no water is built), the reference's
 * gen_pose_waters output. */
int tmol_lkball_waters(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                       const tmol_lkball_tables* lkb, const void* coords, int32_t dtype,
                       void* waters, void* stream);

/* Backward of the waters: dxyz [NA, 3] += (dWaters/dx)^T dwaters [NA, 4, 3]. */
int tmol_lkball_waters_backward(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                                const tmol_lkball_tables* lkb, const void* coords,
                                int32_t dtype, const void* dwaters, void* dxyz, void* stream);

/* Neighbor pass for lk_ball (spheres + work list, as tmol_ljlk_neighbors,
 * without the LJ+LK atom records): blocks within max_dis (the reference's
 * host scalar) of each other.  Rows are full: row g lists every neighbor
 * block of g in its pose (both triangles; the row capacity is the most
 * blocks of any pose, as for tmol_ljlk_neighbors); ragged stacks need
 * pose_block_start.  dense_neighbors receives the upper triangle (the
 * reference's flags). */
int tmol_lkball_neighbors(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                          const void* coords, int32_t dtype, double max_dis,
                          const int32_t* tiles, int64_t n_tiles, void* workspace,
                          int64_t ws_bytes, int64_t row_capacity, int32_t* dense_neighbors,
                          void* stream);

/* Score the work list.
 *   flags 0: energies [4, P] (lk_ball_iso, lk_ball, lk_bridge,
 *     lk_bridge_uncpl); BLOCK_PAIR (padded stacks; ENOTSUP for ragged):
 *     [4, P, B, B], upper triangle.
 *   WGRAD (with weights = dTdV of the energies' shape): dxyz [NA, 3] and
 *     dwaters [NA, 4, 3] of sum_t w[t] V_t (zeroed first unless ACCUMULATE);
 *     with NO_ENERGY the backward alone.  The full coordinate gradient adds
 *     tmol_lkball_waters_backward of dwaters. */
int tmol_lkball_score(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                      const tmol_lkball_tables* lkb, const void* coords, const void* waters,
                      int32_t dtype, void* workspace, int64_t ws_bytes, int64_t row_capacity,
                      int32_t flags, const void* weights, void* energies, void* dxyz,
                      void* dwaters, int32_t n_wave_groups, void* stream);

/* Census of the lk_ball work list: census[4] (device uint64) = candidate
 * (polar, occluder) pairs, pairs at separation <= 3, evaluated pairs (within
 * max_dis, polar atom with waters), of which with the water bridge
 * evaluated; for algorithmic FLOP accounting. */
int tmol_lkball_census(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                       const tmol_lkball_tables* lkb, const void* coords, const void* waters,
                       int32_t dtype, void* workspace, int64_t ws_bytes, int64_t row_capacity,
                       unsigned long long* census, void* stream);

/* ---------------------------------------------------------------------
 * hbond on the same scheduler.  Replaces, for its pose path, the
 * reference's ops
 *   tmol_hbond::gen_hbond_bases     tmol/score/hbond/potentials/compiled.ops.cpp:930-996
 *                                   (GenerateHBondBases::forward,
 *                                   gen_hbond_bases.impl.hh:380-468)
 *   tmol_hbond::hbond_pose_scores   compiled.ops.cpp:28-375, 723-825 (HBondPoseScoresOp
 *                                   forward / backward; HBondPoseScoreDispatch::forward,
 *                                   hbond_pose_score.impl.hh:422-720, backward :860-1000)
 * The stack and block-type structs are the LJ+LK ones (block types: n_atoms,
 * n_conn, conn_atoms).  hbond's own tables (device pointers):
 *   don: per block type its polar hydrogens in atom order, at most 32,
 *     TMOL_HBOND_DON_WIDTH int32 each: {H, D (heavy atom), donor type, 0};
 *   acc: its acceptors in atom order, at most 32, TMOL_HBOND_ACC_WIDTH
 *     int32 each: {A, B, B0, acceptor type, hybridization (1 sp2, 2 sp3,
 *     3 ring), 0, 0, 0}, bases -1 where missing, all atoms of the block;
 *   base_src: [T, Amax, 3] the source atom of each derived slot of each atom
 *     (slot 0: D of a polar H; 1, 2: B, B0 of an acceptor), -1 a missing
 *     base, -2 no role;
 *   path_distance: [T, Amax, Amax] intra-block bond separations;
 *   the potential's tables as the reference stacks them: pair_params
 *     [nD, nA, 3] {hybridization, acceptor weight, donor weight} and
 *     global_params [6] {sp2_range_span, sp2_BAH180_rise, sp2_outer_width,
 *     sp3_softmax_fade, threshold_distance, max_ha_dis} in the coordinates'
 *     dtype, pair_polys [nD, nA, 3, 16] float64 (per polynomial c10..c0, pad,
 *     xmin, xmax, min_val, max_val; A-H distance, cos BAH, AHD angle).
 * Block types of more than 32 atoms are refused (ENOTSUP). */
#define TMOL_HBOND_DON_WIDTH 4
#define TMOL_HBOND_ACC_WIDTH 8
typedef struct {
  int32_t max_n_don;             /* Dmax (<= 32) */
  int32_t max_n_acc;             /* Xmax (<= 32) */
  int32_t n_don_types;           /* nD */
  int32_t n_acc_types;           /* nA */
  const int32_t* n_don;          /* [T] */
  const int32_t* n_acc;          /* [T] */
  const int32_t* don;            /* [T, Dmax, 4] */
  const int32_t* acc;            /* [T, Xmax, 8] */
  const int32_t* base_src;       /* [T, Amax, 3] */
  const int32_t* path_distance;  /* [T, Amax, Amax] */
  const void* pair_params;       /* [nD, nA, 3] */
  const double* pair_polys;      /* [nD, nA, 48] */
  const void* global_params;     /* [6] */
} tmol_hbond_tables;

/* Derived coordinates [NA, 3, 3] and their source atoms [NA, 3] (int32):
 * the reference's gen_hbond_bases output (NaN / -1 for atoms with no role;
 * (0, 0, 0) / -1 for a base that does not exist). */
int tmol_hbond_bases(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                     const tmol_hbond_tables* hb, const void* coords, int32_t dtype,
                     void* derived_coords, int32_t* derived_atom_inds, void* stream);

/* Neighbor pass for hbond (spheres + half-triangle work list, as
 * tmol_ljlk_neighbors without the LJ+LK atom records): blocks within
 * max_dis (the reference's 5.5) of each other; row g1 lists g2 >= g1 of its
 * pose.  Workspace: tmol_ljlk_workspace_bytes. */
int tmol_hbond_neighbors(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                         const void* coords, int32_t dtype, double max_dis,
                         const int32_t* tiles, int64_t n_tiles, void* workspace,
                         int64_t ws_bytes, int64_t row_capacity, int32_t* dense_neighbors,
                         void* stream);

/* Score the work list.
 *   flags 0: energies [P]; BLOCK_PAIR (padded stacks; ENOTSUP for ragged):
 *     [P, B, B], upper triangle.
 *   WGRAD (with weights = dTdV of the energies' shape): dxyz [NA, 3] of
 *     sum w V through D, H, A, B, B0 (zeroed first unless ACCUMULATE); with
 *     NO_ENERGY the backward alone. */
int tmol_hbond_score(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                     const tmol_hbond_tables* hb, const void* coords,
                     const void* derived_coords, const int32_t* derived_atom_inds,
                     int32_t dtype, void* workspace, int64_t ws_bytes, int64_t row_capacity,
                     int32_t flags, const void* weights, void* energies, void* dxyz,
                     int32_t n_wave_groups, void* stream);

/* Census of the hbond work list: census[3] (device uint64) = candidate
 * donor x acceptor pairs, pairs below separation 5, evaluated pairs (H-A
 * within max_ha_dis); for algorithmic FLOP accounting. */
int tmol_hbond_census(const tmol_ljlk_stack* stack, const tmol_ljlk_block_types* bt,
                      const tmol_hbond_tables* hb, const void* coords,
                      const void* derived_coords, const int32_t* derived_atom_inds,
                      int32_t dtype, void* workspace, int64_t ws_bytes, int64_t row_capacity,
                      unsigned long long* census, void* stream);

/* ---------------------------------------------------------------------
 * Missing leaf atoms of a pose stack (canonical-form ingest).  Replaces the
 * reference's op
 *   tmol_io::gen_pose_leaf_atoms   tmol/io/details/compiled/compiled.ops.cpp:19-222
 *                                  (PoseLeafAtomGen forward / backward;
 *                                  GeneratePoseLeafAtoms, gen_pose_leaf_atoms.impl.hh:25-336)
 * The arguments are the op's, in the padded pose layout (device pointers):
 * atoms given in block_atom_missing = 0 are copied, a missing atom is built
 * from three ancestors (atom index in the block, or connection + bonds from
 * it into the connected block) and its ideal (phi, theta, d), with a backup
 * ancestor set where the first is incomplete; NaN when neither resolves.
 * The table indices are not checked here (the caller validates them, as
 * tmol_amd/ingest.py does).  ENOTSUP: max_n_block_atoms > 2048, or
 * P * B * Amax >= 2^31, or P * NA * 3 >= 2^31. */
typedef struct {
  int32_t n_poses;                            /* P */
  int32_t max_n_pose_atoms;                   /* NA */
  int32_t max_n_blocks;                       /* B */
  int32_t max_n_block_atoms;                  /* Amax */
  int32_t max_n_conn;                         /* C */
  const uint8_t* block_atom_missing;          /* [P, B, Amax] */
  const uint8_t* pose_atom_missing;           /* [P, NA] */
  const int32_t* block_coord_offset;          /* [P, B] */
  const int32_t* block_type;                  /* [P, B], -1 = none */
  const int32_t* inter_block_connections;     /* [P, B, C, 2] */
  const int32_t* bt_n_atoms;                  /* [T] */
  const int32_t* bt_atom_downstream_of_conn;  /* [T, C, Amax] */
  const int32_t* bt_atom_ancestors;           /* [T, Amax, 3, 3] */
  const void* bt_atom_icoors;                 /* [T, Amax, 3] phi, theta, d (coords dtype) */
  const int32_t* bt_atom_ancestors_backup;    /* [T, Amax, 3, 3] */
  const void* bt_atom_icoors_backup;          /* [T, Amax, 3] */
} tmol_leaf_atoms_args;

/* new_coords [P, NA, 3] from orig_coords [P, NA, 3]. */
int tmol_leaf_atoms_build(const tmol_leaf_atoms_args* args, const void* orig_coords,
                          int32_t dtype, void* new_coords, void* stream);

/* dorig [P, NA, 3] (overwritten) from dnew [P, NA, 3]: the backward of
 * tmol_leaf_atoms_build. */
int tmol_leaf_atoms_backward(const tmol_leaf_atoms_args* args, const void* orig_coords,
                             int32_t dtype, const void* dnew, void* dorig, void* stream);

/* ---------------------------------------------------------------------
 * Segmented L-BFGS history products (float64), for the minimizer that
 * drives this op (no op counterpart: the reference's LBFGS_Armijo,
 * tmol/optimization/lbfgs_armijo.py:6-58, does this algebra in torch).  The
 * history is [n_segments, history, seg_len] (rows contiguous); rows
 * 0 .. n_used - 1 are used.
 *   tmol_lbfgs_dots:    out[k, p, i] = H[p, i, :] . vecs[k, p, :], k < n_vecs (<= 2)
 *   tmol_lbfgs_combine: out[p, :] = g[p, :] + sum_i c1[p, i] H1[p, i, :]
 *                                             + c2[p, i] H2[p, i, :]   (c: [P, n_used]) */
int tmol_lbfgs_dots(int32_t n_segments, int32_t history, int64_t seg_len, int32_t n_used,
                    const double* H, const double* vecs, int32_t n_vecs, double* out,
                    void* stream);
int tmol_lbfgs_combine(int32_t n_segments, int32_t history, int64_t seg_len, int32_t n_used,
                       const double* g, const double* H1, const double* c1, const double* H2,
                       const double* c2, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TMOL_LJLK_H */
