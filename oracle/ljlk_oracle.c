/* CPU oracle for the LJ+LK pose-score op -- TEST INFRASTRUCTURE ONLY.
 * Float and double instantiations of ljlk_oracle_impl.h, plus the integer
 * count-pair dump.  Built by oracle/Makefile with -O2 -ffp-contract=off
 * (no fused multiply-add, like the reference's x86-64 CPU build). */
#include <math.h>
#include <stdint.h>

typedef struct {
  int32_t n_poses, max_n_blocks;
  int32_t n_block_types, max_n_block_atoms, max_n_tiles, max_n_conn;
  const int32_t* rot_coord_offset;    /* [P*B] */
  const int32_t* block_type;          /* [P*B], -1 = padding */
  const int32_t* min_bondsep;         /* [P,B,B] */
  const int32_t* inter_block_bondsep; /* [P,B,B,C,C] */
  const int32_t* bt_n_atoms;          /* [T] */
  const int32_t* bt_n_heavy_in_tile;  /* [T, nt] */
  const int32_t* bt_heavy_in_tile;    /* [T, 32*nt] */
  const int32_t* bt_atom_types;       /* [T, Am] */
  const int32_t* bt_n_conn;           /* [T] */
  const int32_t* bt_conn_atoms;       /* [T, C] */
  const int32_t* bt_path_distance;    /* [T, Am, Am] */
  const int32_t* bt_is_ligand;        /* [T] */
} oracle_stack;

#define REAL float
#define RSUF _f
#define SQRT sqrtf
#define EXP expf
#define FLOOR floorf
#include "ljlk_oracle_impl.h"
#undef REAL
#undef RSUF
#undef SQRT
#undef EXP
#undef FLOOR

#define REAL double
#define RSUF _d
#define SQRT sqrt
#define EXP exp
#define FLOOR floor
#include "ljlk_oracle_impl.h"
#undef REAL
#undef RSUF
#undef SQRT
#undef EXP
#undef FLOOR

/* integer count-pair separation for one atom pair of a block pair
 * (b1 != b2: interres rule; b1 == b2: intra-block path distance) */
int oracle_pair_sep(const oracle_stack* st, int p, int b1, int b2, int a1,
                    int a2) {
  int B = st->max_n_blocks, Am = st->max_n_block_atoms;
  int bt1 = st->block_type[p * B + b1], bt2 = st->block_type[p * B + b2];
  if (bt1 < 0 || bt2 < 0) return -1;
  if (b1 == b2)
    return st->bt_path_distance[((int64_t)bt1 * Am + a1) * Am + a2];
  return inter_sep_f(st, p, b1, b2, bt1, bt2, a1, a2);
}

/* Rotamer-pair dispatch list from block neighbor flags
 * (rot_neighbor_indices_from_block_neighbors, sphere_overlap.impl.hh:516-610):
 * cells (p, b1, b2) in row-major order; a flagged diagonal cell contributes
 * its rotamers' self pairs (r, r), an off-diagonal one all n1 * n2 pairs,
 * rot1-major.  Returns the number of entries.  With indices != NULL (room
 * for n entries: [3, n]) the list is written; entries of a cell whose
 * rotamer offset is negative stay -1, as in the reference. */
int64_t oracle_rot_dispatch(int P, int B, const int32_t* neighbors,
                            const int32_t* n_rots_for_block,
                            const int32_t* rot_offset_for_block,
                            int32_t* indices, int64_t n) {
  int64_t total = 0;
  for (int p = 0; p < P; ++p)
    for (int b1 = 0; b1 < B; ++b1)
      for (int b2 = 0; b2 < B; ++b2) {
        if (!neighbors[((int64_t)p * B + b1) * B + b2]) continue;
        int64_t nr1 = n_rots_for_block[p * B + b1];
        int64_t nr2 = n_rots_for_block[p * B + b2];
        int64_t cnt = b1 == b2 ? nr1 : nr1 * nr2;
        if (indices) {
          int off1 = rot_offset_for_block[p * B + b1];
          int off2 = rot_offset_for_block[p * B + b2];
          for (int64_t k = 0; k < cnt; ++k) {
            int64_t e = total + k;
            int32_t v0 = -1, v1 = -1, v2 = -1;
            if (off1 >= 0 && off2 >= 0) {
              v0 = p;
              if (b1 == b2) {
                v1 = off1 + (int32_t)k;
                v2 = v1;
              } else {
                v1 = off1 + (int32_t)(k / nr2);
                v2 = off2 + (int32_t)(k % nr2);
              }
            }
            indices[e] = v0;
            indices[n + e] = v1;
            indices[2 * n + e] = v2;
          }
        }
        total += cnt;
      }
  return total;
}
