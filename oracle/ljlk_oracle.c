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
