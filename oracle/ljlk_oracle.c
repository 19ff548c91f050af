/* CPU oracle for the LJ+LK pose-score op -- TEST INFRASTRUCTURE ONLY.
 * Float and double instantiations of ljlk_oracle_impl.h, elec_oracle_impl.h,
 * lkball_oracle_impl.h and hbond_oracle_impl.h, plus the integer
 * count-pair dump.  Built by oracle/Makefile with -O2 -ffp-contract=off
 * (no fused multiply-add, like the reference's x86-64 CPU build). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

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

/* eps() of the elec potential at a float argument (potentials.hh:43-48 with
 * Real = float): the reference evaluates it this way for max_dis and min_dis
 * in both precisions; the 0.5 literal lifts the product to double */
static float elec_eps_float(float dist, float D, float D0, float S) {
  double a = 0.5 * (D - D0);
  float poly = 2 + 2 * dist * S + dist * dist * S * S;
  return (float)(D - a * poly * expf(-dist * S));
}

#define REAL float
#define RSUF _f
#define SQRT sqrtf
#define EXP expf
#define FLOOR floorf
#define LOG logf
#define COS cosf
#define SIN sinf
#define ATAN2 atan2f
#define ACOS acosf
#define FMIN fminf
#define FMAX fmaxf
#include "ljlk_oracle_impl.h"
#include "elec_oracle_impl.h"
#include "lkball_oracle_impl.h"
#include "hbond_oracle_impl.h"
#undef REAL
#undef RSUF
#undef SQRT
#undef EXP
#undef FLOOR
#undef LOG
#undef COS
#undef SIN
#undef ATAN2
#undef ACOS
#undef FMIN
#undef FMAX

#define REAL double
#define RSUF _d
#define SQRT sqrt
#define EXP exp
#define FLOOR floor
#define LOG log
#define COS cos
#define SIN sin
#define ATAN2 atan2
#define ACOS acos
#define FMIN fmin
#define FMAX fmax
#include "ljlk_oracle_impl.h"
#include "elec_oracle_impl.h"
#include "lkball_oracle_impl.h"
#include "hbond_oracle_impl.h"
#undef REAL
#undef RSUF
#undef SQRT
#undef EXP
#undef FLOOR
#undef LOG
#undef COS
#undef SIN
#undef ATAN2
#undef ACOS
#undef FMIN
#undef FMAX

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

/* Separations of every atom pair of block pair (p, b1, b2) into out
 * [n1, n2] (row-major; n1, n2 = the blocks' atom counts). */
void oracle_block_pair_seps(const oracle_stack* st, int p, int b1, int b2, int32_t* out) {
  int B = st->max_n_blocks;
  int bt1 = st->block_type[p * B + b1], bt2 = st->block_type[p * B + b2];
  if (bt1 < 0 || bt2 < 0) return;
  int n1 = st->bt_n_atoms[bt1], n2 = st->bt_n_atoms[bt2];
  for (int a1 = 0; a1 < n1; ++a1)
    for (int a2 = 0; a2 < n2; ++a2) out[a1 * n2 + a2] = oracle_pair_sep(st, p, b1, b2, a1, a2);
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

/* inter_block_bondsep [P, B, B, C, C] and min_bondsep [P, B, B] from block
 * connectivity, restating the reference's pose-stack construction
 * (tmol/pose/pose_stack_builder.py:459-599, 1230-1349;
 * tmol/score/bond_dependent_term.py:91-100):
 *  - one graph node per connection point of every real block (numbered
 *    block by block);
 *  - two connection points of one block: the path distance between their
 *    atoms; two points joined by an inter-residue connection: 1; otherwise
 *    6 (no edge); the weight of a node pair is the smaller of its two
 *    directions;
 *  - shortest paths capped at 6 (the reference runs a cutoff-6 Dijkstra,
 *    tmol/pose/compiled/apsp.cpu.cpp; Floyd-Warshall here, which gives the
 *    same capped distances);
 *  - entries of non-existent connection pairs stay 6; min_bondsep is the
 *    minimum over all [C, C] entries.
 * irc: inter-residue connections [P, B, C, 2] = (block, connection), -1 = none. */
void oracle_inter_block_bondsep(int P, int B, int C, int Amax, const int32_t* block_type,
                                const int32_t* irc, const int32_t* bt_n_conn,
                                const int32_t* bt_conn_atoms,
                                const int32_t* bt_path_distance, int32_t* ibs,
                                int32_t* min_sep) {
  const int CUT = 6;
  int* off = (int*)malloc(sizeof(int) * (B + 1));
  for (int p = 0; p < P; ++p) {
    int n = 0;
    for (int b = 0; b < B; ++b) {
      off[b] = n;
      int bt = block_type[p * B + b];
      n += bt >= 0 ? bt_n_conn[bt] : 0;
    }
    off[B] = n;
    int* W = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n * n : 1));
    for (int64_t k = 0; k < (int64_t)n * n; ++k) W[k] = CUT;
    for (int b = 0; b < B; ++b) {
      int bt = block_type[p * B + b];
      if (bt < 0) continue;
      int nc = bt_n_conn[bt];
      for (int c1 = 0; c1 < nc; ++c1)
        for (int c2 = 0; c2 < nc; ++c2) {
          int a1 = bt_conn_atoms[bt * C + c1], a2 = bt_conn_atoms[bt * C + c2];
          W[(off[b] + c1) * n + off[b] + c2] =
              bt_path_distance[((int64_t)bt * Amax + a1) * Amax + a2];
        }
      for (int c = 0; c < C; ++c) {
        const int32_t* e = irc + (((int64_t)p * B + b) * C + c) * 2;
        if (e[0] < 0) continue;
        W[(off[b] + c) * n + off[e[0]] + e[1]] = 1;
      }
    }
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        int w = W[i * n + j] < W[j * n + i] ? W[i * n + j] : W[j * n + i];
        W[i * n + j] = W[j * n + i] = w;
      }
    for (int i = 0; i < n; ++i) W[i * n + i] = 0;
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < n; ++i) {
        int wik = W[i * n + k];
        if (wik >= CUT) continue;
        for (int j = 0; j < n; ++j) {
          int v = wik + W[k * n + j];
          if (v < W[i * n + j]) W[i * n + j] = v;
        }
      }
    for (int b1 = 0; b1 < B; ++b1)
      for (int b2 = 0; b2 < B; ++b2) {
        int m = CUT;
        int bt1 = block_type[p * B + b1], bt2 = block_type[p * B + b2];
        int nc1 = bt1 >= 0 ? bt_n_conn[bt1] : 0, nc2 = bt2 >= 0 ? bt_n_conn[bt2] : 0;
        for (int c1 = 0; c1 < C; ++c1)
          for (int c2 = 0; c2 < C; ++c2) {
            int v = CUT;
            if (c1 < nc1 && c2 < nc2) {
              v = W[(off[b1] + c1) * n + off[b2] + c2];
              if (v > CUT) v = CUT;
            }
            ibs[((((int64_t)p * B + b1) * B + b2) * C + c1) * C + c2] = v;
            if (v < m) m = v;
          }
        min_sep[((int64_t)p * B + b1) * B + b2] = m;
      }
    free(W);
  }
  free(off);
}
