// Shared definitions of the LJ+LK device code: constants, workspace layout,
// views of the stack / block-type tables, wave helpers, count-pair lookups.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tmol_ljlk.h"
#include "ljlk_math.hh"

namespace tmol_ljlk {

constexpr int TILE = 32;         // atoms per block tile
constexpr int MAX_CONN = 4;      // inter-block connections per block type
constexpr int WAVE = 64;         // CDNA wavefront
constexpr int WAVES_PER_WG = 4;  // scoring workgroups: 4 independent waves
constexpr int NBR_TILE = 64;     // neighbor pass: 64 x 64 block tiles
constexpr int REC = 8;           // Reals per prepared (block type, atom) record
constexpr int INTRA_MAX = TILE * (TILE - 1) / 2;  // intra-block atom pairs of a 32-atom type

// prepared per (block type, atom) record:
//   [0] lj_radius  [1] sqrt(lj_wdepth)  [2] K = dg / (2 pi^1.5 lambda)
//   [3] 1 / lambda^2  [4] lk_volume  [5] flags (F_* bits, F_HEAVY)  [6] 1
// per (block type, atom, connection): path distance from the connection atom
constexpr int F_HEAVY = 32;

// score modes (compile-time template bits); the public flags plus
constexpr int M_DVDX = TMOL_LJLK_DVDX;
constexpr int M_WGRAD = TMOL_LJLK_WGRAD;
constexpr int M_BP = TMOL_LJLK_BLOCK_PAIR;
constexpr int M_NOE = TMOL_LJLK_NO_ENERGY;
constexpr int M_CENSUS = 64;  // branch census instead of energies (FLOP accounting)
constexpr int M_ELEC = 256;   // elec scored in the same pass (fourth term)
constexpr int M_DET = TMOL_LJLK_DETERMINISTIC;  // fixed-point global sums (float only)

// census slots (counts of enumerated atom pairs)
enum : int {
  C_LJ_PAIRS = 0, C_LJ_CP0, C_LJ_FAR, C_LJ_LIN, C_LJ_MID, C_LJ_SPL,
  C_LK_PAIRS, C_LK_CP0, C_LK_FAR, C_LK_CONST, C_LK_CLOSE, C_LK_MID, C_LK_SPL,
  C_ITEMS, C_INTRA_ITEMS, C_N
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// number of set bits of `mask` below this lane
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// ------------------------------------------------------------------ workspace

struct WsLayout {
  int64_t counters;   // uint32[16]: [0] block pairs found, [1] row overflow
  int64_t cursors;    // uint32[8 * 64]: row-queue cursors, one per 256-byte line
  int64_t energies;   // double[4 * P] (atr, rep, lk, elec)
  int64_t spheres;    // Real[4 * G]
  int64_t records;    // Real[T * Amax * REC]
  int64_t conn_pd;    // uint8[T * Amax * MAX_CONN]: path distance atom -> connection atom
  int64_t intra;      // int32[T] + uint16[T * INTRA_MAX]: weighted intra-block pairs per type
  int64_t e_pd;       // uint32[T * Amax]: elec representative path distances, packed
  int64_t e_intra;    // int32[T] + uint16[T * INTRA_MAX]: elec's intra-block pairs per type
  int64_t row_count;  // uint32[G]
  int64_t row_items;  // int32[G * row_cap]
  int64_t fixed;      // DETERMINISTIC: int64[det_planes * NA * 3] fixed-point gradients
  int64_t total;
};

inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

inline WsLayout ws_layout(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                          int64_t row_cap, int dtype, int det_planes = 0) {
  int64_t rs = dtype == TMOL_LJLK_F64 ? 8 : 4;
  int64_t G = s->n_blocks_total;
  WsLayout L;
  int64_t o = 0;
  L.counters = o;
  o = align256(o + 64);
  L.cursors = o;
  o = align256(o + 8 * 256);
  L.energies = o;
  o = align256(o + 8 * 4 * int64_t(s->n_poses));
  L.spheres = o;
  o = align256(o + rs * 4 * G);
  L.records = o;
  o = align256(o + rs * REC * int64_t(bt->n_block_types) * bt->max_n_atoms);
  L.conn_pd = o;
  o = align256(o + int64_t(bt->n_block_types) * MAX_CONN * bt->max_n_atoms);
  L.intra = o;
  o = align256(o + 4 * int64_t(bt->n_block_types) + 2 * int64_t(bt->n_block_types) * INTRA_MAX);
  L.e_pd = o;
  o = align256(o + 4 * int64_t(bt->n_block_types) * bt->max_n_atoms);
  L.e_intra = o;
  o = align256(o + 4 * int64_t(bt->n_block_types) + 2 * int64_t(bt->n_block_types) * INTRA_MAX);
  L.row_count = o;
  o = align256(o + 4 * G);
  L.row_items = o;
  o = align256(o + 4 * G * row_cap);
  L.fixed = o;
  o = align256(o + 8 * int64_t(det_planes) * 3 * s->n_atoms_total);
  L.total = o;
  return L;
}

// ------------------------------------------------------------------ views

struct StackView {
  int layout, P, B, G, C;
  const int32_t* block_atom_offset;
  const int32_t* block_type;
  const int32_t* block_pose;
  const int32_t* min_bondsep;
  const int32_t* inter_block_bondsep;
  const int32_t* bonded_start;  // ragged: [G + 1]
  const int32_t* bonded;        // ragged: [n, BONDED_W] forward bonded table

  __device__ __forceinline__ int pose_of(int g) const {
    return layout == TMOL_LJLK_PADDED ? g / B : block_pose[g];
  }
};

struct BTView {
  int T, Amax, nt, C;
  const int32_t* n_atoms;
  const int32_t* n_heavy_in_tile;
  const int32_t* heavy_in_tile;
  const int32_t* atom_types;
  const int32_t* n_conn;
  const int32_t* conn_atoms;
  const int32_t* path_distance;
  const int32_t* is_ligand;
};

// work list: row g1 holds the g2 >= g1 neighbor blocks of g1 in the same pose
struct WorkView {
  const uint32_t* row_count;  // [G]
  const int32_t* row_items;   // [G * row_cap]
  int64_t row_cap;
};

inline StackView make_stack_view(const tmol_ljlk_stack* s, int C) {
  StackView v;
  v.layout = s->layout;
  v.P = s->n_poses;
  v.B = s->max_n_blocks;
  v.G = s->n_blocks_total;
  v.C = C;
  v.block_atom_offset = s->block_atom_offset;
  v.block_type = s->block_type;
  v.block_pose = s->block_pose;
  v.min_bondsep = s->min_bondsep;
  v.inter_block_bondsep = s->inter_block_bondsep;
  v.bonded_start = s->bonded_start;
  v.bonded = s->bonded;
  return v;
}

inline BTView make_bt_view(const tmol_ljlk_block_types* b) {
  BTView v;
  v.T = b->n_block_types;
  v.Amax = b->max_n_atoms;
  v.nt = b->max_n_tiles;
  v.C = b->max_n_conn;
  v.n_atoms = b->n_atoms;
  v.n_heavy_in_tile = b->n_heavy_in_tile;
  v.heavy_in_tile = b->heavy_in_tile;
  v.atom_types = b->atom_types;
  v.n_conn = b->n_conn;
  v.conn_atoms = b->conn_atoms;
  v.path_distance = b->path_distance;
  v.is_ligand = b->is_ligand;
  return v;
}

// Block-pair outputs and weights (BLOCK_PAIR mode):
//   padded: [3, P, B, B] at (pose, b1, b2), the reference's layout;
//   ragged: [3, G, row_cap] at (g1, k), k = the pair's position in row g1 of
//           the work list (one entry per scored block pair).
__device__ __forceinline__ int64_t bp_plane(const StackView& s, const WorkView& w) {
  return s.layout == TMOL_LJLK_PADDED ? int64_t(s.P) * s.B * s.B : int64_t(s.G) * w.row_cap;
}
__device__ __forceinline__ int64_t bp_index(const StackView& s, const WorkView& w, int g1, int g2,
                                            int64_t k) {
  if (s.layout == TMOL_LJLK_PADDED) {
    const int p = g1 / s.B;
    return (int64_t(p) * s.B + (g1 - p * s.B)) * s.B + (g2 - p * s.B);
  }
  return int64_t(g1) * w.row_cap + k;
}

constexpr int BONDED_W = TMOL_LJLK_BONDED_WIDTH;

// ragged: the forward bonded-table row of (lo, hi), lo < hi, or null when the
// two blocks are more than 5 bonds apart (rows are sorted by g2 and short:
// a handful per block)
__device__ __forceinline__ const int32_t* bonded_row(const StackView& s, int lo, int hi) {
  const int k0 = s.bonded_start[lo], k1 = s.bonded_start[lo + 1];
  // rows are sorted by g2: most pairs are past the row's last entry
  if (k1 == k0 || s.bonded[int64_t(k1 - 1) * BONDED_W] < hi) return nullptr;
  for (int k = k0; k < k1; ++k) {
    const int g = s.bonded[int64_t(k) * BONDED_W];
    if (g == hi) return s.bonded + int64_t(k) * BONDED_W;
    if (g > hi) break;
  }
  return nullptr;
}

// bond separation between connection c1 of g1 and c2 of g2 (g1 != g2)
__device__ __forceinline__ int conn_sep(const StackView& s, int g1, int g2, int c1, int c2) {
  if (s.layout == TMOL_LJLK_PADDED) {
    int b2 = g2 - (g2 / s.B) * s.B;
    int64_t idx = ((int64_t(g1) * s.B + b2) * s.C + c1) * s.C + c2;
    return s.inter_block_bondsep[idx];
  }
  if (g2 < g1) {
    const int t = g1; g1 = g2; g2 = t;
    const int u = c1; c1 = c2; c2 = u;
  }
  const int32_t* e = bonded_row(s, g1, g2);
  if (!e) return 6;
  return (int)(((uint32_t)e[4 + c1] >> (8 * c2)) & 255u);
}

// minimum bond separation over connection pairs (min_block_bondsep)
__device__ __forceinline__ int min_block_sep(const StackView& s, int g1, int g2, int nc1,
                                             int nc2) {
  if (s.layout == TMOL_LJLK_PADDED) {
    int b2 = g2 - (g2 / s.B) * s.B;
    return s.min_bondsep[int64_t(g1) * s.B + b2];
  }
  const int32_t* e = g1 < g2 ? bonded_row(s, g1, g2) : bonded_row(s, g2, g1);
  return e ? e[1] : 6;
}

// Inter-block count-pair separation from LDS tables (impl.hh:47-67,
// count_pair.hh:162-207): min over connection pairs of
// pd1[c1][a1] + cs[c1][c2] + pd2[c2][a2], starting at 5, only when the
// blocks' minimum separation is <= 4; both ligand fragments -> 3/4 become 5.
__device__ __forceinline__ int inter_pair_sep(const uint8_t (*pd1)[TILE],
                                              const uint8_t (*pd2)[TILE],
                                              const uint8_t* cs, int min_sep, int nc1,
                                              int nc2, int a1, int a2, bool lig) {
  int sep = min_sep;
  if (sep <= 4) {
    int best = 5;
    for (int c1 = 0; c1 < nc1; ++c1) {
      int p1 = pd1[c1][a1];
      for (int c2 = 0; c2 < nc2; ++c2) {
        int v = p1 + cs[c1 * MAX_CONN + c2] + pd2[c2][a2];
        best = v < best ? v : best;
      }
    }
    sep = best;
  }
  if (lig && (sep == 3 || sep == 4)) sep = 5;
  return sep;
}

// Accumulator for wave-private LDS sums (atom gradients, block-pair
// energies).  On gfx950 a float LDS atomic (ds_add_f32) costs ~190 CU cycles
// per wave instruction against ~6 for a 64-bit integer one (ds_add_u64), so
// float contributions are added as 64-bit fixed point with 2^-28 resolution:
// exact and independent of summation order.  The encoding rounds to nearest
// via the 1.5 * 2^52 magic constant and saturates a single contribution at
// +-2^22 (4.2e6, orders of magnitude above any pair term); sums may reach
// 2^35.  Double contributions accumulate as doubles.
template <typename Real>
struct LdsAcc;

template <>
struct LdsAcc<float> {
  using T = unsigned long long;
  __device__ static __forceinline__ T enc(float v) {
    double d = double(v) * 268435456.0;  // 2^28
    d = fmin(fmax(d, -1125899906842624.0), 1125899906842624.0);  // +-2^50
    d = d + 6755399441055744.0;                                   // 1.5 * 2^52
    return (T)(__double_as_longlong(d) - 0x4338000000000000LL);
  }
  __device__ static __forceinline__ float dec(T x) {
    return float(double((long long)x) * (1.0 / 268435456.0));
  }
};

template <>
struct LdsAcc<double> {
  using T = double;
  __device__ static __forceinline__ T enc(double v) { return v; }
  __device__ static __forceinline__ double dec(T x) { return x; }
};

// M_DET: per-pose energies as 64-bit fixed point with 2^-24 kcal/mol
// resolution (range +-2^39), summed with integer atomics (order-independent)
constexpr double DET_E_SCALE = 16777216.0;  // 2^24
__device__ __forceinline__ unsigned long long det_energy_enc(double v) {
  v = fmin(fmax(v * DET_E_SCALE, -9.0e18), 9.0e18);
  return (unsigned long long)__double2ll_rn(v);
}

// one lane's running energy sum, accumulated in FP64 (each FP32 pair term is
// added exactly; FP64 adds run at the FP32 rate on MI355X)
struct LaneSum {
  double s = 0.0;
  template <typename Real>
  __device__ __forceinline__ void add(Real v) {
    s += double(v);
  }
  __device__ __forceinline__ double value() const { return s; }
};

// compensated (Neumaier) accumulation of one lane's pair energies
template <typename Real>
struct KSum {
  Real s = Real(0), c = Real(0);
  __device__ __forceinline__ void add(Real v) {
    Real t = s + v;
    c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
    s = t;
  }
  __device__ __forceinline__ double value() const { return double(s) + double(c); }
};

template <typename Real>
struct ScoreArgs {
  StackView s;
  BTView bt;
  WorkView w;
  const Real* coords;
  const Real* spheres;  // [G, 4] centre, radius (from the neighbor pass)
  const Real* records;
  const uint8_t* conn_pd;
  const int32_t* intra_count;   // [T] weighted intra-block pairs per type
  const uint16_t* intra_pairs;  // [T, INTRA_MAX]: a1 | a2 << 5 | (separation 4) << 10
  const Real* global_params;
  const uint32_t* counters;
  uint32_t* row_cursor;        // [8 * QSTRIDE] RowQueue cursors (workspace)
  int row_grab;                // rows per queue grab (0: QROWS)
  const Real* weights;         // [3, P] or [3, P, B, B]
  double* energies64;          // [3, P] whole pose
  Real* energies_bp;           // [3, P, B, B] block pair
  Real* dvdx;                  // [3, NA, 3]
  Real* grad;                  // [NA, 3]
  unsigned long long* fixed;   // M_DET: fixed-point gradient planes [NG, NA, 3]
  int64_t n_atoms;
  unsigned long long* census;  // [16] with M_CENSUS
  // M_ELEC: elec's tables (charges [T, Amax]; packed representative path
  // distances and intra-block pairs prepared in the workspace) and constants
  const Real* e_charge;
  const uint32_t* e_pd;
  const int32_t* e_intra_count;
  const uint16_t* e_intra_pairs;
  ElecConst<Real> ek;
};

template <typename Real>
__device__ __forceinline__ GlobalParams<Real> load_global_params(const Real* g) {
  GlobalParams<Real> gp;
  gp.max_dis = g[0];
  gp.dlin_factor = g[1];
  gp.hbond_dis = g[2];
  gp.hbond_oh_dis = g[3];
  gp.hbond_hdis = g[4];
  gp.spline_start = gp.max_dis - Real(1.5);
  gp.cpoly_cap = gp.max_dis - Real(0.1);
  gp.max_dis2 = gp.max_dis * gp.max_dis;
  return gp;
}

// Dynamic row scheduling for the persistent scoring kernels.  Rows [0, n)
// are split into 8 contiguous ranges, one per XCD (workgroups are dispatched
// round-robin over the 8 XCDs of MI355X, so workgroup b runs on XCD b % 8 and
// the pose-ordered rows of its range share that XCD's L2).  A wave takes rows
// of its own range through an atomic cursor and, once that range is
// exhausted, rows of the other ranges: no wave idles while rows remain.  Rows
// are handed out one at a time (a row is tens of microseconds of one wave's
// work, so larger grabs lengthen the tail) and each cursor sits on its own
// 256-byte line (same-line L2 atomics serialize).  The cursors are zeroed by
// a fill kernel before every scoring launch.
// The rows taken per cursor grab (`grab`, default 1) is a launch parameter the
// autotuner may raise.
constexpr int QROWS = 1;
constexpr int QSTRIDE = 64;  // uint32 between cursors

struct RowQueue {
  uint32_t* cursor;
  int64_t n;
  int64_t cur, end;
  int xcd, left, grab;
  __device__ RowQueue(uint32_t* c, int64_t n_rows, int rows_per_grab = QROWS)
      : cursor(c), n(n_rows), cur(0), end(0), xcd(int(blockIdx.x & 7)), left(8),
        grab(rows_per_grab > 0 ? rows_per_grab : QROWS) {}
  // next row of this wave (wave-uniform), -1 when every row is taken
  __device__ int64_t next() {
    if (cur < end) return cur++;
    while (left > 0) {
      const int64_t lo = n * xcd / 8, hi = n * (xcd + 1) / 8;
      uint32_t k = 0;
      if (lane_id() == 0) k = atomicAdd(&cursor[xcd * QSTRIDE], uint32_t(grab));
      k = __builtin_amdgcn_readfirstlane(k);
      if (lo + int64_t(k) < hi) {
        cur = lo + int64_t(k);
        end = cur + grab < hi ? cur + grab : hi;
        return cur++;
      }
      xcd = (xcd + 1) & 7;
      --left;
    }
    return -1;
  }
};

}  // namespace tmol_ljlk
