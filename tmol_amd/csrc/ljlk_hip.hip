// MI355X (gfx950) LJ+LK pose scoring: neighbor pass, scoring kernel,
// backward combine, and the C ABI declared in include/tmol_ljlk.h.
//
// Device work per call (reference: ljlk_pose_score.impl.hh:1123-1155):
//   1. prepare + spheres   one thread per block type atom / per block
//   2. neighbors           64x64 block tiles, wave ballot compaction into a
//                          global work list of block pairs (g1 <= g2)
//   3. score               persistent waves walk the work list; each wave
//                          stages one block pair in LDS, evaluates LJ over all
//                          atom pairs and LK over heavy-atom pairs, reduces
//                          energies (compensated FP32 per lane, FP64 across
//                          the wave and across block pairs) and scatters
//                          gradients through wave-private LDS accumulators
//                          (one global atomic per atom, term and component
//                          per block pair)
//   4. finalize            FP64 per-pose totals -> output dtype
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "../../include/tmol_ljlk.h"
#include "ljlk_math.hh"

namespace tmol_ljlk {

constexpr int TILE = 32;
constexpr int MAX_CONN = 4;
constexpr int WAVE = 64;
constexpr int WAVES_PER_WG = 4;
constexpr int NBR_TILE = 64;
constexpr int REC = 8;  // Reals per prepared atom record

// prepared per (block type, atom) record
//   [0] lj_radius  [1] sqrt(lj_wdepth)  [2] K = dg / (2 pi^1.5 lambda)
//   [3] 1 / lambda^2  [4] lk_volume  [5] flags  [6] 1 if the type is real
// per (block type, connection, atom): path distance from the connection atom

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }

template <typename Real>
__device__ __forceinline__ Real wave_sum(Real v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// ------------------------------------------------------------------ layout

struct WsLayout {
  int64_t counters;   // uint32[16]
  int64_t energies;   // double[3 * P]
  int64_t spheres;    // Real[4 * G]
  int64_t records;    // Real[T * Amax * REC]
  int64_t conn_pd;    // uint8[T * MAX_CONN * Amax]
  int64_t items;      // int2[capacity]
  int64_t total;
};

static int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

static WsLayout ws_layout(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                          int64_t capacity, int dtype) {
  int64_t rs = dtype == TMOL_LJLK_F64 ? 8 : 4;
  WsLayout L;
  int64_t o = 0;
  L.counters = o;
  o = align256(o + 64);
  L.energies = o;
  o = align256(o + 8 * 3 * int64_t(s->n_poses));
  L.spheres = o;
  o = align256(o + rs * 4 * int64_t(s->n_blocks_total));
  L.records = o;
  o = align256(o + rs * REC * int64_t(bt->n_block_types) * bt->max_n_atoms);
  L.conn_pd = o;
  o = align256(o + int64_t(bt->n_block_types) * MAX_CONN * bt->max_n_atoms);
  L.items = o;
  o = align256(o + 8 * capacity);
  L.total = o;
  return L;
}

// ------------------------------------------------------------------ views

struct StackView {
  int layout, P, B, G;
  const int32_t* block_atom_offset;
  const int32_t* block_type;
  const int32_t* block_pose;
  const int32_t* min_bondsep;
  const int32_t* inter_block_bondsep;
  const int32_t* block_conn_pos;
  const int32_t* block_segment;
  int C;

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

static StackView make_stack_view(const tmol_ljlk_stack* s, int C) {
  StackView v;
  v.layout = s->layout;
  v.P = s->n_poses;
  v.B = s->max_n_blocks;
  v.G = s->n_blocks_total;
  v.block_atom_offset = s->block_atom_offset;
  v.block_type = s->block_type;
  v.block_pose = s->block_pose;
  v.min_bondsep = s->min_bondsep;
  v.inter_block_bondsep = s->inter_block_bondsep;
  v.block_conn_pos = s->block_conn_pos;
  v.block_segment = s->block_segment;
  v.C = C;
  return v;
}

static BTView make_bt_view(const tmol_ljlk_block_types* b) {
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

// bond separation between connection c1 of g1 and c2 of g2 (g1 != g2)
__device__ __forceinline__ int conn_sep(const StackView& s, int g1, int g2, int c1,
                                        int c2) {
  if (s.layout == TMOL_LJLK_PADDED) {
    int b2 = g2 - (g2 / s.B) * s.B;
    int64_t idx = ((int64_t(g1) * s.B + b2) * s.C + c1) * s.C + c2;
    return s.inter_block_bondsep[idx];
  } else {
    int p1 = s.block_conn_pos[int64_t(g1) * s.C + c1];
    int p2 = s.block_conn_pos[int64_t(g2) * s.C + c2];
    if (p1 < 0 || p2 < 0 || s.block_segment[g1] != s.block_segment[g2]) return 6;
    int d = p1 > p2 ? p1 - p2 : p2 - p1;
    return d < 6 ? d : 6;
  }
}

__device__ __forceinline__ int min_block_sep(const StackView& s, int g1, int g2,
                                             int nc1, int nc2) {
  if (s.layout == TMOL_LJLK_PADDED) {
    int b2 = g2 - (g2 / s.B) * s.B;
    return s.min_bondsep[int64_t(g1) * s.B + b2];
  }
  int m = 6;
  for (int c1 = 0; c1 < nc1; ++c1)
    for (int c2 = 0; c2 < nc2; ++c2) {
      int v = conn_sep(s, g1, g2, c1, c2);
      m = v < m ? v : m;
    }
  return m;
}

// ------------------------------------------------------------------ prepare

template <typename Real>
__global__ void prepare_records_kernel(BTView bt, const Real* __restrict__ type_params,
                                       int n_atom_types, Real* __restrict__ records,
                                       uint8_t* __restrict__ conn_pd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = bt.T * bt.Amax;
  if (i >= n) return;
  int t = i / bt.Amax, a = i - t * bt.Amax;
  int at = bt.atom_types[i];
  Real* r = records + int64_t(i) * REC;
  if (at < 0 || at >= n_atom_types || a >= bt.n_atoms[t]) {
    for (int k = 0; k < REC; ++k) r[k] = Real(0);
  } else {
    const Real* tp = type_params + int64_t(at) * 10;
    Real lj_radius = tp[0], wdepth = tp[1], dg = tp[2], lam = tp[3], vol = tp[4];
    int flags = 0;
    if (tp[5] != Real(0)) flags |= F_DONOR;
    if (tp[5] != Real(0) && tp[6] != Real(0)) flags |= F_OH_DONOR;
    if (tp[7] != Real(0)) flags |= F_POLARH;
    if (tp[8] != Real(0)) flags |= F_ACCEPTOR;
    if (tp[9] != Real(0)) flags |= F_CARBON_LK;
    const Real pi_pow1p5 = Real(5.56832799683f);  // lk_isotropic.hh:40
    r[0] = lj_radius;
    r[1] = sqrt(wdepth);
    r[2] = dg / (Real(2) * pi_pow1p5 * lam);
    r[3] = Real(1) / (lam * lam);
    r[4] = vol;
    r[5] = Real(flags);
    r[6] = Real(1);
    r[7] = Real(0);
  }
  int nc = bt.n_conn[t];
  for (int c = 0; c < MAX_CONN; ++c) {
    uint8_t v = 6;
    if (c < nc && a < bt.n_atoms[t]) {
      int ca = bt.conn_atoms[t * bt.C + c];
      v = (uint8_t)bt.path_distance[(int64_t(t) * bt.Amax + ca) * bt.Amax + a];
    }
    conn_pd[(int64_t(t) * MAX_CONN + c) * bt.Amax + a] = v;
  }
}

// ------------------------------------------------------------------ spheres

// Bounding sphere per block, bit-identical to the reference's CPU pass
// (sphere_overlap.impl.hh:109-196 with nt = 1): coordinates summed in atom
// order, divided by n, radius = sqrt(max d^2) with ((dx*dx + dy*dy) + dz*dz);
// no contraction, IEEE division and square root.
template <typename Real>
__global__ void spheres_kernel(StackView s, BTView bt, const Real* __restrict__ coords,
                               Real* __restrict__ spheres) {
#pragma clang fp contract(off)
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= s.G) return;
  int t = s.block_type[g];
  Real* out = spheres + int64_t(g) * 4;
  if (t < 0) {
    out[0] = out[1] = out[2] = out[3] = Real(0);
    return;
  }
  int n = bt.n_atoms[t];
  const Real* c = coords + int64_t(s.block_atom_offset[g]) * 3;
  Real cx = 0, cy = 0, cz = 0;
  for (int a = 0; a < n; ++a) {
    cx = cx + c[3 * a + 0];
    cy = cy + c[3 * a + 1];
    cz = cz + c[3 * a + 2];
  }
  Real rn = Real(n);
  cx = cx / rn;
  cy = cy / rn;
  cz = cz / rn;
  Real d2max = 0;
  for (int a = 0; a < n; ++a) {
    Real dx = c[3 * a + 0] - cx;
    Real dy = c[3 * a + 1] - cy;
    Real dz = c[3 * a + 2] - cz;
    Real dxx = dx * dx;
    Real dyy = dy * dy;
    Real dzz = dz * dz;
    Real d2 = (dxx + dyy) + dzz;
    d2max = d2 > d2max ? d2 : d2max;
  }
  out[0] = cx;
  out[1] = cy;
  out[2] = cz;
  out[3] = sqrt(d2max);
}

// ------------------------------------------------------------------ neighbors

// One workgroup per 64x64 tile of a pose's block-pair upper triangle.  Wave w
// tests rows 16w..16w+15 against the 64 columns (one lane per column), keeps
// the 16 ballot masks in scalar registers, reserves space for all of its hits
// with one atomic and writes them in row order.
template <typename Real>
__global__ __launch_bounds__(256) void neighbors_kernel(
    StackView s, const int4* __restrict__ tiles, int64_t n_tiles,
    const Real* __restrict__ spheres, Real reach, int2* __restrict__ items,
    uint32_t* __restrict__ counters, int64_t capacity, int32_t* __restrict__ dense) {
#pragma clang fp contract(off)
  __shared__ Real sph_r[NBR_TILE][4];
  __shared__ Real sph_c[NBR_TILE][4];
  __shared__ int typ_r[NBR_TILE];
  __shared__ int typ_c[NBR_TILE];
  int64_t tile_id = blockIdx.x;
  if (tile_id >= n_tiles) return;
  int4 tl = tiles[tile_id];
  int row0 = tl.x, col0 = tl.y, end = tl.z;
  int tid = threadIdx.x;
  if (tid < NBR_TILE) {
    int g = row0 + tid;
    bool ok = g < end;
    for (int k = 0; k < 4; ++k) sph_r[tid][k] = ok ? spheres[int64_t(g) * 4 + k] : Real(0);
    typ_r[tid] = ok ? s.block_type[g] : -1;
  } else if (tid < 2 * NBR_TILE) {
    int j = tid - NBR_TILE;
    int g = col0 + j;
    bool ok = g < end;
    for (int k = 0; k < 4; ++k) sph_c[j][k] = ok ? spheres[int64_t(g) * 4 + k] : Real(0);
    typ_c[j] = ok ? s.block_type[g] : -1;
  }
  __syncthreads();
  int lane = lane_id();
  int wave = tid / WAVE;
  int g2 = col0 + lane;
  Real c0 = sph_c[lane][0], c1 = sph_c[lane][1], c2 = sph_c[lane][2], c3 = sph_c[lane][3];
  bool col_ok = typ_c[lane] >= 0;
  uint64_t masks[16];
  int total = 0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    int lr = wave * 16 + r;
    int g1 = row0 + lr;
    bool ok = col_ok && typ_r[lr] >= 0 && g1 <= g2;
    Real dx = sph_r[lr][0] - c0;
    Real dy = sph_r[lr][1] - c1;
    Real dz = sph_r[lr][2] - c2;
    Real dxx = dx * dx;
    Real dyy = dy * dy;
    Real dzz = dz * dz;
    Real d2 = (dxx + dyy) + dzz;
    Real thr = (sph_r[lr][3] + c3) + reach;
    Real thr2 = thr * thr;
    bool hit = ok && (d2 < thr2);
    masks[r] = __ballot(hit);
    total += __popcll(masks[r]);
  }
  if (total == 0) return;
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(counters, (uint32_t)total);
  base = __shfl(base, 0, WAVE);
  uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t run = base;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    uint64_t m = masks[r];
    if ((m >> lane) & 1ull) {
      uint32_t idx = run + __popcll(m & below);
      int g1 = row0 + wave * 16 + r;
      if (idx < capacity) items[idx] = make_int2(g1, g2);
      if (dense) {
        int b2 = g2 - (g2 / s.B) * s.B;
        dense[int64_t(g1) * s.B + b2] = 1;
      }
    }
    run += __popcll(m);
  }
  if (lane == 0 && uint64_t(base) + total > uint64_t(capacity)) counters[1] = 1;
}

// ------------------------------------------------------------------ scoring

template <typename Real>
struct WaveLDS {
  Real x[2 * TILE], y[2 * TILE], z[2 * TILE];
  Real rad[2 * TILE], sw[2 * TILE], K[2 * TILE], il2[2 * TILE], vol[2 * TILE];
  int flg[2 * TILE];
  Real grad[3][3][2 * TILE];  // [term][component][slot]
  uint8_t heavy[2][TILE];
  uint8_t pd[2][MAX_CONN][TILE];
  uint8_t cs[MAX_CONN * MAX_CONN];
  int n_heavy[2];
};

template <typename Real>
struct ScoreArgs {
  StackView s;
  BTView bt;
  const Real* coords;
  const Real* records;
  const uint8_t* conn_pd;
  const Real* global_params;
  const int2* items;
  const uint32_t* counters;
  int64_t capacity;
  const Real* weights;   // [3, P] or [3, P, B, B]
  double* energies64;    // [3, P] whole pose
  Real* energies_bp;     // [3, P, B, B] block pair
  Real* dvdx;            // [3, NA, 3]
  Real* grad;            // [NA, 3]
  int64_t n_atoms;
  unsigned long long* census;  // [16] with M_CENSUS
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

// Per block-pair count-pair separation for tile slots (a1, a2).
// Inter-block: min over connection pairs of pd1 + conn_sep + pd2 starting at
// 5, only when the blocks' minimum separation is <= 4 (impl.hh:47-67,
// count_pair.hh:162-207); both ligand fragments -> 3/4 become 5.
template <typename Real>
__device__ __forceinline__ int inter_pair_sep(const WaveLDS<Real>& L, int min_sep,
                                              int nc1, int nc2, int a1, int a2,
                                              bool lig) {
  int sep = min_sep;
  if (sep <= 4) {
    int best = 5;
    for (int c1 = 0; c1 < nc1; ++c1) {
      int p1 = L.pd[0][c1][a1];
      for (int c2 = 0; c2 < nc2; ++c2) {
        int v = p1 + L.cs[c1 * MAX_CONN + c2] + L.pd[1][c2][a2];
        best = v < best ? v : best;
      }
    }
    sep = best;
  }
  if (lig && (sep == 3 || sep == 4)) sep = 5;
  return sep;
}

template <typename Real>
__device__ __forceinline__ void grad_add(WaveLDS<Real>& L, int term, int slot,
                                         Real gx, Real gy, Real gz) {
  atomicAdd(&L.grad[term][0][slot], gx);
  atomicAdd(&L.grad[term][1][slot], gy);
  atomicAdd(&L.grad[term][2][slot], gz);
}

// flags
constexpr int M_DVDX = TMOL_LJLK_DVDX;
constexpr int M_WGRAD = TMOL_LJLK_WGRAD;
constexpr int M_BP = TMOL_LJLK_BLOCK_PAIR;
constexpr int M_NOE = TMOL_LJLK_NO_ENERGY;
constexpr int M_CENSUS = 64;  // branch census instead of energies (FLOP accounting)

// census slots (counts of enumerated atom pairs)
enum : int {
  C_LJ_PAIRS = 0, C_LJ_CP0, C_LJ_FAR, C_LJ_LIN, C_LJ_MID, C_LJ_SPL,
  C_LK_PAIRS, C_LK_CP0, C_LK_FAR, C_LK_CONST, C_LK_CLOSE, C_LK_MID, C_LK_SPL,
  C_ITEMS, C_INTRA_ITEMS, C_N
};

template <typename Real, int MODE>
__global__ __launch_bounds__(256) void score_kernel(ScoreArgs<Real> A) {
  constexpr bool DVDX = MODE & M_DVDX;
  constexpr bool WGRAD = MODE & M_WGRAD;
  constexpr bool BP = MODE & M_BP;
  constexpr bool CENSUS = MODE & M_CENSUS;
  constexpr bool NOE = (MODE & M_NOE) || CENSUS;
  constexpr bool DERIV = (DVDX || WGRAD) && !CENSUS;
  uint32_t cnt[C_N];
#pragma unroll
  for (int k = 0; k < C_N; ++k) cnt[k] = 0;
  __shared__ WaveLDS<Real> lds_all[WAVES_PER_WG];
  const int lane = lane_id();
  const int wave_in_wg = threadIdx.x / WAVE;
  WaveLDS<Real>& L = lds_all[wave_in_wg];

  GlobalParams<Real> gp;
  gp.max_dis = A.global_params[0];
  gp.dlin_factor = A.global_params[1];
  gp.hbond_dis = A.global_params[2];
  gp.hbond_oh_dis = A.global_params[3];
  gp.hbond_hdis = A.global_params[4];
  gp.spline_start = gp.max_dis - Real(1.5);
  gp.cpoly_cap = gp.max_dis - Real(0.1);
  gp.max_dis2 = gp.max_dis * gp.max_dis;

  int64_t n_items = A.counters[0];
  if (n_items > A.capacity) n_items = A.capacity;
  // XCD-aware split: workgroups are dealt round-robin to the 8 XCDs; give
  // each XCD a contiguous slice of the (pose-ordered) work list so block data
  // is reused in that XCD's L2.
  const int n_wg = gridDim.x;
  const int n_xcd = n_wg >= 8 ? 8 : 1;
  const int xcd = blockIdx.x % n_xcd;
  const int wg_in_xcd = blockIdx.x / n_xcd;
  const int wg_per_xcd = (n_wg - xcd + n_xcd - 1) / n_xcd;
  const int64_t waves_here = int64_t(wg_per_xcd) * WAVES_PER_WG;
  const int64_t lo = n_items * xcd / n_xcd;
  const int64_t hi = n_items * (xcd + 1) / n_xcd;
  const StackView& s = A.s;
  const BTView& bt = A.bt;

  for (int64_t it = lo + int64_t(wg_in_xcd) * WAVES_PER_WG + wave_in_wg; it < hi;
       it += waves_here) {
    int2 item = A.items[it];
    const int g1 = item.x, g2 = item.y;
    const int t1 = s.block_type[g1], t2 = s.block_type[g2];
    const int p = s.pose_of(g1);
    const int n1 = bt.n_atoms[t1], n2 = bt.n_atoms[t2];
    const int64_t o1 = s.block_atom_offset[g1], o2 = s.block_atom_offset[g2];
    const bool intra = g1 == g2;
    const int nc1 = bt.n_conn[t1], nc2 = bt.n_conn[t2];
    const bool lig = bt.is_ligand[t1] && bt.is_ligand[t2];
    int min_sep = intra ? 0 : min_block_sep(s, g1, g2, nc1, nc2);
    const bool need_cp = !intra && min_sep <= 4;
    if (need_cp) {
      if (lane < MAX_CONN * MAX_CONN) {
        int c1 = lane / MAX_CONN, c2 = lane % MAX_CONN;
        L.cs[lane] = (c1 < nc1 && c2 < nc2) ? (uint8_t)conn_sep(s, g1, g2, c1, c2) : 6;
      }
    }
    Real w[3] = {Real(0), Real(0), Real(0)};
    if (WGRAD) {
      if (BP) {
        int b1 = g1 - p * s.B, b2 = g2 - p * s.B;
        int64_t PBB = int64_t(s.P) * s.B * s.B;
        int64_t base = (int64_t(p) * s.B + b1) * s.B + b2;
        for (int t = 0; t < 3; ++t) w[t] = A.weights[t * PBB + base];
      } else {
        for (int t = 0; t < 3; ++t) w[t] = A.weights[t * s.P + p];
      }
    }
    KSum<Real> e_atr, e_rep, e_lk;
    if (CENSUS && lane == 0) {
      cnt[C_ITEMS] += 1;
      cnt[C_INTRA_ITEMS] += intra ? 1 : 0;
    }
    const int nti1 = (n1 - 1) / TILE + 1, nti2 = (n2 - 1) / TILE + 1;
    for (int ti = 0; ti < nti1; ++ti) {
      for (int tj = intra ? ti : 0; tj < nti2; ++tj) {
        const int s1 = ti * TILE, s2 = tj * TILE;
        const int m1 = min(TILE, n1 - s1), m2 = min(TILE, n2 - s2);
        const bool same_tile = intra && ti == tj;
        wave_sync();
        // ---- stage both tiles: slots 0..31 = block 1 tile, 32..63 = block 2
        {
          int slot = lane;
          int half = slot >= TILE;
          int a = slot - half * TILE;
          int tt = half ? t2 : t1;
          int m = half ? m2 : m1;
          int st = half ? s2 : s1;
          int64_t o = half ? o2 : o1;
          bool ok = a < m && !(half && same_tile);
          if (ok) {
            const Real* xyz = A.coords + (o + st + a) * 3;
            L.x[slot] = xyz[0];
            L.y[slot] = xyz[1];
            L.z[slot] = xyz[2];
            const Real* r = A.records + (int64_t(tt) * bt.Amax + st + a) * REC;
            L.rad[slot] = r[0];
            L.sw[slot] = r[1];
            L.K[slot] = r[2];
            L.il2[slot] = r[3];
            L.vol[slot] = r[4];
            L.flg[slot] = int(r[5]);
            if (need_cp) {
              for (int c = 0; c < MAX_CONN; ++c)
                L.pd[half][c][a] = A.conn_pd[(int64_t(tt) * MAX_CONN + c) * bt.Amax + st + a];
            }
          }
          if (a < TILE) {
            int hv = bt.heavy_in_tile[int64_t(tt) * TILE * bt.nt + (half ? tj : ti) * TILE + a];
            L.heavy[half][a] = (uint8_t)(hv < 0 ? 0 : hv);
          }
          if (a == 0) L.n_heavy[half] = bt.n_heavy_in_tile[tt * bt.nt + (half ? tj : ti)];
          if (DERIV) {
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
              for (int k = 0; k < 3; ++k) L.grad[t][k][slot] = Real(0);
          }
        }
        wave_sync();
        // slot offsets of the two atom sets
        const int base2 = same_tile ? 0 : TILE;
        // ---- LJ: all atom pairs, diagonal order (rows = smaller set)
        {
          const bool swap = m1 > m2;
          const int mr = swap ? m2 : m1;  // rows
          const int mc = swap ? m1 : m2;  // columns
          const int q = WAVE / mr, rstep = WAVE - q * mr;
          int r = lane % mr, tcol = lane / mr;
          for (; tcol < mc; ) {
            int c = r + tcol;
            c = c >= mc ? c - mc : c;
            int a1 = swap ? c : r;  // slot within tile 1
            int a2 = swap ? r : c;  // slot within tile 2
            bool ok = !same_tile || (a1 < a2);
            int i1 = a1, i2 = base2 + a2;
            if (ok) {
              int sep;
              if (intra) {
                sep = bt.path_distance[(int64_t(t1) * bt.Amax + s1 + a1) * bt.Amax + s2 + a2];
              } else {
                sep = inter_pair_sep(L, min_sep, nc1, nc2, a1, a2, lig);
              }
              Real dx = L.x[i1] - L.x[i2], dy = L.y[i1] - L.y[i2], dz = L.z[i1] - L.z[i2];
              Real d2 = dx * dx + dy * dy + dz * dz;
              if (CENSUS) {
                cnt[C_LJ_PAIRS] += 1;
                if (sep < 4) {
                  cnt[C_LJ_CP0] += 1;
                } else if (!(d2 < gp.max_dis2)) {
                  cnt[C_LJ_FAR] += 1;
                } else {
                  Real dist = fast_sqrt(d2);
                  Real sigma = lj_sigma(L.flg[i1], L.flg[i2], L.rad[i1], L.rad[i2], gp);
                  Real d_lin = sigma * gp.dlin_factor;
                  Real cmin = sigma > gp.spline_start ? sigma : gp.spline_start;
                  cmin = cmin > gp.cpoly_cap ? gp.cpoly_cap : cmin;
                  cnt[dist < d_lin ? C_LJ_LIN : (dist < cmin ? C_LJ_MID : C_LJ_SPL)] += 1;
                }
              } else if (sep >= 4 && d2 < gp.max_dis2) {
                Real dist = fast_sqrt(d2);
                int f1 = L.flg[i1], f2 = L.flg[i2];
                Real sigma = lj_sigma(f1, f2, L.rad[i1], L.rad[i2], gp);
                Real eps = L.sw[i1] * L.sw[i2];
                Real wcp = cp_weight<Real>(sep);
                Real atr, rep, dlj;
                lj_pair<Real, DERIV>(dist, sigma, eps, wcp, gp, atr, rep, dlj);
                e_atr.add(atr);
                e_rep.add(rep);
                if (DERIV && dlj != Real(0) && dist > Real(0)) {
                  Real inv_d = fast_rcp(dist);
                  bool inside = dist < sigma;
                  if (DVDX) {
                    Real gs = dlj * inv_d;
                    int term = inside ? 1 : 0;
                    grad_add(L, term, i1, gs * dx, gs * dy, gs * dz);
                    grad_add(L, term, i2, -gs * dx, -gs * dy, -gs * dz);
                  } else {
                    Real gs = dlj * inv_d * (inside ? w[1] : w[0]);
                    grad_add(L, 0, i1, gs * dx, gs * dy, gs * dz);
                    grad_add(L, 0, i2, -gs * dx, -gs * dy, -gs * dz);
                  }
                }
              }
            }
            r += rstep;
            tcol += q;
            if (r >= mr) {
              r -= mr;
              tcol += 1;
            }
          }
        }
        // ---- LK: heavy-atom pairs
        {
          const int h1 = L.n_heavy[0];
          const int h2 = same_tile ? h1 : L.n_heavy[1];
          const int hb = same_tile ? 0 : 1;
          if (h1 > 0 && h2 > 0) {
            const bool swap = h1 > h2;
            const int mr = swap ? h2 : h1;
            const int mc = swap ? h1 : h2;
            const int q = WAVE / mr, rstep = WAVE - q * mr;
            int r = lane % mr, tcol = lane / mr;
            for (; tcol < mc; ) {
              int c = r + tcol;
              c = c >= mc ? c - mc : c;
              int k1 = swap ? c : r;
              int k2 = swap ? r : c;
              bool ok = !same_tile || (k1 < k2);
              if (ok) {
                int a1 = L.heavy[0][k1];
                int a2 = L.heavy[hb][k2];
                int i1 = a1, i2 = base2 + a2;
                int sep;
                if (intra) {
                  sep = bt.path_distance[(int64_t(t1) * bt.Amax + s1 + a1) * bt.Amax + s2 + a2];
                } else {
                  sep = inter_pair_sep(L, min_sep, nc1, nc2, a1, a2, lig);
                }
                Real dx = L.x[i1] - L.x[i2], dy = L.y[i1] - L.y[i2], dz = L.z[i1] - L.z[i2];
                Real d2 = dx * dx + dy * dy + dz * dz;
                if (CENSUS) {
                  cnt[C_LK_PAIRS] += 1;
                  if (sep < 4) {
                    cnt[C_LK_CP0] += 1;
                  } else if (!(d2 < gp.max_dis2)) {
                    cnt[C_LK_FAR] += 1;
                  } else {
                    Real dist = fast_sqrt(d2);
                    int f1 = L.flg[i1], f2 = L.flg[i2];
                    Real sigma = lj_sigma(f1, f2, L.rad[i1], L.rad[i2], gp);
                    bool cc = (f1 & F_CARBON_LK) && (f2 & F_CARBON_LK);
                    Real d_min, cdmin, cdmax;
                    lk_knots(sigma, cc, d_min, cdmin, cdmax);
                    cnt[dist < cdmin ? C_LK_CONST
                                     : (dist < cdmax ? C_LK_CLOSE
                                                     : (dist < gp.spline_start ? C_LK_MID : C_LK_SPL))] += 1;
                  }
                } else if (sep >= 4 && d2 < gp.max_dis2) {
                  Real dist = fast_sqrt(d2);
                  int f1 = L.flg[i1], f2 = L.flg[i2];
                  Real sigma = lj_sigma(f1, f2, L.rad[i1], L.rad[i2], gp);
                  bool cc = (f1 & F_CARBON_LK) && (f2 & F_CARBON_LK);
                  Real d_min, cdmin, cdmax;
                  lk_knots(sigma, cc, d_min, cdmin, cdmax);
                  Real v1, dv1, v2, dv2;
                  lk_dir<Real, DERIV>(dist, d_min, cdmin, cdmax, -L.vol[i2] * L.K[i1], L.rad[i1],
                                      L.il2[i1], gp, v1, dv1);
                  lk_dir<Real, DERIV>(dist, d_min, cdmin, cdmax, -L.vol[i1] * L.K[i2], L.rad[i2],
                                      L.il2[i2], gp, v2, dv2);
                  Real wcp = cp_weight<Real>(sep);
                  e_lk.add(wcp * (v1 + v2));
                  if (DERIV && dist > Real(0)) {
                    Real dlk = wcp * (dv1 + dv2);
                    if (dlk != Real(0)) {
                      Real gs = dlk * fast_rcp(dist);
                      if (WGRAD) gs *= w[2];
                      int term = DVDX ? 2 : 0;
                      grad_add(L, term, i1, gs * dx, gs * dy, gs * dz);
                      grad_add(L, term, i2, -gs * dx, -gs * dy, -gs * dz);
                    }
                  }
                }
              }
              r += rstep;
              tcol += q;
              if (r >= mr) {
                r -= mr;
                tcol += 1;
              }
            }
          }
        }
        wave_sync();
        // ---- flush this tile pair's gradients: one atomic per nonzero value
        if (DERIV) {
          int slot = lane;
          int half = slot >= TILE;
          int a = slot - half * TILE;
          int m = half ? m2 : m1;
          bool ok = a < m && !(half && same_tile);
          if (ok) {
            int64_t atom = (half ? o2 + s2 : o1 + s1) + a;
            if (DVDX) {
#pragma unroll
              for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                  Real v = L.grad[t][k][slot];
                  if (v != Real(0))
                    unsafeAtomicAdd(&A.dvdx[(int64_t(t) * A.n_atoms + atom) * 3 + k], v);
                }
            } else {
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                Real v = L.grad[0][k][slot];
                if (v != Real(0)) unsafeAtomicAdd(&A.grad[atom * 3 + k], v);
              }
            }
          }
        }
      }
    }
    if (!NOE) {
      double ea = wave_sum(e_atr.value());
      double er = wave_sum(e_rep.value());
      double el = wave_sum(e_lk.value());
      if (lane == 0) {
        if (BP) {
          int b1 = g1 - p * s.B, b2 = g2 - p * s.B;
          int64_t PBB = int64_t(s.P) * s.B * s.B;
          int64_t base = (int64_t(p) * s.B + b1) * s.B + b2;
          A.energies_bp[base] = Real(ea);
          A.energies_bp[PBB + base] = Real(er);
          A.energies_bp[2 * PBB + base] = Real(el);
        } else {
          unsafeAtomicAdd(&A.energies64[p], ea);
          unsafeAtomicAdd(&A.energies64[s.P + p], er);
          unsafeAtomicAdd(&A.energies64[2 * s.P + p], el);
        }
      }
    }
  }
  if (CENSUS) {
#pragma unroll
    for (int k = 0; k < C_N; ++k) {
      uint32_t v = cnt[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
      if (lane == 0 && v) atomicAdd(&A.census[k], (unsigned long long)v);
    }
  }
}

// ------------------------------------------------------------------ small kernels

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, int64_t n, T v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename Real>
__global__ void finalize_kernel(const double* __restrict__ e64, Real* __restrict__ out,
                                int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = Real(e64[i]);
}

template <typename Real>
__global__ void combine_kernel(const Real* __restrict__ dvdx, const Real* __restrict__ gs,
                               const int32_t* __restrict__ atom_pose, int64_t n_atoms,
                               int n_poses, Real* __restrict__ grad) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_atoms * 3) return;
  int64_t a = i / 3;
  int p = atom_pose[a];
  Real v = gs[p] * dvdx[i];
  v += gs[n_poses + p] * dvdx[n_atoms * 3 + i];
  v += gs[2 * n_poses + p] * dvdx[2 * n_atoms * 3 + i];
  grad[i] = v;
}

// Count-pair separations of the listed block pairs, computed with the same
// staging and device function as score_kernel (one wave per item).
template <typename Real>
__global__ __launch_bounds__(64) void debug_sep_kernel(StackView s, BTView bt,
                                                       const uint8_t* __restrict__ conn_pd,
                                                       const int32_t* __restrict__ items,
                                                       int64_t n_items, uint8_t* __restrict__ out) {
  __shared__ WaveLDS<Real> L;
  int64_t it = blockIdx.x;
  if (it >= n_items) return;
  const int lane = lane_id();
  int g1 = items[2 * it], g2 = items[2 * it + 1];
  int t1 = s.block_type[g1], t2 = s.block_type[g2];
  int n1 = min(TILE, bt.n_atoms[t1]), n2 = min(TILE, bt.n_atoms[t2]);
  int nc1 = bt.n_conn[t1], nc2 = bt.n_conn[t2];
  bool lig = bt.is_ligand[t1] && bt.is_ligand[t2];
  bool intra = g1 == g2;
  int min_sep = intra ? 0 : min_block_sep(s, g1, g2, nc1, nc2);
  if (lane < MAX_CONN * MAX_CONN) {
    int c1 = lane / MAX_CONN, c2 = lane % MAX_CONN;
    L.cs[lane] = (c1 < nc1 && c2 < nc2 && !intra) ? (uint8_t)conn_sep(s, g1, g2, c1, c2) : 6;
  }
  {
    int half = lane >= TILE;
    int a = lane - half * TILE;
    int tt = half ? t2 : t1;
    for (int c = 0; c < MAX_CONN; ++c)
      L.pd[half][c][a] = conn_pd[(int64_t(tt) * MAX_CONN + c) * bt.Amax + min(a, bt.Amax - 1)];
  }
  wave_sync();
  for (int k = lane; k < TILE * TILE; k += WAVE) {
    int a1 = k / TILE, a2 = k % TILE;
    uint8_t v = 255;
    if (a1 < n1 && a2 < n2) {
      if (intra)
        v = (uint8_t)bt.path_distance[(int64_t(t1) * bt.Amax + a1) * bt.Amax + a2];
      else
        v = (uint8_t)inter_pair_sep(L, min_sep, nc1, nc2, a1, a2, lig);
    }
    out[it * TILE * TILE + k] = v;
  }
}

// ------------------------------------------------------------------ host side

static int check(hipError_t e) { return e == hipSuccess ? TMOL_LJLK_OK : TMOL_LJLK_EHIP; }

static int validate(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt, int dtype) {
  if (!s || !bt) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (bt->dtype != dtype) return TMOL_LJLK_EINVAL;
  if (bt->max_n_conn > MAX_CONN || bt->max_n_conn < 0) return TMOL_LJLK_ENOTSUP;
  if (bt->n_block_types <= 0 || bt->max_n_atoms <= 0) return TMOL_LJLK_EINVAL;
  if (bt->max_n_tiles != (bt->max_n_atoms - 1) / TILE + 1) return TMOL_LJLK_EINVAL;
  if (s->n_poses <= 0 || s->n_blocks_total <= 0) return TMOL_LJLK_EINVAL;
  if (s->layout == TMOL_LJLK_PADDED) {
    if (int64_t(s->n_poses) * s->max_n_blocks != s->n_blocks_total) return TMOL_LJLK_EINVAL;
    if (!s->min_bondsep || !s->inter_block_bondsep) return TMOL_LJLK_EINVAL;
  } else if (s->layout == TMOL_LJLK_RAGGED) {
    if (!s->block_pose || !s->block_conn_pos || !s->block_segment) return TMOL_LJLK_EINVAL;
  } else {
    return TMOL_LJLK_EINVAL;
  }
  if (!s->block_atom_offset || !s->block_type) return TMOL_LJLK_EINVAL;
  if (s->n_atoms_total >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  return TMOL_LJLK_OK;
}

template <typename Real>
static int run_neighbors(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                         const void* coords, double max_dis, const int32_t* tiles,
                         int64_t n_tiles, void* ws, int64_t capacity, int32_t* dense,
                         hipStream_t stream, const WsLayout& Lw) {
  StackView s = make_stack_view(st, btp->max_n_conn);
  BTView bt = make_bt_view(btp);
  char* w = (char*)ws;
  uint32_t* counters = (uint32_t*)(w + Lw.counters);
  Real* spheres = (Real*)(w + Lw.spheres);
  Real* records = (Real*)(w + Lw.records);
  uint8_t* conn_pd = (uint8_t*)(w + Lw.conn_pd);
  int2* items = (int2*)(w + Lw.items);
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, stream, counters, 16,
                     (uint32_t)0);
  int nrec = bt.T * bt.Amax;
  hipLaunchKernelGGL(prepare_records_kernel<Real>, dim3((nrec + 255) / 256), dim3(256), 0,
                     stream, bt, (const Real*)btp->type_params, btp->n_atom_types, records,
                     conn_pd);
  hipLaunchKernelGGL(spheres_kernel<Real>, dim3((s.G + 255) / 256), dim3(256), 0, stream, s,
                     bt, (const Real*)coords, spheres);
  if (dense) {
    int64_t n = int64_t(s.P) * s.B * s.B;
    hipLaunchKernelGGL(fill_kernel<int32_t>, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)),
                       dim3(256), 0, stream, dense, n, 0);
  }
  if (n_tiles > 0)
    hipLaunchKernelGGL(neighbors_kernel<Real>, dim3((unsigned)n_tiles), dim3(256), 0, stream, s,
                       (const int4*)tiles, n_tiles, spheres, (Real)max_dis, items, counters,
                       capacity, dense);
  return check(hipGetLastError());
}

template <typename Real, int MODE>
static void launch_score(const ScoreArgs<Real>& a, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL((score_kernel<Real, MODE>), dim3(n_wg), dim3(256), 0, stream, a);
}

template <typename Real>
static int run_score(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                     const void* coords, void* ws, int64_t capacity, int flags,
                     const void* weights, void* energies, void* dvdx, void* grad,
                     int n_wave_groups, hipStream_t stream, const WsLayout& Lw) {
  ScoreArgs<Real> a;
  a.s = make_stack_view(st, btp->max_n_conn);
  a.bt = make_bt_view(btp);
  char* w = (char*)ws;
  a.coords = (const Real*)coords;
  a.records = (const Real*)(w + Lw.records);
  a.conn_pd = (const uint8_t*)(w + Lw.conn_pd);
  a.global_params = (const Real*)btp->global_params;
  a.items = (const int2*)(w + Lw.items);
  a.counters = (const uint32_t*)(w + Lw.counters);
  a.capacity = capacity;
  a.weights = (const Real*)weights;
  a.energies64 = (double*)(w + Lw.energies);
  a.energies_bp = (Real*)energies;
  a.dvdx = (Real*)dvdx;
  a.grad = (Real*)grad;
  a.n_atoms = st->n_atoms_total;
  const bool bp = flags & TMOL_LJLK_BLOCK_PAIR;
  const bool noe = flags & TMOL_LJLK_NO_ENERGY;
  const bool acc = flags & TMOL_LJLK_ACCUMULATE;
  const int P = st->n_poses;
  const int64_t NA3 = st->n_atoms_total * 3;
  auto fill = [&](Real* p, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<Real>, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)),
                       dim3(256), 0, stream, p, n, Real(0));
  };
  if (!noe) {
    if (bp) {
      fill((Real*)energies, 3 * int64_t(P) * st->max_n_blocks * st->max_n_blocks);
    } else {
      hipLaunchKernelGGL(fill_kernel<double>, dim3((3 * P + 255) / 256), dim3(256), 0, stream,
                         a.energies64, int64_t(3) * P, 0.0);
    }
  }
  if ((flags & TMOL_LJLK_DVDX) && !acc) fill((Real*)dvdx, 3 * NA3);
  if ((flags & TMOL_LJLK_WGRAD) && !acc) fill((Real*)grad, NA3);
  int n_wg = n_wave_groups > 0 ? n_wave_groups : 2048;
  int mode = flags & (TMOL_LJLK_DVDX | TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR | TMOL_LJLK_NO_ENERGY);
  switch (mode) {
    case 0: launch_score<Real, 0>(a, n_wg, stream); break;
    case M_DVDX: launch_score<Real, M_DVDX>(a, n_wg, stream); break;
    case M_WGRAD: launch_score<Real, M_WGRAD>(a, n_wg, stream); break;
    case M_WGRAD | M_NOE: launch_score<Real, M_WGRAD | M_NOE>(a, n_wg, stream); break;
    case M_BP: launch_score<Real, M_BP>(a, n_wg, stream); break;
    case M_BP | M_WGRAD | M_NOE: launch_score<Real, M_BP | M_WGRAD | M_NOE>(a, n_wg, stream); break;
    case M_BP | M_WGRAD: launch_score<Real, M_BP | M_WGRAD>(a, n_wg, stream); break;
    default: return TMOL_LJLK_ENOTSUP;
  }
  if (!noe && !bp)
    hipLaunchKernelGGL(finalize_kernel<Real>, dim3((3 * P + 255) / 256), dim3(256), 0, stream,
                       (const double*)a.energies64, (Real*)energies, 3 * P);
  return check(hipGetLastError());
}

}  // namespace tmol_ljlk

using namespace tmol_ljlk;

extern "C" {

int64_t tmol_ljlk_count_neighbor_tiles(const int32_t* pose_block_start_host, int32_t n_poses) {
  int64_t n = 0;
  for (int p = 0; p < n_poses; ++p) {
    int64_t nb = pose_block_start_host[p + 1] - pose_block_start_host[p];
    int64_t t = (nb + NBR_TILE - 1) / NBR_TILE;
    n += t * (t + 1) / 2;
  }
  return n;
}

int tmol_ljlk_fill_neighbor_tiles(const int32_t* pbs, int32_t n_poses, int32_t* tiles) {
  int64_t k = 0;
  for (int p = 0; p < n_poses; ++p) {
    int start = pbs[p], end = pbs[p + 1];
    int t = (end - start + NBR_TILE - 1) / NBR_TILE;
    for (int i = 0; i < t; ++i)
      for (int j = i; j < t; ++j) {
        tiles[4 * k + 0] = start + i * NBR_TILE;
        tiles[4 * k + 1] = start + j * NBR_TILE;
        tiles[4 * k + 2] = end;
        tiles[4 * k + 3] = p;
        ++k;
      }
  }
  return TMOL_LJLK_OK;
}

int64_t tmol_ljlk_workspace_bytes(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                                  int64_t item_capacity, int32_t dtype) {
  if (!s || !bt || item_capacity < 0) return -1;
  return ws_layout(s, bt, item_capacity, dtype).total;
}

int tmol_ljlk_work_list(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                        void* workspace, int64_t capacity, int32_t dtype, int32_t** items_out) {
  if (!s || !bt || !items_out) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, capacity, dtype);
  *items_out = (int32_t*)((char*)workspace + L.items);
  return TMOL_LJLK_OK;
}

int tmol_ljlk_neighbors(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                        const void* coords, int32_t dtype, double max_dis, const int32_t* tiles,
                        int64_t n_tiles, void* workspace, int64_t ws_bytes,
                        int64_t capacity, int32_t* dense, void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!coords || !workspace || (n_tiles > 0 && !tiles) || capacity < 0) return TMOL_LJLK_EINVAL;
  if (n_tiles >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  if (capacity >= (int64_t(1) << 32)) return TMOL_LJLK_ENOTSUP;
  if (dense && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_neighbors<float>(s, bt, coords, max_dis, tiles, n_tiles, workspace, capacity,
                                dense, st, L);
  return run_neighbors<double>(s, bt, coords, max_dis, tiles, n_tiles, workspace, capacity,
                               dense, st, L);
}

int tmol_ljlk_score(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                    const void* coords, int32_t dtype, void* workspace, int64_t ws_bytes,
                    int64_t capacity, int32_t flags, const void* weights, void* energies,
                    void* dvdx, void* grad, int32_t n_wave_groups, void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!coords || !workspace) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  if (!(flags & TMOL_LJLK_NO_ENERGY) && !energies) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_DVDX) && !dvdx) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_WGRAD) && (!grad || !weights)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_DVDX) && (flags & TMOL_LJLK_WGRAD)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_BLOCK_PAIR) && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_ENOTSUP;
  if ((flags & TMOL_LJLK_BLOCK_PAIR) && (flags & TMOL_LJLK_DVDX)) return TMOL_LJLK_ENOTSUP;
  if (n_wave_groups < 0 || n_wave_groups > (1 << 20)) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_score<float>(s, bt, coords, workspace, capacity, flags, weights, energies, dvdx,
                            grad, n_wave_groups, st, L);
  return run_score<double>(s, bt, coords, workspace, capacity, flags, weights, energies, dvdx,
                           grad, n_wave_groups, st, L);
}

int tmol_ljlk_census(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                     const void* coords, int32_t dtype, void* workspace, int64_t ws_bytes,
                     int64_t capacity, uint64_t* census, void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!coords || !workspace || !census) return TMOL_LJLK_EINVAL;
  WsLayout Lw = ws_layout(s, bt, capacity, dtype);
  if (ws_bytes < Lw.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, st,
                     (unsigned long long*)census, 16, 0ull);
  char* w = (char*)workspace;
  if (dtype == TMOL_LJLK_F32) {
    ScoreArgs<float> a{};
    a.s = make_stack_view(s, bt->max_n_conn);
    a.bt = make_bt_view(bt);
    a.coords = (const float*)coords;
    a.records = (const float*)(w + Lw.records);
    a.conn_pd = (const uint8_t*)(w + Lw.conn_pd);
    a.global_params = (const float*)bt->global_params;
    a.items = (const int2*)(w + Lw.items);
    a.counters = (const uint32_t*)(w + Lw.counters);
    a.capacity = capacity;
    a.n_atoms = s->n_atoms_total;
    a.census = (unsigned long long*)census;
    launch_score<float, M_CENSUS>(a, 2048, st);
  } else {
    return TMOL_LJLK_ENOTSUP;
  }
  return check(hipGetLastError());
}

int tmol_ljlk_backward_combine(const void* dvdx, const void* grad_scores,
                               const int32_t* atom_pose, int64_t n_atoms, int32_t n_poses,
                               int32_t dtype, void* grad, void* stream) {
  if (!dvdx || !grad_scores || !atom_pose || !grad || n_atoms < 0 || n_poses <= 0)
    return TMOL_LJLK_EINVAL;
  int64_t n = n_atoms * 3;
  if (n == 0) return TMOL_LJLK_OK;
  unsigned nb = (unsigned)((n + 255) / 256);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    hipLaunchKernelGGL(combine_kernel<float>, dim3(nb), dim3(256), 0, st, (const float*)dvdx,
                       (const float*)grad_scores, atom_pose, n_atoms, n_poses, (float*)grad);
  else if (dtype == TMOL_LJLK_F64)
    hipLaunchKernelGGL(combine_kernel<double>, dim3(nb), dim3(256), 0, st, (const double*)dvdx,
                       (const double*)grad_scores, atom_pose, n_atoms, n_poses, (double*)grad);
  else
    return TMOL_LJLK_EINVAL;
  return check(hipGetLastError());
}

int tmol_ljlk_read_counters(const void* workspace, int64_t* out, void* stream) {
  if (!workspace || !out) return TMOL_LJLK_EINVAL;
  uint32_t c[2];
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(c, workspace, sizeof(c), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return TMOL_LJLK_EHIP;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return TMOL_LJLK_EHIP;
  out[0] = c[0];
  out[1] = c[1];
  return TMOL_LJLK_OK;
}

int tmol_ljlk_debug_pair_separations(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                                     const void* workspace, int64_t capacity, int32_t dtype,
                                     const int32_t* items, int64_t n_items, uint8_t* out,
                                     void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!workspace || !items || !out || n_items < 0) return TMOL_LJLK_EINVAL;
  if (n_items == 0) return TMOL_LJLK_OK;
  if (n_items >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  WsLayout L = ws_layout(s, bt, capacity, dtype);
  const uint8_t* conn_pd = (const uint8_t*)((const char*)workspace + L.conn_pd);
  hipLaunchKernelGGL(debug_sep_kernel<float>, dim3((unsigned)n_items), dim3(64), 0,
                     (hipStream_t)stream, make_stack_view(s, bt->max_n_conn), make_bt_view(bt),
                     conn_pd, items, n_items, out);
  return check(hipGetLastError());
}

int tmol_ljlk_device_cu_count(int32_t* cu) {
  if (!cu) return TMOL_LJLK_EINVAL;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TMOL_LJLK_EHIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return TMOL_LJLK_EHIP;
  *cu = prop.multiProcessorCount;
  return TMOL_LJLK_OK;
}

}  // extern "C"
