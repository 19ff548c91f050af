// Neighbor pass and small kernels: per-block-type atom records, bounding
// spheres, block-pair culling into the row-major work list, fills,
// finalization, whole-pose backward, count-pair debug dump.
#pragma once

#include "ljlk_common.hh"

namespace tmol_ljlk {

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
    // heavy atom: listed in its tile's heavy-atom list
    int tile = a / TILE, ta = a - tile * TILE;
    int nh = bt.n_heavy_in_tile[t * bt.nt + tile];
    for (int k = 0; k < nh; ++k)
      if (bt.heavy_in_tile[int64_t(t) * TILE * bt.nt + tile * TILE + k] == ta) flags |= F_HEAVY;
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
    conn_pd[(int64_t(t) * bt.Amax + a) * MAX_CONN + c] = v;
  }
}

// Intra-block pairs of every block type with at most 32 atoms that carry
// count-pair weight (path distance >= 4; the others score exactly zero),
// a1 < a2 in the order (a1, a2), one wave per type (lanes over a2, ballot
// compaction keeps the order):
//   pairs[t * INTRA_MAX + k] = a1 | a2 << 5 | (path distance == 4) << 10
__global__ __launch_bounds__(64) void prepare_intra_kernel(BTView bt, int32_t* __restrict__ count,
                                                           uint16_t* __restrict__ pairs) {
  const int t = blockIdx.x;
  if (t >= bt.T) return;
  const int lane = threadIdx.x;
  const int n = bt.n_atoms[t];
  int k = 0;
  if (n <= TILE)
    for (int a1 = 0; a1 < n; ++a1) {
      const int a2 = a1 + 1 + lane;
      int sep = 0;
      if (a2 < n) sep = bt.path_distance[(int64_t(t) * bt.Amax + a1) * bt.Amax + a2];
      const bool keep = a2 < n && sep >= 4;
      const uint64_t m = __ballot(keep);
      if (keep)
        pairs[int64_t(t) * INTRA_MAX + k + lanes_below(m)] =
            (uint16_t)(a1 | (a2 << 5) | ((sep == 4) << 10));
      k += __popcll(m);
    }
  if (lane == 0) count[t] = k;
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

// One workgroup per 64-block chunk of a pose: the diagonal tiles of the tile
// list do the work, the other tiles return at once.  The chunk's rows are
// tested against every column chunk from itself to the end of the pose, in
// column order.  A column chunk is skipped when the bounding boxes of the
// two chunks' spheres (centres grown by radii) are farther apart than reach
// + 1e-3 A: then every sphere pair is, and no test can pass.  Wave w owns
// rows 16w..16w+15 and appends to them in column order with counters in
// registers, so the lists need no atomics and come out sorted by g2.
// The pair test is the reference's (sphere_overlap.impl.hh:287-351):
// d^2 < ((r1 + r2) + reach)^2 with its association, no contraction.
template <typename Real>
__device__ __forceinline__ void chunk_box(const Real (*sph)[4], const int* typ, int lane,
                                          Real* lo, Real* hi) {
  const bool ok = typ[lane] >= 0;
  const Real big = Real(1e30);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Real l = ok ? sph[lane][k] - sph[lane][3] : big;
    Real h = ok ? sph[lane][k] + sph[lane][3] : -big;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      l = fmin(l, __shfl_xor(l, o, WAVE));
      h = fmax(h, __shfl_xor(h, o, WAVE));
    }
    lo[k] = l;
    hi[k] = h;
  }
}

// FULL: every neighbor g2 of the pose goes to row g1 (both triangles; the
// rows of a gather kernel that owns its block's atoms), columns from the
// pose's first block; pose_start (ragged) gives it.  dense still receives
// the upper triangle only.
template <typename Real, bool FULL = false>
__global__ __launch_bounds__(256) void neighbor_rows_kernel(
    StackView s, const int4* __restrict__ tiles, int64_t n_tiles,
    const Real* __restrict__ spheres, Real reach, uint32_t* __restrict__ row_count,
    int32_t* __restrict__ row_items, int64_t row_cap, uint32_t* __restrict__ counters,
    int32_t* __restrict__ dense, const int32_t* __restrict__ pose_start = nullptr) {
#pragma clang fp contract(off)
  __shared__ Real sph_r[NBR_TILE][4];
  __shared__ Real sph_c[NBR_TILE][4];
  __shared__ int typ_r[NBR_TILE];
  __shared__ int typ_c[NBR_TILE];
  const int64_t tile_id = blockIdx.x;
  if (tile_id >= n_tiles) return;
  const int4 tl = tiles[tile_id];
  if (tl.x != tl.y) return;
  const int row0 = tl.x, end = tl.z;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid / WAVE;
  if (tid < NBR_TILE) {
    const int g = row0 + tid;
    const bool ok = g < end;
    for (int k = 0; k < 4; ++k) sph_r[tid][k] = ok ? spheres[int64_t(g) * 4 + k] : Real(0);
    typ_r[tid] = ok ? s.block_type[g] : -1;
  }
  __syncthreads();
  Real rlo[3], rhi[3];
  chunk_box(sph_r, typ_r, lane, rlo, rhi);
  const Real pad = reach + Real(1e-3);
  uint32_t cnt[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) cnt[r] = 0;
  bool overflow = false;
  const int col_begin =
      !FULL ? row0 : (s.layout == TMOL_LJLK_PADDED ? tl.w * s.B : pose_start[tl.w]);
  for (int col0 = col_begin; col0 < end; col0 += NBR_TILE) {
    __syncthreads();  // the previous column chunk is no longer read
    if (tid < NBR_TILE) {
      const int g = col0 + tid;
      const bool ok = g < end;
      for (int k = 0; k < 4; ++k) sph_c[tid][k] = ok ? spheres[int64_t(g) * 4 + k] : Real(0);
      typ_c[tid] = ok ? s.block_type[g] : -1;
    }
    __syncthreads();
    Real clo[3], chi[3];
    chunk_box(sph_c, typ_c, lane, clo, chi);
    Real gap2 = Real(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const Real gk = fmax(Real(0), fmax(clo[k] - rhi[k], rlo[k] - chi[k]));
      gap2 += gk * gk;
    }
    if (gap2 > pad * pad) continue;  // uniform across the workgroup
    const int g2 = col0 + lane;
    const Real c0 = sph_c[lane][0], c1 = sph_c[lane][1], c2 = sph_c[lane][2], c3 = sph_c[lane][3];
    const bool col_ok = typ_c[lane] >= 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int lr = wave * 16 + r;
      const int g1 = row0 + lr;
      const bool ok = col_ok && typ_r[lr] >= 0 && (FULL || g1 <= g2);
      const Real dx = sph_r[lr][0] - c0;
      const Real dy = sph_r[lr][1] - c1;
      const Real dz = sph_r[lr][2] - c2;
      const Real dxx = dx * dx;
      const Real dyy = dy * dy;
      const Real dzz = dz * dz;
      const Real d2 = (dxx + dyy) + dzz;
      const Real thr = (sph_r[lr][3] + c3) + reach;
      const Real thr2 = thr * thr;
      const bool hit = ok && (d2 < thr2);
      const uint64_t m = __ballot(hit);
      if (m == 0) continue;
      if (hit) {
        const int64_t idx = int64_t(cnt[r]) + lanes_below(m);
        if (idx < row_cap)
          row_items[int64_t(g1) * row_cap + idx] = g2;
        else
          overflow = true;
        if (dense && g1 <= g2) {
          const int b2 = g2 - (g2 / s.B) * s.B;
          dense[int64_t(g1) * s.B + b2] = 1;
        }
      }
      cnt[r] += __popcll(m);
    }
  }
  uint32_t total = 0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int g1 = row0 + wave * 16 + r;
    if (g1 < end && lane == 0) row_count[g1] = cnt[r];
    total += cnt[r];
  }
  if (__ballot(overflow) != 0 && lane == 0) atomicOr(&counters[1], 1u);
  if (total > 0 && lane == 0) atomicAdd(&counters[0], total);
}

// ------------------------------------------------------------------ small kernels

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, int64_t n, T v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename Real>
__global__ void finalize_kernel(const double* __restrict__ e64, Real* __restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = Real(e64[i]);
}

// DETERMINISTIC: fixed-point per-pose energies -> output dtype
template <typename Real>
__global__ void finalize_fixed_kernel(const unsigned long long* __restrict__ e, Real* __restrict__ out,
                                      int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = Real(double((long long)e[i]) * (1.0 / DET_E_SCALE));
}

// DETERMINISTIC: fixed-point gradient planes -> output (added to it when
// accumulate), one rounding per component
template <typename Real>
__global__ void fixed_to_real_kernel(const unsigned long long* __restrict__ fx,
                                     Real* __restrict__ out, int64_t n, bool accumulate) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) {
    const Real v = LdsAcc<Real>::dec(fx[i]);
    out[i] = accumulate ? out[i] + v : v;
  }
}

// whole-pose backward: grad[a, k] = sum_t gs[t, pose(a)] * dvdx[t, a, k]
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

// Count-pair separations of listed block pairs (first tiles), computed with
// the scoring kernels' count-pair function (one wave per item).
__global__ __launch_bounds__(64) void debug_sep_kernel(StackView s, BTView bt,
                                                       const uint8_t* __restrict__ conn_pd,
                                                       const int32_t* __restrict__ items,
                                                       int64_t n_items, uint8_t* __restrict__ out) {
  __shared__ uint8_t pd[2][MAX_CONN][TILE];
  __shared__ uint8_t cs[MAX_CONN * MAX_CONN];
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
    cs[lane] = (c1 < nc1 && c2 < nc2 && !intra) ? (uint8_t)conn_sep(s, g1, g2, c1, c2) : 6;
  }
  {
    int half = lane >= TILE;
    int a = lane - half * TILE;
    int tt = half ? t2 : t1;
    for (int c = 0; c < MAX_CONN; ++c)
      pd[half][c][a] = conn_pd[(int64_t(tt) * bt.Amax + min(a, bt.Amax - 1)) * MAX_CONN + c];
  }
  __syncthreads();
  for (int k = lane; k < TILE * TILE; k += WAVE) {
    int a1 = k / TILE, a2 = k % TILE;
    uint8_t v = 255;
    if (a1 < n1 && a2 < n2) {
      if (intra)
        v = (uint8_t)bt.path_distance[(int64_t(t1) * bt.Amax + a1) * bt.Amax + a2];
      else
        v = (uint8_t)inter_pair_sep(pd[0], pd[1], cs, min_sep, nc1, nc2, a1, a2, lig);
    }
    out[it * TILE * TILE + k] = v;
  }
}

}  // namespace tmol_ljlk
