// Missing leaf atoms of a pose stack (hydrogens, carbonyl oxygens): each is
// built from three resolved ancestors and its ideal internal coordinates,
// with a backup ancestor set where the first is absent (gen_pose_leaf_atoms).
//
// Reference: tmol/io/details/compiled/gen_pose_leaf_atoms.impl.hh:25-336,
// leaf_atoms.hh:11-100; tmol/score/common/uaid_util.hh:110-205 (atoms named
// across inter-block connections), gen_coord.hh (build_coordinate).
#pragma once

#include "ljlk_common.hh"
#include "lkball.hh"

namespace tmol_ljlk {

// padded pose layout of the op boundary
struct LeafArgs {
  int P, NA, B, Amax, C;
  const uint8_t* block_missing;   // [P, B, Amax] atom not given
  const uint8_t* pose_missing;    // [P, NA]
  const int32_t* block_offset;    // [P, B]
  const int32_t* block_type;      // [P, B]
  const int32_t* irc;             // [P, B, C, 2] inter-block connections
  const int32_t* n_atoms;         // [T]
  const int32_t* downstream;      // [T, C, Amax] atoms downstream of each connection
  const int32_t* anc;             // [T, Amax, 3, 3] (atom, conn, bonds from conn)
  const int32_t* anc_backup;
  const void* geom;               // [T, Amax, 3] phi, theta, d
  const void* geom_backup;
};

// pose-wide index of an unresolved atom id of block b (-1: none)
__device__ __forceinline__ int leaf_resolve(const LeafArgs& a, int p, int b, const int32_t* u) {
  if (u[0] != -1) return a.block_offset[p * a.B + b] + u[0];
  if (u[1] == -1) return -1;
  const int32_t* c = a.irc + ((int64_t(p) * a.B + b) * a.C + u[1]) * 2;
  const int ob = c[0];
  if (ob == -1) return -1;
  const int obt = a.block_type[p * a.B + ob];
  const int idx = a.downstream[(int64_t(obt) * a.C + c[1]) * a.Amax + u[2]];
  if (idx < 0) return -1;
  return a.block_offset[p * a.B + ob] + idx;
}

// the three ancestors and geometry (phi, theta, d) of atom slot `at` of a
// block of type bt: the primary set, else the backup, else none (false)
template <typename Real>
__device__ __forceinline__ bool leaf_ancestors(const LeafArgs& a, int p, int b, int bt, int at,
                                               int* anc, Real* geom) {
  const int64_t row = int64_t(bt) * a.Amax + at;
  const int32_t* sets[2] = {a.anc + row * 9, a.anc_backup + row * 9};
  const Real* geoms[2] = {(const Real*)a.geom + row * 3, (const Real*)a.geom_backup + row * 3};
  for (int k = 0; k < 2; ++k) {
    bool ok = true;
    for (int j = 0; j < 3; ++j) {
      anc[j] = leaf_resolve(a, p, b, sets[k] + 3 * j);
      ok = ok && anc[j] != -1 && !a.pose_missing[int64_t(p) * a.NA + anc[j]];
    }
    if (ok) {
      geom[0] = geoms[k][0];
      geom[1] = geoms[k][1];
      geom[2] = geoms[k][2];
      return true;
    }
  }
  return false;
}

template <typename Real>
__device__ __forceinline__ bool any_nan(Vec3<Real> v) {
  return v.x != v.x || v.y != v.y || v.z != v.z;
}

// Launch shape of all three kernels: one wave per workgroup (its barriers
// are wave barriers, reached by the whole workgroup whatever the wave's
// trip count).  Wave w owns the LEAF_BPW blocks q = LEAF_BPW w ... of the flat
// (pose, block) grid: it loads their types, keeps the real ones (a ballot)
// and takes them G = 64 / Amax at a time (Amax <= 64: lane = g * Amax + atom
// slot), or one at a time with lanes striding over the atom slots (Amax >
// 64).  Padding blocks cost one load per 64.
constexpr int LEAF_WAVE = 64;
// blocks whose types one wave loads: fewer than 64 keeps one or two passes
// per wave, so the dependent loads of many waves overlap
constexpr int LEAF_BPW = 16;
constexpr int LEAF_MAX_ATOMS = 2048;  // larger block types: ENOTSUP

struct LeafPass {
  int g;       // this lane's block of the pass (-1: none)
  int q;       // its flat (pose, block) index
  int bt;      // its block type
  int at0;     // first atom slot of this lane
  int stride;  // atom-slot stride (past Amax when one slot per lane)
};

__device__ __forceinline__ int leaf_blocks_per_pass(const LeafArgs& a) {
  return a.Amax <= LEAF_WAVE ? LEAF_WAVE / a.Amax : 1;
}

// this lane's block in the pass that takes the lowest blocks of `real`
__device__ __forceinline__ LeafPass leaf_pass(const LeafArgs& a, uint64_t real, int64_t q0,
                                              int bt_lane, int lane) {
  const int G = leaf_blocks_per_pass(a);
  LeafPass r;
  const bool small = a.Amax <= LEAF_WAVE;
  const int g = small ? lane / a.Amax : 0;
  r.at0 = small ? lane - g * a.Amax : lane;
  r.stride = small ? (1 << 30) : LEAF_WAVE;  // small: one slot per lane
  uint64_t m = real;
  for (int k = 0; k < g && m; ++k) m &= m - 1;
  const int pos = (g < G && m) ? __ffsll((unsigned long long)m) - 1 : -1;
  const int bt = __shfl(bt_lane, pos < 0 ? 0 : pos, LEAF_WAVE);
  r.g = pos < 0 ? -1 : g;
  r.q = pos < 0 ? -1 : int(q0 + pos);
  r.bt = pos < 0 ? -1 : bt;
  return r;
}

__device__ __forceinline__ uint64_t leaf_drop(const LeafArgs& a, uint64_t real) {
  const int G = leaf_blocks_per_pass(a);
  for (int k = 0; k < G && real; ++k) real &= real - 1;
  return real;
}

// given atoms are copied, missing ones built (NaN when no complete ancestor
// set exists)
template <typename Real>
__global__ __launch_bounds__(LEAF_WAVE) void leaf_atoms_kernel(LeafArgs a, const Real* orig,
                                                               Real* out) {
  const int lane = threadIdx.x;
  const int64_t nblk = int64_t(a.P) * a.B;
  for (int64_t w = blockIdx.x; w * LEAF_BPW < nblk; w += gridDim.x) {
    const int64_t q0 = w * LEAF_BPW;
    const int bt_lane = lane < LEAF_BPW && q0 + lane < nblk ? a.block_type[q0 + lane] : -1;
    uint64_t real = __ballot(bt_lane >= 0);
    while (real) {
      const LeafPass r = leaf_pass(a, real, q0, bt_lane, lane);
      real = leaf_drop(a, real);
      if (r.g < 0) continue;
      const int p = r.q / a.B, b = r.q - p * a.B;
      const int n_at = a.n_atoms[r.bt];
      const int64_t off = int64_t(p) * a.NA + a.block_offset[r.q];
      for (int at = r.at0; at < n_at; at += r.stride) {
        const int64_t dst = off + at;
        Real* o = out + 3 * dst;
        if (!a.block_missing[int64_t(r.q) * a.Amax + at]) {
          o[0] = orig[3 * dst];
          o[1] = orig[3 * dst + 1];
          o[2] = orig[3 * dst + 2];
          continue;
        }
        int anc[3];
        Real g[3];
        Vec3<Real> W = {Real(NAN), Real(NAN), Real(NAN)};
        if (leaf_ancestors(a, p, b, r.bt, at, anc, g)) {
          const Real* base = orig + 3 * int64_t(p) * a.NA;
          const Vec3<Real> A = v3(base + 3 * anc[0]), B = v3(base + 3 * anc[1]),
                           C = v3(base + 3 * anc[2]);
          if (!any_nan(A) && !any_nan(B) && !any_nan(C)) {
            Vec3<Real> e0, e1, e2;
            water_frame(A, B, C, e0, e1, e2);
            const WaterCoef<Real> c = water_coef(g[2], g[1], g[0]);
            W = {A.x + c.k0 * e0.x + c.k1 * e1.x + c.k2 * e2.x,
                 A.y + c.k0 * e0.y + c.k1 * e1.y + c.k2 * e2.y,
                 A.z + c.k0 * e0.z + c.k1 * e1.z + c.k2 * e2.z};
          }
        }
        o[0] = W.x;
        o[1] = W.y;
        o[2] = W.z;
      }
    }
  }
}

// the ancestors of a built atom (pose-local indices) with, for each, its
// atom slot in the same block or -1 when named across a connection; false
// when no complete ancestor set exists
template <typename Real>
__device__ __forceinline__ bool leaf_ancestor_slots(const LeafArgs& a, int p, int b, int bt,
                                                    int at, int* anc, int* slot, Real* g) {
  const int64_t row = int64_t(bt) * a.Amax + at;
  const int32_t* sets[2] = {a.anc + row * 9, a.anc_backup + row * 9};
  const Real* geoms[2] = {(const Real*)a.geom + row * 3, (const Real*)a.geom_backup + row * 3};
  for (int k = 0; k < 2; ++k) {
    bool ok = true;
    for (int j = 0; j < 3; ++j) {
      const int32_t* u = sets[k] + 3 * j;
      anc[j] = leaf_resolve(a, p, b, u);
      slot[j] = u[0];
      ok = ok && anc[j] != -1 && !a.pose_missing[int64_t(p) * a.NA + anc[j]];
    }
    if (ok) {
      g[0] = geoms[k][0];
      g[1] = geoms[k][1];
      g[2] = geoms[k][2];
      return true;
    }
  }
  return false;
}

// the ancestor shares of the atom at slot `at` of block q if it is built
// (false otherwise, or without ancestors or with an absent coordinate; with
// REMOTE_ONLY also when every ancestor is in the same block)
template <typename Real, bool REMOTE_ONLY>
__device__ __forceinline__ bool leaf_shares(const LeafArgs& a, const Real* orig,
                                            const Real* dnew, int q, int bt, int at, int* anc,
                                            int* slot, Vec3<Real>* gs) {
  if (!a.block_missing[int64_t(q) * a.Amax + at]) return false;
  const int p = q / a.B, b = q - p * a.B;
  Real g[3];
  if (!leaf_ancestor_slots(a, p, b, bt, at, anc, slot, g)) return false;
  if (REMOTE_ONLY && slot[0] >= 0 && slot[1] >= 0 && slot[2] >= 0) return false;
  const int64_t base = int64_t(p) * a.NA;
  const Vec3<Real> A = v3(orig + 3 * (base + anc[0])), B = v3(orig + 3 * (base + anc[1])),
                   C = v3(orig + 3 * (base + anc[2]));
  if (any_nan(A) || any_nan(B) || any_nan(C)) return false;
  const int64_t dst = base + a.block_offset[q] + at;
  build_coord_vjp(A, B, C, water_coef(g[2], g[1], g[0]), v3(dnew + 3 * dst), gs[0], gs[1],
                  gs[2]);
  return true;
}

// Backward, first kernel: a pass's blocks sum the shares their built atoms
// send to ancestors in the same block in LDS (dynamic, 3 x max(64, Amax)
// per wave), then write each atom once: its own gradient if given, plus the
// sum.  Shares sent across a connection are added by the second kernel,
// with global atomics, after the first has written every atom.
template <typename Real>
__global__ __launch_bounds__(LEAF_WAVE) void leaf_atoms_backward_local_kernel(
    LeafArgs a, const Real* orig, const Real* dnew, Real* dorig) {
  extern __shared__ unsigned char leaf_lds[];
  Real* acc = (Real*)leaf_lds;
  const int lane = threadIdx.x;
  const int W = a.Amax <= LEAF_WAVE ? LEAF_WAVE : a.Amax;  // entries per component
  const int64_t nblk = int64_t(a.P) * a.B;
  for (int64_t w = blockIdx.x; w * LEAF_BPW < nblk; w += gridDim.x) {
    const int64_t q0 = w * LEAF_BPW;
    const int bt_lane = lane < LEAF_BPW && q0 + lane < nblk ? a.block_type[q0 + lane] : -1;
    uint64_t real = __ballot(bt_lane >= 0);
    while (real) {
      const LeafPass r = leaf_pass(a, real, q0, bt_lane, lane);
      real = leaf_drop(a, real);
      for (int i = lane; i < 3 * W; i += LEAF_WAVE) acc[i] = Real(0);
      __syncthreads();
      const int n_at = r.g < 0 ? 0 : a.n_atoms[r.bt];
      const int gbase = r.g < 0 ? 0 : r.g * a.Amax;
      for (int at = r.at0; at < n_at; at += r.stride) {
        int anc[3], slot[3];
        Vec3<Real> gs[3];
        if (!leaf_shares<Real, false>(a, orig, dnew, r.q, r.bt, at, anc, slot, gs)) continue;
        for (int j = 0; j < 3; ++j) {
          if (slot[j] < 0) continue;
          const int e = gbase + slot[j];
          atomicAdd(&acc[e], gs[j].x);
          atomicAdd(&acc[W + e], gs[j].y);
          atomicAdd(&acc[2 * W + e], gs[j].z);
        }
      }
      __syncthreads();
      if (r.g >= 0) {
        const int p = r.q / a.B;
        const int64_t off = int64_t(p) * a.NA + a.block_offset[r.q];
        for (int at = r.at0; at < n_at; at += r.stride) {
          const int64_t dst = off + at;
          const int e = gbase + at;
          const bool given = !a.block_missing[int64_t(r.q) * a.Amax + at];
          dorig[3 * dst] = (given ? dnew[3 * dst] : Real(0)) + acc[e];
          dorig[3 * dst + 1] = (given ? dnew[3 * dst + 1] : Real(0)) + acc[W + e];
          dorig[3 * dst + 2] = (given ? dnew[3 * dst + 2] : Real(0)) + acc[2 * W + e];
        }
      }
      __syncthreads();
    }
  }
}

template <typename Real>
__global__ __launch_bounds__(LEAF_WAVE) void leaf_atoms_backward_remote_kernel(
    LeafArgs a, const Real* orig, const Real* dnew, Real* dorig) {
  const int lane = threadIdx.x;
  const int64_t nblk = int64_t(a.P) * a.B;
  for (int64_t w = blockIdx.x; w * LEAF_BPW < nblk; w += gridDim.x) {
    const int64_t q0 = w * LEAF_BPW;
    const int bt_lane = lane < LEAF_BPW && q0 + lane < nblk ? a.block_type[q0 + lane] : -1;
    uint64_t real = __ballot(bt_lane >= 0);
    while (real) {
      const LeafPass r = leaf_pass(a, real, q0, bt_lane, lane);
      real = leaf_drop(a, real);
      if (r.g < 0) continue;
      const int n_at = a.n_atoms[r.bt];
      const int64_t base = int64_t(r.q / a.B) * a.NA;
      for (int at = r.at0; at < n_at; at += r.stride) {
        int anc[3], slot[3];
        Vec3<Real> gs[3];
        if (!leaf_shares<Real, true>(a, orig, dnew, r.q, r.bt, at, anc, slot, gs)) continue;
        for (int j = 0; j < 3; ++j) {
          if (slot[j] >= 0) continue;
          Real* d = dorig + 3 * (base + anc[j]);
          atomicAdd(&d[0], gs[j].x);
          atomicAdd(&d[1], gs[j].y);
          atomicAdd(&d[2], gs[j].z);
        }
      }
    }
  }
}

}  // namespace tmol_ljlk
