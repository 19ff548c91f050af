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

// one thread per (pose, block, atom slot): given atoms are copied, missing
// ones built (NaN when no complete ancestor set exists)
template <typename Real>
__global__ __launch_bounds__(256) void leaf_atoms_kernel(LeafArgs a, const Real* orig,
                                                         Real* out) {
  const int64_t n = int64_t(a.P) * a.B * a.Amax;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int p = int(t / (int64_t(a.B) * a.Amax));
    const int64_t r = t - int64_t(p) * a.B * a.Amax;
    const int b = int(r / a.Amax), at = int(r - int64_t(b) * a.Amax);
    const int bt = a.block_type[p * a.B + b];
    if (bt < 0 || at >= a.n_atoms[bt]) continue;
    const int64_t dst = int64_t(p) * a.NA + a.block_offset[p * a.B + b] + at;
    Real* o = out + 3 * dst;
    if (!a.block_missing[t]) {
      o[0] = orig[3 * dst];
      o[1] = orig[3 * dst + 1];
      o[2] = orig[3 * dst + 2];
      continue;
    }
    int anc[3];
    Real g[3];
    Vec3<Real> W = {Real(NAN), Real(NAN), Real(NAN)};
    if (leaf_ancestors(a, p, b, bt, at, anc, g)) {
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

// backward: the gradient of a given atom passes through; that of a built
// atom goes to its three ancestors (atomics: an ancestor serves several
// built atoms and is itself given)
template <typename Real>
__global__ __launch_bounds__(256) void leaf_atoms_backward_kernel(LeafArgs a, const Real* orig,
                                                                  const Real* dnew,
                                                                  Real* dorig) {
  const int64_t n = int64_t(a.P) * a.B * a.Amax;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int p = int(t / (int64_t(a.B) * a.Amax));
    const int64_t r = t - int64_t(p) * a.B * a.Amax;
    const int b = int(r / a.Amax), at = int(r - int64_t(b) * a.Amax);
    const int bt = a.block_type[p * a.B + b];
    if (bt < 0 || at >= a.n_atoms[bt]) continue;
    const int64_t dst = int64_t(p) * a.NA + a.block_offset[p * a.B + b] + at;
    const Vec3<Real> gw = v3(dnew + 3 * dst);
    if (!a.block_missing[t]) {
      atomicAdd(&dorig[3 * dst], gw.x);
      atomicAdd(&dorig[3 * dst + 1], gw.y);
      atomicAdd(&dorig[3 * dst + 2], gw.z);
      continue;
    }
    int anc[3];
    Real g[3];
    if (!leaf_ancestors(a, p, b, bt, at, anc, g)) continue;
    const int64_t base = int64_t(p) * a.NA;
    const Vec3<Real> A = v3(orig + 3 * (base + anc[0])), B = v3(orig + 3 * (base + anc[1])),
                     C = v3(orig + 3 * (base + anc[2]));
    if (any_nan(A) || any_nan(B) || any_nan(C)) continue;
    Vec3<Real> gA, gB, gC;
    build_coord_vjp(A, B, C, water_coef(g[2], g[1], g[0]), gw, gA, gB, gC);
    const Vec3<Real> gs[3] = {gA, gB, gC};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Real* d = dorig + 3 * (base + anc[j]);
      atomicAdd(&d[0], gs[j].x);
      atomicAdd(&d[1], gs[j].y);
      atomicAdd(&d[2], gs[j].z);
    }
  }
}

}  // namespace tmol_ljlk
