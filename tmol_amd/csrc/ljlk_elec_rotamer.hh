// Rotamer-pair elec scoring for packing (reference: ElecRotamerScoreDispatch,
// tmol/score/elec/potentials/elec_pose_score.impl.hh:1132-1700, on the
// rotamer dispatch of sphere_overlap.impl.hh:23-107, 455-610).
//
// The dispatch list and work units are the LJ+LK rotamer ones (rotamer and
// block spheres, neighbor cells, scans, fill) built with elec's max_dis;
// only the scoring kernel differs:
//   inter     the lanes of a wave hold the atoms of a pack of consecutive b2
//             rotamers (as many as fit in 64 lanes); for each b1 rotamer of
//             the unit, lanes whose atom lies beyond the b1 rotamer's sphere
//             radius + max_dis drop out (a pack with none left skips the
//             rotamer), and the rotamer is staged in LDS (x, y, z, charge and
//             the packed representative path distances to its connection
//             atoms) and read as wave-uniform broadcasts.  Uncharged b1 atoms
//             are skipped (wave-uniform), uncharged lanes never evaluate.
//             Each rotamer pair's energy is a segmented lane reduction.
//   intra     a rotamer with itself: the block type's prepared table of
//             atom pairs a1 < a2 of representative separation >= 4 with both
//             charges non-zero, one pair per lane.
// Block types of at most 32 atoms (the intra table's encoding).
#pragma once

#include "ljlk_elec.hh"
#include "ljlk_rotamer.hh"

namespace tmol_ljlk {

template <typename Real>
struct ElecRotArgs {
  RotView v;
  BTView bt;  // n_atoms, n_conn, conn_atoms, is_ligand (elec's table)
  const Real* coords;
  const Real* rot_spheres;      // [R, 4] centre, radius (from the neighbor pass)
  const Real* charge;           // [T, Amax]
  const uint32_t* epd;          // [T, Amax] packed representative path distances
  const int32_t* intra_count;   // [T]
  const uint16_t* intra_pairs;  // [T, INTRA_MAX]: a1 | a2 << 5 | (separation 4) << 10
  ElecConst<Real> k;
  const RotUnit* units;
  int64_t n_units;
  int64_t n_pairs;
  uint32_t* cursor;   // [8 * QSTRIDE] queue cursors
  Real* energies;     // [n_pairs]
  const Real* dTdV;   // [n_pairs] (backward)
  Real* grad;         // [NA, 3] (backward)
};

// wave-private LDS: the staged atoms of the current b1 rotamer
template <typename Real>
struct ElecRotLDS {
  alignas(4 * sizeof(Real)) Real pos[TILE][4];  // x, y, z, charge
  uint32_t pdw[TILE];                 // packed representative path distances
  uint8_t cs[MAX_CONN * MAX_CONN];    // connection-pair bond separations of the cell
  uint8_t mark[WAVE];                 // lane starts a rotamer of the pack
};

template <typename Real>
__device__ __forceinline__ void load_elec_rot_atom(ElecAtom<Real>& R, const ElecRotArgs<Real>& A,
                                                   int t, int o, int a) {
  const Real* xyz = A.coords + (int64_t(o) + a) * 3;
  R.x = xyz[0];
  R.y = xyz[1];
  R.z = xyz[2];
  const int64_t ta = int64_t(t) * A.bt.Amax + a;
  R.q = A.charge[ta];
  R.pd = A.epd[ta];
}

// stage the n (<= 32) atoms of a rotamer (type t, first atom o) into LDS
template <typename Real>
__device__ __forceinline__ void stage_elec_rot(ElecRotLDS<Real>& L, const ElecRotArgs<Real>& A,
                                               int t, int o, int n, int lane) {
  using V = typename V4<Real>::T;
  wave_sync();
  if (lane < n) {
    ElecAtom<Real> R;
    load_elec_rot_atom(R, A, t, o, lane);
    V p;
    p.x = R.x;
    p.y = R.y;
    p.z = R.z;
    p.w = R.q;
    *(V*)L.pos[lane] = p;
    L.pdw[lane] = R.pd;
  }
  wave_sync();
}

// energy of one charged pair at squared distance d2 with count-pair weight
// wcp; BWD: also the coefficient gs of (x1 - x2) for atom 1 (atom 2 gets the
// negative), scaled by the pair's dTdV weight w
template <typename Real, bool BWD>
__device__ __forceinline__ Real elec_rot_pair(Real d2, Real qq, Real wcp, Real w,
                                              const ElecConst<Real>& k, Real& gs) {
  const Real inv_d = fast_rsqrt(d2);
  const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);
  Real v, dv;
  elec_unit<Real, BWD>(dist, k, v, dv);
  gs = Real(0);
  if (BWD && dist > Real(0)) gs = w * (wcp * (qq * dv)) * inv_d;
  return wcp * (qq * v);
}

// a1's gradient: sum over the wave, one atomic per component
template <typename Real>
__device__ __forceinline__ void elec_rot_flush_a1(Real* grad, int64_t atom, Real gx, Real gy,
                                                  Real gz, int lane) {
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  if (lane < 3) {
    const Real v = lane == 0 ? gx : (lane == 1 ? gy : gz);
    if (v != Real(0)) unsafeAtomicAdd(&grad[atom * 3 + lane], v);
  }
}

template <typename Real>
__device__ __forceinline__ void elec_rot_zero(const ElecRotArgs<Real>& A, int64_t e0, int64_t n,
                                              int lane) {
  for (int64_t i = lane; i < n; i += WAVE) A.energies[e0 + i] = Real(0);
}

// inter-block unit: rotamers [r1_first, +r1_count) of g1 x every rotamer of g2
template <typename Real, bool BWD>
__device__ void elec_rot_inter(const ElecRotArgs<Real>& A, ElecRotLDS<Real>& L, int g1, int g2,
                               int b2, int r1_first, int r1_count, int64_t entry, int lane) {
  using V = typename V4<Real>::T;
  const RotView& v = A.v;
  const ElecConst<Real>& k = A.k;
  const int nr2 = uni(v.n_rots_for_block[g2]), off2 = uni(v.rot_offset_for_block[g2]);
  const int min_sep = uni(v.min_bondsep[int64_t(g1) * v.B + b2]);
  const bool cp = min_sep <= 4;
  if (cp) {
    wave_sync();
    if (lane < MAX_CONN * MAX_CONN) {
      const int c1 = lane / MAX_CONN, c2 = lane - c1 * MAX_CONN;
      L.cs[lane] = (c1 < v.C && c2 < v.C)
                       ? (uint8_t)v.inter_block_bondsep[((int64_t(g1) * v.B + b2) * v.C + c1) *
                                                            v.C + c2]
                       : (uint8_t)6;
    }
    wave_sync();
  }
  for (int ps = 0; ps < nr2;) {
    // pack: consecutive rotamers ps.. whose atoms fit in 64 lanes (every
    // block type has at most 32 atoms, so a pack holds at least two)
    const int rl = ps + lane;
    int nat = 0;
    if (rl < nr2) {
      const int t = v.block_type_for_rot[off2 + rl];
      nat = t >= 0 ? A.bt.n_atoms[t] : 0;
    }
    int incl = nat;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int t = __shfl_up(incl, o, WAVE);
      if (lane >= o) incl += t;
    }
    const int kp = __popcll(__ballot(rl < nr2 && incl <= WAVE));
    const int excl = incl - nat;
    wave_sync();
    L.mark[lane] = 0;
    wave_sync();
    if (lane < kp && nat > 0) L.mark[excl] = (uint8_t)(lane + 1);
    wave_sync();
    const uint64_t starts = __ballot(L.mark[lane] != 0);
    const int total = lane_val(incl, kp - 1);
    const uint64_t below = starts & ((2ull << lane) - 1ull);
    const int start_pos = below ? 63 - __builtin_clzll(below) : -1;
    const int slot = start_pos >= 0 ? int(L.mark[start_pos]) - 1 : -1;
    const int slot_c = slot < 0 ? 0 : slot;
    const int seg_start = __shfl(excl, slot_c, WAVE);
    const int seg_nat = __shfl(nat, slot_c, WAVE);
    const int a2 = lane - seg_start;
    const bool last = lane == seg_start + seg_nat - 1;
    bool active = lane < total && slot >= 0;
    const int r2 = off2 + ps + slot_c;
    const int t2 = active ? v.block_type_for_rot[r2] : 0;
    const int o2 = active ? v.rot_coord_offset[r2] : 0;
    active = active && t2 >= 0 && a2 < seg_nat;
    const int nc2 = active ? A.bt.n_conn[t2] : 0;
    const bool lig2 = active && A.bt.is_ligand[t2];
    ElecAtom<Real> R2;
    R2.x = R2.y = R2.z = R2.q = Real(0);
    R2.pd = 0;
    if (active) load_elec_rot_atom(R2, A, t2, o2, a2);
    const bool charged2 = active && R2.q != Real(0);
    Real g2x = 0, g2y = 0, g2z = 0;
    for (int i = 0; i < r1_count; ++i) {
      const int r1 = r1_first + i;
      const int t1 = uni(v.block_type_for_rot[r1]);
      if (t1 < 0) {
        if (!BWD) elec_rot_zero(A, entry + int64_t(i) * nr2 + ps, kp, lane);
        continue;
      }
      if (!BWD && lane < kp && nat == 0) A.energies[entry + int64_t(i) * nr2 + ps + lane] = Real(0);
      const int o1 = uni(v.rot_coord_offset[r1]);
      const int n1 = uni(A.bt.n_atoms[t1]);
      const int nc1 = uni(A.bt.n_conn[t1]);
      const bool lig = lig2 && A.bt.is_ligand[t1];
      const int64_t e = entry + int64_t(i) * nr2 + ps + slot_c;
      const Real w = (BWD && active) ? A.dTdV[e] : Real(0);
      Real ee = 0;
      // cull: a lane's atom farther than r1's radius + max_dis from r1's
      // centre reaches no atom of r1 (1e-3 A margin for rounding)
      const V sph = *(const V*)(A.rot_spheres + 4 * int64_t(r1));
      const Real cx = R2.x - sph.x, cy = R2.y - sph.y, cz = R2.z - sph.z;
      const Real reach = sph.w + k.hpe + Real(1e-3);
      const bool near = charged2 && cx * cx + cy * cy + cz * cz < reach * reach;
      const int n1n = __ballot(near) != 0 ? n1 : 0;  // wave-uniform
      if (n1n > 0) stage_elec_rot(L, A, t1, o1, n1, lane);
      for (int j = 0; j < n1n; ++j) {
        const V p1 = *(const V*)L.pos[j];
        if (p1.w == Real(0)) continue;  // uncharged b1 atom: every pair is 0
        const Real dx = p1.x - R2.x, dy = p1.y - R2.y, dz = p1.z - R2.z;
        const Real d2 = dx * dx + dy * dy + dz * dz;
        bool in = near && d2 < k.cull2;
        if (__ballot(in) == 0) continue;
        Real wcp = Real(1);
        if (cp) {
          const int sep = packed_pair_sep(L.pdw[j], R2.pd, L.cs, min_sep, nc1, nc2, lig);
          wcp = sep > 4 ? Real(1) : (sep == 4 ? Real(0.2) : Real(0));
          in = in && sep >= 4;
        }
        Real gs = Real(0);
        if (in) {
          const Real E = elec_rot_pair<Real, BWD>(d2, p1.w * R2.q, wcp, w, k, gs);
          if (!BWD) ee += E;
        }
        if (BWD) {
          g2x -= gs * dx;
          g2y -= gs * dy;
          g2z -= gs * dz;
          if (__ballot(gs != Real(0)) != 0)
            elec_rot_flush_a1(A.grad, int64_t(o1) + j, gs * dx, gs * dy, gs * dz, lane);
        }
      }
      if (!BWD) {
        ee = segment_scan(ee, seg_start, lane);
        if (active && last) A.energies[e] = ee;
      }
    }
    if (BWD && active) {
      const int64_t at = (int64_t(o2) + a2) * 3;
      if (g2x != Real(0)) unsafeAtomicAdd(&A.grad[at + 0], g2x);
      if (g2y != Real(0)) unsafeAtomicAdd(&A.grad[at + 1], g2y);
      if (g2z != Real(0)) unsafeAtomicAdd(&A.grad[at + 2], g2z);
    }
    ps += kp;
  }
}

// intra-block unit: each rotamer r of [r1_first, +r1_count) with itself,
// the block type's prepared pair table one pair per lane
template <typename Real, bool BWD>
__device__ void elec_rot_intra(const ElecRotArgs<Real>& A, ElecRotLDS<Real>& L, int r1_first,
                               int r1_count, int64_t entry, int lane) {
  using V = typename V4<Real>::T;
  const RotView& v = A.v;
  const ElecConst<Real>& k = A.k;
  for (int i = 0; i < r1_count; ++i) {
    const int r = r1_first + i;
    const int t = uni(v.block_type_for_rot[r]);
    const int64_t e = entry + i;
    if (t < 0) {
      if (!BWD && lane == 0) A.energies[e] = Real(0);
      continue;
    }
    const int o = uni(v.rot_coord_offset[r]);
    const int na = uni(A.bt.n_atoms[t]);
    const int n_ip = uni(A.intra_count[t]);
    const uint16_t* ip = A.intra_pairs + int64_t(t) * INTRA_MAX;
    const Real w = BWD ? A.dTdV[e] : Real(0);
    stage_elec_rot(L, A, t, o, na, lane);
    Real ee = 0;
    for (int q0 = 0; q0 < n_ip; q0 += WAVE) {
      if (q0 + lane < n_ip) {
        const int en = ip[q0 + lane];
        const int a1 = en & 31, a2 = (en >> 5) & 31;
        const Real wcp = (en >> 10) & 1 ? Real(0.2) : Real(1);
        const V p1 = *(const V*)L.pos[a1], p2 = *(const V*)L.pos[a2];
        const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
        const Real d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < k.cull2) {
          Real gs;
          const Real E = elec_rot_pair<Real, BWD>(d2, p1.w * p2.w, wcp, w, k, gs);
          if (!BWD) ee += E;
          if (BWD && gs != Real(0)) {
            const int64_t i1 = (int64_t(o) + a1) * 3, i2 = (int64_t(o) + a2) * 3;
            const Real gx = gs * dx, gy = gs * dy, gz = gs * dz;
            if (gx != Real(0)) {
              unsafeAtomicAdd(&A.grad[i1 + 0], gx);
              unsafeAtomicAdd(&A.grad[i2 + 0], -gx);
            }
            if (gy != Real(0)) {
              unsafeAtomicAdd(&A.grad[i1 + 1], gy);
              unsafeAtomicAdd(&A.grad[i2 + 1], -gy);
            }
            if (gz != Real(0)) {
              unsafeAtomicAdd(&A.grad[i1 + 2], gz);
              unsafeAtomicAdd(&A.grad[i2 + 2], -gz);
            }
          }
        }
      }
    }
    if (!BWD) {
      ee = wave_sum(ee);
      if (lane == 0) A.energies[e] = ee;
    }
  }
}

// Persistent elec rotamer-pair kernel: 4 independent waves per workgroup,
// units from the per-XCD queues.  BWD false: energies [n_pairs] (every entry
// written, zero for pairs that are not scored); BWD true:
// grad [NA, 3] += sum_i dTdV[i] dV(pair i)/dx.
template <typename Real, bool BWD>
__global__ __launch_bounds__(256) void elec_rotamer_kernel(ElecRotArgs<Real> A) {
  __shared__ ElecRotLDS<Real> lds[WAVES_PER_WG];
  const int lane = lane_id();
  ElecRotLDS<Real>& L = lds[threadIdx.x / WAVE];
  RowQueue queue(A.cursor, A.n_units);
  for (int64_t u = queue.next(); u >= 0; u = queue.next()) {
    const RotUnit* up = A.units + u;
    const int g1 = uni(up->g1), b2 = uni(up->b2);
    const int r1_first = uni(up->r1_first), r1_count = uni(up->r1_count);
    const int64_t entry = uni64(up->entry);
    const int p = g1 / A.v.B, b1 = g1 - p * A.v.B;
    if (r1_count <= 0) {
      if (!BWD && r1_count < 0)
        elec_rot_zero(A, entry,
                      int64_t(-r1_count) *
                          (b1 == b2 ? 1 : uni(A.v.n_rots_for_block[p * A.v.B + b2])),
                      lane);
      continue;
    }
    if (b1 == b2)
      elec_rot_intra<Real, BWD>(A, L, r1_first, r1_count, entry, lane);
    else
      elec_rot_inter<Real, BWD>(A, L, g1, p * A.v.B + b2, b2, r1_first, r1_count, entry, lane);
  }
}

}  // namespace tmol_ljlk
