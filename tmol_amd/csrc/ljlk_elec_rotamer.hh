// Rotamer-pair elec scoring for packing (reference: ElecRotamerScoreDispatch,
// tmol/score/elec/potentials/elec_pose_score.impl.hh:1132-1700, on the
// rotamer dispatch of sphere_overlap.impl.hh:23-107, 455-610).
//
// The dispatch list and work units are the LJ+LK rotamer ones (rotamer and
// block spheres, neighbor cells, scans, fill) built with elec's max_dis;
// only the scoring kernel differs:
//   inter     culled and compacted ("energies and gradients, culled" below):
//             per b1 rotamer, every b2 atom is tested against the b1
//             rotamer's sphere and charged survivors fill 64-lane chunks;
//             pair tests read the b1 atoms from registers and compact the
//             pairs within max_dis into a FIFO evaluated 64 at a time into
//             fixed-point LDS accumulators (energy per chunk lane, or the
//             gradient of both atoms)
//   intra     a rotamer with itself: the block type's prepared table of
//             atom pairs a1 < a2 of representative separation >= 4 with both
//             charges non-zero, one pair per lane.
// Block types of at most 32 atoms (the intra table's encoding).
#pragma once

#include <type_traits>

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
  uint8_t mark[WAVE];                 // candidate lane that starts a b2 rotamer
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

template <typename Real>
__device__ __forceinline__ void elec_rot_zero(const ElecRotArgs<Real>& A, int64_t e0, int64_t n,
                                              int lane) {
  for (int64_t i = lane; i < n; i += WAVE) A.energies[e0 + i] = Real(0);
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

// ------------------------------------------------------------------ energies and gradients, culled

// Culling and compaction (the structure of the LJ+LK rotamer energy kernel),
// for the energies and for the dTdV-weighted backward.  Per b1 rotamer r1 of
// a unit:
//   cull      every atom of b2's rotamers (the candidates, 64 at a time in
//             rotamer order, over windows of up to 64 b2 rotamers) is tested
//             against r1's sphere; charged survivors are compacted, in order,
//             into a 64-lane chunk
//   phase A   for a full (or the last) chunk: r1's atoms come from the
//             registers of the lanes that staged them (readlane), each lane
//             tests its chunk atom; pairs within max_dis of non-zero
//             count-pair weight are ballot-compacted into a FIFO
//   phase B   64 FIFO entries at a time, one per lane, into 64-bit
//             fixed-point LDS accumulators (exact, order-independent): the
//             chunk lane's energy, or (backward) the gradient of the b1 atom
//             and of the chunk atom, the pair weighted by its dTdV
//   fold      forward: at the end of a chunk, a segmented scan adds the
//             lanes' sums per rotamer into the window's per-rotamer sums;
//             backward: the chunk atoms' gradients are flushed (one atomic
//             per non-zero component)
//   write     forward: after the window, one store per rotamer pair (zero
//             for pairs with no surviving atoms); backward: the b1 atoms'
//             gradients are flushed

template <typename Real, bool BWD>
struct ElecRotLDS2 {
  using T = typename LdsAcc<Real>::T;
  ElecRotLDS<Real> r1;  // staged atoms of the b1 rotamer, cell separations, marks
  alignas(4 * sizeof(Real)) Real pos2[WAVE][4];  // chunk: x, y, z, charge
  int meta2[WAVE];       // n_conn | ligand << 3 | slot << 16
  uint32_t pd2[WAVE];    // packed representative path distances
  uint16_t fifo[2 * WAVE];  // j | chunk lane << 5 | (separation 4) << 11
  T lacc[BWD ? 1 : WAVE];   // forward: energy per chunk lane
  Real sum[BWD ? 1 : WAVE];  // forward: per rotamer of the window
  T acc1[BWD ? 3 : 1][BWD ? TILE : 1];  // backward: gradient of the b1 rotamer's atoms
  T acc2[BWD ? 3 : 1][BWD ? WAVE : 1];  // backward: gradient of the chunk's atoms
  int idx2[BWD ? WAVE : 1];             // backward: chunk lane's atom
  Real wt2[BWD ? WAVE : 1];             // backward: dTdV of the chunk lane's rotamer pair
};

template <typename Real, bool BWD>
__device__ __forceinline__ void elec_chunk_eval(ElecRotLDS2<Real, BWD>& L, int consume,
                                                const ElecConst<Real>& k, int lane) {
  using V = typename V4<Real>::T;
  using Acc = LdsAcc<Real>;
  using T = typename Acc::T;
  if (lane < consume) {
    const int en = L.fifo[lane];
    const int j = en & 31, l2 = (en >> 5) & 63;
    const Real wcp = (en >> 11) & 1 ? Real(0.2) : Real(1);
    const V p1 = *(const V*)L.r1.pos[j], p2 = *(const V*)L.pos2[l2];
    const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
    Real gs;
    if constexpr (!BWD) {
      const Real E = elec_rot_pair<Real, false>(dx * dx + dy * dy + dz * dz, p1.w * p2.w, wcp,
                                                Real(0), k, gs);
      atomicAdd(&L.lacc[l2], Acc::enc(E));
    } else {
      elec_rot_pair<Real, true>(dx * dx + dy * dy + dz * dz, p1.w * p2.w, wcp, L.wt2[l2], k, gs);
      if (gs != Real(0)) {
        const T ex = Acc::enc(gs * dx), ey = Acc::enc(gs * dy), ez = Acc::enc(gs * dz);
        atomicAdd(&L.acc1[0][j], ex);
        atomicAdd(&L.acc1[1][j], ey);
        atomicAdd(&L.acc1[2][j], ez);
        atomicAdd(&L.acc2[0][l2], T(0) - ex);
        atomicAdd(&L.acc2[1][l2], T(0) - ey);
        atomicAdd(&L.acc2[2][l2], T(0) - ez);
      }
    }
  }
}

// phase A + B over the chunk's nb lanes against r1 (n1 atoms held in the
// registers R1 of lanes 0..n1-1), then (forward) the fold into the window's
// sums or (backward) the flush of the chunk atoms' gradient
template <typename Real, bool BWD>
__device__ void elec_chunk(ElecRotLDS2<Real, BWD>& L, const ElecConst<Real>& k, int nb, int n1,
                           bool cp, int min_sep, int nc1, bool lig1, const ElecAtom<Real>& R1,
                           Real* grad, int lane) {
  using V = typename V4<Real>::T;
  using Acc = LdsAcc<Real>;
  wave_sync();
  const bool valid = lane < nb;
  V p2 = {};
  int meta = 0;
  uint32_t pd2 = 0;
  if (valid) {
    p2 = *(const V*)L.pos2[lane];
    meta = L.meta2[lane];
    pd2 = L.pd2[lane];
  }
  const int nc2 = meta & 7;
  const bool lig = lig1 && ((meta >> 3) & 1);
  if constexpr (!BWD)
    L.lacc[lane] = 0;
  else
    L.acc2[0][lane] = L.acc2[1][lane] = L.acc2[2][lane] = 0;
  int n_f = 0;
  for (int j = 0; j < n1; ++j) {
    if (lane_val(R1.q, j) == Real(0)) continue;  // uncharged b1 atom
    const Real dx = lane_val(R1.x, j) - p2.x, dy = lane_val(R1.y, j) - p2.y,
               dz = lane_val(R1.z, j) - p2.z;
    bool in = valid && dx * dx + dy * dy + dz * dz < k.cull2;
    if (__ballot(in) == 0) continue;
    bool w02 = false;
    if (cp) {
      const int sep = packed_pair_sep((uint32_t)lane_val((int)R1.pd, j), pd2, L.r1.cs, min_sep,
                                      nc1, nc2, lig);
      in = in && sep >= 4;
      w02 = sep == 4;
    }
    const uint64_t m = __ballot(in);
    if (in) L.fifo[n_f + lanes_below(m)] = (uint16_t)(j | (lane << 5) | (w02 ? (1 << 11) : 0));
    n_f += __popcll(m);
    if (n_f >= WAVE) {
      wave_sync();
      elec_chunk_eval(L, WAVE, k, lane);
      fifo_shift(L.fifo, WAVE, n_f, lane);
      n_f -= WAVE;
    }
  }
  wave_sync();
  if (n_f > 0) elec_chunk_eval(L, n_f, k, lane);
  wave_sync();
  if constexpr (BWD) {
    if (valid) {
      const int64_t at = int64_t(L.idx2[lane]) * 3;
      const Real gx = Acc::dec(L.acc2[0][lane]), gy = Acc::dec(L.acc2[1][lane]),
                 gz = Acc::dec(L.acc2[2][lane]);
      if (gx != Real(0)) unsafeAtomicAdd(&grad[at + 0], gx);
      if (gy != Real(0)) unsafeAtomicAdd(&grad[at + 1], gy);
      if (gz != Real(0)) unsafeAtomicAdd(&grad[at + 2], gz);
    }
    wave_sync();
    return;
  }
  // fold: lanes of one rotamer (slot) are contiguous
  const int slot = meta >> 16;
  const int prev = lane > 0 ? L.meta2[lane - 1] >> 16 : -1;
  const uint64_t starts = __ballot(valid && (lane == 0 || slot != prev));
  const uint64_t below = starts & ((2ull << lane) - 1ull);
  const int seg_start = below ? 63 - __builtin_clzll(below) : 0;
  const bool last =
      valid && (lane == nb - 1 || (L.meta2[lane + 1 < WAVE ? lane + 1 : lane] >> 16) != slot);
  if constexpr (!BWD) {
    Real e = valid ? Acc::dec(L.lacc[lane]) : Real(0);
    e = segment_scan(e, seg_start, lane);
    if (last) L.sum[slot] += e;
  }
  wave_sync();
}

// inter-block unit, culled and compacted: energies (forward) or the
// dTdV-weighted gradient (backward)
template <typename Real, bool BWD>
__device__ void elec_rot_inter_chunked(const ElecRotArgs<Real>& A, ElecRotLDS2<Real, BWD>& L,
                                       int g1, int g2, int b2, int r1_first, int r1_count,
                                       int64_t entry, int lane) {
  using V = typename V4<Real>::T;
  using Acc = LdsAcc<Real>;
  const RotView& v = A.v;
  const ElecConst<Real>& k = A.k;
  const int nr2 = uni(v.n_rots_for_block[g2]), off2 = uni(v.rot_offset_for_block[g2]);
  const int min_sep = uni(v.min_bondsep[int64_t(g1) * v.B + b2]);
  const bool cp = min_sep <= 4;
  if (cp) {
    wave_sync();
    if (lane < MAX_CONN * MAX_CONN) {
      const int c1 = lane / MAX_CONN, c2 = lane - c1 * MAX_CONN;
      L.r1.cs[lane] = (c1 < v.C && c2 < v.C)
                          ? (uint8_t)v.inter_block_bondsep[((int64_t(g1) * v.B + b2) * v.C + c1) *
                                                               v.C + c2]
                          : (uint8_t)6;
    }
    wave_sync();
  }
  for (int w0 = 0; w0 < nr2; w0 += WAVE) {
    const int wn = min(WAVE, nr2 - w0);
    // the window's rotamers (lane = rotamer): type, first atom, atom count,
    // first candidate; candidates fetch theirs by lane shuffle
    int nat = 0, t2l = -1, o2l = 0;
    if (lane < wn) {
      t2l = v.block_type_for_rot[off2 + w0 + lane];
      o2l = v.rot_coord_offset[off2 + w0 + lane];
      nat = t2l >= 0 ? A.bt.n_atoms[t2l] : 0;
    }
    int incl = nat;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int t = __shfl_up(incl, o, WAVE);
      if (lane >= o) incl += t;
    }
    const int excl = incl - nat;
    const int total = lane_val(incl, wn - 1);
    for (int i = 0; i < r1_count; ++i) {
      const int r1 = r1_first + i;
      const int t1 = uni(v.block_type_for_rot[r1]);
      if (t1 < 0) {
        if (!BWD) elec_rot_zero(A, entry + int64_t(i) * nr2 + w0, wn, lane);
        continue;
      }
      const int o1 = uni(v.rot_coord_offset[r1]);
      const int n1 = uni(A.bt.n_atoms[t1]);
      const int nc1 = uni(A.bt.n_conn[t1]);
      const bool lig1 = A.bt.is_ligand[t1];
      const V sph = *(const V*)(A.rot_spheres + 4 * int64_t(r1));
      const Real reach = sph.w + k.hpe + Real(1e-3);
      const Real reach2 = reach * reach;
      ElecAtom<Real> R1;
      R1.x = R1.y = R1.z = R1.q = Real(0);
      R1.pd = 0;
      wave_sync();
      if (lane < n1) {
        load_elec_rot_atom(R1, A, t1, o1, lane);
        V p;
        p.x = R1.x;
        p.y = R1.y;
        p.z = R1.z;
        p.w = R1.q;
        *(V*)L.r1.pos[lane] = p;
      }
      if constexpr (!BWD)
        L.sum[lane] = 0;
      else if (lane < TILE)
        L.acc1[0][lane] = L.acc1[1][lane] = L.acc1[2][lane] = 0;
      int nb = 0;  // surviving candidates in the chunk (wave-uniform)
      for (int c0 = 0; c0 < total; c0 += WAVE) {
        // candidate c = c0 + lane: its rotamer (slot) and atom
        const int c = c0 + lane;
        const uint64_t before = __ballot(lane < wn && excl <= c0 && nat > 0);
        const int carry = before ? 63 - __builtin_clzll(before) : 0;
        wave_sync();
        L.r1.mark[lane] = 0;
        wave_sync();
        if (lane < wn && nat > 0 && excl > c0 && excl < c0 + WAVE)
          L.r1.mark[excl - c0] = (uint8_t)(lane + 1);
        wave_sync();
        const uint64_t st = __ballot(L.r1.mark[lane] != 0);
        const uint64_t bl = st & ((2ull << lane) - 1ull);
        const int slot = bl ? int(L.r1.mark[63 - __builtin_clzll(bl)]) - 1 : carry;
        const bool cand = c < total;
        const int a2 = c - __shfl(excl, slot, WAVE);
        const int t2 = __shfl(t2l, slot, WAVE);
        const int o2 = __shfl(o2l, slot, WAVE);
        ElecAtom<Real> C2;
        bool keep = false;
        if (cand) {
          load_elec_rot_atom(C2, A, t2, o2, a2);
          const Real ex = C2.x - sph.x, ey = C2.y - sph.y, ez = C2.z - sph.z;
          keep = C2.q != Real(0) && ex * ex + ey * ey + ez * ez < reach2;
        }
        const uint64_t mk = __ballot(keep);
        const int nk = __popcll(mk);
        if (nk == 0) continue;
        if (nb + nk > WAVE) {
          elec_chunk(L, k, nb, n1, cp, min_sep, nc1, lig1, R1, A.grad, lane);
          nb = 0;
        }
        if (keep) {
          const int pos = nb + lanes_below(mk);
          V p;
          p.x = C2.x;
          p.y = C2.y;
          p.z = C2.z;
          p.w = C2.q;
          *(V*)L.pos2[pos] = p;
          L.meta2[pos] = A.bt.n_conn[t2] | ((A.bt.is_ligand[t2] ? 1 : 0) << 3) | (slot << 16);
          L.pd2[pos] = C2.pd;
          if constexpr (BWD) {
            L.idx2[pos] = o2 + a2;
            L.wt2[pos] = A.dTdV[entry + int64_t(i) * nr2 + w0 + slot];
          }
        }
        nb += nk;
      }
      if (nb > 0) elec_chunk(L, k, nb, n1, cp, min_sep, nc1, lig1, R1, A.grad, lane);
      wave_sync();
      if constexpr (!BWD) {
        if (lane < wn) A.energies[entry + int64_t(i) * nr2 + w0 + lane] = L.sum[lane];
      } else {
        for (int q = lane; q < 3 * n1; q += WAVE) {
          const int a = q / 3, c = q - 3 * a;
          const Real gv = Acc::dec(L.acc1[c][a]);
          if (gv != Real(0)) unsafeAtomicAdd(&A.grad[(int64_t(o1) + a) * 3 + c], gv);
        }
      }
    }
  }
}

// Persistent elec rotamer-pair kernel: 4 independent waves per workgroup,
// units from the per-XCD queues.  BWD false: energies [n_pairs] (every entry
// written, zero for pairs that are not scored); BWD true: grad [NA, 3] +=
// sum_i dTdV[i] dV(pair i)/dx.  Float builds are held to 7 waves per SIMD:
// the kernel is latency-bound, and the few spilled registers cost less than
// the waves gained (forward 5.76 -> 4.77 ms on the 16 x 300-residue
// repacking set).
template <typename Real, bool BWD>
__global__ __launch_bounds__(256, sizeof(Real) == 4 ? 7 : 2) void elec_rotamer_kernel(
    ElecRotArgs<Real> A) {
  using LDS = ElecRotLDS2<Real, BWD>;
  __shared__ LDS lds[WAVES_PER_WG];
  const int lane = lane_id();
  LDS& L = lds[threadIdx.x / WAVE];
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
      elec_rot_intra<Real, BWD>(A, L.r1, r1_first, r1_count, entry, lane);
    else
      elec_rot_inter_chunked<Real, BWD>(A, L, g1, p * A.v.B + b2, b2, r1_first, r1_count, entry,
                                        lane);
  }
}

}  // namespace tmol_ljlk
