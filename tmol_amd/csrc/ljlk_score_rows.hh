// Fast scoring kernel for blocks of at most 32 atoms (all canonical residue
// types): one wave walks one row of the work list -- block g1 against every
// neighbor block g2 >= g1 of its pose -- in two phases per block pair.
//
//   stage     g1's atoms once per row (LDS slots 0..31); each g2 into one of
//             two ring slots (32..63, 64..95)
//   cull      atoms of g2 farther than r1 + max_dis from g1's sphere centre,
//             and atoms of g1 farther than r2 + max_dis from g2's, cannot
//             reach any atom of the other block: the pair loop runs over the
//             remaining atoms only (exact: the triangle inequality, with a
//             1e-3 A margin for rounding)
//   phase A   remaining atom pairs: squared distance and count pair only;
//             pairs within max_dis with non-zero weight are ballot-compacted
//             into a wave-private FIFO of LJ entries (and of LK entries when
//             both atoms are heavy)
//   phase B   whenever a FIFO holds a full wave of entries (or a ring slot must
//             be reused, or the row ends): 64 entries, one per lane, every
//             lane busy, block pairs mixed freely
//
// Energies accumulate per lane across the whole row (one pose).  Gradients
// accumulate in LDS per atom slot as 64-bit fixed point (integer LDS atomics)
// -- R copies selected by lane so that lanes sharing an atom rarely hit the
// same address in one instruction -- and reach
// global memory once per atom, term and component when a ring slot is
// recycled (g2) or the row ends (g1).  Loop bounds and FIFO counters are
// wave-uniform scalars.
#pragma once

#include "ljlk_common.hh"

namespace tmol_ljlk {

constexpr int RSLOTS = 96;  // 32 (g1) + 2 x 32 (ring of g2 blocks)
constexpr int LIST = 256;   // FIFO capacity (entries)

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uni64(int64_t x) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(x & 0xffffffffu));
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int lane_val(int v, int j) { return __builtin_amdgcn_readlane(v, j); }

// a / b for 0 <= a < 2^10, 1 <= b <= 64 through the float reciprocal: (a + 1/2) / b
// lies at least 1/(2b) from every integer, far beyond the reciprocal's error,
// so the truncation is exact (a handful of VALU instead of the integer
// division sequence)
__device__ __forceinline__ int small_div(int a, int b) {
  return int((float(a) + 0.5f) * __builtin_amdgcn_rcpf(float(b)));
}
__device__ __forceinline__ float lane_val(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_val(double v, int j) {
  const long long b = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(b & 0xffffffffll), j);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((unsigned long long)b >> 32), j);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// four Reals moved as one 16-byte (float) or 32-byte (double) access
template <typename Real>
struct V4;
template <>
struct V4<float> {
  using T = float4;
};
template <>
struct V4<double> {
  using T = double4;
};

// one lane's atom of a block, held in registers between its global load and
// its LDS staging: coordinates, prepared record, connection path distances
template <typename Real>
struct AtomRegs {
  Real x, y, z;
  typename V4<Real>::T r0;  // lj_radius, sqrt(wdepth), K, 1/lambda^2
  Real vol;
  int flags;
  uint32_t pd;  // byte c: path distance to connection atom c
  Real q;       // elec: partial charge (M_ELEC)
  uint32_t epd; // elec: representative path distances, packed (M_ELEC)
};

template <typename Real, bool ELEC = false>
__device__ __forceinline__ void load_atom(AtomRegs<Real>& R, const ScoreArgs<Real>& A, int t,
                                          int o, int a) {
  using V = typename V4<Real>::T;
  const Real* xyz = A.coords + (int64_t(o) + a) * 3;
  R.x = xyz[0];
  R.y = xyz[1];
  R.z = xyz[2];
  const Real* rec = A.records + (int64_t(t) * A.bt.Amax + a) * REC;
  R.r0 = *(const V*)rec;
  const V r1 = *(const V*)(rec + 4);
  R.vol = r1.x;
  R.flags = int(r1.y);
  R.pd = *(const uint32_t*)(A.conn_pd + (int64_t(t) * A.bt.Amax + a) * MAX_CONN);
  if (ELEC) {
    R.q = A.e_charge[int64_t(t) * A.bt.Amax + a];
    R.epd = A.e_pd[int64_t(t) * A.bt.Amax + a];
  }
}

// Count-pair separation of an inter-block atom pair from packed path
// distances (impl.hh:47-67, count_pair.hh:162-207): min over connection pairs
// of pd1[c1] + cs[c1][c2] + pd2[c2], starting at 5, only when the blocks'
// minimum separation is <= 4; both ligand fragments -> 3/4 become 5.
__device__ __forceinline__ int packed_pair_sep(uint32_t pd1, uint32_t pd2, const uint8_t* cs,
                                               int min_sep, int nc1, int nc2, bool lig) {
  int sep = min_sep;
  if (sep <= 4) {
    int best = 5;
    for (int c1 = 0; c1 < nc1; ++c1) {
      const int p1 = (pd1 >> (8 * c1)) & 255;
      for (int c2 = 0; c2 < nc2; ++c2) {
        const int v = p1 + cs[c1 * MAX_CONN + c2] + ((pd2 >> (8 * c2)) & 255);
        best = v < best ? v : best;
      }
    }
    sep = best;
  }
  if (lig && (sep == 3 || sep == 4)) sep = 5;
  return sep;
}

template <typename Real, int MODE>
struct RowCtx {
  static constexpr bool DVDX = MODE & M_DVDX;
  static constexpr bool WGRAD = MODE & M_WGRAD;
  static constexpr bool BP = MODE & M_BP;
  static constexpr bool NOE = MODE & M_NOE;
  static constexpr bool DERIV = DVDX || WGRAD;
  static constexpr bool ELEC = MODE & M_ELEC;  // whole-pose energies / WGRAD only
  static constexpr bool DET = MODE & M_DET;    // fixed-point global sums (float)
  static constexpr int NG = DVDX ? 3 : (WGRAD ? 1 : 0);
  // gradient accumulator copies: fewer when the planes are many or wide
  static constexpr int R = 1;
};

template <typename Real, int MODE>
struct RowLDS {
  using C = RowCtx<Real, MODE>;
  using T = typename LdsAcc<Real>::T;
  T grad[C::R][C::NG > 0 ? C::NG : 1][3][C::DERIV ? RSLOTS : 1];
  T ebp[C::BP ? 3 : 1][3];  // block-pair energies: ring 0, ring 1, intra
  alignas(4 * sizeof(Real)) Real pos[RSLOTS][4];  // x, y, z, lj_radius
  alignas(4 * sizeof(Real)) Real prm[RSLOTS][4];  // sqrt(lj_wdepth), K, 1/lambda^2, lk_volume
  uint8_t flg[RSLOTS];  // F_* bits and F_HEAVY
  uint16_t lj[LIST];
  uint16_t lk[LIST];
  // culled atoms of the current block pair, compacted: x, y, z, atom index
  alignas(4 * sizeof(Real)) Real c1[TILE][4];  // of g1
  alignas(4 * sizeof(Real)) Real c2[TILE][4];  // of g2
  uint32_t pdw[3][TILE];               // packed path distances: [0] g1, [1 + r] ring slot r
  uint8_t cs[2][MAX_CONN * MAX_CONN];  // per ring slot
  Real wt[C::BP ? 3 : 1][3];           // block-pair weights:  ring 0, ring 1, intra
  // M_ELEC: charge per slot, elec's packed path distances, elec FIFO
  Real qch[C::ELEC ? RSLOTS : 1];
  uint32_t epdw[C::ELEC ? 3 : 1][C::ELEC ? TILE : 1];
  uint16_t el[C::ELEC ? LIST : 1];
};

// entry: a1 slot (5 bits) | slot2 (7 bits) | weight-0.2 flag (bit 12)
__device__ __forceinline__ uint16_t make_entry(int a1, int s2, bool w02) {
  return (uint16_t)(a1 | (s2 << 5) | (w02 ? (1 << 12) : 0));
}

// owner of an entry's second slot: 0/1 ring slot, 2 = g1 (intra)
__device__ __forceinline__ int entry_owner(int s2) { return s2 < TILE ? 2 : (s2 - TILE) >> 5; }

template <typename Real, int MODE, typename LDS>
__device__ __forceinline__ void grad_scatter(LDS& L, int term, int s1, int s2,
                                             Real gx, Real gy, Real gz, int lane) {
  using C = RowCtx<Real, MODE>;
  using Acc = LdsAcc<Real>;
  using T = typename Acc::T;
  const int c = lane & (C::R - 1);
  const T ex = Acc::enc(gx), ey = Acc::enc(gy), ez = Acc::enc(gz);
  atomicAdd(&L.grad[c][term][0][s1], ex);
  atomicAdd(&L.grad[c][term][1][s1], ey);
  atomicAdd(&L.grad[c][term][2][s1], ez);
  atomicAdd(&L.grad[c][term][0][s2], T(0) - ex);
  atomicAdd(&L.grad[c][term][1][s2], T(0) - ey);
  atomicAdd(&L.grad[c][term][2][s2], T(0) - ez);
}

// Evaluate LJ entries [0, consume) of the FIFO, one per lane.
template <typename Real, int MODE, typename LDS>
__device__ __forceinline__ void eval_lj(LDS& L, int consume,
                                        const GlobalParams<Real>& gp, const Real* wrow,
                                        LaneSum& e_atr, LaneSum& e_rep, int lane) {
  using C = RowCtx<Real, MODE>;
  for (int i0 = 0; i0 < consume; i0 += WAVE) {
    const int i = i0 + lane;
    if (i < consume) {
      const int e = L.lj[i];
      const int s1 = e & 31, s2 = (e >> 5) & 127;
      const Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
      using V = typename V4<Real>::T;
      const V p1 = *(const V*)L.pos[s1], p2 = *(const V*)L.pos[s2];
      const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
      const Real d2 = dx * dx + dy * dy + dz * dz;
      const Real inv_d = fast_rsqrt(d2);
      const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);  // coincident atoms: d = 0
      const Real sigma = lj_sigma(L.flg[s1], L.flg[s2], p1.w, p2.w, gp);
      const Real eps = L.prm[s1][0] * L.prm[s2][0];
      Real atr, rep, dlj;
      lj_pair<Real, C::DERIV>(dist, sigma, eps, wcp, gp, atr, rep, dlj);
      if (!C::NOE) {
        if (C::BP) {
          const int ow = entry_owner(s2);
          atomicAdd(&L.ebp[ow][0], LdsAcc<Real>::enc(atr));
          atomicAdd(&L.ebp[ow][1], LdsAcc<Real>::enc(rep));
        } else {
          e_atr.add(atr);
          e_rep.add(rep);
        }
      }
      if (C::DERIV && dlj != Real(0) && dist > Real(0)) {
        const bool inside = dist < sigma;
        Real gs = dlj * inv_d;
        int term = 0;
        if (C::DVDX)
          term = inside ? 1 : 0;
        else
          gs *= C::BP ? L.wt[entry_owner(s2)][inside ? 1 : 0] : wrow[inside ? 1 : 0];
        grad_scatter<Real, MODE>(L, term, s1, s2, gs * dx, gs * dy, gs * dz, lane);
      }
    }
  }
}

template <typename Real, int MODE, typename LDS>
__device__ __forceinline__ void eval_lk(LDS& L, int consume,
                                        const GlobalParams<Real>& gp, const Real* wrow,
                                        LaneSum& e_lk, int lane) {
  using C = RowCtx<Real, MODE>;
  for (int i0 = 0; i0 < consume; i0 += WAVE) {
    const int i = i0 + lane;
    if (i < consume) {
      const int e = L.lk[i];
      const int s1 = e & 31, s2 = (e >> 5) & 127;
      const Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
      using V = typename V4<Real>::T;
      const V p1 = *(const V*)L.pos[s1], p2 = *(const V*)L.pos[s2];
      const V q1 = *(const V*)L.prm[s1], q2 = *(const V*)L.prm[s2];
      const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
      const Real d2 = dx * dx + dy * dy + dz * dz;
      const Real inv_d = fast_rsqrt(d2);
      const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);  // coincident atoms: d = 0
      const int f1 = L.flg[s1], f2 = L.flg[s2];
      const Real sigma = lj_sigma(f1, f2, p1.w, p2.w, gp);
      const bool cc = (f1 & F_CARBON_LK) && (f2 & F_CARBON_LK);
      Real d_min, cdmin, cdmax;
      lk_knots(sigma, cc, d_min, cdmin, cdmax);
      Real va, dva, vb, dvb;
      lk_dir<Real, C::DERIV>(dist, d_min, cdmin, cdmax, -q2.w * q1.y, p1.w, q1.z, gp, va, dva);
      lk_dir<Real, C::DERIV>(dist, d_min, cdmin, cdmax, -q1.w * q2.y, p2.w, q2.z, gp, vb, dvb);
      const Real lk = wcp * (va + vb);
      if (!C::NOE) {
        if (C::BP)
          atomicAdd(&L.ebp[entry_owner(s2)][2], LdsAcc<Real>::enc(lk));
        else
          e_lk.add(lk);
      }
      if (C::DERIV && dist > Real(0)) {
        const Real dlk = wcp * (dva + dvb);
        if (dlk != Real(0)) {
          Real gs = dlk * inv_d;
          if (C::WGRAD) gs *= C::BP ? L.wt[entry_owner(s2)][2] : wrow[2];
          grad_scatter<Real, MODE>(L, C::DVDX ? 2 : 0, s1, s2, gs * dx, gs * dy, gs * dz, lane);
        }
      }
    }
  }
}

// Elec entries [0, consume) of the FIFO (M_ELEC), one per lane: the
// potential per unit charge product (elec_unit), times q1 q2 and the
// count-pair weight; the gradient, times the pose's elec weight, joins the
// same LDS plane as LJ and LK.
template <typename Real, int MODE>
__device__ __forceinline__ void eval_el(RowLDS<Real, MODE>& L, int consume,
                                        const ElecConst<Real>& k, Real wel, LaneSum& e_el,
                                        int lane) {
  using C = RowCtx<Real, MODE>;
  using V = typename V4<Real>::T;
  for (int i0 = 0; i0 < consume; i0 += WAVE) {
    const int i = i0 + lane;
    if (i < consume) {
      const int e = L.el[i];
      const int s1 = e & 31, s2 = (e >> 5) & 127;
      const Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
      const V p1 = *(const V*)L.pos[s1], p2 = *(const V*)L.pos[s2];
      const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
      const Real d2 = dx * dx + dy * dy + dz * dz;
      const Real inv_d = fast_rsqrt(d2);
      const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);
      Real v, dv;
      elec_unit<Real, C::DERIV>(dist, k, v, dv);
      const Real qq = L.qch[s1] * L.qch[s2];
      if (!C::NOE) e_el.add(wcp * (qq * v));
      if (C::DERIV && dist > Real(0)) {
        const Real dE = wcp * (qq * dv);
        if (dE != Real(0)) {
          const Real gs = dE * inv_d * wel;
          grad_scatter<Real, MODE>(L, 0, s1, s2, gs * dx, gs * dy, gs * dz, lane);
        }
      }
    }
  }
}

// move the unconsumed tail [consumed, n) of a FIFO to its front (rem < 64)
__device__ __forceinline__ void fifo_shift(uint16_t* list, int consumed, int n, int lane) {
  const int rem = n - consumed;
  if (consumed == 0 || rem <= 0) return;
  uint16_t v = 0;
  if (lane < rem) v = list[consumed + lane];
  wave_sync();
  if (lane < rem) list[lane] = v;
  wave_sync();
}

// flush gradients of slots [slot0, slot0 + n) to atoms [atom0, atom0 + n) and
// zero them.  Lane l takes component l % 3 of atom l / 3, so one instruction
// covers the x, y, z of 21 atoms as one contiguous run of the [NA, 3] array:
// fewer L2 atomic transactions than one component per instruction.
template <typename Real, int MODE>
__device__ __forceinline__ void flush_slots(RowLDS<Real, MODE>& L, const ScoreArgs<Real>& A,
                                            int slot0, int n, int64_t atom0, int lane) {
  using C = RowCtx<Real, MODE>;
  if (!C::DERIV) return;
  for (int i = lane; i < 3 * n; i += WAVE) {
    const int a = i / 3, k = i - 3 * a;
    const int slot = slot0 + a;
#pragma unroll
    for (int t = 0; t < C::NG; ++t) {
      typename LdsAcc<Real>::T acc = 0;
#pragma unroll
      for (int c = 0; c < C::R; ++c) {
        acc += L.grad[c][t][k][slot];
        L.grad[c][t][k][slot] = 0;
      }
      if (C::DET) {
        // the exact fixed-point sum goes to global memory as it is: integer
        // adds in any order give the same total
        if (acc != 0) atomicAdd(&A.fixed[(int64_t(t) * A.n_atoms + atom0) * 3 + i], acc);
        continue;
      }
      const Real v = LdsAcc<Real>::dec(acc);
      if (v != Real(0)) {
        if (C::DVDX)
          unsafeAtomicAdd(&A.dvdx[(int64_t(t) * A.n_atoms + atom0) * 3 + i], v);
        else
          unsafeAtomicAdd(&A.grad[atom0 * 3 + i], v);
      }
    }
  }
}

// stage one lane's atom into LDS slot `slot`
template <typename Real, int MODE>
__device__ __forceinline__ void stage_atom(RowLDS<Real, MODE>& L, int slot, int pd_row, int a,
                                           const AtomRegs<Real>& R) {
  using V = typename V4<Real>::T;
  V p, q;
  p.x = R.x;
  p.y = R.y;
  p.z = R.z;
  p.w = R.r0.x;
  q.x = R.r0.y;
  q.y = R.r0.z;
  q.z = R.r0.w;
  q.w = R.vol;
  *(V*)L.pos[slot] = p;
  *(V*)L.prm[slot] = q;
  L.flg[slot] = (uint8_t)R.flags;
  L.pdw[pd_row][a] = R.pd;
  if (RowCtx<Real, MODE>::ELEC) {
    L.qch[slot] = R.q;
    L.epdw[pd_row][a] = R.epd;
  }
}

// Occupancy: by default 5 waves per SIMD without gradients (the trimmed LDS
// layout fits 5 workgroups per CU); 4 with them, where a 96-register cap would
// spill and the LDS pipe, busy with the gradient atomics, gains nothing from
// more waves.  OCC > 0 compiles a variant for OCC waves per SIMD (the
// register budget follows), one of the autotuner's choices.
template <int MODE, int OCC>
constexpr int rows_waves_per_simd() {
  return OCC > 0 ? OCC : ((MODE & (M_DVDX | M_WGRAD)) ? 4 : 5);
}

template <typename Real, int MODE, int OCC = 0>
__global__ __launch_bounds__(256, (rows_waves_per_simd<MODE, OCC>())) void score_rows_kernel(
    ScoreArgs<Real> A) {
  using C = RowCtx<Real, MODE>;
  using V = typename V4<Real>::T;
  __shared__ RowLDS<Real, MODE> lds_all[WAVES_PER_WG];
  const int lane = lane_id();
  RowLDS<Real, MODE>& L = lds_all[uni(threadIdx.x / WAVE)];
  const GlobalParams<Real> gp = load_global_params(A.global_params);
  const StackView& s = A.s;
  const BTView& bt = A.bt;
  RowQueue queue(A.row_cursor, s.G, A.row_grab);
  const Real cull_pad = gp.max_dis + Real(1e-3);
  const ElecConst<Real> ek = A.ek;

  if (C::DERIV) {
    auto* g = &L.grad[0][0][0][0];
    for (int i = lane; i < C::R * (C::NG > 0 ? C::NG : 1) * 3 * RSLOTS; i += WAVE) g[i] = 0;
  }
  wave_sync();

  for (int64_t row = queue.next(); row >= 0; row = queue.next()) {
    const int n_items = uni((int)min<int64_t>(A.w.row_count[row], A.w.row_cap));
    if (n_items == 0) continue;
    const int g1 = (int)row;
    const int t1 = uni(s.block_type[g1]);
    const int n1 = uni(bt.n_atoms[t1]);
    const int o1 = uni(s.block_atom_offset[g1]);
    const int p = uni(s.pose_of(g1));
    const int nc1 = uni(bt.n_conn[t1]);
    const bool lig1 = bt.is_ligand[t1] != 0;
    const V sph1 = *(const V*)(A.spheres + 4 * int64_t(g1));
    Real wrow[3] = {Real(0), Real(0), Real(0)};
    if (C::WGRAD && !C::BP) {
      wrow[0] = A.weights[p];
      wrow[1] = A.weights[s.P + p];
      wrow[2] = A.weights[2 * s.P + p];
    }
    const Real wel = (C::ELEC && C::WGRAD) ? A.weights[3 * s.P + p] : Real(0);
    // ---- stage g1 into slots 0..31
    bool heavy1 = false;
    if (lane < n1) {
      AtomRegs<Real> R;
      load_atom<Real, C::ELEC>(R, A, t1, o1, lane);
      stage_atom(L, lane, 0, lane, R);
      heavy1 = (R.flags & F_HEAVY) != 0;
    }
    const uint32_t hmask1 = (uint32_t)__ballot(heavy1);
    if (C::BP && lane < 9) (&L.ebp[0][0])[lane] = 0;
    LaneSum e_atr, e_rep, e_lk, e_el;
    // FIFO state (wave-uniform)
    int n_lj = 0, n_lk = 0;      // pending entries
    int seq_lj = 0, seq_lk = 0;  // entries appended in this row
    int done_lj = 0, done_lk = 0;
    int end_lj0 = 0, end_lj1 = 0, end_lk0 = 0, end_lk1 = 0;
    int n_el = 0, seq_el = 0, done_el = 0, end_el0 = 0, end_el1 = 0;  // M_ELEC
    bool used0 = false, used1 = false;
    int o2s0 = 0, o2s1 = 0;
    int n2s0 = 0, n2s1 = 0;
    int64_t bps0 = 0, bps1 = 0, bp_intra = -1;  // block-pair output indices
    const int64_t PBB = C::BP ? bp_plane(s, A.w) : 0;
    int ring = 0;

    auto drain = [&](bool all) {
      const int c1 = all ? n_lj : (n_lj & ~(WAVE - 1));
      const int c2 = all ? n_lk : (n_lk & ~(WAVE - 1));
      const int c3 = C::ELEC ? (all ? n_el : (n_el & ~(WAVE - 1))) : 0;
      if (c1 > 0) eval_lj<Real, MODE>(L, c1, gp, wrow, e_atr, e_rep, lane);
      if (c2 > 0) eval_lk<Real, MODE>(L, c2, gp, wrow, e_lk, lane);
      if (C::ELEC && c3 > 0) eval_el<Real, MODE>(L, c3, ek, wel, e_el, lane);
      wave_sync();
      fifo_shift(L.lj, c1, n_lj, lane);
      fifo_shift(L.lk, c2, n_lk, lane);
      if (C::ELEC) fifo_shift(L.el, c3, n_el, lane);
      n_lj -= c1;
      done_lj += c1;
      n_lk -= c2;
      done_lk += c2;
      n_el -= c3;
      done_el += c3;
    };
    // some FIFO holds a full wave / could overflow with `room` more entries
    auto full = [&]() { return n_lj >= WAVE || n_lk >= WAVE || (C::ELEC && n_el >= WAVE); };
    auto near_full = [&](int room) {
      return n_lj > LIST - room || n_lk > LIST - room || (C::ELEC && n_el > LIST - room);
    };
    auto retire = [&](int r) {
      // all entries of ring slot r are consumed: flush its g2 atoms
      flush_slots<Real, MODE>(L, A, TILE + TILE * r, r ? n2s1 : n2s0, r ? o2s1 : o2s0, lane);
      if (C::BP && !C::NOE && lane < 3) {
        A.energies_bp[lane * PBB + (r ? bps1 : bps0)] = LdsAcc<Real>::dec(L.ebp[r][lane]);
        L.ebp[r][lane] = 0;
      }
      if (r)
        used1 = false;
      else
        used0 = false;
    };

    for (int c0 = 0; c0 < n_items; c0 += WAVE) {
      // ---- per-item descriptors, one item per lane
      const int n_chunk = min(WAVE, n_items - c0);
      // d_pk: n_atoms | n_conn << 8 | is_ligand << 12 | min block separation << 16
      int d_g2 = g1, d_t2 = t1, d_o2 = 0, d_pk = 0;
      V d_sph;
      d_sph.x = d_sph.y = d_sph.z = d_sph.w = Real(0);
      if (lane < n_chunk) {
        d_g2 = A.w.row_items[row * A.w.row_cap + c0 + lane];
        d_t2 = s.block_type[d_g2];
        d_o2 = s.block_atom_offset[d_g2];
        d_sph = *(const V*)(A.spheres + 4 * int64_t(d_g2));
        const int nc = bt.n_conn[d_t2];
        const int ms = d_g2 == g1 ? 0 : min_block_sep(s, g1, d_g2, nc1, nc);
        d_pk = bt.n_atoms[d_t2] | (nc << 8) | ((bt.is_ligand[d_t2] != 0) << 12) | (ms << 16);
      }
      // atoms of the chunk's first block, then always one block ahead
      AtomRegs<Real> pf;
      {
        const int t = lane_val(d_t2, 0), o = lane_val(d_o2, 0), n = lane_val(d_pk, 0) & 255;
        if (lane < n) load_atom<Real, C::ELEC>(pf, A, t, o, lane);
      }
      for (int j = 0; j < n_chunk; ++j) {
        const int g2 = lane_val(d_g2, j);
        const int pk = lane_val(d_pk, j);
        const int min_sep = pk >> 16;
        const AtomRegs<Real> cur = pf;
        if (g2 == g1) {
          // ---- intra-block pairs a1 < a2, separation = path distance
          if (C::BP) bp_intra = bp_index(s, A.w, g1, g1, c0 + j);
          if (C::BP && C::WGRAD && lane < 3) L.wt[2][lane] = A.weights[lane * PBB + bp_intra];
          if (j + 1 < n_chunk) {
            const int t = lane_val(d_t2, j + 1), o = lane_val(d_o2, j + 1);
            const int n = lane_val(d_pk, j + 1) & 255;
            if (lane < n) load_atom<Real, C::ELEC>(pf, A, t, o, lane);
          }
          wave_sync();
          // only the type's pairs of non-zero weight (prepared table)
          const int n_ip = uni(A.intra_count[t1]);
          const uint16_t* ip = A.intra_pairs + int64_t(t1) * INTRA_MAX;
          for (int k0 = 0; k0 < n_ip; k0 += WAVE) {
            if (near_full(WAVE)) drain(false);
            bool in = false, both_heavy = false, w02 = false;
            int a1 = 0, a2 = 0;
            if (k0 + lane < n_ip) {
              const int e = ip[k0 + lane];
              a1 = e & 31;
              a2 = (e >> 5) & 31;
              const V pa = *(const V*)L.pos[a1], pb = *(const V*)L.pos[a2];
              const Real dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
              in = dx * dx + dy * dy + dz * dz < gp.max_dis2;
              w02 = (e >> 10) & 1;
              both_heavy = in && ((hmask1 >> a1) & (hmask1 >> a2) & 1u);
            }
            const uint64_t m = __ballot(in);
            if (in) L.lj[n_lj + lanes_below(m)] = make_entry(a1, a2, w02);
            n_lj += __popcll(m);
            seq_lj += __popcll(m);
            const uint64_t mh = __ballot(both_heavy);
            if (both_heavy) L.lk[n_lk + lanes_below(mh)] = make_entry(a1, a2, w02);
            n_lk += __popcll(mh);
            seq_lk += __popcll(mh);
            wave_sync();
          }
          if (C::ELEC) {
            // elec's own intra-block table (representative separations,
            // charged atoms only)
            const int n_ie = uni(A.e_intra_count[t1]);
            const uint16_t* ie = A.e_intra_pairs + int64_t(t1) * INTRA_MAX;
            for (int k0 = 0; k0 < n_ie; k0 += WAVE) {
              if (near_full(WAVE)) drain(false);
              bool in = false, w02 = false;
              int a1 = 0, a2 = 0;
              if (k0 + lane < n_ie) {
                const int e = ie[k0 + lane];
                a1 = e & 31;
                a2 = (e >> 5) & 31;
                const V pa = *(const V*)L.pos[a1], pb = *(const V*)L.pos[a2];
                const Real dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
                in = dx * dx + dy * dy + dz * dz < ek.cull2;
                w02 = (e >> 10) & 1;
              }
              const uint64_t m = __ballot(in);
              if (in) L.el[n_el + lanes_below(m)] = make_entry(a1, a2, w02);
              n_el += __popcll(m);
              seq_el += __popcll(m);
              wave_sync();
            }
          }
        } else {
          // ---- inter-block pairs through ring slot `ring`
          const int r = ring;
          ring ^= 1;
          const int n2 = pk & 255;
          const int o2 = lane_val(d_o2, j);
          const int nc2 = (pk >> 8) & 15;
          const bool lig = lig1 && ((pk >> 12) & 1);
          const Real c2x = lane_val(d_sph.x, j), c2y = lane_val(d_sph.y, j);
          const Real c2z = lane_val(d_sph.z, j), c2r = lane_val(d_sph.w, j);
          const int sb = TILE + TILE * r;
          const bool need_cp = min_sep <= 4;
          const int64_t bpi = C::BP ? bp_index(s, A.w, g1, g2, c0 + j) : 0;
          // cull: g2 atoms within reach of g1's sphere (lanes 0..31) and g1
          // atoms within reach of g2's sphere (lanes 32..63)
          V mine;
          mine.x = mine.y = mine.z = mine.w = Real(0);
          bool near = false;
          if (lane < n2) {
            mine.x = cur.x;
            mine.y = cur.y;
            mine.z = cur.z;
            mine.w = Real(lane);
            const Real ex = cur.x - sph1.x, ey = cur.y - sph1.y, ez = cur.z - sph1.z;
            const Real reach = sph1.w + cull_pad;
            near = ex * ex + ey * ey + ez * ez < reach * reach;
          } else if (lane >= TILE && lane - TILE < n1) {
            mine = *(const V*)L.pos[lane - TILE];
            mine.w = Real(lane - TILE);
            const Real ex = mine.x - c2x, ey = mine.y - c2y, ez = mine.z - c2z;
            const Real reach = c2r + cull_pad;
            near = ex * ex + ey * ey + ez * ez < reach * reach;
          }
          const uint64_t mnear = __ballot(near);
          const uint32_t hmask2 = (uint32_t)__ballot(lane < n2 && (cur.flags & F_HEAVY));
          const uint32_t m2 = (uint32_t)mnear, m1 = (uint32_t)(mnear >> 32);
          const int n1c = __popc(m1), n2c = __popc(m2);
          if (n1c == 0 || n2c == 0) {
            // no atom pair within reach: nothing staged, nothing to flush
            if (j + 1 < n_chunk) {
              const int t = lane_val(d_t2, j + 1), o = lane_val(d_o2, j + 1);
              const int n = lane_val(d_pk, j + 1) & 255;
              if (lane < n) load_atom<Real, C::ELEC>(pf, A, t, o, lane);
            }
            ring = r;
            if (full()) drain(false);
            continue;
          }
          if (r ? used1 : used0) {
            const int end_lj = r ? end_lj1 : end_lj0;
            const int end_lk = r ? end_lk1 : end_lk0;
            const int end_el = r ? end_el1 : end_el0;
            if (end_lj > done_lj || end_lk > done_lk || (C::ELEC && end_el > done_el)) drain(true);
            retire(r);
          }
          // global reads this item needs now, before the next prefetch
          if (need_cp && lane >= TILE && lane < TILE + MAX_CONN * MAX_CONN) {
            const int qq = lane - TILE;
            const int cc1 = qq / MAX_CONN, cc2 = qq % MAX_CONN;
            L.cs[r][qq] = (cc1 < nc1 && cc2 < nc2) ? (uint8_t)conn_sep(s, g1, g2, cc1, cc2) : 6;
          }
          if (C::BP && C::WGRAD && lane < 3) L.wt[r][lane] = A.weights[lane * PBB + bpi];
          if (lane < n2) stage_atom(L, sb + lane, 1 + r, lane, cur);
          if (near) {
            if (lane < TILE)
              *(V*)L.c2[__builtin_amdgcn_mbcnt_lo(m2, 0)] = mine;
            else
              *(V*)L.c1[__builtin_amdgcn_mbcnt_hi(m1, 0)] = mine;
          }
          if (j + 1 < n_chunk) {
            const int t = lane_val(d_t2, j + 1), o = lane_val(d_o2, j + 1);
            const int n = lane_val(d_pk, j + 1) & 255;
            if (lane < n) load_atom<Real, C::ELEC>(pf, A, t, o, lane);
          }
          if (r) {
            used1 = true; o2s1 = o2; n2s1 = n2; bps1 = bpi;
          } else {
            used0 = true; o2s0 = o2; n2s0 = n2; bps0 = bpi;
          }
          wave_sync();
          // phase A over the culled atoms, two pair indices per lane and
          // iteration: k = lane + 128 it (+ 64), t = k / n2c, u = k % n2c
          // ie / we: the elec entry (M_ELEC) -- charged atoms within elec's
          // max_dis, separation from the representative path distances
          auto test_pair = [&](int t, int u, bool& in, bool& hv, bool& w02, int& a1, int& a2,
                               bool& ie, bool& we) {
            in = hv = w02 = ie = we = false;
            a1 = a2 = 0;
            if (t < n1c) {
              const V P1 = *(const V*)L.c1[t], P2 = *(const V*)L.c2[u];
              a1 = int(P1.w);
              a2 = int(P2.w);
              const Real dx = P1.x - P2.x, dy = P1.y - P2.y, dz = P1.z - P2.z;
              const Real d2 = dx * dx + dy * dy + dz * dz;
              const int sep = need_cp ? packed_pair_sep(L.pdw[0][a1], L.pdw[1 + r][a2], L.cs[r],
                                                        min_sep, nc1, nc2, lig)
                                      : min_sep;
              in = sep >= 4 && d2 < gp.max_dis2;
              w02 = sep == 4;
              hv = in && ((hmask1 >> a1) & (hmask2 >> a2) & 1u);
              if (C::ELEC) {
                const int sep_e = need_cp ? packed_pair_sep(L.epdw[0][a1], L.epdw[1 + r][a2],
                                                            L.cs[r], min_sep, nc1, nc2, lig)
                                          : min_sep;
                ie = sep_e >= 4 && d2 < ek.cull2 && L.qch[a1] != Real(0) &&
                     L.qch[sb + a2] != Real(0);
                we = sep_e == 4;
              }
            }
          };
          auto append = [&](bool in, bool hv, bool w02, int a1, int a2, bool ie, bool we) {
            const uint64_t m = __ballot(in);
            if (in) L.lj[n_lj + lanes_below(m)] = make_entry(a1, sb + a2, w02);
            n_lj += __popcll(m);
            seq_lj += __popcll(m);
            const uint64_t mh = __ballot(hv);
            if (hv) L.lk[n_lk + lanes_below(mh)] = make_entry(a1, sb + a2, w02);
            n_lk += __popcll(mh);
            seq_lk += __popcll(mh);
            if (C::ELEC) {
              const uint64_t me = __ballot(ie);
              if (ie) L.el[n_el + lanes_below(me)] = make_entry(a1, sb + a2, we);
              n_el += __popcll(me);
              seq_el += __popcll(me);
            }
          };
          const int total = n1c * n2c;
          const int q = small_div(WAVE, n2c), rstep = WAVE - q * n2c;
          int t = small_div(lane, n2c), u = lane - t * n2c;
          for (int k0 = 0; k0 < total; k0 += 2 * WAVE) {
            if (near_full(2 * WAVE)) drain(false);
            int uB = u + rstep, tB = t + q;
            if (uB >= n2c) {
              uB -= n2c;
              tB += 1;
            }
            bool inA, hvA, wA, inB, hvB, wB, ieA, weA, ieB, weB;
            int a1A, a2A, a1B, a2B;
            test_pair(t, u, inA, hvA, wA, a1A, a2A, ieA, weA);
            test_pair(tB, uB, inB, hvB, wB, a1B, a2B, ieB, weB);
            append(inA, hvA, wA, a1A, a2A, ieA, weA);
            append(inB, hvB, wB, a1B, a2B, ieB, weB);
            u = uB + rstep;
            t = tB + q;
            if (u >= n2c) {
              u -= n2c;
              t += 1;
            }
            wave_sync();
          }
          if (r) {
            end_lj1 = seq_lj;
            end_lk1 = seq_lk;
            end_el1 = seq_el;
          } else {
            end_lj0 = seq_lj;
            end_lk0 = seq_lk;
            end_el0 = seq_el;
          }
        }
        // evaluate whatever fills whole waves
        if (full()) drain(false);
      }
    }
    // ---- row end: evaluate the rest, flush everything
    drain(true);
    if (used0) retire(0);
    if (used1) retire(1);
    flush_slots<Real, MODE>(L, A, 0, n1, o1, lane);
    if (C::BP && !C::NOE && lane < 3 && bp_intra >= 0) {
      A.energies_bp[lane * PBB + bp_intra] = LdsAcc<Real>::dec(L.ebp[2][lane]);
      L.ebp[2][lane] = 0;
    }
    if (!C::NOE && !C::BP) {
      const double ea = wave_sum(e_atr.value());
      const double er = wave_sum(e_rep.value());
      const double el = wave_sum(e_lk.value());
      if (lane == 0 && C::DET) {
        auto* e = (unsigned long long*)A.energies64;
        atomicAdd(&e[p], det_energy_enc(ea));
        atomicAdd(&e[s.P + p], det_energy_enc(er));
        atomicAdd(&e[2 * s.P + p], det_energy_enc(el));
      } else if (lane == 0) {
        unsafeAtomicAdd(&A.energies64[p], ea);
        unsafeAtomicAdd(&A.energies64[s.P + p], er);
        unsafeAtomicAdd(&A.energies64[2 * s.P + p], el);
      }
      if (C::ELEC) {
        const double ee = wave_sum(e_el.value());
        if (lane == 0) unsafeAtomicAdd(&A.energies64[3 * s.P + p], ee);
      }
    }
    wave_sync();
  }
}

}  // namespace tmol_ljlk
