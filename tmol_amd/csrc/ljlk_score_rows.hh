// Fast scoring kernel for blocks of at most 32 atoms (all canonical residue
// types): one wave walks one row of the work list -- block g1 against every
// neighbor block g2 >= g1 of its pose -- in two phases per block pair.
//
//   stage     g1's atoms once per row (LDS slots 0..31); each g2 into one of
//             two ring slots (32..63, 64..95), its loads issued one block pair
//             ahead
//   phase A   all atom pairs of the block pair: squared distance and count
//             pair only; pairs within max_dis with non-zero weight are
//             ballot-compacted into a wave-private FIFO of LJ entries (and of
//             LK entries when both atoms are heavy)
//   phase B   whenever a FIFO holds a full wave of entries (or a ring slot must
//             be reused, or the row ends): evaluate 64 entries, one per lane,
//             with every lane busy, block pairs mixed freely
//
// Energies accumulate per lane across the whole row (one pose); gradients
// accumulate in LDS per atom slot and reach global memory once per atom,
// term and component when a ring slot is recycled (g2) or the row ends (g1).
#pragma once

#include "ljlk_common.hh"

namespace tmol_ljlk {

constexpr int RSLOTS = 96;   // 32 (g1) + 2 x 32 (ring of g2 blocks)
constexpr int LIST = 256;    // FIFO capacity (entries)

template <typename Real, int NG>
struct RowLDS {
  Real pos[RSLOTS][4];  // x, y, z, lj_radius
  Real prm[RSLOTS][4];  // sqrt(lj_wdepth), K, 1/lambda^2, lk_volume
  int flg[RSLOTS];
  Real grad[NG > 0 ? NG : 1][3][RSLOTS];
  uint16_t lj[LIST];
  uint16_t lk[LIST];
  uint8_t pd[3][MAX_CONN][TILE];      // [0] g1, [1 + r] ring slot r
  uint8_t cs[2][MAX_CONN * MAX_CONN];  // per ring slot
  Real ebp[3][3];                      // block-pair energies: ring 0, ring 1, intra
  Real wt[3][3];                       // block-pair weights:  ring 0, ring 1, intra
};

// entry: a1 (5 bits) | slot2 (7 bits) | weight-0.2 flag (bit 12)
__device__ __forceinline__ uint16_t make_entry(int a1, int s2, bool w02) {
  return (uint16_t)(a1 | (s2 << 5) | (w02 ? (1 << 12) : 0));
}

// owner of an entry's second slot: 0/1 ring slot, 2 = g1 (intra)
__device__ __forceinline__ int entry_owner(int s2) { return s2 < TILE ? 2 : (s2 - TILE) >> 5; }

template <typename Real, int MODE>
struct RowCtx {
  static constexpr bool DVDX = MODE & M_DVDX;
  static constexpr bool WGRAD = MODE & M_WGRAD;
  static constexpr bool BP = MODE & M_BP;
  static constexpr bool NOE = MODE & M_NOE;
  static constexpr bool DERIV = DVDX || WGRAD;
  static constexpr int NG = DVDX ? 3 : (WGRAD ? 1 : 0);
};

// Evaluate LJ entries [0, n) of the FIFO in batches of 64 (the last one may be
// partial when `all`); returns the number consumed.
template <typename Real, int MODE>
__device__ __forceinline__ int drain_lj(RowLDS<Real, RowCtx<Real, MODE>::NG>& L, int n, bool all,
                                        const GlobalParams<Real>& gp, const Real* wrow,
                                        KSum<Real>& e_atr, KSum<Real>& e_rep, int lane) {
  using C = RowCtx<Real, MODE>;
  int consume = all ? n : (n & ~(WAVE - 1));
  for (int i0 = 0; i0 < consume; i0 += WAVE) {
    int i = i0 + lane;
    if (i < consume) {
      int e = L.lj[i];
      int s1 = e & 31, s2 = (e >> 5) & 127;
      Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
      Real x1 = L.pos[s1][0], y1 = L.pos[s1][1], z1 = L.pos[s1][2], r1 = L.pos[s1][3];
      Real x2 = L.pos[s2][0], y2 = L.pos[s2][1], z2 = L.pos[s2][2], r2 = L.pos[s2][3];
      Real dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
      Real d2 = dx * dx + dy * dy + dz * dz;
      Real dist = fast_sqrt(d2);
      int f1 = L.flg[s1], f2 = L.flg[s2];
      Real sigma = lj_sigma(f1, f2, r1, r2, gp);
      Real eps = L.prm[s1][0] * L.prm[s2][0];
      Real atr, rep, dlj;
      lj_pair<Real, C::DERIV>(dist, sigma, eps, wcp, gp, atr, rep, dlj);
      if (!C::NOE) {
        if (C::BP) {
          int ow = entry_owner(s2);
          atomicAdd(&L.ebp[ow][0], atr);
          atomicAdd(&L.ebp[ow][1], rep);
        } else {
          e_atr.add(atr);
          e_rep.add(rep);
        }
      }
      if (C::DERIV && dlj != Real(0) && dist > Real(0)) {
        bool inside = dist < sigma;
        Real gs = dlj * fast_rcp(dist);
        int term = 0;
        if (C::DVDX) {
          term = inside ? 1 : 0;
        } else {
          Real wv;
          if (C::BP)
            wv = L.wt[entry_owner(s2)][inside ? 1 : 0];
          else
            wv = wrow[inside ? 1 : 0];
          gs *= wv;
        }
        Real gx = gs * dx, gy = gs * dy, gz = gs * dz;
        atomicAdd(&L.grad[term][0][s1], gx);
        atomicAdd(&L.grad[term][1][s1], gy);
        atomicAdd(&L.grad[term][2][s1], gz);
        atomicAdd(&L.grad[term][0][s2], -gx);
        atomicAdd(&L.grad[term][1][s2], -gy);
        atomicAdd(&L.grad[term][2][s2], -gz);
      }
    }
  }
  return consume;
}

template <typename Real, int MODE>
__device__ __forceinline__ int drain_lk(RowLDS<Real, RowCtx<Real, MODE>::NG>& L, int n, bool all,
                                        const GlobalParams<Real>& gp, const Real* wrow,
                                        KSum<Real>& e_lk, int lane) {
  using C = RowCtx<Real, MODE>;
  int consume = all ? n : (n & ~(WAVE - 1));
  for (int i0 = 0; i0 < consume; i0 += WAVE) {
    int i = i0 + lane;
    if (i < consume) {
      int e = L.lk[i];
      int s1 = e & 31, s2 = (e >> 5) & 127;
      Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
      Real x1 = L.pos[s1][0], y1 = L.pos[s1][1], z1 = L.pos[s1][2], r1 = L.pos[s1][3];
      Real x2 = L.pos[s2][0], y2 = L.pos[s2][1], z2 = L.pos[s2][2], r2 = L.pos[s2][3];
      Real dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
      Real d2 = dx * dx + dy * dy + dz * dz;
      Real dist = fast_sqrt(d2);
      int f1 = L.flg[s1], f2 = L.flg[s2];
      Real sigma = lj_sigma(f1, f2, r1, r2, gp);
      bool cc = (f1 & F_CARBON_LK) && (f2 & F_CARBON_LK);
      Real d_min, cdmin, cdmax;
      lk_knots(sigma, cc, d_min, cdmin, cdmax);
      Real K1 = L.prm[s1][1], il1 = L.prm[s1][2], v1 = L.prm[s1][3];
      Real K2 = L.prm[s2][1], il2 = L.prm[s2][2], v2 = L.prm[s2][3];
      Real va, dva, vb, dvb;
      lk_dir<Real, C::DERIV>(dist, d_min, cdmin, cdmax, -v2 * K1, r1, il1, gp, va, dva);
      lk_dir<Real, C::DERIV>(dist, d_min, cdmin, cdmax, -v1 * K2, r2, il2, gp, vb, dvb);
      Real lk = wcp * (va + vb);
      if (!C::NOE) {
        if (C::BP)
          atomicAdd(&L.ebp[entry_owner(s2)][2], lk);
        else
          e_lk.add(lk);
      }
      if (C::DERIV && dist > Real(0)) {
        Real dlk = wcp * (dva + dvb);
        if (dlk != Real(0)) {
          Real gs = dlk * fast_rcp(dist);
          int term = C::DVDX ? 2 : 0;
          if (C::WGRAD) gs *= C::BP ? L.wt[entry_owner(s2)][2] : wrow[2];
          Real gx = gs * dx, gy = gs * dy, gz = gs * dz;
          atomicAdd(&L.grad[term][0][s1], gx);
          atomicAdd(&L.grad[term][1][s1], gy);
          atomicAdd(&L.grad[term][2][s1], gz);
          atomicAdd(&L.grad[term][0][s2], -gx);
          atomicAdd(&L.grad[term][1][s2], -gy);
          atomicAdd(&L.grad[term][2][s2], -gz);
        }
      }
    }
  }
  return consume;
}

// move the unconsumed tail [consumed, n) of a FIFO to its front
__device__ __forceinline__ void fifo_shift(uint16_t* list, int consumed, int n, int lane) {
  int rem = n - consumed;
  if (consumed == 0 || rem <= 0) return;
  // rem < 64 except after a partial-free drain of a full list; loop for safety
  for (int b = 0; b < rem; b += WAVE) {
    uint16_t v = 0;
    if (b + lane < rem) v = list[consumed + b + lane];
    wave_sync();
    if (b + lane < rem) list[b + lane] = v;
    wave_sync();
  }
}

// flush gradients of slots [slot0, slot0 + n) to atoms [atom0, atom0 + n), zero them
template <typename Real, int MODE>
__device__ __forceinline__ void flush_slots(RowLDS<Real, RowCtx<Real, MODE>::NG>& L,
                                            const ScoreArgs<Real>& A, int slot0, int n,
                                            int64_t atom0, int lane) {
  using C = RowCtx<Real, MODE>;
  if (!C::DERIV) return;
  if (lane < n) {
    int slot = slot0 + lane;
    int64_t atom = atom0 + lane;
#pragma unroll
    for (int t = 0; t < C::NG; ++t)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        Real v = L.grad[t][k][slot];
        if (v != Real(0)) {
          if (C::DVDX)
            unsafeAtomicAdd(&A.dvdx[(int64_t(t) * A.n_atoms + atom) * 3 + k], v);
          else
            unsafeAtomicAdd(&A.grad[atom * 3 + k], v);
        }
        L.grad[t][k][slot] = Real(0);
      }
  }
}

template <typename Real, int MODE>
__global__ __launch_bounds__(256) void score_rows_kernel(ScoreArgs<Real> A) {
  using C = RowCtx<Real, MODE>;
  __shared__ RowLDS<Real, C::NG> lds_all[WAVES_PER_WG];
  const int lane = lane_id();
  RowLDS<Real, C::NG>& L = lds_all[threadIdx.x / WAVE];
  const GlobalParams<Real> gp = load_global_params(A.global_params);
  const StackView& s = A.s;
  const BTView& bt = A.bt;
  const XcdSlice sl = xcd_slice(s.G);

  // zero the gradient accumulators once; flushes re-zero what they read
  if (C::DERIV) {
    for (int i = lane; i < (C::NG > 0 ? C::NG : 1) * 3 * RSLOTS; i += WAVE)
      (&L.grad[0][0][0])[i] = Real(0);
  }
  wave_sync();

  for (int64_t row = sl.first; row < sl.end; row += sl.stride) {
    int n_items = (int)min<int64_t>(A.w.row_count[row], A.w.row_cap);
    if (n_items == 0) continue;
    const int g1 = (int)row;
    const int t1 = s.block_type[g1];
    const int n1 = bt.n_atoms[t1];
    const int64_t o1 = s.block_atom_offset[g1];
    const int p = s.pose_of(g1);
    const int nc1 = bt.n_conn[t1];
    const bool lig1 = bt.is_ligand[t1] != 0;
    const int b1 = g1 - p * s.B;  // block index within the pose (padded layout)
    Real wrow[3] = {Real(0), Real(0), Real(0)};
    if (C::WGRAD && !C::BP) {
      wrow[0] = A.weights[p];
      wrow[1] = A.weights[s.P + p];
      wrow[2] = A.weights[2 * s.P + p];
    }
    // ---- stage g1 into slots 0..31
    if (lane < TILE) {
      int a = lane;
      if (a < n1) {
        const Real* xyz = A.coords + (o1 + a) * 3;
        const Real* r = A.records + (int64_t(t1) * bt.Amax + a) * REC;
        L.pos[a][0] = xyz[0];
        L.pos[a][1] = xyz[1];
        L.pos[a][2] = xyz[2];
        L.pos[a][3] = r[0];
        L.prm[a][0] = r[1];
        L.prm[a][1] = r[2];
        L.prm[a][2] = r[3];
        L.prm[a][3] = r[4];
        L.flg[a] = int(r[5]);
        for (int c = 0; c < MAX_CONN; ++c)
          L.pd[0][c][a] = A.conn_pd[(int64_t(t1) * MAX_CONN + c) * bt.Amax + a];
      }
    }
    if (C::BP && lane < 9) (&L.ebp[0][0])[lane] = Real(0);
    KSum<Real> e_atr, e_rep, e_lk;
    // FIFO state (wave-uniform)
    int n_lj = 0, n_lk = 0;          // pending entries
    int64_t seq_lj = 0, seq_lk = 0;  // entries ever appended (this row)
    int64_t done_lj = 0, done_lk = 0;
    int64_t end_lj[2] = {0, 0}, end_lk[2] = {0, 0};  // last seq + 1 per ring slot
    bool slot_used[2] = {false, false};
    int64_t slot_o2[2] = {0, 0};
    int slot_n2[2] = {0, 0};
    int slot_b2[2] = {0, 0};
    int ring = 0;

    auto drain = [&](bool all) {
      int c1 = drain_lj<Real, MODE>(L, n_lj, all, gp, wrow, e_atr, e_rep, lane);
      wave_sync();
      fifo_shift(L.lj, c1, n_lj, lane);
      n_lj -= c1;
      done_lj += c1;
      int c2 = drain_lk<Real, MODE>(L, n_lk, all, gp, wrow, e_lk, lane);
      wave_sync();
      fifo_shift(L.lk, c2, n_lk, lane);
      n_lk -= c2;
      done_lk += c2;
    };
    auto retire_slot = [&](int r) {
      // all entries of ring slot r are consumed: flush its g2 atoms
      flush_slots<Real, MODE>(L, A, TILE + TILE * r, slot_n2[r], slot_o2[r], lane);
      if (C::BP && !C::NOE && lane < 3) {
        int64_t PBB = int64_t(s.P) * s.B * s.B;
        int64_t base = (int64_t(p) * s.B + b1) * s.B + slot_b2[r];
        A.energies_bp[lane * PBB + base] = L.ebp[r][lane];
        L.ebp[r][lane] = Real(0);
      }
      slot_used[r] = false;
    };

    for (int c0 = 0; c0 < n_items; c0 += WAVE) {
      // ---- per-item descriptors, one item per lane
      const int n_chunk = min(WAVE, n_items - c0);
      int d_g2 = g1, d_t2 = t1, d_n2 = 0, d_min = 6;
      int64_t d_o2 = 0;
      if (lane < n_chunk) {
        d_g2 = A.w.row_items[row * A.w.row_cap + c0 + lane];
        d_t2 = s.block_type[d_g2];
        d_o2 = s.block_atom_offset[d_g2];
        d_n2 = bt.n_atoms[d_t2];
        d_min = d_g2 == g1 ? 0 : min_block_sep(s, g1, d_g2, nc1, bt.n_conn[d_t2]);
      }
      for (int j = 0; j < n_chunk; ++j) {
        const int g2 = __shfl(d_g2, j, WAVE);
        const int t2 = __shfl(d_t2, j, WAVE);
        const int n2 = __shfl(d_n2, j, WAVE);
        const int min_sep = __shfl(d_min, j, WAVE);
        const int64_t o2 = __shfl(d_o2, j, WAVE);
        if (g2 == g1) {
          // ---- intra-block pairs a1 < a2, separation = path distance
          if (C::BP && C::WGRAD && lane < 3) {
            int64_t PBB = int64_t(s.P) * s.B * s.B;
            L.wt[2][lane] = A.weights[lane * PBB + (int64_t(p) * s.B + b1) * s.B + b1];
          }
          wave_sync();
          const int total = n1 * n1;
          for (int k0 = 0; k0 < total; k0 += WAVE) {
            if (n_lj > LIST - WAVE || n_lk > LIST - WAVE) drain(false);
            int k = k0 + lane;
            int a1 = k / n1, a2 = k - (k / n1) * n1;
            bool in = false, both_heavy = false, w02 = false;
            if (k < total && a1 < a2) {
              int sep = bt.path_distance[(int64_t(t1) * bt.Amax + a1) * bt.Amax + a2];
              Real dx = L.pos[a1][0] - L.pos[a2][0];
              Real dy = L.pos[a1][1] - L.pos[a2][1];
              Real dz = L.pos[a1][2] - L.pos[a2][2];
              Real d2 = dx * dx + dy * dy + dz * dz;
              in = sep >= 4 && d2 < gp.max_dis2;
              w02 = sep == 4;
              both_heavy = in && (L.flg[a1] & L.flg[a2] & F_HEAVY);
            }
            uint64_t m = __ballot(in);
            if (in) L.lj[n_lj + lanes_below(m)] = make_entry(a1, a2, w02);
            n_lj += __popcll(m);
            seq_lj += __popcll(m);
            uint64_t mh = __ballot(both_heavy);
            if (both_heavy) L.lk[n_lk + lanes_below(mh)] = make_entry(a1, a2, w02);
            n_lk += __popcll(mh);
            seq_lk += __popcll(mh);
            wave_sync();
          }
        } else {
          // ---- inter-block pairs through ring slot `ring`
          const int r = ring;
          ring ^= 1;
          if (slot_used[r]) {
            if (end_lj[r] > done_lj || end_lk[r] > done_lk) drain(true);
            retire_slot(r);
          }
          const int sb = TILE + TILE * r;
          const bool need_cp = min_sep <= 4;
          const int nc2 = bt.n_conn[t2];
          const bool lig = lig1 && bt.is_ligand[t2] != 0;
          if (lane < n2) {
            int a = lane;
            const Real* xyz = A.coords + (o2 + a) * 3;
            const Real* rr = A.records + (int64_t(t2) * bt.Amax + a) * REC;
            L.pos[sb + a][0] = xyz[0];
            L.pos[sb + a][1] = xyz[1];
            L.pos[sb + a][2] = xyz[2];
            L.pos[sb + a][3] = rr[0];
            L.prm[sb + a][0] = rr[1];
            L.prm[sb + a][1] = rr[2];
            L.prm[sb + a][2] = rr[3];
            L.prm[sb + a][3] = rr[4];
            L.flg[sb + a] = int(rr[5]);
            if (need_cp)
              for (int c = 0; c < MAX_CONN; ++c)
                L.pd[1 + r][c][a] = A.conn_pd[(int64_t(t2) * MAX_CONN + c) * bt.Amax + a];
          }
          if (need_cp && lane >= TILE && lane < TILE + MAX_CONN * MAX_CONN) {
            int q = lane - TILE;
            int c1 = q / MAX_CONN, c2 = q % MAX_CONN;
            L.cs[r][q] = (c1 < nc1 && c2 < nc2) ? (uint8_t)conn_sep(s, g1, g2, c1, c2) : 6;
          }
          const int b2 = g2 - p * s.B;
          if (C::BP && C::WGRAD && lane < 3) {
            int64_t PBB = int64_t(s.P) * s.B * s.B;
            L.wt[r][lane] = A.weights[lane * PBB + (int64_t(p) * s.B + b1) * s.B + b2];
          }
          slot_used[r] = true;
          slot_o2[r] = o2;
          slot_n2[r] = n2;
          slot_b2[r] = b2;
          wave_sync();
          // phase A: n1 x n2 pairs, k = lane + 64 it, a1 = k / n2, a2 = k % n2
          const int total = n1 * n2;
          const int q = WAVE / n2, rstep = WAVE - q * n2;
          int a2 = lane % n2, a1 = lane / n2;
          for (int k0 = 0; k0 < total; k0 += WAVE) {
            if (n_lj > LIST - WAVE || n_lk > LIST - WAVE) drain(false);
            bool in = false, both_heavy = false, w02 = false;
            if (a1 < n1) {
              Real dx = L.pos[a1][0] - L.pos[sb + a2][0];
              Real dy = L.pos[a1][1] - L.pos[sb + a2][1];
              Real dz = L.pos[a1][2] - L.pos[sb + a2][2];
              Real d2 = dx * dx + dy * dy + dz * dz;
              int sep = need_cp ? inter_pair_sep(L.pd[0], L.pd[1 + r], L.cs[r], min_sep, nc1,
                                                 nc2, a1, a2, lig)
                                : min_sep;
              in = sep >= 4 && d2 < gp.max_dis2;
              w02 = sep == 4;
              both_heavy = in && (L.flg[a1] & L.flg[sb + a2] & F_HEAVY);
            }
            uint64_t m = __ballot(in);
            if (in) L.lj[n_lj + lanes_below(m)] = make_entry(a1, sb + a2, w02);
            n_lj += __popcll(m);
            seq_lj += __popcll(m);
            uint64_t mh = __ballot(both_heavy);
            if (both_heavy) L.lk[n_lk + lanes_below(mh)] = make_entry(a1, sb + a2, w02);
            n_lk += __popcll(mh);
            seq_lk += __popcll(mh);
            a2 += rstep;
            a1 += q;
            if (a2 >= n2) {
              a2 -= n2;
              a1 += 1;
            }
            wave_sync();
          }
          end_lj[r] = seq_lj;
          end_lk[r] = seq_lk;
        }
        // evaluate whatever fills whole waves
        if (n_lj >= WAVE || n_lk >= WAVE) drain(false);
      }
    }
    // ---- row end: evaluate the rest, flush everything
    drain(true);
    for (int r = 0; r < 2; ++r)
      if (slot_used[r]) retire_slot(r);
    flush_slots<Real, MODE>(L, A, 0, n1, o1, lane);
    if (C::BP && !C::NOE && lane < 3) {
      int64_t PBB = int64_t(s.P) * s.B * s.B;
      int64_t base = (int64_t(p) * s.B + b1) * s.B + b1;
      A.energies_bp[lane * PBB + base] = L.ebp[2][lane];
      L.ebp[2][lane] = Real(0);
    }
    if (!C::NOE && !C::BP) {
      double ea = wave_sum(e_atr.value());
      double er = wave_sum(e_rep.value());
      double el = wave_sum(e_lk.value());
      if (lane == 0) {
        unsafeAtomicAdd(&A.energies64[p], ea);
        unsafeAtomicAdd(&A.energies64[s.P + p], er);
        unsafeAtomicAdd(&A.energies64[2 * s.P + p], el);
      }
    }
    wave_sync();
  }
}

}  // namespace tmol_ljlk
