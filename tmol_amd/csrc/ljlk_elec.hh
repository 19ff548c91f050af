// Elec (sigmoidal-dielectric Coulomb) on the LJ+LK scheduler: the same
// neighbor pass and row-major work list (built with elec's max_dis), the
// same persistent row queue, and a row kernel with the structure of
// score_rows_kernel -- g1 staged once per row, g2 blocks through two ring
// slots, sphere cull, phase A compaction of atom pairs within max_dis and of
// non-zero count-pair weight into a wave FIFO, phase B evaluation of full
// waves, 64-bit fixed-point LDS gradients flushed once per atom.
//
// Reference: tmol/score/elec/potentials/elec_pose_score.impl.hh:383-856
// (forward), :858-1125 (block-pair backward), potentials.hh:31-201 (the
// potential), elec.hh:406-548 (atom pair + gradient), elec_energy_term.py
// (count pair through representative atoms).
//
// Count pair: inter-block separations use path distances from connection
// atoms to each atom's representative (packed per (type, atom) as one byte
// per connection, capped at 6 -- values past 4 all weigh 1); intra-block
// pairs come from a prepared table of pairs a1 < a2 whose representative
// separation is >= 4 and whose charges are both non-zero.  Atoms with zero
// charge are dropped at the cull: the reference returns exactly 0 for them.
#pragma once

#include "ljlk_common.hh"
#include "ljlk_score_rows.hh"

namespace tmol_ljlk {

// one (block type, atom) of the elec tables, held by a lane between its
// global load and its LDS staging
template <typename Real>
struct ElecAtom {
  Real x, y, z, q;
  uint32_t pd;  // byte c: representative path distance from connection atom c
};

// elec census slots (counts of atom pairs; FLOP accounting)
enum : int {
  EC_ITEMS = 0, EC_INTRA_ITEMS, EC_PAIRS, EC_CHARGED, EC_CP0,
  EC_FLAT, EC_LOW, EC_COULOMB, EC_HIGH, EC_BEYOND, EC_N
};

template <typename Real>
struct ElecArgs {
  StackView s;
  BTView bt;  // n_atoms, n_conn, conn_atoms (is_ligand: elec's table)
  WorkView w;
  const Real* coords;
  const Real* spheres;
  const Real* charge;           // [T, Amax]
  const uint32_t* epd;          // [T, Amax] packed representative path distances
  const int32_t* intra_count;   // [T]
  const uint16_t* intra_pairs;  // [T, INTRA_MAX]
  ElecConst<Real> k;
  uint32_t* row_cursor;
  const Real* weights;   // [P] or [P, B, B] / [G, row_cap] (block pair)
  double* energies64;    // [P]
  Real* energies_bp;     // [P, B, B] / [G, row_cap]
  Real* dvdx;            // [NA, 3]
  Real* grad;            // [NA, 3]
  int64_t n_atoms;
  unsigned long long* census;  // [16] with M_CENSUS
};

// Per block type: packed inter-block path distances and the intra-block
// table (one wave per type, ballot-compacted in (a1, a2) order).
template <typename Real>
__global__ __launch_bounds__(64) void elec_prepare_kernel(BTView bt, const Real* charge,
                                                          const int32_t* inter_rpd,
                                                          const int32_t* intra_rpd,
                                                          uint32_t* epd, int32_t* count,
                                                          uint16_t* pairs) {
  const int t = blockIdx.x;
  const int lane = lane_id();
  const int n = bt.n_atoms[t];
  const int nc = bt.n_conn[t];
  const int64_t A = bt.Amax;
  for (int a = lane; a < bt.Amax; a += WAVE) {
    uint32_t pk = 0;
    for (int c = 0; c < MAX_CONN; ++c) {
      int v = 6;
      if (a < n && c < nc) {
        const int ca = bt.conn_atoms[t * bt.C + c];
        if (ca >= 0 && ca < bt.Amax) {
          v = inter_rpd[(int64_t(t) * A + ca) * A + a];
          v = v < 0 ? 0 : (v > 6 ? 6 : v);
        }
      }
      pk |= uint32_t(v) << (8 * c);
    }
    epd[int64_t(t) * A + a] = pk;
  }
  int k = 0;
  for (int a1 = 0; a1 < n; ++a1) {
    const int a2 = a1 + 1 + lane;
    bool keep = false;
    int sep = 0;
    if (a2 < n) {
      sep = intra_rpd[(int64_t(t) * A + a1) * A + a2];
      keep = sep >= 4 && charge[int64_t(t) * A + a1] != Real(0) &&
             charge[int64_t(t) * A + a2] != Real(0);
    }
    const uint64_t m = __ballot(keep);
    if (keep) pairs[int64_t(t) * INTRA_MAX + k + lanes_below(m)] =
        (uint16_t)(a1 | (a2 << 5) | ((sep == 4) << 10));
    k += __popcll(m);
  }
  if (lane == 0) count[t] = k;
}

template <typename Real, int MODE>
struct ElecLDS {
  using C = RowCtx<Real, MODE>;
  using T = typename LdsAcc<Real>::T;
  T grad[3][C::DERIV ? RSLOTS : 1];
  T ebp[3];  // block-pair energies: ring 0, ring 1, intra
  alignas(4 * sizeof(Real)) Real pos[RSLOTS][4];  // x, y, z, charge
  uint16_t fifo[LIST];
  alignas(4 * sizeof(Real)) Real c1[TILE][4];  // culled atoms of g1: x, y, z, index
  alignas(4 * sizeof(Real)) Real c2[TILE][4];  // of g2
  uint32_t pdw[3][TILE];
  uint8_t cs[2][MAX_CONN * MAX_CONN];
  Real wt[3];
};

// per-lane census counters (M_CENSUS): evaluated pairs by potential region
struct ElecCensus {
  uint32_t flat = 0, low = 0, coulomb = 0, high = 0, beyond = 0;
};

template <typename Real, int MODE>
__device__ __forceinline__ void eval_elec(ElecLDS<Real, MODE>& L, int consume,
                                          const ElecConst<Real>& k, Real wrow, LaneSum& e_sum,
                                          ElecCensus& cen, int lane) {
  using C = RowCtx<Real, MODE>;
  using Acc = LdsAcc<Real>;
  using V = typename V4<Real>::T;
  for (int i0 = 0; i0 < consume; i0 += WAVE) {
    const int i = i0 + lane;
    if (i < consume) {
      const int e = L.fifo[i];
      const int s1 = e & 31, s2 = (e >> 5) & 127;
      const Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
      const V p1 = *(const V*)L.pos[s1], p2 = *(const V*)L.pos[s2];
      const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
      const Real d2 = dx * dx + dy * dy + dz * dz;
      const Real inv_d = fast_rsqrt(d2);
      const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);
      if (MODE & M_CENSUS) {
        if (dist >= k.hpe)
          ++cen.beyond;
        else if (dist < k.lps)
          ++cen.flat;
        else if (dist < k.lpe)
          ++cen.low;
        else if (dist < k.hps)
          ++cen.coulomb;
        else
          ++cen.high;
        continue;
      }
      Real v, dv;
      elec_unit<Real, C::DERIV>(dist, k, v, dv);
      const Real qq = p1.w * p2.w;
      const Real E = wcp * (qq * v);
      if (!C::NOE) {
        if (C::BP)
          atomicAdd(&L.ebp[entry_owner(s2)], Acc::enc(E));
        else
          e_sum.add(E);
      }
      if (C::DERIV && dist > Real(0)) {
        const Real dE = wcp * (qq * dv);
        if (dE != Real(0)) {
          Real gs = dE * inv_d;
          if (C::WGRAD) gs *= C::BP ? L.wt[entry_owner(s2)] : wrow;
          const typename Acc::T ex = Acc::enc(gs * dx), ey = Acc::enc(gs * dy),
                                ez = Acc::enc(gs * dz);
          atomicAdd(&L.grad[0][s1], ex);
          atomicAdd(&L.grad[1][s1], ey);
          atomicAdd(&L.grad[2][s1], ez);
          atomicAdd(&L.grad[0][s2], typename Acc::T(0) - ex);
          atomicAdd(&L.grad[1][s2], typename Acc::T(0) - ey);
          atomicAdd(&L.grad[2][s2], typename Acc::T(0) - ez);
        }
      }
    }
  }
}

// flush slots [slot0, slot0 + n) to atoms [atom0, atom0 + n): x, y, z of 21
// atoms per instruction as one contiguous run of the [NA, 3] output
template <typename Real, int MODE>
__device__ __forceinline__ void elec_flush(ElecLDS<Real, MODE>& L, const ElecArgs<Real>& A,
                                           int slot0, int n, int64_t atom0, int lane) {
  using C = RowCtx<Real, MODE>;
  if (!C::DERIV) return;
  Real* out = C::DVDX ? A.dvdx : A.grad;
  for (int i = lane; i < 3 * n; i += WAVE) {
    const int a = i / 3, c = i - 3 * a;
    const Real v = LdsAcc<Real>::dec(L.grad[c][slot0 + a]);
    L.grad[c][slot0 + a] = 0;
    if (v != Real(0)) unsafeAtomicAdd(&out[atom0 * 3 + i], v);
  }
}

template <typename Real>
__device__ __forceinline__ void load_elec_atom(ElecAtom<Real>& R, const ElecArgs<Real>& A,
                                               int t, int o, int a) {
  const Real* xyz = A.coords + (int64_t(o) + a) * 3;
  R.x = xyz[0];
  R.y = xyz[1];
  R.z = xyz[2];
  const int64_t ta = int64_t(t) * A.bt.Amax + a;
  R.q = A.charge[ta];
  R.pd = A.epd[ta];
}

template <typename Real, int MODE>
__device__ __forceinline__ void stage_elec_atom(ElecLDS<Real, MODE>& L, int slot, int pd_row,
                                                int a, const ElecAtom<Real>& R) {
  using V = typename V4<Real>::T;
  V p;
  p.x = R.x;
  p.y = R.y;
  p.z = R.z;
  p.w = R.q;
  *(V*)L.pos[slot] = p;
  L.pdw[pd_row][a] = R.pd;
}

// Float WGRAD (the weighted energy + gradient step) runs at 5 waves per
// SIMD with a few spilled registers: 2.15 against 2.33 ms at 4 on the
// 512-pose batch; the other modes keep 4.
template <typename Real, int MODE>
__global__ __launch_bounds__(256, (sizeof(Real) == 4 && MODE == M_WGRAD) ? 5 : 4) void
elec_rows_kernel(ElecArgs<Real> A) {
  using C = RowCtx<Real, MODE>;
  using V = typename V4<Real>::T;
  __shared__ ElecLDS<Real, MODE> lds_all[WAVES_PER_WG];
  const int lane = lane_id();
  ElecLDS<Real, MODE>& L = lds_all[uni(threadIdx.x / WAVE)];
  const ElecConst<Real> k = A.k;
  const StackView& s = A.s;
  const BTView& bt = A.bt;
  RowQueue queue(A.row_cursor, s.G);
  const Real cull_pad = k.hpe + Real(1e-3);
  constexpr bool CEN = MODE & M_CENSUS;
  ElecCensus cen;
  unsigned long long c_items = 0, c_intra = 0, c_pairs = 0, c_charged = 0, c_cp0 = 0;

  if (C::DERIV)
    for (int i = lane; i < 3 * RSLOTS; i += WAVE) (&L.grad[0][0])[i] = 0;
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
    const Real wrow = (C::WGRAD && !C::BP) ? A.weights[p] : Real(0);
    bool q1 = false;
    if (lane < n1) {
      ElecAtom<Real> R;
      load_elec_atom(R, A, t1, o1, lane);
      stage_elec_atom(L, lane, 0, lane, R);
      q1 = R.q != Real(0);
    }
    const int nq1 = __popcll(__ballot(q1));
    if (C::BP && lane < 3) L.ebp[lane] = 0;
    LaneSum e_sum;
    int n_f = 0, seq = 0, done = 0, end0 = 0, end1 = 0;
    bool used0 = false, used1 = false;
    int o2s0 = 0, o2s1 = 0, n2s0 = 0, n2s1 = 0;
    int64_t bps0 = 0, bps1 = 0, bp_intra = -1;
    const int64_t PBB = C::BP ? bp_plane(s, A.w) : 0;
    int ring = 0;

    auto drain = [&](bool all) {
      const int c = all ? n_f : (n_f & ~(WAVE - 1));
      if (c > 0) eval_elec<Real, MODE>(L, c, k, wrow, e_sum, cen, lane);
      wave_sync();
      fifo_shift(L.fifo, c, n_f, lane);
      n_f -= c;
      done += c;
    };
    auto retire = [&](int r) {
      elec_flush<Real, MODE>(L, A, TILE + TILE * r, r ? n2s1 : n2s0, r ? o2s1 : o2s0, lane);
      if (C::BP && !C::NOE && lane == 0) {
        A.energies_bp[r ? bps1 : bps0] = LdsAcc<Real>::dec(L.ebp[r]);
        L.ebp[r] = 0;
      }
      if (r)
        used1 = false;
      else
        used0 = false;
    };

    for (int c0 = 0; c0 < n_items; c0 += WAVE) {
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
      ElecAtom<Real> pf;
      pf.x = pf.y = pf.z = pf.q = Real(0);
      pf.pd = 0;
      {
        const int t = lane_val(d_t2, 0), o = lane_val(d_o2, 0), n = lane_val(d_pk, 0) & 255;
        if (lane < n) load_elec_atom(pf, A, t, o, lane);
      }
      for (int j = 0; j < n_chunk; ++j) {
        const int g2 = lane_val(d_g2, j);
        const int pk = lane_val(d_pk, j);
        const int min_sep = pk >> 16;
        const ElecAtom<Real> cur = pf;
        if (j + 1 < n_chunk) {
          const int t = lane_val(d_t2, j + 1), o = lane_val(d_o2, j + 1);
          const int n = lane_val(d_pk, j + 1) & 255;
          if (lane < n) load_elec_atom(pf, A, t, o, lane);
        }
        if (g2 == g1) {
          // ---- intra-block pairs from the prepared table
          if (C::BP) bp_intra = bp_index(s, A.w, g1, g1, c0 + j);
          if (C::BP && C::WGRAD && lane == 0) L.wt[2] = A.weights[bp_intra];
          wave_sync();
          const int n_ip = uni(A.intra_count[t1]);
          const uint16_t* ip = A.intra_pairs + int64_t(t1) * INTRA_MAX;
          if (CEN) {
            c_intra += 1;
            c_pairs += uint64_t(n1) * (n1 - 1) / 2;
            c_charged += uint64_t(nq1) * (nq1 - 1) / 2;
            c_cp0 += uint64_t(nq1) * (nq1 - 1) / 2 - n_ip;
          }
          for (int q0 = 0; q0 < n_ip; q0 += WAVE) {
            if (n_f > LIST - WAVE) drain(false);
            bool in = false, w02 = false;
            int a1 = 0, a2 = 0;
            if (q0 + lane < n_ip) {
              const int e = ip[q0 + lane];
              a1 = e & 31;
              a2 = (e >> 5) & 31;
              const V pa = *(const V*)L.pos[a1], pb = *(const V*)L.pos[a2];
              const Real dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
              in = dx * dx + dy * dy + dz * dz < k.cull2;
              w02 = (e >> 10) & 1;
            }
            const uint64_t m = __ballot(in);
            if (in) L.fifo[n_f + lanes_below(m)] = make_entry(a1, a2, w02);
            n_f += __popcll(m);
            seq += __popcll(m);
            wave_sync();
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
          // cull: charged g2 atoms within reach of g1's sphere (lanes 0..31),
          // charged g1 atoms within reach of g2's sphere (lanes 32..63)
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
            near = cur.q != Real(0) && ex * ex + ey * ey + ez * ez < reach * reach;
          } else if (lane >= TILE && lane - TILE < n1) {
            const V P1 = *(const V*)L.pos[lane - TILE];
            mine.x = P1.x;
            mine.y = P1.y;
            mine.z = P1.z;
            mine.w = Real(lane - TILE);
            const Real ex = P1.x - c2x, ey = P1.y - c2y, ez = P1.z - c2z;
            const Real reach = c2r + cull_pad;
            near = P1.w != Real(0) && ex * ex + ey * ey + ez * ez < reach * reach;
          }
          const uint64_t mnear = __ballot(near);
          const uint32_t m2 = (uint32_t)mnear, m1 = (uint32_t)(mnear >> 32);
          const int n1c = __popc(m1), n2c = __popc(m2);
          if (CEN) {
            const int nq2 = __popcll(__ballot(lane < n2 && cur.q != Real(0)));
            c_items += 1;
            c_pairs += uint64_t(n1) * n2;
            c_charged += uint64_t(nq1) * nq2;
          }
          if (n1c == 0 || n2c == 0) {
            ring = r;
            if (n_f >= WAVE) drain(false);
            continue;
          }
          if (r ? used1 : used0) {
            if ((r ? end1 : end0) > done) drain(true);
            retire(r);
          }
          if (need_cp && lane >= TILE && lane < TILE + MAX_CONN * MAX_CONN) {
            const int qq = lane - TILE;
            const int cc1 = qq / MAX_CONN, cc2 = qq % MAX_CONN;
            L.cs[r][qq] = (cc1 < nc1 && cc2 < nc2) ? (uint8_t)conn_sep(s, g1, g2, cc1, cc2) : 6;
          }
          if (C::BP && C::WGRAD && lane == 0) L.wt[r] = A.weights[bpi];
          if (lane < n2) stage_elec_atom(L, sb + lane, 1 + r, lane, cur);
          if (near) {
            if (lane < TILE)
              *(V*)L.c2[__builtin_amdgcn_mbcnt_lo(m2, 0)] = mine;
            else
              *(V*)L.c1[__builtin_amdgcn_mbcnt_hi(m1, 0)] = mine;
          }
          if (r) {
            used1 = true; o2s1 = o2; n2s1 = n2; bps1 = bpi;
          } else {
            used0 = true; o2s0 = o2; n2s0 = n2; bps0 = bpi;
          }
          wave_sync();
          auto test_pair = [&](int t, int u, bool& in, bool& w02, int& a1, int& a2) {
            in = w02 = false;
            a1 = a2 = 0;
            bool cp0 = false;
            if (t < n1c) {
              const V P1 = *(const V*)L.c1[t], P2 = *(const V*)L.c2[u];
              a1 = int(P1.w);
              a2 = int(P2.w);
              const Real dx = P1.x - P2.x, dy = P1.y - P2.y, dz = P1.z - P2.z;
              const int sep = need_cp ? packed_pair_sep(L.pdw[0][a1], L.pdw[1 + r][a2], L.cs[r],
                                                        min_sep, nc1, nc2, lig)
                                      : min_sep;
              in = sep >= 4 && dx * dx + dy * dy + dz * dz < k.cull2;
              w02 = sep == 4;
              cp0 = sep < 4;
            }
            if (CEN) c_cp0 += __popcll(__ballot(cp0));
          };
          auto append = [&](bool in, bool w02, int a1, int a2) {
            const uint64_t m = __ballot(in);
            if (in) L.fifo[n_f + lanes_below(m)] = make_entry(a1, sb + a2, w02);
            n_f += __popcll(m);
            seq += __popcll(m);
          };
          const int total = n1c * n2c;
          const int q = small_div(WAVE, n2c), rstep = WAVE - q * n2c;
          int t = small_div(lane, n2c), u = lane - t * n2c;
          for (int k0 = 0; k0 < total; k0 += 2 * WAVE) {
            if (n_f > LIST - 2 * WAVE) drain(false);
            int uB = u + rstep, tB = t + q;
            if (uB >= n2c) {
              uB -= n2c;
              tB += 1;
            }
            bool inA, wA, inB, wB;
            int a1A, a2A, a1B, a2B;
            test_pair(t, u, inA, wA, a1A, a2A);
            test_pair(tB, uB, inB, wB, a1B, a2B);
            append(inA, wA, a1A, a2A);
            append(inB, wB, a1B, a2B);
            u = uB + rstep;
            t = tB + q;
            if (u >= n2c) {
              u -= n2c;
              t += 1;
            }
            wave_sync();
          }
          if (r)
            end1 = seq;
          else
            end0 = seq;
        }
        if (n_f >= WAVE) drain(false);
      }
    }
    drain(true);
    if (used0) retire(0);
    if (used1) retire(1);
    elec_flush<Real, MODE>(L, A, 0, n1, o1, lane);
    if (C::BP && !C::NOE && lane == 0 && bp_intra >= 0) {
      A.energies_bp[bp_intra] = LdsAcc<Real>::dec(L.ebp[2]);
      L.ebp[2] = 0;
    }
    if (!C::NOE && !C::BP && !CEN) {
      const double e = wave_sum(e_sum.value());
      if (lane == 0) unsafeAtomicAdd(&A.energies64[p], e);
    }
    wave_sync();
  }
  if (CEN) {
    const unsigned long long f = wave_sum((unsigned long long)cen.flat);
    const unsigned long long lo = wave_sum((unsigned long long)cen.low);
    const unsigned long long co = wave_sum((unsigned long long)cen.coulomb);
    const unsigned long long hi = wave_sum((unsigned long long)cen.high);
    const unsigned long long be = wave_sum((unsigned long long)cen.beyond);
    if (lane == 0) {
      atomicAdd(&A.census[EC_ITEMS], c_items);
      atomicAdd(&A.census[EC_INTRA_ITEMS], c_intra);
      atomicAdd(&A.census[EC_PAIRS], c_pairs);
      atomicAdd(&A.census[EC_CHARGED], c_charged);
      atomicAdd(&A.census[EC_CP0], c_cp0);
      atomicAdd(&A.census[EC_FLAT], f);
      atomicAdd(&A.census[EC_LOW], lo);
      atomicAdd(&A.census[EC_COULOMB], co);
      atomicAdd(&A.census[EC_HIGH], hi);
      atomicAdd(&A.census[EC_BEYOND], be);
    }
  }
}

}  // namespace tmol_ljlk
