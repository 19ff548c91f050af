// Streaming scoring kernel (blocks of at most 32 atoms; whole-pose energies
// and the weighted gradient): one wave walks one row of the work list -- block
// g1 against its neighbor blocks g2 > g1 -- but treats the row's g2 atoms as
// one stream instead of block pair by block pair.
//
//   descriptors  64 neighbor blocks at a time, one per lane: type, atom
//                offset, atom count, minimum bond separation and, for blocks
//                within 4 bonds, the connection separations; an exclusive scan
//                of the atom counts places each block's atoms in the stream
//   raw chunk    64 consecutive stream atoms, one per lane: the lane finds its
//                block from per-chunk start markers and a ballot, loads the
//                atom and culls it against g1's sphere (an atom farther than
//                r1 + max_dis from g1's centre reaches no atom of g1); the
//                survivors are staged into the current test chunk
//   test chunk   32 survivors held in LDS ring slots (three chunks of 32 live
//                at a time); lanes 0..31 and 32..63 hold the same 32 atoms
//                and test them against g1 atoms t and t + 1: squared
//                distance and, only in chunks holding atoms within 4 bonds
//                of g1, the count-pair separation from precombined path
//                bytes.  In-range pairs of non-zero weight are
//                ballot-compacted into the LJ (and LK) FIFOs
//   phase B      64 FIFO entries at a time, one per lane (the evaluators of
//                the row kernel); gradients accumulate in LDS as 64-bit fixed
//                point and a ring chunk is flushed to global memory when its
//                slots are reused, g1 at row end
//
// Per block pair the row kernel pays staging, a cull of both blocks and
// bookkeeping; here that work is per 64 atoms of the stream, and the pair
// tests read g2 atoms from registers and g1 atoms by LDS broadcast.  The
// separation of a pair is min(5, min_c1 pd1[c1] + q[c1]) with
// q[c1] = min_c2 cs[c1][c2] + pd2[c2] combined once per g2 atom (absent
// connections carry path distance 6, so their sums never reach below 5), the
// count-pair rule of impl.hh:47-67 / count_pair.hh:162-207.
#pragma once

#include "ljlk_score_rows.hh"

namespace tmol_ljlk {

constexpr int SCH = 32;                // atoms per test chunk
constexpr int SRING = 3 * SCH;         // ring slots
constexpr int SSLOTS = TILE + SRING;   // g1 slots 0..31, ring 32..127
constexpr int SLIST = 128;             // FIFO capacity (drained at 64)
constexpr uint32_t Q_NONE = 0x64646464u;  // q bytes of blocks more than 4 bonds apart (100)
constexpr uint32_t Q_HEAVY = 0x80000000u; // bit 7 of byte 3: heavy atom
constexpr uint32_t Q_LIG = 0x00800000u;   // bit 7 of byte 2: ligand-fragment block

// Copies of the LDS gradient accumulators, chosen by lane: LDS atomics on one
// address serialize, and consecutive FIFO entries share their g1 atom (and,
// less often, their ring atom).  g1 slots: G1COPIES copies (lane % G1COPIES);
// ring slots: RINGCOPIES copies ((lane / 2) % RINGCOPIES).  Flushes add the
// copies.
#ifndef TMOL_STREAM_G1COPIES
#define TMOL_STREAM_G1COPIES 2
#endif
#ifndef TMOL_STREAM_RINGCOPIES
#define TMOL_STREAM_RINGCOPIES 1
#endif
constexpr int G1COPIES = TMOL_STREAM_G1COPIES;
constexpr int RINGCOPIES = TMOL_STREAM_RINGCOPIES;

template <typename Real, int MODE>
struct StreamLDS {
  using C = RowCtx<Real, MODE>;
  using T = typename LdsAcc<Real>::T;
  static constexpr int NGP = C::NG > 0 ? C::NG : 1;
  T gring[RINGCOPIES][NGP][3][C::DERIV ? SRING : 1];  // ring slots TILE.. (index - TILE)
  T g1acc[G1COPIES][NGP][3][C::DERIV ? TILE : 1];     // g1 slots 0..31
  T ebp[1][3];  // block-pair modes run the row kernel; kept for the evaluators
  Real wt[1][3];
  alignas(4 * sizeof(Real)) Real pos[SSLOTS][4];  // x, y, z, lj_radius
  alignas(4 * sizeof(Real)) Real prm[SSLOTS][4];  // sqrt(lj_wdepth), K, 1/lambda^2, lk_volume
  alignas(4 * sizeof(Real)) Real tst[SCH][4];  // test chunk atoms: x, y, z
  uint32_t tq[SCH];      // test chunk atoms: connection bytes q | Q_HEAVY | Q_LIG
  uint8_t flg[SSLOTS];
  uint16_t lj[SLIST];
  uint16_t lk[SLIST];
  int32_t atom[SRING];   // global atom of each ring slot
  uint32_t pd1[TILE];    // g1 atoms' connection path distances, byte per connection
  uint8_t mark[WAVE];    // raw chunk: a block's first atom sits at this lane
};

// smallest of the four 7-bit byte fields of v, capped at 5 (the count-pair
// separation from packed sums)
__device__ __forceinline__ int min_sep_bytes(uint32_t v) {
  const int b0 = v & 127, b1 = (v >> 8) & 127, b2 = (v >> 16) & 127, b3 = (v >> 24) & 127;
  const int m = min(min(b0, b1), min(b2, b3));
  return m < 5 ? m : 5;
}

// q[c1] = min_c2 (cs[c1][c2] + pd2[c2]) as bytes, cs row c1 packed (byte c2)
__device__ __forceinline__ uint32_t combine_q(const uint32_t cs[MAX_CONN], uint32_t pd2) {
  uint32_t q = 0;
#pragma unroll
  for (int c1 = 0; c1 < MAX_CONN; ++c1) {
    const uint32_t s = cs[c1] + pd2;  // bytes <= 12: no carries
    const uint32_t b0 = s & 255, b1 = (s >> 8) & 255, b2 = (s >> 16) & 255, b3 = s >> 24;
    const uint32_t m = min(min(b0, b1), min(b2, b3));
    q |= m << (8 * c1);
  }
  return q;
}

// packed connection separations cs[c1] (byte c2) of the pair (g1, g2), g1 < g2
__device__ __forceinline__ void conn_rows(const StackView& s, int g1, int g2, int nc1, int nc2,
                                          uint32_t cs[MAX_CONN]) {
#pragma unroll
  for (int c1 = 0; c1 < MAX_CONN; ++c1) cs[c1] = 0x06060606u;
  if (s.layout == TMOL_LJLK_PADDED) {
    for (int c1 = 0; c1 < nc1; ++c1) {
      uint32_t r = 0x06060606u;
      for (int c2 = 0; c2 < nc2; ++c2) {
        const uint32_t v = (uint32_t)min(conn_sep(s, g1, g2, c1, c2), 6);
        r = (r & ~(255u << (8 * c2))) | (v << (8 * c2));
      }
      cs[c1] = r;
    }
    return;
  }
  const int32_t* e = bonded_row(s, g1, g2);
  if (!e) return;
  for (int c1 = 0; c1 < nc1; ++c1) cs[c1] = (uint32_t)e[4 + c1];
}

// gradient of an entry's pair: + to the g1 atom's slot s1, - to slot s2
// (a ring slot, or a g1 slot for intra-block pairs), into the lane's copies
template <typename Real, int MODE>
__device__ __forceinline__ void grad_scatter(StreamLDS<Real, MODE>& L, int term, int s1, int s2,
                                             Real gx, Real gy, Real gz, int lane) {
  using Acc = LdsAcc<Real>;
  using T = typename Acc::T;
  const int c1 = lane & (G1COPIES - 1);
  const T ex = Acc::enc(gx), ey = Acc::enc(gy), ez = Acc::enc(gz);
  T* b1 = &L.g1acc[c1][term][0][s1];
  const bool g2_in_g1 = s2 < TILE;
  T* b2 = g2_in_g1 ? &L.g1acc[c1][term][0][s2]
                   : &L.gring[(lane >> 1) & (RINGCOPIES - 1)][term][0][s2 - TILE];
  const int k2 = g2_in_g1 ? TILE : SRING;
  atomicAdd(b1, ex);
  atomicAdd(b1 + TILE, ey);
  atomicAdd(b1 + 2 * TILE, ez);
  atomicAdd(b2, T(0) - ex);
  atomicAdd(b2 + k2, T(0) - ey);
  atomicAdd(b2 + 2 * k2, T(0) - ez);
}

// flush the gradients of ring chunk r (n atoms, slots TILE + SCH r ...) to
// their atoms and zero them; lane l takes component l % 3 of atom l / 3
template <typename Real, int MODE>
__device__ __forceinline__ void flush_chunk(StreamLDS<Real, MODE>& L, const ScoreArgs<Real>& A,
                                            int r, int n, int lane) {
  using C = RowCtx<Real, MODE>;
  if (!C::DERIV) return;
  for (int i = lane; i < 3 * n; i += WAVE) {
    const int a = i / 3, k = i - 3 * a;
    const int slot = SCH * r + a;
    const int64_t atom = L.atom[SCH * r + a];
#pragma unroll
    for (int t = 0; t < C::NG; ++t) {
      typename LdsAcc<Real>::T acc = 0;
#pragma unroll
      for (int c = 0; c < RINGCOPIES; ++c) {
        acc += L.gring[c][t][k][slot];
        L.gring[c][t][k][slot] = 0;
      }
      if (C::DET) {
        if (acc != 0) atomicAdd(&A.fixed[(int64_t(t) * A.n_atoms + atom) * 3 + k], acc);
        continue;
      }
      const Real v = LdsAcc<Real>::dec(acc);
      if (v != Real(0)) {
        if (C::DVDX)
          unsafeAtomicAdd(&A.dvdx[(int64_t(t) * A.n_atoms + atom) * 3 + k], v);
        else
          unsafeAtomicAdd(&A.grad[atom * 3 + k], v);
      }
    }
  }
}

// g1's atoms [0, n) (slots 0..n-1, contiguous atoms from atom0)
template <typename Real, int MODE>
__device__ __forceinline__ void flush_g1(StreamLDS<Real, MODE>& L, const ScoreArgs<Real>& A, int n,
                                         int64_t atom0, int lane) {
  using C = RowCtx<Real, MODE>;
  if (!C::DERIV) return;
  for (int i = lane; i < 3 * n; i += WAVE) {
    const int a = i / 3, k = i - 3 * a;
#pragma unroll
    for (int t = 0; t < C::NG; ++t) {
      typename LdsAcc<Real>::T acc = 0;
#pragma unroll
      for (int c = 0; c < G1COPIES; ++c) {
        acc += L.g1acc[c][t][k][a];
        L.g1acc[c][t][k][a] = 0;
      }
      if (C::DET) {
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

template <int MODE, int OCC>
constexpr int stream_waves_per_simd() {
  return OCC > 0 ? OCC : 4;
}

template <typename Real, int MODE, int OCC = 0>
__global__ __launch_bounds__(256, (stream_waves_per_simd<MODE, OCC>())) void score_stream_kernel(
    ScoreArgs<Real> A) {
  using C = RowCtx<Real, MODE>;
  using V = typename V4<Real>::T;
  static_assert(!C::BP && !C::ELEC, "block-pair and elec modes run the row kernel");
  __shared__ StreamLDS<Real, MODE> lds_all[WAVES_PER_WG];
  const int lane = lane_id();
  StreamLDS<Real, MODE>& L = lds_all[uni(threadIdx.x / WAVE)];
  const GlobalParams<Real> gp = load_global_params(A.global_params);
  const StackView& s = A.s;
  const BTView& bt = A.bt;
  RowQueue queue(A.row_cursor, s.G, A.row_grab);
  const Real cull_pad = gp.max_dis + Real(1e-3);
  const int u = lane & (SCH - 1), half = lane >> 5;

  if (C::DERIV) {
    auto* g = &L.gring[0][0][0][0];
    for (int i = lane; i < RINGCOPIES * L.NGP * 3 * SRING; i += WAVE) g[i] = 0;
    auto* gb = &L.g1acc[0][0][0][0];
    for (int i = lane; i < G1COPIES * L.NGP * 3 * TILE; i += WAVE) gb[i] = 0;
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
    if (C::WGRAD) {
      wrow[0] = A.weights[p];
      wrow[1] = A.weights[s.P + p];
      wrow[2] = A.weights[2 * s.P + p];
    }
    const Real reach1 = sph1.w + cull_pad;
    const Real reach1_2 = reach1 * reach1;
    // ---- stage g1 into slots 0..31
    bool heavy1 = false;
    if (lane < n1) {
      AtomRegs<Real> R;
      load_atom<Real, false>(R, A, t1, o1, lane);
      V pp, qq;
      pp.x = R.x; pp.y = R.y; pp.z = R.z; pp.w = R.r0.x;
      qq.x = R.r0.y; qq.y = R.r0.z; qq.z = R.r0.w; qq.w = R.vol;
      *(V*)L.pos[lane] = pp;
      *(V*)L.prm[lane] = qq;
      L.flg[lane] = (uint8_t)R.flags;
      L.pd1[lane] = R.pd;
      heavy1 = (R.flags & F_HEAVY) != 0;
    }
    const uint32_t hmask1 = (uint32_t)__ballot(heavy1);
    LaneSum e_atr, e_rep, e_lk;
    int n_lj = 0, n_lk = 0, seq_lj = 0, seq_lk = 0, done_lj = 0, done_lk = 0;
    int end_lj[3] = {0, 0, 0}, end_lk[3] = {0, 0, 0};
    int cnt[3] = {0, 0, 0};  // atoms held by each ring chunk (0: free)
    int cur = 0;             // ring chunk being filled
    int fill = 0;            // its atoms so far
    bool cur_cp = false, cur_lig = false;  // it holds atoms within 4 bonds / ligand atoms
    wave_sync();

    auto drain = [&](bool all) {
      const int c1 = all ? n_lj : (n_lj & ~(WAVE - 1));
      const int c2 = all ? n_lk : (n_lk & ~(WAVE - 1));
      if (c1 > 0) eval_lj<Real, MODE>(L, c1, gp, wrow, e_atr, e_rep, lane);
      if (c2 > 0) eval_lk<Real, MODE>(L, c2, gp, wrow, e_lk, lane);
      wave_sync();
      fifo_shift(L.lj, c1, n_lj, lane);
      fifo_shift(L.lk, c2, n_lk, lane);
      n_lj -= c1;
      done_lj += c1;
      n_lk -= c2;
      done_lk += c2;
    };
    auto drain_full = [&]() {
      if (n_lj >= WAVE || n_lk >= WAVE) drain(false);
    };
    auto append = [&](bool in, bool hv, int e) {
      const uint64_t m = __ballot(in);
      if (in) L.lj[n_lj + lanes_below(m)] = (uint16_t)e;
      n_lj += __popcll(m);
      seq_lj += __popcll(m);
      const uint64_t mh = __ballot(hv);
      if (hv) L.lk[n_lk + lanes_below(mh)] = (uint16_t)e;
      n_lk += __popcll(mh);
      seq_lk += __popcll(mh);
    };
    // ring chunk r no longer referenced: evaluate what still points at it,
    // flush its atoms
    auto retire = [&](int r) {
      if (cnt[r] == 0) return;
      if (end_lj[r] > done_lj || end_lk[r] > done_lk) drain(true);
      flush_chunk<Real, MODE>(L, A, r, cnt[r], lane);
      cnt[r] = 0;
    };
    // phase A of the current chunk (fill atoms), then the next ring chunk
    auto test_chunk = [&]() {
      const int r = cur;
      const int n = fill;
      const V tr = *(const V*)L.tst[u];
      const uint32_t qb = L.tq[u];
      const bool heavy2 = (qb & Q_HEAVY) != 0;
      const bool lig2 = (qb & Q_LIG) != 0;
      const bool ok_u = u < n;
      const int ebase = (TILE + SCH * r + u) << 5;
      for (int t = 0; t < n1; t += 2) {
        const int ta = t + half;
        const V P1 = *(const V*)L.pos[ta < n1 ? ta : t];
        const Real dx = P1.x - tr.x, dy = P1.y - tr.y, dz = P1.z - tr.z;
        const Real d2 = dx * dx + dy * dy + dz * dz;
        bool in = ok_u && ta < n1 && d2 < gp.max_dis2;
        bool w02 = false;
        if (cur_cp) {
          int sep = min_sep_bytes(qb + L.pd1[ta < n1 ? ta : t]);
          if (cur_lig && lig1 && lig2 && (sep == 3 || sep == 4)) sep = 5;
          in = in && sep >= 4;
          w02 = sep == 4;
        }
        const bool hv = in && heavy2 && ((hmask1 >> ta) & 1u);
        append(in, hv, ebase | ta | (w02 ? (1 << 12) : 0));
        wave_sync();
        drain_full();
      }
      end_lj[r] = seq_lj;
      end_lk[r] = seq_lk;
      cnt[r] = n;
      cur = cur == 2 ? 0 : cur + 1;
      fill = 0;
      cur_cp = cur_lig = false;
      retire(cur);  // the chunk two before: its entries are consumed by now
    };

    int k0 = 0;
    if (uni(A.w.row_items[row * A.w.row_cap]) == g1) {
      // ---- intra-block pairs a1 < a2 of non-zero weight (prepared table)
      const int n_ip = uni(A.intra_count[t1]);
      const uint16_t* ip = A.intra_pairs + int64_t(t1) * INTRA_MAX;
      for (int j0 = 0; j0 < n_ip; j0 += WAVE) {
        bool in = false, hv = false, w02 = false;
        int a1 = 0, a2 = 0;
        if (j0 + lane < n_ip) {
          const int e = ip[j0 + lane];
          a1 = e & 31;
          a2 = (e >> 5) & 31;
          const V pa = *(const V*)L.pos[a1], pb = *(const V*)L.pos[a2];
          const Real dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
          in = dx * dx + dy * dy + dz * dz < gp.max_dis2;
          w02 = (e >> 10) & 1;
          hv = in && ((hmask1 >> a1) & (hmask1 >> a2) & 1u);
        }
        append(in, hv, a1 | (a2 << 5) | (w02 ? (1 << 12) : 0));
        wave_sync();
        drain_full();
      }
      k0 = 1;
    }

    for (int c0 = k0; c0 < n_items; c0 += WAVE) {
      // ---- descriptors: one neighbor block per lane
      const int nb = min(WAVE, n_items - c0);
      int d_t2 = 0, d_o2 = 0, d_n2 = 0, d_info = 0;
      uint32_t cs[MAX_CONN] = {Q_NONE, Q_NONE, Q_NONE, Q_NONE};
      if (lane < nb) {
        const int g2 = A.w.row_items[row * A.w.row_cap + c0 + lane];
        if (g2 != g1) {
          d_t2 = s.block_type[g2];
          d_o2 = s.block_atom_offset[g2];
          d_n2 = bt.n_atoms[d_t2];
          const int nc2 = bt.n_conn[d_t2];
          const int ms = min_block_sep(s, g1, g2, nc1, nc2);
          const bool lig2 = lig1 && bt.is_ligand[d_t2] != 0;
          d_info = (ms <= 4 ? 1 : 0) | (lig2 ? 2 : 0) | (d_n2 << 8);
          if (ms <= 4) conn_rows(s, g1, g2, nc1, nc2, cs);
        }
      }
      // exclusive scan of the atom counts: the stream position of each block
      int incl = d_n2;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += y;
      }
      const int d_pre = incl - d_n2;
      const int total = uni(__shfl(incl, WAVE - 1, WAVE));

      for (int q0 = 0; q0 < total; q0 += WAVE) {
        // ---- raw chunk: lane -> (block, atom) through start markers
        L.mark[lane] = 0;
        wave_sync();
        if (d_n2 > 0 && d_pre >= q0 && d_pre < q0 + WAVE) L.mark[d_pre - q0] = 1;
        wave_sync();
        const uint64_t M = __ballot(L.mark[lane] != 0);
        // blocks with atoms that start at or before q0, less one: the block
        // holding stream position q0 (in lane order over blocks with atoms)
        const int kfirst = __popcll(__ballot(d_n2 > 0 && d_pre <= q0)) - 1;
        const int rank = kfirst + lanes_below(M) + int((M >> lane) & 1ull) - int(M & 1ull);
        // every listed block but g1 itself (row position 0, skipped) has
        // atoms, so the rank-th block with atoms is descriptor lane `rank`
        const int k = rank;
        const int q = q0 + lane;
        const bool valid = q < total;  // (and a within its block, checked below)
        const int pre_k = __shfl(d_pre, k, WAVE);
        const int o2 = __shfl(d_o2, k, WAVE), t2 = __shfl(d_t2, k, WAVE);
        const int info = __shfl(d_info, k, WAVE);
        const int a = q - pre_k;
        AtomRegs<Real> R;
        bool near = false;
        if (valid && a >= 0 && a < (info >> 8)) {
          load_atom<Real, false>(R, A, t2, o2, a);
          const Real ex = R.x - sph1.x, ey = R.y - sph1.y, ez = R.z - sph1.z;
          near = ex * ex + ey * ey + ez * ez < reach1_2;
        }
        const uint64_t mnear = __ballot(near);
        const int n_surv = __popcll(mnear);
        if (n_surv == 0) continue;
        // combined connection bytes q (count pair) for atoms within 4 bonds
        uint32_t qv = Q_NONE;
        const uint64_t mcp = __ballot(near && (info & 1));
        if (mcp) {
          uint32_t csk[MAX_CONN];
#pragma unroll
          for (int c1 = 0; c1 < MAX_CONN; ++c1) csk[c1] = (uint32_t)__shfl((int)cs[c1], k, WAVE);
          if (info & 1) qv = combine_q(csk, R.pd);
        }
        qv |= (near && (R.flags & F_HEAVY)) ? Q_HEAVY : 0u;
        qv |= (info & 2) ? Q_LIG : 0u;
        const bool any_lig = __ballot(near && (info & 2)) != 0;
        const int rk = lanes_below(mnear);
        // stage the survivors into test chunks of 32
        int done = 0;
        while (done < n_surv) {
          const int take = min(n_surv - done, SCH - fill);
          const bool mine = near && rk >= done && rk < done + take;
          if (mine) {
            const int pos = fill + rk - done;
            const int slot = TILE + SCH * cur + pos;
            V pp, pq;
            pp.x = R.x; pp.y = R.y; pp.z = R.z; pp.w = R.r0.x;
            pq.x = R.r0.y; pq.y = R.r0.z; pq.z = R.r0.w; pq.w = R.vol;
            *(V*)L.pos[slot] = pp;
            *(V*)L.prm[slot] = pq;
            *(V*)L.tst[pos] = pp;
            L.tq[pos] = qv;
            L.flg[slot] = (uint8_t)R.flags;
            L.atom[SCH * cur + pos] = o2 + a;
          }
          cur_cp = cur_cp || (__ballot(mine && (info & 1)) != 0);
          cur_lig = cur_lig || (any_lig && __ballot(mine && (info & 2)) != 0);
          fill += take;
          done += take;
          wave_sync();
          if (fill == SCH) test_chunk();
        }
      }
    }
    // ---- row end: test the partial chunk, evaluate the rest, flush
    if (fill > 0) test_chunk();
    drain(true);
    for (int r = 0; r < 3; ++r)
      if (cnt[r] > 0) {
        flush_chunk<Real, MODE>(L, A, r, cnt[r], lane);
        cnt[r] = 0;
      }
    flush_g1<Real, MODE>(L, A, n1, o1, lane);
    if (!C::NOE) {
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
    }
    wave_sync();
  }
}

}  // namespace tmol_ljlk
