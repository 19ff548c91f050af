// General scoring kernel: any block size (multi-tile blocks), all modes
// including the branch census.  One wave per block pair at a time; pairs
// enumerated in diagonal order, gradients pre-reduced in wave-private LDS.
// Used for block types with more than 32 atoms and for the census; the
// rows kernel (ljlk_score_rows.hh) is the fast path.
#pragma once

#include "ljlk_common.hh"

namespace tmol_ljlk {

template <typename Real>
struct GenLDS {
  typename LdsAcc<Real>::T grad[3][3][2 * TILE];  // [term][component][slot]
  Real x[2 * TILE], y[2 * TILE], z[2 * TILE];
  Real rad[2 * TILE], sw[2 * TILE], K[2 * TILE], il2[2 * TILE], vol[2 * TILE];
  int flg[2 * TILE];
  uint8_t heavy[2][TILE];
  uint8_t pd[2][MAX_CONN][TILE];
  uint8_t cs[MAX_CONN * MAX_CONN];
  int n_heavy[2];
};

// +g to slot i1, -g to slot i2 (exact negation of the encoded value)
template <typename Real>
__device__ __forceinline__ void grad_pair(GenLDS<Real>& L, int term, int i1, int i2,
                                          Real gx, Real gy, Real gz) {
  using Acc = LdsAcc<Real>;
  using T = typename Acc::T;
  const T ex = Acc::enc(gx), ey = Acc::enc(gy), ez = Acc::enc(gz);
  atomicAdd(&L.grad[term][0][i1], ex);
  atomicAdd(&L.grad[term][1][i1], ey);
  atomicAdd(&L.grad[term][2][i1], ez);
  atomicAdd(&L.grad[term][0][i2], T(0) - ex);
  atomicAdd(&L.grad[term][1][i2], T(0) - ey);
  atomicAdd(&L.grad[term][2][i2], T(0) - ez);
}

template <typename Real, int MODE>
__global__ __launch_bounds__(256) void score_general_kernel(ScoreArgs<Real> A) {
  constexpr bool DVDX = MODE & M_DVDX;
  constexpr bool WGRAD = MODE & M_WGRAD;
  constexpr bool BP = MODE & M_BP;
  constexpr bool CENSUS = MODE & M_CENSUS;
  constexpr bool NOE = (MODE & M_NOE) || CENSUS;
  constexpr bool DERIV = (DVDX || WGRAD) && !CENSUS;
  uint32_t cnt[C_N];
#pragma unroll
  for (int k = 0; k < C_N; ++k) cnt[k] = 0;
  __shared__ GenLDS<Real> lds_all[WAVES_PER_WG];
  const int lane = lane_id();
  const int wave_in_wg = threadIdx.x / WAVE;
  GenLDS<Real>& L = lds_all[wave_in_wg];

  const GlobalParams<Real> gp = load_global_params(A.global_params);
  const StackView& s = A.s;
  const BTView& bt = A.bt;
  RowQueue queue(A.row_cursor, s.G);

  for (int64_t row = queue.next(); row >= 0; row = queue.next()) {
  uint32_t row_n = A.w.row_count[row];
  if (row_n > A.w.row_cap) row_n = (uint32_t)A.w.row_cap;
  for (uint32_t kk = 0; kk < row_n; ++kk) {
    const int g1 = (int)row;
    const int g2 = A.w.row_items[row * A.w.row_cap + kk];
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
        const int64_t PBB = bp_plane(s, A.w), base = bp_index(s, A.w, g1, g2, kk);
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
                L.pd[half][c][a] = A.conn_pd[(int64_t(tt) * bt.Amax + st + a) * MAX_CONN + c];
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
              for (int k = 0; k < 3; ++k) L.grad[t][k][slot] = typename LdsAcc<Real>::T(0);
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
                sep = inter_pair_sep(L.pd[0], L.pd[1], L.cs, min_sep, nc1, nc2, a1, a2, lig);
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
                    grad_pair(L, term, i1, i2, gs * dx, gs * dy, gs * dz);
                  } else {
                    Real gs = dlj * inv_d * (inside ? w[1] : w[0]);
                    grad_pair(L, 0, i1, i2, gs * dx, gs * dy, gs * dz);
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
                  sep = inter_pair_sep(L.pd[0], L.pd[1], L.cs, min_sep, nc1, nc2, a1, a2, lig);
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
                      grad_pair(L, term, i1, i2, gs * dx, gs * dy, gs * dz);
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
                  Real v = LdsAcc<Real>::dec(L.grad[t][k][slot]);
                  if (v != Real(0))
                    unsafeAtomicAdd(&A.dvdx[(int64_t(t) * A.n_atoms + atom) * 3 + k], v);
                }
            } else {
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                Real v = LdsAcc<Real>::dec(L.grad[0][k][slot]);
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
          const int64_t PBB = bp_plane(s, A.w), base = bp_index(s, A.w, g1, g2, kk);
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


}  // namespace tmol_ljlk
