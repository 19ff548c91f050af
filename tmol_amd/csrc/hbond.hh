// hbond on the LJ+LK scheduler: the donor / acceptor base atoms of each
// polar atom (gen_hbond_bases), the five-atom hbond potential and the pose
// score over the half-triangle neighbor work list (hbond_pose_scores).
//
// Reference: tmol/score/hbond/potentials/gen_hbond_bases.impl.hh:200-468,
// hbond_pose_score.impl.hh:47-200, 430-1000, hbond.hh:640-913,
// potentials.hh:36-508; tmol/score/common/geom.hh, polynomial.hh.
#pragma once

#include "ljlk_common.hh"
#include "lkball.hh"

namespace tmol_ljlk {

constexpr int HB_POLY = 16;     // c10..c0, pad, xmin, xmax, min_val, max_val (double)
constexpr int HB_DW = TMOL_HBOND_DON_WIDTH;
constexpr int HB_AW = TMOL_HBOND_ACC_WIDTH;
constexpr int HB_CHUNK = 64;    // work-list items staged per pass
constexpr int HB_FIFO = 128;    // passing donor x acceptor pairs awaiting evaluation
constexpr int HB_MAX_SEP = 4;   // pairs below separation 5 score 0

// ------------------------------------------------------------------ bases

// derived [NA, 3, 3] / inds [NA, 3]: slot 0 the donor heavy atom of a polar
// hydrogen, slots 1, 2 the bases B, B0 of an acceptor.  src [T, Amax, 3]
// holds each slot's source atom in the block: -1 a missing base ((0, 0, 0),
// index -1), -2 no role (left as filled: NaN, -1).
template <typename Real>
__global__ __launch_bounds__(256) void hb_bases_kernel(StackView s, const int32_t* n_atoms,
                                                       const int32_t* src, int Amax,
                                                       const Real* coords, Real* derived,
                                                       int32_t* inds) {
  const int64_t total = int64_t(s.G) * Amax;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t g = idx / Amax;
    const int a = int(idx - g * Amax);
    const int bt = s.block_type[g];
    if (bt < 0 || a >= n_atoms[bt]) continue;
    const int64_t off = s.block_atom_offset[g];
    const int64_t at = off + a;
    const int32_t* sr = src + (int64_t(bt) * Amax + a) * 3;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int x = sr[k];
      if (x == -2) continue;
      Real* d = derived + 9 * at + 3 * k;
      if (x < 0) {
        d[0] = d[1] = d[2] = Real(0);
        continue;
      }
      const int64_t xa = off + x;
      d[0] = coords[3 * xa];
      d[1] = coords[3 * xa + 1];
      d[2] = coords[3 * xa + 2];
      inds[3 * at + k] = int32_t(xa);
    }
  }
}

// ------------------------------------------------------------------ potential

template <typename Real>
__device__ __forceinline__ Vec3<Real> add3(Vec3<Real> a, Vec3<Real> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename Real>
__device__ __forceinline__ Real norm3(Vec3<Real> a) {
  return sqrt(dot(a, a));
}

// bounded degree-10 polynomial, evaluated in double (polynomial.hh
// bound_poly<11, double>): below xmin min_val, above xmax max_val
template <typename Real, bool DERIV>
__device__ __forceinline__ Real hb_poly(Real x, const double* P, Real& dv) {
  const double xd = double(x);
  double v, d = 0.0;
  if (xd < P[12]) {
    v = P[14];
  } else if (xd > P[13]) {
    v = P[15];
  } else {
    v = P[0];
    d = P[0];
#pragma unroll
    for (int i = 1; i < 10; ++i) {
      v = v * xd + P[i];
      if (DERIV) d = d * xd + v;
    }
    v = v * xd + P[10];
  }
  if (DERIV) dv = Real(d);
  return Real(v);
}

// the angle between vectors a and b, 2 atan2(|a x b|, |a||b| + a.b), and
// its gradients (geom.hh interior_angle)
template <typename Real, bool DERIV>
__device__ __forceinline__ Real hb_interior(Vec3<Real> a, Vec3<Real> b, Vec3<Real>& da,
                                            Vec3<Real>& db) {
  const Vec3<Real> cr = cross(a, b);
  const Real n2 = dot(cr, cr);
  Vec3<Real> z = cr;
  if (n2 > Real(0)) {
    const Real n = sqrt(n2);
    z = {cr.x / n, cr.y / n, cr.z / n};
  }
  const Real an = norm3(a), bn = norm3(b);
  const Real v = Real(2) * atan2(dot(cr, z), an * bn + dot(a, b));
  if (DERIV) {
    const Vec3<Real> ua = {a.x / an, a.y / an, a.z / an}, ub = {b.x / bn, b.y / bn, b.z / bn};
    const Vec3<Real> ta = cross(ua, z), tb = cross(ub, z);
    da = {ta.x / an, ta.y / an, ta.z / an};
    db = {-tb.x / bn, -tb.y / bn, -tb.z / bn};
  }
  return v;
}

// angle at B of A-B-C (geom.hh pt_interior_angle)
template <typename Real, bool DERIV>
__device__ __forceinline__ Real hb_pt_angle(Vec3<Real> A, Vec3<Real> B, Vec3<Real> C,
                                            Vec3<Real>& dA, Vec3<Real>& dB, Vec3<Real>& dC) {
  const Real v = hb_interior<Real, DERIV>(sub(A, B), sub(C, B), dA, dC);
  if (DERIV) dB = {-(dA.x + dC.x), -(dA.y + dC.y), -(dA.z + dC.z)};
  return v;
}

// dihedral I-J-K-L with the Blondel-Karplus gradients (geom.hh dihedral_angle);
// the cosine is clamped with fmin / fmax, which pass a NaN over
template <typename Real, bool DERIV>
__device__ __forceinline__ Real hb_dihedral(Vec3<Real> I, Vec3<Real> J, Vec3<Real> K,
                                            Vec3<Real> L, Vec3<Real>* d) {
  const Vec3<Real> F = sub(I, J), G = sub(J, K), H = sub(L, K);
  const Vec3<Real> A = cross(F, G), B = cross(H, G);
  const Real sign = dot(G, cross(A, B)) >= Real(0) ? Real(-1) : Real(1);
  const Real c = fmax(Real(-1), fmin(dot(A, B) / (norm3(A) * norm3(B)), Real(1)));
  const Real v = sign * acos(c);
  if (DERIV) {
    const Real gn = norm3(G), aa = dot(A, A), bb = dot(B, B);
    const Real fg = dot(F, G), hg = dot(H, G);
    const Real ka = gn / aa, kb = gn / bb, fa = fg / (aa * gn), hb = hg / (bb * gn);
    d[0] = scale(A, -ka);
    d[1] = {ka * A.x + fa * A.x - hb * B.x, ka * A.y + fa * A.y - hb * B.y,
            ka * A.z + fa * A.z - hb * B.z};
    d[2] = {-kb * B.x - fa * A.x + hb * B.x, -kb * B.y - fa * A.y + hb * B.y,
            -kb * B.z - fa * A.z + hb * B.z};
    d[3] = scale(B, kb);
  }
  return v;
}

// polynomial of cos(angle between H - A and A - B) (potentials.hh
// _BAH_angle_base_form_V_dV); gradients on B, A, H
template <typename Real, bool DERIV>
__device__ __forceinline__ Real hb_bah_base(Vec3<Real> B, Vec3<Real> A, Vec3<Real> H,
                                            const double* P, Vec3<Real>* g) {
  const Vec3<Real> ah = sub(H, A), ba = sub(A, B);
  const Real an = norm3(ah), bn = norm3(ba), abn = an * bn;
  const Real c = dot(ah, ba) / abn;
  Real dp;
  const Real v = hb_poly<Real, DERIV>(c, P, dp);
  if (DERIV) {
    const Real ia = -c / (an * an), ib = -c / (bn * bn);
    const Vec3<Real> cA = {ia * ah.x + ba.x / abn, ia * ah.y + ba.y / abn, ia * ah.z + ba.z / abn};
    const Vec3<Real> cB = {ib * ba.x + ah.x / abn, ib * ba.y + ah.y / abn, ib * ba.z + ah.z / abn};
    g[0] = scale(cB, -dp);
    g[1] = scale(sub(cB, cA), dp);
    g[2] = scale(cA, dp);
  }
  return v;
}

// the sp2 chi term (potentials.hh sp2chi_energy_V_dV): d = BAH180_rise,
// m = range_span, l = outer_width
template <typename Real, bool DERIV>
__device__ __forceinline__ Real hb_sp2chi(Real ang, Real chi, Real d, Real m, Real l, Real& da,
                                          Real& dc) {
  const Real pi = Real(3.14159265358979323846);
  const Real H = Real(0.5) * (cos(Real(2) * chi) + Real(1));
  da = dc = Real(0);
  if (double(ang) > double(pi) * 2.0 / 3.0) {
    const Real F = d / Real(2) * cos(Real(3) * (pi - ang)) + d / Real(2) - Real(0.5);
    const Real G = d - Real(0.5);
    if (DERIV) {
      da = Real(1.5) * d * sin(Real(3) * ang) * cos(chi) * cos(chi);
      dc = Real(0.5) * d * (cos(Real(3) * ang) + Real(1)) * sin(Real(2) * chi);
    }
    return H * F + (Real(1) - H) * G;
  }
  if (double(ang) >= double(pi) * (2.0 / 3.0 - double(l))) {
    const Real outer = cos(pi - (pi * Real(2) / Real(3) - ang) / l);
    const Real F = m / Real(2) * outer + m / Real(2) - Real(0.5);
    const Real G = (m - d) / Real(2) * outer + (m - d) / Real(2) + d - Real(0.5);
    if (DERIV) {
      const Real u = (ang - Real(2) * pi / Real(3)) / l;
      da = Real(0.5) * (-d * sin(chi) * sin(chi) + m) * sin(u) / l;
      dc = Real(0.5) * d * (cos(u) + Real(1)) * sin(Real(2) * chi);
    }
    return H * F + (Real(1) - H) * G;
  }
  return m - Real(0.5);
}

// The hbond energy of donor H (heavy atom D) and acceptor A (bases B, B0)
// (potentials.hh hbond_score::V_dV): A-H distance, A-H-D angle and B-A-H
// angle polynomials, the sp2 chi term, times the donor x acceptor weight
// (rounded to float), cut to 0 above 0.1 and faded over [-0.1, 0.1].
//   pp: [acceptor hybridization, acceptor weight, donor weight]
//   P: the pair's three polynomials (A-H distance, cos BAH, AHD angle)
//   gp: [sp2_range_span, sp2_BAH180_rise, sp2_outer_width, sp3_softmax_fade, ...]
//   g: gradients of D, H, A, B, B0
template <typename Real, bool DERIV>
__device__ __forceinline__ Real hb_energy(Vec3<Real> D, Vec3<Real> H, Vec3<Real> A,
                                          Vec3<Real> B, Vec3<Real> B0, const Real* pp,
                                          const double* P, const Real* gp, Vec3<Real>* g) {
  const Vec3<Real> zero = {Real(0), Real(0), Real(0)};
  Vec3<Real> gD = zero, gH = zero, gA = zero, gB = zero, gB0 = zero;
  Real E;
  {  // A-H distance
    const Vec3<Real> v = sub(A, H);
    const Real dist = norm3(v);
    Real dp;
    E = hb_poly<Real, DERIV>(dist, P, dp);
    if (DERIV && dist != Real(0)) {
      const Real k = dp / dist;
      gA = scale(v, k);
      gH = scale(v, -k);
    }
  }
  {  // A-H-D angle
    Vec3<Real> a, h, d;
    const Real ang = hb_pt_angle<Real, DERIV>(A, H, D, a, h, d);
    Real dp;
    E += hb_poly<Real, DERIV>(ang, P + 2 * HB_POLY, dp);
    if (DERIV) {
      gA = add3(gA, scale(a, dp));
      gH = add3(gH, scale(h, dp));
      gD = scale(d, dp);
    }
  }
  const int hyb = int(pp[0]);
  const double* Pb = P + HB_POLY;
  if (hyb == 1 || hyb == 3) {  // sp2: B; ring: the midpoint of B and B0
    const Vec3<Real> Bx = hyb == 1 ? B : scale(add3(B, B0), Real(0.5));
    Vec3<Real> t[3];
    E += hb_bah_base<Real, DERIV>(Bx, A, H, Pb, t);
    if (DERIV) {
      if (hyb == 1) {
        gB = t[0];
      } else {
        gB = scale(t[0], Real(0.5));
        gB0 = gB;
      }
      gA = add3(gA, t[1]);
      gH = add3(gH, t[2]);
    }
  } else if (hyb == 2) {  // sp3: soft-max of the B and B0 forms
    Vec3<Real> t[3], t0[3];
    const Real x = hb_bah_base<Real, DERIV>(B, A, H, Pb, t);
    const Real x0 = hb_bah_base<Real, DERIV>(B0, A, H, Pb, t0);
    const Real f = gp[3];
    const Real e1 = exp(x * f), e0 = exp(x0 * f);
    E += log(e1 + e0) / f;
    if (DERIV) {
      const Real w1 = e1 / (e1 + e0), w0 = e0 / (e1 + e0);
      gB = scale(t[0], w1);
      gB0 = scale(t0[0], w0);
      gA = add3(gA, add3(scale(t[1], w1), scale(t0[1], w0)));
      gH = add3(gH, add3(scale(t[2], w1), scale(t0[2], w0)));
    }
  }
  if (hyb == 1) {  // B0-B-A-H chi
    Vec3<Real> bb, ba, bh, dd[4];
    const Real bah = hb_pt_angle<Real, DERIV>(B, A, H, bb, ba, bh);
    const Real chi = hb_dihedral<Real, DERIV>(B0, B, A, H, dd);
    Real da, dc;
    E += hb_sp2chi<Real, DERIV>(bah, chi, gp[1], gp[0], gp[2], da, dc);
    if (DERIV) {
      gB0 = add3(gB0, scale(dd[0], dc));
      gB = add3(gB, add3(scale(dd[1], dc), scale(bb, da)));
      gA = add3(gA, add3(scale(dd[2], dc), scale(ba, da)));
      gH = add3(gH, add3(scale(dd[3], dc), scale(bh, da)));
    }
  }
  const float w = float(pp[1] * pp[2]);
  E *= Real(w);
  Real slope = Real(w);
  if (E > Real(0.1)) {
    E = Real(0);
    slope = Real(0);
  } else if (E > Real(-0.1)) {
    const Real E0 = E;
    E = Real(-0.025) + Real(0.5) * E - Real(2.5) * E * E;
    slope = Real(w) * (Real(-5) * E0 + Real(0.5));
  }
  if (DERIV) {
    g[0] = scale(gD, slope);
    g[1] = scale(gH, slope);
    g[2] = scale(gA, slope);
    g[3] = scale(gB, slope);
    g[4] = scale(gB0, slope);
  }
  return E;
}

// ------------------------------------------------------------------ pose score

template <typename Real>
struct HbArgs {
  StackView s;
  WorkView w;
  const Real* coords;
  const Real* derived;     // [NA, 3, 3]
  const int32_t* dinds;    // [NA, 3]
  const int32_t* n_conn;   // [T]
  const int32_t* conn_atoms;  // [T, C]
  const int32_t* n_don;    // [T]
  const int32_t* n_acc;    // [T]
  const int32_t* don;      // [T, Dmax, HB_DW]
  const int32_t* acc;      // [T, Xmax, HB_AW]
  int Dmax, Xmax, Am;
  const int32_t* path_distance;  // [T, Am, Am]
  const Real* pair_params;       // [nD, nA, 3]
  const double* polys;           // [nD, nA, 48]
  int n_acc_types;
  const Real* gp;                // [6]
  const Real* spheres;           // [G, 4] block spheres of the neighbor pass
  uint32_t* row_cursor;
  const Real* weights;           // dTdV: [P] or block pairs
  double* energies64;            // [P]
  Real* energies_bp;             // block pairs
  Real* dxyz;                    // [NA, 3]
  unsigned long long* census;    // [3]
};

// a work-list item of the row: the other block, its type, first atom,
// donor / acceptor / connection counts, minimum separation from the row's
// block and its candidate donor x acceptor pairs
struct HbItem {
  int g2, bt2, off2, nd2, na2, nc2, msep, cand;
};

struct HbLDS {
  HbItem it[HB_CHUNK];
  int32_t cend[HB_CHUNK];  // inclusive prefix sums of cand
  int32_t qh[HB_FIFO], qa[HB_FIFO], qt[HB_FIFO], qb[HB_FIFO];  // H, A, types, other block
};

// evaluate n (<= 64) FIFO entries from head
template <typename Real, int MODE>
__device__ __forceinline__ void hb_phase_b(HbLDS& L, const HbArgs<Real>& a, int g1, int head,
                                           int n, int lane, Real w_row, LaneSum& esum,
                                           unsigned long long& n_eval) {
  constexpr bool DERIV = (MODE & M_WGRAD) != 0;
  constexpr bool BP = (MODE & M_BP) != 0;
  constexpr bool NOE = (MODE & M_NOE) != 0;
  constexpr bool CENSUS = (MODE & M_CENSUS) != 0;
  if (lane >= n) return;
  const int e = (head + lane) & (HB_FIFO - 1);
  const int64_t h = L.qh[e], x = L.qa[e];
  const int ty = L.qt[e];
  const int dt = ty & 0xffff, at = ty >> 16;
  if (CENSUS) {
    ++n_eval;
    return;
  }
  const int64_t pi = int64_t(dt) * a.n_acc_types + at;
  const Vec3<Real> D = v3(a.derived + 9 * h), H = v3(a.coords + 3 * h), A = v3(a.coords + 3 * x);
  const Vec3<Real> B = v3(a.derived + 9 * x + 3), B0 = v3(a.derived + 9 * x + 6);
  Vec3<Real> g[5];
  const Real E = hb_energy<Real, DERIV>(D, H, A, B, B0, a.pair_params + 3 * pi,
                                        a.polys + 3 * HB_POLY * pi, a.gp, g);
  int64_t cell = 0;
  if (BP) {
    const int g2 = L.qb[e];
    const int lo = g1 < g2 ? g1 : g2, hi = g1 < g2 ? g2 : g1;
    const int p = lo / a.s.B;
    cell = (int64_t(p) * a.s.B + (lo - p * a.s.B)) * a.s.B + (hi - p * a.s.B);
    if (!NOE && E != Real(0)) atomicAdd(&a.energies_bp[cell], E);
  } else if (!NOE) {
    esum.add(E);
  }
  if (DERIV) {
    const Real w = BP ? a.weights[cell] : w_row;
    const int64_t atoms[5] = {a.dinds[3 * h], h, x, a.dinds[3 * x + 1], a.dinds[3 * x + 2]};
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      if (atoms[j] < 0) continue;
      const Real gx = w * g[j].x, gy = w * g[j].y, gz = w * g[j].z;
      if (gx != Real(0)) atomicAdd(&a.dxyz[3 * atoms[j]], gx);
      if (gy != Real(0)) atomicAdd(&a.dxyz[3 * atoms[j] + 1], gy);
      if (gz != Real(0)) atomicAdd(&a.dxyz[3 * atoms[j] + 2], gz);
    }
  }
}

// Persistent waves over the rows of the half-triangle work list (row g1
// holds its neighbor blocks g2 >= g1, g1 itself for the intra-block pairs).
// Per pass, up to 64 items are staged as records in LDS (an item whose block
// spheres are farther apart than max_ha_dis gets no candidates); the wave
// then walks
// their donor x acceptor candidates 64 at a time (an item found by binary
// search over the prefix sums): count pair (inter-block pairs: the
// reference's lookup with the donor's and acceptor's list positions as atom
// indices; intra-block pairs: the atoms' path distance), then the H-A
// distance against max_ha_dis.  Passing pairs are compacted into an LDS
// FIFO and evaluated 64 at a time.  Energies: per-pose FP64 sums (one
// atomic per row) or block-pair cells; gradients through the derived atoms'
// indices with global atomics (few pairs pass).
// waves per SIMD the scoring kernel is compiled for: it is latency-bound
// (the per-item loads form dependent chains), so occupancy pays until the
// register cap spills.  Measured on the 512-pose batch: gradients 3 / 4 / 5 /
// 6 waves 1.07 / 0.91 / 1.03 / 1.10 ms (5 and 6 spill); energies 4 / 5 / 6 / 8
// waves 0.74 / 0.68 / 0.67 / 0.67 ms (6 and 8 spill)
#ifndef TMOL_HB_WAVES_GRAD
#define TMOL_HB_WAVES_GRAD 4
#endif
#ifndef TMOL_HB_WAVES_ENERGY
#define TMOL_HB_WAVES_ENERGY 5
#endif
template <int MODE>
constexpr int hb_waves_per_simd() {
  return (MODE & M_WGRAD) ? TMOL_HB_WAVES_GRAD : TMOL_HB_WAVES_ENERGY;
}

template <typename Real, int MODE>
__global__ __launch_bounds__(256, hb_waves_per_simd<MODE>()) void hb_score_kernel(HbArgs<Real> a) {
  constexpr bool DERIV = (MODE & M_WGRAD) != 0;
  constexpr bool BP = (MODE & M_BP) != 0;
  constexpr bool NOE = (MODE & M_NOE) != 0;
  constexpr bool CENSUS = (MODE & M_CENSUS) != 0;
  __shared__ HbLDS lds_all[WAVES_PER_WG];
  const int wave = threadIdx.x >> 6, lane = lane_id();
  HbLDS& L = lds_all[wave];
  const Real max_ha = a.gp[5];
  unsigned long long n_cand = 0, n_cp = 0, n_eval = 0;
  RowQueue rq(a.row_cursor, a.s.G);
  for (int64_t row = rq.next(); row >= 0; row = rq.next()) {
    const int g1 = int(row);
    const int bt1 = a.s.block_type[g1];
    if (bt1 < 0) continue;
    const int n_items = int(min<int64_t>(a.w.row_count[g1], a.w.row_cap));
    const int nd1 = a.n_don[bt1], na1 = a.n_acc[bt1];
    if (n_items == 0 || (nd1 == 0 && na1 == 0)) continue;
    const int pose = a.s.pose_of(g1);
    const int64_t off1 = a.s.block_atom_offset[g1];
    const int nc1 = a.n_conn[bt1];
    const int32_t* items = a.w.row_items + int64_t(g1) * a.w.row_cap;
    const Real w_row = (DERIV && !BP) ? a.weights[pose] : Real(0);
    LaneSum esum;
    int head = 0, tail = 0;
    for (int k0 = 0; k0 < n_items; k0 += HB_CHUNK) {
      const int n = n_items - k0 < HB_CHUNK ? n_items - k0 : HB_CHUNK;
      int cand = 0;
      if (lane < n) {
        HbItem it;
        it.g2 = items[k0 + lane];
        it.bt2 = a.s.block_type[it.g2];
        it.off2 = a.s.block_atom_offset[it.g2];
        it.nd2 = it.bt2 >= 0 ? a.n_don[it.bt2] : 0;
        it.na2 = it.bt2 >= 0 ? a.n_acc[it.bt2] : 0;
        it.nc2 = it.bt2 >= 0 ? a.n_conn[it.bt2] : 0;
        it.msep = 6;
        if (it.g2 == g1) {
          cand = nd1 * na1;
        } else if (it.bt2 >= 0) {
          // blocks whose spheres stay farther apart than max_ha_dis (with
          // a margin for rounding) hold no pair within it
          const Real* s1 = a.spheres + 4 * int64_t(g1);
          const Real* s2 = a.spheres + 4 * int64_t(it.g2);
          const Vec3<Real> dc = {s1[0] - s2[0], s1[1] - s2[1], s1[2] - s2[2]};
          if (norm3(dc) <= s1[3] + s2[3] + max_ha + Real(1e-3)) {
            it.msep = min_block_sep(a.s, g1, it.g2, nc1, it.nc2);
            cand = nd1 * it.na2 + na1 * it.nd2;
          }
        }
        it.cand = cand;
        L.it[lane] = it;
      }
      // inclusive prefix sum of the candidate counts
      int incl = cand;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const int v = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += v;
      }
      if (lane < n) L.cend[lane] = incl;
      const int total = __shfl(incl, WAVE - 1, WAVE);
      if (CENSUS) n_cand += uint64_t(lane == 0 ? total : 0);
      wave_sync();
      for (int base = 0; base < total; base += WAVE) {
        const int c = base + lane;
        bool pass = false;
        int32_t qh = 0, qa = 0, qt = 0, qb = 0;
        if (c < total) {
          int lo = 0, hi = n - 1;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (L.cend[mid] > c)
              hi = mid;
            else
              lo = mid + 1;
          }
          const HbItem& it = L.it[lo];
          const int pp = c - (lo ? L.cend[lo - 1] : 0);
          int hl, al, dt, at, sep;
          int64_t ho, ao;
          if (it.g2 == g1) {
            const int i = pp / na1, j = pp - i * na1;
            const int32_t* dr = a.don + (int64_t(bt1) * a.Dmax + i) * HB_DW;
            const int32_t* ar = a.acc + (int64_t(bt1) * a.Xmax + j) * HB_AW;
            hl = dr[0];
            dt = dr[2];
            al = ar[0];
            at = ar[3];
            ho = ao = off1;
            sep = a.path_distance[(int64_t(bt1) * a.Am + hl) * a.Am + al];
          } else {
            const int nA = nd1 * it.na2;
            int r1, r2;  // positions in g1's and g2's lists (the count-pair lookup)
            if (pp < nA) {  // g1 donates, g2 accepts
              const int i = pp / it.na2, j = pp - i * it.na2;
              const int32_t* dr = a.don + (int64_t(bt1) * a.Dmax + i) * HB_DW;
              const int32_t* ar = a.acc + (int64_t(it.bt2) * a.Xmax + j) * HB_AW;
              hl = dr[0];
              dt = dr[2];
              al = ar[0];
              at = ar[3];
              ho = off1;
              ao = it.off2;
              r1 = i;
              r2 = j;
            } else {  // g2 donates, g1 accepts
              const int q = pp - nA;
              const int i = q / na1, j = q - i * na1;
              const int32_t* dr = a.don + (int64_t(it.bt2) * a.Dmax + i) * HB_DW;
              const int32_t* ar = a.acc + (int64_t(bt1) * a.Xmax + j) * HB_AW;
              hl = dr[0];
              dt = dr[2];
              al = ar[0];
              at = ar[3];
              ho = it.off2;
              ao = off1;
              r1 = j;
              r2 = i;
            }
            sep = it.msep;
            if (sep <= HB_MAX_SEP) {
              int best = HB_MAX_SEP + 1;
              for (int c1 = 0; c1 < nc1; ++c1) {
                const int ca1 = a.conn_atoms[bt1 * a.s.C + c1];
                const int pd1 = a.path_distance[(int64_t(bt1) * a.Am + ca1) * a.Am + r1];
                for (int c2 = 0; c2 < it.nc2; ++c2) {
                  const int ca2 = a.conn_atoms[it.bt2 * a.s.C + c2];
                  const int pd2 = a.path_distance[(int64_t(it.bt2) * a.Am + ca2) * a.Am + r2];
                  const int v = pd1 + conn_sep(a.s, g1, it.g2, c1, c2) + pd2;
                  best = v < best ? v : best;
                }
              }
              sep = best;
            }
          }
          if (sep > HB_MAX_SEP) {
            const int64_t hg = ho + hl, ag = ao + al;
            const Vec3<Real> d = sub(v3(a.coords + 3 * hg), v3(a.coords + 3 * ag));
            pass = norm3(d) < max_ha;
            qh = int32_t(hg);
            qa = int32_t(ag);
            qt = dt | (at << 16);
            qb = it.g2;
          } else if (CENSUS) {
            ++n_cp;
          }
        }
        const uint64_t mask = __ballot(pass);
        if (pass) {
          const int e = (tail + lanes_below(mask)) & (HB_FIFO - 1);
          L.qh[e] = qh;
          L.qa[e] = qa;
          L.qt[e] = qt;
          L.qb[e] = qb;
        }
        tail += __popcll(mask);
        if (tail - head >= WAVE) {
          wave_sync();
          hb_phase_b<Real, MODE>(L, a, g1, head, WAVE, lane, w_row, esum, n_eval);
          head += WAVE;
        }
      }
      wave_sync();
    }
    while (tail > head) {
      const int n = tail - head < WAVE ? tail - head : WAVE;
      wave_sync();
      hb_phase_b<Real, MODE>(L, a, g1, head, n, lane, w_row, esum, n_eval);
      head += n;
    }
    wave_sync();
    if (!BP && !NOE && !CENSUS) {
      const double v = wave_sum(esum.value());
      if (lane == 0 && v != 0.0) atomicAdd(&a.energies64[pose], v);
    }
  }
  if (CENSUS) {
    const unsigned long long c1 = wave_sum(n_cp), c2 = wave_sum(n_eval);
    if (lane == 0) {
      atomicAdd(&a.census[0], n_cand);
      atomicAdd(&a.census[1], c1);
      atomicAdd(&a.census[2], c2);
    }
  }
}

}  // namespace tmol_ljlk
