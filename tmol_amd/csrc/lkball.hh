// lk_ball on the LJ+LK scheduler: water generation, its backward, and the
// pose score over the neighbor pass's work list.
//
// Reference: tmol/score/lk_ball/potentials/gen_pose_waters.impl.hh:28-432
// (waters and their derivatives), water.hh:29-98, :323-595 (donor and
// acceptor water builders), tmol/score/common/gen_coord.hh (build_coordinate),
// lk_ball_pose_score.impl.hh:358-700 (forward) and its block-pair /
// whole-pose backward, lk_ball.hh:42-494 (lk_fraction, lk_bridge_fraction,
// lk_ball_score), :1095-1361 (polar x occluder pairs).
//
// Waters: one thread per (block, water recipe).  A recipe names the water's
// atom, its slot and the atoms its frame is built from (donor: D, H;
// acceptor: A, B, B0, the bases found on the host by the reference's walk
// over the all-bonds table).  The acceptor water's backward is a
// vector-Jacobian product through the frame's normalisations and cross
// products.
//
// Score: persistent waves take rows of the work list (row g1: its neighbor
// blocks g2 >= g1).  A wave stages g1's occluders (heavy atoms, polar ones
// first) with coordinates, waters and parameters in LDS once per row, then
// each g2's; lanes run over the (polar, occluder) pairs of the block pair in
// both directions (intra-block: ordered pairs of distinct atoms).  Pairs
// beyond max_dis or at count-pair separation <= 3 score zero.  Gradients
// (weighted by dTdV) accumulate per staged atom in LDS -- coordinates and the
// four waters, 15 sums -- and are flushed to global memory once per block
// pair (g2) and once per row (g1).
#pragma once

#include "ljlk_common.hh"
#include "ljlk_math.hh"

namespace tmol_ljlk {

constexpr int LKB_OMAX = 32;  // occluders (heavy atoms) of one block the score kernel stages
constexpr int LKB_RW = TMOL_LKBALL_RECIPE_WIDTH;

template <typename Real>
__device__ __forceinline__ Real lkb_log(Real x);
template <>
__device__ __forceinline__ float lkb_log<float>(float x) {
  return __logf(x);
}
template <>
__device__ __forceinline__ double lkb_log<double>(double x) {
  return log(x);
}

// ------------------------------------------------------------------ waters

template <typename Real>
struct Vec3 {
  Real x, y, z;
};

template <typename Real>
__device__ __forceinline__ Vec3<Real> v3(const Real* p) {
  return {p[0], p[1], p[2]};
}
template <typename Real>
__device__ __forceinline__ Vec3<Real> sub(Vec3<Real> a, Vec3<Real> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename Real>
__device__ __forceinline__ Vec3<Real> cross(Vec3<Real> a, Vec3<Real> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename Real>
__device__ __forceinline__ Real dot(Vec3<Real> a, Vec3<Real> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename Real>
__device__ __forceinline__ Vec3<Real> scale(Vec3<Real> a, Real s) {
  return {a.x * s, a.y * s, a.z * s};
}

// frame of build_coordinate: e0 = (A - B)/|A - B|, e2 = e0 x (C - B)
// normalised (a fixed perpendicular when A, B, C are collinear), e1 = e2 x e0
template <typename Real>
__device__ __forceinline__ Real water_frame(Vec3<Real> A, Vec3<Real> B, Vec3<Real> C,
                                            Vec3<Real>& e0, Vec3<Real>& e1, Vec3<Real>& e2) {
  Vec3<Real> ab = sub(A, B);
  e0 = scale(ab, Real(1) / sqrt(dot(ab, ab)));
  e2 = cross(e0, sub(C, B));
  Real n = sqrt(dot(e2, e2));
  const Real nraw = n;
  if (n == Real(0)) {
    Vec3<Real> c1 = (e0.x == Real(1) || e0.x == Real(-1)) ? Vec3<Real>{0, 1, 0}
                                                           : Vec3<Real>{1, 0, 0};
    e2 = cross(e0, c1);
    n = sqrt(dot(e2, e2));
  }
  e2 = scale(e2, Real(1) / n);
  e1 = cross(e2, e0);
  return nraw;
}

// the water of one recipe row; P, G, GG its frame atoms' coordinates
// (acceptors: A, B, B0; ring: B replaced by (B + B0) / 2)
template <typename Real>
struct WaterCoef {
  Real k0, k1, k2;
};

template <typename Real>
__device__ __forceinline__ WaterCoef<Real> water_coef(Real dist, Real angle, Real tors) {
  const Real pi = Real(3.14159265358979323846);
  const Real s = dist * sin(pi - angle);
  return {dist * cos(pi - angle), s * cos(tors), s * sin(tors)};
}

template <typename Real>
__global__ __launch_bounds__(256) void lkb_waters_kernel(StackView s, const int32_t* n_recipes,
                                                         const int32_t* recipes, int Rmax,
                                                         const Real* coords, const Real* wg,
                                                         const Real* tors, Real* waters) {
  const int64_t n = int64_t(s.G) * Rmax;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int g = int(t / Rmax), r = int(t - int64_t(g) * Rmax);
    const int bt = s.block_type[g];
    if (bt < 0 || r >= n_recipes[bt]) continue;
    const int32_t* e = recipes + (int64_t(bt) * Rmax + r) * LKB_RW;
    const int64_t off = s.block_atom_offset[g];
    const int kind = e[2];
    const Real dist = wg[0];
    Vec3<Real> W;
    const Vec3<Real> P = v3(coords + 3 * (off + e[3]));
    const Vec3<Real> G = v3(coords + 3 * (off + e[4]));
    if (kind == 0) {  // donor: D + dist (H - D) / |H - D|
      Vec3<Real> v = sub(G, P);
      Real inv = Real(1) / sqrt(dot(v, v));
      W = {P.x + dist * (v.x * inv), P.y + dist * (v.y * inv), P.z + dist * (v.z * inv)};
    } else {
      const Vec3<Real> GG = v3(coords + 3 * (off + e[5]));
      Vec3<Real> B = G;
      if (kind == 3) B = {(G.x + GG.x) / Real(2), (G.y + GG.y) / Real(2), (G.z + GG.z) / Real(2)};
      Vec3<Real> e0, e1, e2;
      water_frame(P, B, GG, e0, e1, e2);
      const WaterCoef<Real> c = water_coef(dist, wg[kind], tors[(kind - 1) * 4 + e[6]]);
      W = {(e0.x * c.k0 + e1.x * c.k1 + e2.x * c.k2) + P.x,
           (e0.y * c.k0 + e1.y * c.k1 + e2.y * c.k2) + P.y,
           (e0.z * c.k0 + e1.z * c.k1 + e2.z * c.k2) + P.z};
    }
    Real* out = waters + ((off + e[0]) * 4 + e[1]) * 3;
    out[0] = W.x;
    out[1] = W.y;
    out[2] = W.z;
  }
}

// dE/dx += J^T dE/dW for every built water (atomics: a frame atom serves
// several waters)
template <typename Real>
__global__ __launch_bounds__(256) void lkb_waters_backward_kernel(
    StackView s, const int32_t* n_recipes, const int32_t* recipes, int Rmax, const Real* coords,
    const Real* wg, const Real* tors, const Real* dwat, Real* dxyz) {
  const int64_t n = int64_t(s.G) * Rmax;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int g = int(t / Rmax), r = int(t - int64_t(g) * Rmax);
    const int bt = s.block_type[g];
    if (bt < 0 || r >= n_recipes[bt]) continue;
    const int32_t* e = recipes + (int64_t(bt) * Rmax + r) * LKB_RW;
    const int64_t off = s.block_atom_offset[g];
    const Vec3<Real> gw = v3(dwat + ((off + e[0]) * 4 + e[1]) * 3);
    if (gw.x == Real(0) && gw.y == Real(0) && gw.z == Real(0)) continue;
    const int kind = e[2];
    const Real dist = wg[0];
    const int64_t ip = off + e[3], ig = off + e[4];
    const Vec3<Real> P = v3(coords + 3 * ip);
    const Vec3<Real> G = v3(coords + 3 * ig);
    if (kind == 0) {
      // W = D + d u, u = v / |v|, v = H - D: dE/dv = d / |v| (g - u (u . g))
      Vec3<Real> v = sub(G, P);
      Real inv = Real(1) / sqrt(dot(v, v));
      Vec3<Real> u = scale(v, inv);
      Real ug = dot(u, gw);
      Vec3<Real> gv = scale(sub(gw, scale(u, ug)), dist * inv);
      atomicAdd(&dxyz[3 * ig + 0], gv.x);
      atomicAdd(&dxyz[3 * ig + 1], gv.y);
      atomicAdd(&dxyz[3 * ig + 2], gv.z);
      atomicAdd(&dxyz[3 * ip + 0], gw.x - gv.x);
      atomicAdd(&dxyz[3 * ip + 1], gw.y - gv.y);
      atomicAdd(&dxyz[3 * ip + 2], gw.z - gv.z);
      continue;
    }
    const int64_t igg = off + e[5];
    const Vec3<Real> GG = v3(coords + 3 * igg);
    Vec3<Real> B = G;
    if (kind == 3) B = {(G.x + GG.x) / Real(2), (G.y + GG.y) / Real(2), (G.z + GG.z) / Real(2)};
    Vec3<Real> e0, e1, e2;
    const Real nraw = water_frame(P, B, GG, e0, e1, e2);
    // dW/dA = I always; collinear frames pass nothing to B and B0
    Vec3<Real> gA = gw, gB = {0, 0, 0}, gC = {0, 0, 0};
    if (nraw != Real(0)) {
      const WaterCoef<Real> c = water_coef(dist, wg[kind], tors[(kind - 1) * 4 + e[6]]);
      Vec3<Real> g0 = scale(gw, c.k0), g1 = scale(gw, c.k1), g2 = scale(gw, c.k2);
      // e1 = e2 x e0
      Vec3<Real> t1 = cross(e0, g1), t2 = cross(g1, e2);
      g2 = {g2.x + t1.x, g2.y + t1.y, g2.z + t1.z};
      g0 = {g0.x + t2.x, g0.y + t2.y, g0.z + t2.z};
      // e2 = n / |n|, n = e0 x (C - B)
      const Real d2 = dot(e2, g2);
      const Vec3<Real> gn = scale(sub(g2, scale(e2, d2)), Real(1) / nraw);
      const Vec3<Real> cb = sub(GG, B);
      const Vec3<Real> t3 = cross(cb, gn);
      g0 = {g0.x + t3.x, g0.y + t3.y, g0.z + t3.z};
      const Vec3<Real> gcb = cross(gn, e0);
      // e0 = (A - B) / |A - B|
      const Vec3<Real> ab = sub(P, B);
      const Real d0 = dot(e0, g0);
      const Vec3<Real> gab = scale(sub(g0, scale(e0, d0)), Real(1) / sqrt(dot(ab, ab)));
      gA = {gA.x + gab.x, gA.y + gab.y, gA.z + gab.z};
      gB = {-gab.x - gcb.x, -gab.y - gcb.y, -gab.z - gcb.z};
      gC = gcb;
    }
    atomicAdd(&dxyz[3 * ip + 0], gA.x);
    atomicAdd(&dxyz[3 * ip + 1], gA.y);
    atomicAdd(&dxyz[3 * ip + 2], gA.z);
    if (kind == 3) {  // B = (B + B0) / 2: half to each
      gC = {gC.x + gB.x / Real(2), gC.y + gB.y / Real(2), gC.z + gB.z / Real(2)};
      gB = {gB.x / Real(2), gB.y / Real(2), gB.z / Real(2)};
    }
    atomicAdd(&dxyz[3 * ig + 0], gB.x);
    atomicAdd(&dxyz[3 * ig + 1], gB.y);
    atomicAdd(&dxyz[3 * ig + 2], gB.z);
    atomicAdd(&dxyz[3 * igg + 0], gC.x);
    atomicAdd(&dxyz[3 * igg + 1], gC.y);
    atomicAdd(&dxyz[3 * igg + 2], gC.z);
  }
}

// ------------------------------------------------------------------ score

template <typename Real>
struct LkbArgs {
  StackView s;
  BTView bt;  // n_atoms, n_conn, conn_atoms
  WorkView w;
  const Real* coords;         // [NA, 3]
  const Real* waters;         // [NA, 4, 3], NaN where absent
  const int32_t* n_occ;       // [T]
  const int32_t* n_pol;       // [T]
  const int32_t* occ_atoms;   // [T, Omax], polar first
  const Real* occ_params;     // [T, Omax, 9]
  const int32_t* path_distance;  // [T, Amax, Amax]
  int Omax;
  const Real* gp;             // lj_hbond_dis, lj_hbond_OH_donor_dis, lj_hbond_hdis,
                              // lkb_water_dist, max_dis
  uint32_t* row_cursor;
  const Real* weights;        // [4, P] or [4, block-pair plane]
  double* energies64;         // [4, P]
  Real* energies_bp;          // [4, block-pair plane]
  Real* dxyz;                 // [NA, 3]
  Real* dwat;                 // [NA, 4, 3]
};

template <typename Real, bool DERIV>
struct LkbBlock {
  using T = typename LdsAcc<Real>::T;
  Real x[LKB_OMAX], y[LKB_OMAX], z[LKB_OMAX];
  Real w[12][LKB_OMAX];  // waters, component-major; NaN where absent
  Real r[LKB_OMAX];      // lj_radius
  Real K[LKB_OMAX];      // lk_dgfree / (2 pi^1.5 lk_lambda)
  Real il2[LKB_OMAX];    // 1 / lk_lambda^2
  Real vol[LKB_OMAX];    // lk_volume
  int32_t atom[LKB_OMAX];
  uint8_t loc[LKB_OMAX];  // atom index within the block
  uint8_t flg[LKB_OMAX];  // F_DONOR ... F_CARBON_LK
  uint8_t pd[MAX_CONN][LKB_OMAX];  // path distance connection atom -> atom
  T acc[DERIV ? 15 : 1][DERIV ? LKB_OMAX : 1];  // dE/dx (3), dE/dwaters (12)
  int n_occ, n_pol, bt, nc;
};

template <typename Real, bool DERIV>
struct LkbLDS {
  LkbBlock<Real, DERIV> b[2];
  uint8_t cs[MAX_CONN * MAX_CONN];
};

template <typename Real, bool DERIV>
__device__ void lkb_stage(LkbBlock<Real, DERIV>& B, const LkbArgs<Real>& a, int g, int lane) {
  const int bt = a.s.block_type[g];
  const int no = a.n_occ[bt], np = a.n_pol[bt], nc = a.bt.n_conn[bt];
  const int64_t off = a.s.block_atom_offset[g];
  if (lane < no) {
    const int la = a.occ_atoms[int64_t(bt) * a.Omax + lane];
    const int64_t at = off + la;
    B.x[lane] = a.coords[3 * at];
    B.y[lane] = a.coords[3 * at + 1];
    B.z[lane] = a.coords[3 * at + 2];
    const Real nan = __builtin_nan("");
#pragma unroll
    for (int k = 0; k < 12; ++k) B.w[k][lane] = lane < np ? a.waters[12 * at + k] : Real(nan);
    const Real* p = a.occ_params + (int64_t(bt) * a.Omax + lane) * 9;
    const Real pi_pow1p5 = Real(5.56832799683f);  // lk_isotropic.hh:40
    B.r[lane] = p[0];
    B.K[lane] = p[1] / (Real(2) * pi_pow1p5 * p[2]);
    B.il2[lane] = Real(1) / (p[2] * p[2]);
    B.vol[lane] = p[3];
    int f = 0;
    if (p[4] != Real(0)) f |= F_DONOR;
    if (p[4] != Real(0) && p[5] != Real(0)) f |= F_OH_DONOR;
    if (p[6] != Real(0)) f |= F_POLARH;
    if (p[7] != Real(0)) f |= F_ACCEPTOR;
    if (p[8] != Real(0)) f |= F_CARBON_LK;
    B.flg[lane] = uint8_t(f);
    B.atom[lane] = int32_t(at);
    B.loc[lane] = uint8_t(la);
#pragma unroll
    for (int c = 0; c < MAX_CONN; ++c) {
      int v = 6;
      if (c < nc) {
        const int ca = a.bt.conn_atoms[bt * a.bt.C + c];
        v = a.path_distance[(int64_t(bt) * a.bt.Amax + ca) * a.bt.Amax + la];
      }
      B.pd[c][lane] = uint8_t(v);
    }
    if (DERIV)
#pragma unroll
      for (int k = 0; k < 15; ++k) B.acc[k][lane] = 0;
  }
  if (lane == 0) {
    B.n_occ = no;
    B.n_pol = np;
    B.bt = bt;
    B.nc = nc;
  }
}

template <typename Real, bool DERIV>
__device__ void lkb_flush(LkbBlock<Real, DERIV>& B, const LkbArgs<Real>& a, int lane) {
  if (!DERIV) return;
  if (lane < B.n_occ) {
    const int64_t at = B.atom[lane];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const auto v = B.acc[k][lane];
      if (v != 0) atomicAdd(&a.dxyz[3 * at + k], LdsAcc<Real>::dec(v));
    }
    if (lane < B.n_pol)
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        const auto v = B.acc[3 + k][lane];
        if (v != 0) atomicAdd(&a.dwat[12 * at + k], LdsAcc<Real>::dec(v));
      }
  }
}

template <typename Real>
__device__ __forceinline__ void lds_add(typename LdsAcc<Real>::T* p, Real v) {
  if (v != Real(0)) atomicAdd(p, LdsAcc<Real>::enc(v));
}

// one (polar i of block X, occluder j of block Y) pair at separation sep
// (> 3) and distance below max_dis: energies into e4, weighted derivatives
// into the blocks' LDS sums
template <typename Real, bool DERIV>
__device__ __forceinline__ void lkb_pair(LkbBlock<Real, DERIV>& X, int i, LkbBlock<Real, DERIV>& Y,
                                         int j, int sep, Real dx, Real dy, Real dz, Real d2,
                                         const GlobalParams<Real>& g, Real water_dist,
                                         const Real* w4, Real* e4) {
  const Real dist = sqrt(d2);
  const Real wi0 = X.w[0][i];
  if (!(wi0 == wi0)) return;  // polar atom without waters
  const int fi = X.flg[i], fj = Y.flg[j];
  const Real sigma = lj_sigma(fi, fj, X.r[i], Y.r[j], g);
  Real d_min, cdmin, cdmax;
  lk_knots(sigma, bool((fi & F_CARBON_LK) && (fj & F_CARBON_LK)), d_min, cdmin, cdmax);
  Real iso, diso;
  lk_dir<Real, DERIV>(dist, d_min, cdmin, cdmax, -Y.vol[j] * X.K[i], X.r[i], X.il2[i], g, iso,
                      diso);
  const Real wcp = cp_weight<Real>(sep);
  iso *= wcp;
  diso *= wcp;
  const Real Jx = Y.x[j], Jy = Y.y[j], Jz = Y.z[j];
  // lk_fraction (lk_ball.hh:42-124)
  const Real ramp = Real(3.709);
  Real t = Real(1.4 + double(Y.r[j]));
  Real d2_low = t * t - ramp;
  d2_low = d2_low < Real(0) ? Real(0) : d2_low;
  Real fsum = 0, eW[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const Real dx = Jx - X.w[3 * w][i], dy = Jy - X.w[3 * w + 1][i], dz = Jz - X.w[3 * w + 2][i];
    const Real d2 = (dx * dx + dy * dy + dz * dz) - d2_low;
    eW[w] = d2 == d2 ? fast_exp(-d2) : Real(0);
    fsum += eW[w];
  }
  const Real fwd = -lkb_log(fsum);
  Real fr = Real(0), dfr = Real(0);
  if (fwd < Real(0)) {
    fr = Real(1);
  } else if (fwd < ramp) {
    const Real q = fwd / ramp;
    fr = (Real(1) - q * q) * (Real(1) - q * q);
    if (fwd > Real(0)) dfr = Real(-4) * fwd * (ramp * ramp - fwd * fwd) / ((ramp * ramp) * (ramp * ramp));
  }
  // lk_bridge_fraction (lk_ball.hh:126-277), when j is a donor or acceptor
  Real ov = Real(0);
  Real bsum = Real(0), bI = Real(0);  // bI: d(ov)/d|I-J|^2 scale
  Real bW[4][4];
  Real afrac = Real(0), dof = Real(0), ofrac = Real(0);
  const bool bridge = (fj & (F_DONOR | F_ACCEPTOR)) && Y.w[0][j] == Y.w[0][j];
  if (bridge) {
    const Real gap = Real(0.5), width = Real(2.6), aw = Real(2.8 * double(Real(2.6)));
#pragma unroll
    for (int wi = 0; wi < 4; ++wi)
#pragma unroll
      for (int wj = 0; wj < 4; ++wj) {
        const Real dx = X.w[3 * wi][i] - Y.w[3 * wj][j];
        const Real dy = X.w[3 * wi + 1][i] - Y.w[3 * wj + 1][j];
        const Real dz = X.w[3 * wi + 2][i] - Y.w[3 * wj + 2][j];
        const Real d2 = (dx * dx + dy * dy + dz * dz) - gap;
        bW[wi][wj] = d2 == d2 ? fast_exp(-d2) : Real(0);
        bsum += bW[wi][wj];
      }
    const Real wd = -lkb_log(bsum);
    if (wd > width) {
      ofrac = Real(0);
    } else if (wd > Real(0)) {
      const Real q = wd / width;
      ofrac = (Real(1) - q * q) * (Real(1) - q * q);
      dof = Real(-4) * wd * (width * width - wd * wd) / ((width * width) * (width * width));
    } else {
      ofrac = Real(1);
    }
    const Real target = Real(8.0 / 3.0 * double(water_dist * water_dist));
    const Real bd = d2 - target;
    const Real abd = bd < Real(0) ? -bd : bd;
    Real daf = Real(0);
    if (abd <= aw) {
      const Real q = bd / aw;
      afrac = (Real(1) - q * q) * (Real(1) - q * q);
      if (abd > Real(0)) daf = Real(-4) * bd * (aw * aw - bd * bd) / ((aw * aw) * (aw * aw));
    }
    ov = ofrac * afrac;
    bI = ofrac * daf;  // d(ov)/dI = 2 bI (I - J)
  }
  e4[0] += iso;
  e4[1] += iso * fr;
  e4[2] += iso * ov;
  e4[3] += ov / Real(2);
  if (!DERIV) return;
  // weighted derivatives (score types iso, lk_ball, bridge, bridge_uncpl)
  const Real inv = dist > Real(0) ? Real(1) / dist : Real(0);
  const Real cI = (w4[0] + w4[1] * fr + w4[2] * ov) * diso;  // along (I - J) / d
  const Real cB = w4[2] * iso + w4[3] / Real(2);              // on d(ov)
  const Real cF = w4[1] * iso;                                // on d(fr)
  const Real cIJ = cI * inv + cB * bI * Real(2);              // on I - J
  const Real gIx = cIJ * dx, gIy = cIJ * dy, gIz = cIJ * dz;
  Real gJx = -gIx, gJy = -gIy, gJz = -gIz;
  if (cF != Real(0) && dfr != Real(0) && fsum != Real(0)) {
    const Real s = cF * dfr / fsum;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (eW[w] == Real(0)) continue;
      const Real dx = Jx - X.w[3 * w][i], dy = Jy - X.w[3 * w + 1][i], dz = Jz - X.w[3 * w + 2][i];
      const Real c = Real(2) * eW[w] * s;
      gJx += c * dx;
      gJy += c * dy;
      gJz += c * dz;
      lds_add<Real>(&X.acc[3 + 3 * w][i], -c * dx);
      lds_add<Real>(&X.acc[4 + 3 * w][i], -c * dy);
      lds_add<Real>(&X.acc[5 + 3 * w][i], -c * dz);
    }
  }
  if (bridge && cB != Real(0) && dof != Real(0) && bsum != Real(0)) {
    const Real s = cB * afrac * dof / bsum;
#pragma unroll
    for (int wi = 0; wi < 4; ++wi)
#pragma unroll
      for (int wj = 0; wj < 4; ++wj) {
        if (bW[wi][wj] == Real(0)) continue;
        const Real c = Real(2) * bW[wi][wj] * s;
        const Real dx = X.w[3 * wi][i] - Y.w[3 * wj][j];
        const Real dy = X.w[3 * wi + 1][i] - Y.w[3 * wj + 1][j];
        const Real dz = X.w[3 * wi + 2][i] - Y.w[3 * wj + 2][j];
        lds_add<Real>(&X.acc[3 + 3 * wi][i], c * dx);
        lds_add<Real>(&X.acc[4 + 3 * wi][i], c * dy);
        lds_add<Real>(&X.acc[5 + 3 * wi][i], c * dz);
        lds_add<Real>(&Y.acc[3 + 3 * wj][j], -c * dx);
        lds_add<Real>(&Y.acc[4 + 3 * wj][j], -c * dy);
        lds_add<Real>(&Y.acc[5 + 3 * wj][j], -c * dz);
      }
  }
  lds_add<Real>(&X.acc[0][i], gIx);
  lds_add<Real>(&X.acc[1][i], gIy);
  lds_add<Real>(&X.acc[2][i], gIz);
  lds_add<Real>(&Y.acc[0][j], gJx);
  lds_add<Real>(&Y.acc[1][j], gJy);
  lds_add<Real>(&Y.acc[2][j], gJz);
}

template <int MODE>
constexpr bool lkb_deriv() {
  return (MODE & M_WGRAD) != 0;
}

template <typename Real, int MODE>
__global__ __launch_bounds__(256) void lkb_rows_kernel(LkbArgs<Real> a) {
  constexpr bool DERIV = lkb_deriv<MODE>();
  constexpr bool BP = (MODE & M_BP) != 0;
  constexpr bool NOE = (MODE & M_NOE) != 0;
  __shared__ LkbLDS<Real, DERIV> lds_all[WAVES_PER_WG];
  const int wave = threadIdx.x >> 6, lane = lane_id();
  LkbLDS<Real, DERIV>& L = lds_all[wave];
  GlobalParams<Real> g{};
  g.max_dis = a.gp[4];
  g.hbond_dis = a.gp[0];
  g.hbond_oh_dis = a.gp[1];
  g.hbond_hdis = a.gp[2];
  g.spline_start = g.max_dis - Real(1.5);
  g.max_dis2 = g.max_dis * g.max_dis;
  const Real water_dist = a.gp[3];
  const int P = a.s.P;
  const int64_t plane = BP ? bp_plane(a.s, a.w) : int64_t(P);
  RowQueue q(a.row_cursor, a.s.G);
  for (int64_t row = q.next(); row >= 0; row = q.next()) {
    const int g1 = int(row);
    const int n_items = int(min<int64_t>(a.w.row_count[g1], a.w.row_cap));  // count may overflow
    if (n_items == 0) continue;
    if (a.s.block_type[g1] < 0) continue;
    const int pose = a.s.pose_of(g1);
    lkb_stage<Real, DERIV>(L.b[0], a, g1, lane);
    Real w4[4] = {0, 0, 0, 0};
    if (DERIV && !BP)
#pragma unroll
      for (int st = 0; st < 4; ++st) w4[st] = a.weights[int64_t(st) * P + pose];
    double erow[4] = {0, 0, 0, 0};
    for (int k = 0; k < n_items; ++k) {
      const int g2 = a.w.row_items[int64_t(g1) * a.w.row_cap + k];
      const bool intra = g2 == g1;
      int64_t bpi = 0;
      if (BP) {
        bpi = bp_index(a.s, a.w, g1, g2, k);
        if (DERIV)
#pragma unroll
          for (int st = 0; st < 4; ++st) w4[st] = a.weights[int64_t(st) * plane + bpi];
      }
      int min_sep = 6;
      if (!intra) {
        wave_sync();
        lkb_stage<Real, DERIV>(L.b[1], a, g2, lane);
        const int nc1 = a.bt.n_conn[L.b[0].bt];
        const int nc2 = a.bt.n_conn[a.s.block_type[g2]];
        min_sep = min_block_sep(a.s, g1, g2, nc1, nc2);
        if (min_sep <= 4 && lane < MAX_CONN * MAX_CONN) {
          const int c1 = lane / MAX_CONN, c2 = lane - c1 * MAX_CONN;
          L.cs[lane] = uint8_t(c1 < nc1 && c2 < nc2 ? conn_sep(a.s, g1, g2, c1, c2) : 6);
        }
      }
      wave_sync();
      LkbBlock<Real, DERIV>& B1 = L.b[0];
      LkbBlock<Real, DERIV>& B2 = intra ? L.b[0] : L.b[1];
      const int np1 = B1.n_pol, no1 = B1.n_occ, np2 = B2.n_pol, no2 = B2.n_occ;
      const int na = np1 * no2;
      const int n_pairs = intra ? na : na + np2 * no1;
      Real e4[4] = {0, 0, 0, 0};
      for (int p = lane; p < n_pairs; p += WAVE) {
        const bool dir1 = p < na;  // polar in g1, occluder in g2
        const int pp = dir1 ? p : p - na;
        const int nY = dir1 ? no2 : no1;
        const int i = pp / nY, j = pp - i * nY;
        LkbBlock<Real, DERIV>& X = dir1 ? B1 : B2;
        LkbBlock<Real, DERIV>& Y = dir1 ? B2 : B1;
        int sep;
        if (intra) {
          if (i == j) continue;
          sep = a.path_distance[(int64_t(B1.bt) * a.bt.Amax + B1.loc[i]) * a.bt.Amax + B1.loc[j]];
        } else {
          sep = min_sep;
          if (sep <= 4) {
            const int a1 = dir1 ? i : j, a2 = dir1 ? j : i;  // slots in g1, g2
            int best = 5;
            for (int c1 = 0; c1 < B1.nc; ++c1)
              for (int c2 = 0; c2 < B2.nc; ++c2) {
                const int v = B1.pd[c1][a1] + L.cs[c1 * MAX_CONN + c2] + B2.pd[c2][a2];
                best = v < best ? v : best;
              }
            sep = best;
          }
        }
        if (sep <= 3) continue;
        const Real dx = X.x[i] - Y.x[j], dy = X.y[i] - Y.y[j], dz = X.z[i] - Y.z[j];
        const Real d2 = dx * dx + dy * dy + dz * dz;
        if (!(d2 < g.max_dis2)) continue;
        lkb_pair<Real, DERIV>(X, i, Y, j, sep, dx, dy, dz, d2, g, water_dist, w4, e4);
      }
      if (BP && !NOE) {
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const double v = wave_sum(double(e4[st]));
          if (lane == 0) a.energies_bp[int64_t(st) * plane + bpi] = Real(v);
        }
      } else if (!NOE) {
#pragma unroll
        for (int st = 0; st < 4; ++st) erow[st] += double(e4[st]);
      }
      if (!intra) {
        wave_sync();
        lkb_flush<Real, DERIV>(L.b[1], a, lane);
      }
    }
    wave_sync();
    lkb_flush<Real, DERIV>(L.b[0], a, lane);
    if (!BP && !NOE) {
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const double v = wave_sum(erow[st]);
        if (lane == 0 && v != 0.0) atomicAdd(&a.energies64[int64_t(st) * P + pose], v);
      }
    }
    wave_sync();
  }
}

}  // namespace tmol_ljlk
