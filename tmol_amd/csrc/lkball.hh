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
// Score: persistent waves take rows of a full work list (row g: every
// neighbor block of g, both triangles).  A row owns its block's atoms: it
// stages them (coordinates, parameters, waters, path distances) in LDS with
// gradient sums, tests the (polar, occluder) pairs against each neighbor in
// both directions -- g's polar atoms against the neighbor's occluders and
// the neighbor's polar atoms against g's occluders -- four neighbors per pass
// (16 lanes stage each), and queues the pairs within max_dis at count-pair
// separation > 3 in a FIFO that is evaluated 64 pairs at a time.  Each pair
// is thus evaluated in both blocks' rows; a row keeps only its own atoms'
// gradient, and the energy is counted in the row of the lower block (intra-
// block pairs once).  Gradients need no global atomics: every atom is
// written by its own row.
#pragma once

#include "ljlk_common.hh"
#include "ljlk_math.hh"

namespace tmol_ljlk {

constexpr int LKB_OMAX = 32;  // occluders (heavy atoms) of one block the score kernel stages
constexpr int LKB_PMAX = 16;  // polar atoms (with waters) of one block
constexpr int LKB_ITEMS = 4;  // neighbor blocks staged per pass (16 lanes each)
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

// vector-Jacobian product of build_coordinate(A, B, C, dist, angle, tors)
// (the water frame above): gradients gA, gB, gC of A, B, C for the output's
// gradient gw.  dW/dA = I always; collinear frames pass nothing to B and C.
template <typename Real>
__device__ __forceinline__ void build_coord_vjp(Vec3<Real> A, Vec3<Real> B, Vec3<Real> C,
                                                WaterCoef<Real> c, Vec3<Real> gw,
                                                Vec3<Real>& gA, Vec3<Real>& gB,
                                                Vec3<Real>& gC) {
  Vec3<Real> e0, e1, e2;
  const Real nraw = water_frame(A, B, C, e0, e1, e2);
  gA = gw;
  gB = {0, 0, 0};
  gC = {0, 0, 0};
  if (nraw == Real(0)) return;
  Vec3<Real> g0 = scale(gw, c.k0), g1 = scale(gw, c.k1), g2 = scale(gw, c.k2);
  // e1 = e2 x e0
  Vec3<Real> t1 = cross(e0, g1), t2 = cross(g1, e2);
  g2 = {g2.x + t1.x, g2.y + t1.y, g2.z + t1.z};
  g0 = {g0.x + t2.x, g0.y + t2.y, g0.z + t2.z};
  // e2 = n / |n|, n = e0 x (C - B)
  const Real d2 = dot(e2, g2);
  const Vec3<Real> gn = scale(sub(g2, scale(e2, d2)), Real(1) / nraw);
  const Vec3<Real> cb = sub(C, B);
  const Vec3<Real> t3 = cross(cb, gn);
  g0 = {g0.x + t3.x, g0.y + t3.y, g0.z + t3.z};
  const Vec3<Real> gcb = cross(gn, e0);
  // e0 = (A - B) / |A - B|
  const Vec3<Real> ab = sub(A, B);
  const Real d0 = dot(e0, g0);
  const Vec3<Real> gab = scale(sub(g0, scale(e0, d0)), Real(1) / sqrt(dot(ab, ab)));
  gA = {gA.x + gab.x, gA.y + gab.y, gA.z + gab.z};
  gB = {-gab.x - gcb.x, -gab.y - gcb.y, -gab.z - gcb.z};
  gC = gcb;
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
    Vec3<Real> gA, gB, gC;
    build_coord_vjp(P, B, GG, water_coef(dist, wg[kind], tors[(kind - 1) * 4 + e[6]]), gw, gA,
                    gB, gC);
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
  WorkView w;  // full rows: every neighbor block of g (both triangles) in row g
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
  const Real* weights;        // [4, P] or [4, P, B, B]
  double* energies64;         // [4, P]
  Real* energies_bp;          // [4, P, B, B] (padded)
  Real* dxyz;                 // [NA, 3]
  Real* dwat;                 // [NA, 4, 3]
  unsigned long long* census;  // M_CENSUS: [4] candidate pairs, separation <= 3,
                               // evaluated pairs, bridge evaluations
  // per-launch block records (lkb_prep_kernel, in the workspace)
  const int4* info;     // [G] block type, n_occ | n_pol << 8 | n_conn << 16, atom offset,
                        // last bonded partner (ragged; -1 none)
  const float4* opos;   // [G, Omax] occluder coordinates (float)
  const int32_t* oatom; // [G, Omax] occluder atom index
};

// workspace regions of the lk_ball score after the shared layout
struct LkbWs {
  int64_t info, opos, oatom, total;
};
inline LkbWs lkb_ws_layout(const WsLayout& L, int64_t G, int Omax) {
  LkbWs w;
  int64_t o = L.total;
  w.info = o;
  o = align256(o + 16 * G);
  w.opos = o;
  o = align256(o + 16 * G * int64_t(Omax));
  w.oatom = o;
  o = align256(o + 4 * G * int64_t(Omax));
  w.total = o;
  return w;
}

// block records for a launch: one thread per (block, occluder slot)
template <typename Real>
__global__ __launch_bounds__(256) void lkb_prep_kernel(StackView s, const int32_t* n_occ,
                                                       const int32_t* n_pol,
                                                       const int32_t* n_conn,
                                                       const int32_t* occ_atoms, int Omax,
                                                       const Real* coords, int4* info,
                                                       float4* opos, int32_t* oatom) {
  const int64_t n = int64_t(s.G) * Omax;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int g = int(t / Omax), k = int(t - int64_t(g) * Omax);
    const int bt = s.block_type[g];
    const int no = bt >= 0 ? n_occ[bt] : 0;
    const int off = s.block_atom_offset[g];
    if (k == 0) {
      int4 r;
      r.x = bt;
      r.y = bt >= 0 ? (no | (n_pol[bt] << 8) | (n_conn[bt] << 16)) : 0;
      r.z = off;
      r.w = -1;
      if (s.layout == TMOL_LJLK_RAGGED) {
        // the forward bonded table's rows are sorted by partner
        const int k0 = s.bonded_start[g], k1 = s.bonded_start[g + 1];
        if (k1 > k0) r.w = s.bonded[int64_t(k1 - 1) * BONDED_W];
      }
      info[g] = r;
    }
    if (k < no) {
      const int64_t at = off + occ_atoms[int64_t(bt) * Omax + k];
      opos[t] = make_float4(float(coords[3 * at]), float(coords[3 * at + 1]),
                            float(coords[3 * at + 2]), 0.f);
      oatom[t] = int32_t(at);
    }
  }
}

template <int MODE>
constexpr bool lkb_deriv_mode() {
  return (MODE & M_WGRAD) != 0;
}

// one atom of a pair, in registers
template <typename Real>
struct LkbAtom {
  Real x, y, z;
  Real r, K, il2, vol;  // lj_radius, lk_dgfree / (2 pi^1.5 lk_lambda), 1 / lk_lambda^2, lk_volume
  int flg;              // F_DONOR ... F_CARBON_LK
  Real w[12];           // waters, NaN where absent
  const Real* wsrc;     // where w was read from (LDS or global memory; null: none)
};

template <typename Real>
__device__ __forceinline__ void lkb_params(const Real* p, Real& r, Real& K, Real& il2, Real& vol,
                                           int& f) {
  const Real pi_pow1p5 = Real(5.56832799683f);  // lk_isotropic.hh:40
  r = p[0];
  K = p[1] * fast_rcp(Real(2) * pi_pow1p5 * p[2]);
  il2 = fast_rcp(p[2] * p[2]);
  vol = p[3];
  f = 0;
  if (p[4] != Real(0)) f |= F_DONOR;
  if (p[4] != Real(0) && p[5] != Real(0)) f |= F_OH_DONOR;
  if (p[6] != Real(0)) f |= F_POLARH;
  if (p[7] != Real(0)) f |= F_ACCEPTOR;
  if (p[8] != Real(0)) f |= F_CARBON_LK;
}

constexpr int LKB_FIFO = 128;  // compacted pairs awaiting evaluation (two chunks of 64)

// the row's own block (g): its occluders, with the row's gradient sums
template <typename Real, bool DERIV>
struct LkbRow {
  using T = typename LdsAcc<Real>::T;
  alignas(16) Real pos[LKB_OMAX][4];  // x, y, z, lj_radius
  alignas(16) Real prm[LKB_OMAX][4];  // K, 1 / lambda^2, lk_volume, flags
  alignas(16) Real w[LKB_PMAX][12];   // waters of the polar slots, NaN where absent
  int32_t atom[LKB_OMAX];
  uint8_t loc[LKB_OMAX];
  uint8_t pd[MAX_CONN][LKB_OMAX];
  T acc[DERIV ? 3 : 1][DERIV ? LKB_OMAX : 1];    // dE/dx
  T accw[DERIV ? 12 : 1][DERIV ? LKB_PMAX : 1];  // dE/dwaters (polar slots)
};

// a neighbor-row item: the other block, its type, first atom, occluder and
// occluder, polar and connection counts, minimum separation from the row's
// block (6: beyond 4) and its candidate pairs
struct LkbItem {
  int g2, bt2, off2, no2, np2, nc2, msep, cand;
};

// LKB_ITEMS neighbor blocks staged for the pair tests (16 lanes each)
struct LkbNbr {
  LkbItem it[LKB_ITEMS];
  float x[LKB_ITEMS][LKB_OMAX], y[LKB_ITEMS][LKB_OMAX], z[LKB_ITEMS][LKB_OMAX];
  int32_t atom[LKB_ITEMS][LKB_OMAX];
  uint8_t pd[LKB_ITEMS][MAX_CONN][LKB_OMAX];
  uint8_t cs[LKB_ITEMS][MAX_CONN * MAX_CONN];
};

// a compacted pair: the other atom (global index, its block type and slot)
// and meta = slot i of g (5 bits) | dir << 5 (0: g's polar atom against the
// other occluder, 1: the other polar atom against g's occluder) | counted << 6
// (energy of this pair belongs to this row) | sep4 << 7 (count-pair weight
// 0.2) | intra << 8 | slot j of g << 9 (intra pairs: the occluder)
struct LkbFifo {
  int32_t atom[LKB_FIFO];
  int32_t tslot[LKB_FIFO];  // block type * Omax + slot of the other atom
  int32_t blk[LKB_FIFO];    // the other atom's block
  uint32_t meta[LKB_FIFO];
};

template <typename Real, bool DERIV>
struct LkbLDS {
  LkbRow<Real, DERIV> row;
  LkbNbr nb;
  LkbFifo q;
};

template <typename Real, bool DERIV>
__device__ __forceinline__ void lkb_acc(LkbRow<Real, DERIV>& R, int k, int slot, Real v) {
  if (!DERIV || v == Real(0)) return;
  if (k < 3)
    atomicAdd(&R.acc[k][slot], LdsAcc<Real>::enc(v));
  else if (slot < LKB_PMAX)  // waters exist on polar slots only
    atomicAdd(&R.accw[k - 3][slot], LdsAcc<Real>::enc(v));
}

// The four lk_ball energies of polar atom I against occluder J at count-pair
// weight wcp (separation > 3) (lk_ball_score::V / dV, lk_ball.hh:335-494),
// added into e4.  With DERIV the gradient of sum_t w4[t] E_t is added into
// the row's LDS sums: I's coordinates and waters at slot si, J's at slot sj
// (-1: not this row's atom, not computed).  Returns 0 at or beyond max_dis
// (nothing added), 1, or 2 when the water bridge was evaluated.
template <typename Real, bool DERIV>
__device__ __forceinline__ int lkb_eval(const LkbAtom<Real>& I, const LkbAtom<Real>& J, Real wcp,
                                        const GlobalParams<Real>& g, Real water_dist,
                                        const Real* w4, Real* e4, LkbRow<Real, DERIV>& R, int si,
                                        int sj) {
  const Real dx = I.x - J.x, dy = I.y - J.y, dz = I.z - J.z;
  const Real d2 = dx * dx + dy * dy + dz * dz;
  const Real dist = fast_sqrt(d2);
  if (!(dist < g.max_dis)) return 0;  // the reference's cutoff, dist >= max_dis
  const Real sigma = lj_sigma(I.flg, J.flg, I.r, J.r, g);
  Real d_min, cdmin, cdmax;
  lk_knots(sigma, bool((I.flg & F_CARBON_LK) && (J.flg & F_CARBON_LK)), d_min, cdmin, cdmax);
  Real iso, diso;
  lk_dir<Real, DERIV>(dist, d_min, cdmin, cdmax, -J.vol * I.K, I.r, I.il2, g, iso, diso);
  iso *= wcp;
  diso *= wcp;
  // lk_fraction (lk_ball.hh:42-124)
  const Real ramp = Real(3.709), inv_ramp = Real(1) / ramp;
  const Real t = Real(1.4 + double(J.r));
  Real d2_low = t * t - ramp;
  d2_low = d2_low < Real(0) ? Real(0) : d2_low;
  Real fsum = 0, eW[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const Real ex = J.x - I.w[3 * w], ey = J.y - I.w[3 * w + 1], ez = J.z - I.w[3 * w + 2];
    const Real q = (ex * ex + ey * ey + ez * ez) - d2_low;
    eW[w] = q == q ? fast_exp(-q) : Real(0);
    fsum += eW[w];
  }
  const Real fwd = -lkb_log(fsum);
  Real fr = Real(0), dfr = Real(0);
  if (fwd < Real(0)) {
    fr = Real(1);
  } else if (fwd < ramp) {
    const Real q = fwd * inv_ramp;
    fr = (Real(1) - q * q) * (Real(1) - q * q);
    if (fwd > Real(0)) dfr = Real(-4) * fwd * (ramp * ramp - fwd * fwd) * (inv_ramp * inv_ramp) *
                             (inv_ramp * inv_ramp);
  }
  // lk_bridge_fraction (lk_ball.hh:126-277), when J is a donor or acceptor
  // with waters
  Real ov = Real(0), bsum = Real(0), bI = Real(0), afrac = Real(0), dof = Real(0);
  const bool bridge = (J.flg & (F_DONOR | F_ACCEPTOR)) && J.w[0] == J.w[0];
  const Real gap = Real(0.5), width = Real(2.6), aw = Real(2.8 * double(Real(2.6)));
  const Real inv_width = Real(1) / width, inv_aw = Real(1) / aw;
  if (bridge) {
#pragma unroll
    for (int wi = 0; wi < 4; ++wi)
#pragma unroll
      for (int wj = 0; wj < 4; ++wj) {
        const Real ex = I.w[3 * wi] - J.w[3 * wj];
        const Real ey = I.w[3 * wi + 1] - J.w[3 * wj + 1];
        const Real ez = I.w[3 * wi + 2] - J.w[3 * wj + 2];
        const Real q = (ex * ex + ey * ey + ez * ez) - gap;
        bsum += q == q ? fast_exp(-q) : Real(0);
      }
    const Real wd = -lkb_log(bsum);
    Real ofrac;
    if (wd > width) {
      ofrac = Real(0);
    } else if (wd > Real(0)) {
      const Real q = wd * inv_width;
      ofrac = (Real(1) - q * q) * (Real(1) - q * q);
      dof = Real(-4) * wd * (width * width - wd * wd) * (inv_width * inv_width) *
            (inv_width * inv_width);
    } else {
      ofrac = Real(1);
    }
    const Real target = Real(8.0 / 3.0 * double(water_dist * water_dist));
    const Real bd = d2 - target;
    const Real abd = bd < Real(0) ? -bd : bd;
    Real daf = Real(0);
    if (abd <= aw) {
      const Real q = bd * inv_aw;
      afrac = (Real(1) - q * q) * (Real(1) - q * q);
      if (abd > Real(0))
        daf = Real(-4) * bd * (aw * aw - bd * bd) * (inv_aw * inv_aw) * (inv_aw * inv_aw);
    }
    ov = ofrac * afrac;
    bI = ofrac * daf;  // d(ov)/dI = 2 bI (I - J)
  }
  e4[0] += iso;
  e4[1] += iso * fr;
  e4[2] += iso * ov;
  e4[3] += ov / Real(2);
  if (!DERIV) return bridge ? 2 : 1;
  const Real inv = dist > Real(0) ? fast_rcp(dist) : Real(0);
  const Real cI = (w4[0] + w4[1] * fr + w4[2] * ov) * diso;  // along (I - J) / d
  const Real cB = w4[2] * iso + w4[3] / Real(2);              // on d(ov)
  const Real cF = w4[1] * iso;                                // on d(fr)
  const Real cIJ = cI * inv + cB * bI * Real(2);              // on I - J
  Real gJx = -cIJ * dx, gJy = -cIJ * dy, gJz = -cIJ * dz;
  if (si >= 0) {
    lkb_acc(R, 0, si, cIJ * dx);
    lkb_acc(R, 1, si, cIJ * dy);
    lkb_acc(R, 2, si, cIJ * dz);
  }
  if (cF != Real(0) && dfr != Real(0) && fsum != Real(0)) {
    const Real sf = cF * dfr * fast_rcp(fsum);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (eW[w] == Real(0)) continue;
      const Real c = Real(2) * eW[w] * sf;
      const Real ex = J.x - I.w[3 * w], ey = J.y - I.w[3 * w + 1], ez = J.z - I.w[3 * w + 2];
      gJx += c * ex;
      gJy += c * ey;
      gJz += c * ez;
      if (si >= 0) {
        lkb_acc(R, 3 + 3 * w, si, -c * ex);
        lkb_acc(R, 4 + 3 * w, si, -c * ey);
        lkb_acc(R, 5 + 3 * w, si, -c * ez);
      }
    }
  }
  if (sj >= 0) {
    lkb_acc(R, 0, sj, gJx);
    lkb_acc(R, 1, sj, gJy);
    lkb_acc(R, 2, sj, gJz);
  }
  if (bridge && cB != Real(0) && dof != Real(0) && bsum != Real(0)) {
    const Real sb = cB * afrac * dof * fast_rcp(bsum);
    // a rolled loop over I's waters, read again from their source: indexing
    // I.w by the loop counter would place it in scratch memory
#pragma unroll 1
    for (int wi = 0; wi < 4; ++wi) {
      const Real iwx = I.wsrc[3 * wi], iwy = I.wsrc[3 * wi + 1], iwz = I.wsrc[3 * wi + 2];
      Real gx = 0, gy = 0, gz = 0;
#pragma unroll
      for (int wj = 0; wj < 4; ++wj) {
        const Real ex = iwx - J.w[3 * wj];
        const Real ey = iwy - J.w[3 * wj + 1];
        const Real ez = iwz - J.w[3 * wj + 2];
        const Real q = (ex * ex + ey * ey + ez * ez) - gap;
        if (!(q == q)) continue;
        const Real c = Real(2) * fast_exp(-q) * sb;
        gx += c * ex;
        gy += c * ey;
        gz += c * ez;
        if (sj >= 0) {
          lkb_acc(R, 3 + 3 * wj, sj, -c * ex);
          lkb_acc(R, 4 + 3 * wj, sj, -c * ey);
          lkb_acc(R, 5 + 3 * wj, sj, -c * ez);
        }
      }
      if (si >= 0) {
        lkb_acc(R, 3 + 3 * wi, si, gx);
        lkb_acc(R, 4 + 3 * wi, si, gy);
        lkb_acc(R, 5 + 3 * wi, si, gz);
      }
    }
  }
  return bridge ? 2 : 1;
}

template <typename Real, bool DERIV>
__device__ __forceinline__ void lkb_load_row_atom(const LkbRow<Real, DERIV>& R, int i,
                                                  LkbAtom<Real>& A) {
  A.x = R.pos[i][0];
  A.y = R.pos[i][1];
  A.z = R.pos[i][2];
  A.r = R.pos[i][3];
  A.K = R.prm[i][0];
  A.il2 = R.prm[i][1];
  A.vol = R.prm[i][2];
  A.flg = int(R.prm[i][3]);
#pragma unroll
  for (int k = 0; k < 12; ++k) A.w[k] = i < LKB_PMAX ? R.w[i][k] : Real(__builtin_nan(""));
  A.wsrc = i < LKB_PMAX ? &R.w[i][0] : nullptr;
}

template <typename Real>
__device__ __forceinline__ void lkb_load_global_atom(const LkbArgs<Real>& a, int64_t at,
                                                     int tslot, bool polar, LkbAtom<Real>& A) {
  A.x = a.coords[3 * at];
  A.y = a.coords[3 * at + 1];
  A.z = a.coords[3 * at + 2];
  lkb_params(a.occ_params + int64_t(tslot) * 9, A.r, A.K, A.il2, A.vol, A.flg);
  // waters: of the polar partner, or of an occluder that may bridge
  const bool want = polar || (A.flg & (F_DONOR | F_ACCEPTOR));
  const Real nan = __builtin_nan("");
#pragma unroll
  for (int k = 0; k < 12; ++k) A.w[k] = want ? a.waters[12 * at + k] : Real(nan);
  A.wsrc = want ? a.waters + 12 * at : nullptr;
}

// evaluate n (<= 64) entries of the FIFO starting at head (mod LKB_FIFO)
template <typename Real, int MODE>
__device__ __forceinline__ void lkb_phase_b(LkbLDS<Real, lkb_deriv_mode<MODE>()>& L,
                                            const LkbArgs<Real>& a, int g1, int head, int n,
                                            int lane, const GlobalParams<Real>& g,
                                            Real water_dist, const Real* w4row, Real* e4,
                                            uint64_t* cnt) {
  constexpr bool DERIV = lkb_deriv_mode<MODE>();
  constexpr bool BP = (MODE & M_BP) != 0;
  constexpr bool NOE = (MODE & M_NOE) != 0;
  constexpr bool CENSUS = (MODE & M_CENSUS) != 0;
  if (lane >= n) return;
  const int e = (head + lane) & (LKB_FIFO - 1);
  const uint32_t m = L.q.meta[e];
  // block-pair cell (padded layout): (pose, min block, max block)
  int64_t cell = 0;
  Real w4[4];
  if (BP) {
    const int g2 = L.q.blk[e];
    const int lo = g1 < g2 ? g1 : g2, hi = g1 < g2 ? g2 : g1;
    const int p = lo / a.s.B;
    cell = (int64_t(p) * a.s.B + (lo - p * a.s.B)) * a.s.B + (hi - p * a.s.B);
    if (DERIV) {
      const int64_t plane = int64_t(a.s.P) * a.s.B * a.s.B;
#pragma unroll
      for (int st = 0; st < 4; ++st) w4[st] = a.weights[int64_t(st) * plane + cell];
    }
  } else {
#pragma unroll
    for (int st = 0; st < 4; ++st) w4[st] = w4row[st];
  }
  const int i = m & 31, dir = (m >> 5) & 1;
  const bool counted = (m >> 6) & 1, intra = (m >> 8) & 1;
  const Real wcp = (m >> 7) & 1 ? Real(0.2) : Real(1);
  LkbAtom<Real> I, J;
  int si = -1, sj = -1;  // this row's slots on the I and J sides
  if (intra) {
    si = i;
    sj = (m >> 9) & 31;
    lkb_load_row_atom(L.row, si, I);
    lkb_load_row_atom(L.row, sj, J);
  } else if (dir == 0) {
    si = i;
    lkb_load_row_atom(L.row, i, I);
    lkb_load_global_atom(a, L.q.atom[e], L.q.tslot[e], false, J);
  } else {
    sj = i;
    lkb_load_global_atom(a, L.q.atom[e], L.q.tslot[e], true, I);
    lkb_load_row_atom(L.row, i, J);
  }
  if (!(I.w[0] == I.w[0])) return;  // polar atom without waters
  Real e4p[4] = {0, 0, 0, 0};
  const int ev = lkb_eval<Real, DERIV>(I, J, wcp, g, water_dist, w4, e4p, L.row, si, sj);
  if (ev == 0) return;
  if (CENSUS && counted) {
    cnt[2] += 1;
    cnt[3] += ev > 1;
  }
  if (counted && !NOE) {
    if (BP) {
      const int64_t plane = int64_t(a.s.P) * a.s.B * a.s.B;
#pragma unroll
      for (int st = 0; st < 4; ++st)
        if (e4p[st] != Real(0)) atomicAdd(&a.energies_bp[int64_t(st) * plane + cell], e4p[st]);
    } else {
#pragma unroll
      for (int st = 0; st < 4; ++st) e4[st] += e4p[st];
    }
  }
}

// waves per SIMD the kernel is compiled for: 4 with gradients (which
// otherwise take ~200 VGPRs, 2 waves), 5 for energies (measured on the
// 512-pose batch: 4.82 vs 5.33 ms at 5 with gradients; 3.11 vs 3.33 ms at 4
// for energies)
template <int MODE>
constexpr int lkb_waves_per_simd() {
  return (MODE & M_WGRAD) ? 4 : 5;
}

template <typename Real, int MODE>
__global__ __launch_bounds__(256, lkb_waves_per_simd<MODE>()) void lkb_gather_kernel(
    LkbArgs<Real> a) {
  constexpr bool DERIV = lkb_deriv_mode<MODE>();
  constexpr bool BP = (MODE & M_BP) != 0;
  constexpr bool NOE = (MODE & M_NOE) != 0;
  constexpr bool CENSUS = (MODE & M_CENSUS) != 0;
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
  const float max_dis2f = float(g.max_dis2);
  const Real water_dist = a.gp[3];
  const int P = a.s.P;
  uint64_t cnt[4] = {0, 0, 0, 0};
  RowQueue rq(a.row_cursor, a.s.G);
  for (int64_t row = rq.next(); row >= 0; row = rq.next()) {
    const int g1 = int(row);
    const int4 inf1 = a.info[g1];
    const int bt1 = inf1.x;
    if (bt1 < 0) continue;
    const int n_items = int(min<int64_t>(a.w.row_count[g1], a.w.row_cap));  // count may overflow
    if (n_items == 0) continue;
    const int pose = a.s.pose_of(g1);
    const int no1 = inf1.y & 255, np1 = (inf1.y >> 8) & 255, nc1 = inf1.y >> 16;
    const int64_t off1 = inf1.z;
    const int32_t* items = a.w.row_items + int64_t(g1) * a.w.row_cap;
    // the row's block
    if (lane < no1) {
      const int la = a.occ_atoms[int64_t(bt1) * a.Omax + lane];
      const int64_t at = off1 + la;
      L.row.pos[lane][0] = a.coords[3 * at];
      L.row.pos[lane][1] = a.coords[3 * at + 1];
      L.row.pos[lane][2] = a.coords[3 * at + 2];
      int f;
      lkb_params(a.occ_params + (int64_t(bt1) * a.Omax + lane) * 9, L.row.pos[lane][3],
                 L.row.prm[lane][0], L.row.prm[lane][1], L.row.prm[lane][2], f);
      L.row.prm[lane][3] = Real(f);
      const Real nan = __builtin_nan("");
      if (lane < LKB_PMAX)
#pragma unroll
        for (int k = 0; k < 12; ++k) L.row.w[lane][k] = lane < np1 ? a.waters[12 * at + k] : nan;
      L.row.atom[lane] = int32_t(at);
      L.row.loc[lane] = uint8_t(la);
#pragma unroll
      for (int c = 0; c < MAX_CONN; ++c) {
        int v = 6;
        if (c < nc1) {
          const int ca = a.bt.conn_atoms[bt1 * a.bt.C + c];
          v = a.path_distance[(int64_t(bt1) * a.bt.Amax + ca) * a.bt.Amax + la];
        }
        L.row.pd[c][lane] = uint8_t(v);
      }
      if (DERIV) {
#pragma unroll
        for (int k = 0; k < 3; ++k) L.row.acc[k][lane] = 0;
        if (lane < LKB_PMAX)
#pragma unroll
          for (int k = 0; k < 12; ++k) L.row.accw[k][lane] = 0;
      }
    }
    Real w4row[4] = {0, 0, 0, 0};
    if (DERIV && !BP)
#pragma unroll
      for (int st = 0; st < 4; ++st) w4row[st] = a.weights[int64_t(st) * P + pose];
    Real e4[4] = {0, 0, 0, 0};
    int head = 0, tail = 0;  // FIFO positions (mod LKB_FIFO), wave-uniform
    // item k0 + lane of the next pass in lanes < LKB_ITEMS, loaded one pass ahead
    int nxt = lane < LKB_ITEMS && lane < n_items ? items[lane] : -1;
    const int quarter = lane >> 4, ql = lane & 15;
    for (int k0 = 0; k0 < n_items; k0 += LKB_ITEMS) {
      const int my = nxt;
      // this quarter's neighbor slots (ql, ql + 16), loaded without waiting
      // for the block's record
      const int gs = __shfl(my, quarter, WAVE);
      const bool stage = gs >= 0 && gs != g1;
      float4 sp[2];
      int32_t sat[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int sl = ql + 16 * r;
        sp[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        sat[r] = 0;
        if (stage && sl < a.Omax) {
          sp[r] = a.opos[int64_t(gs) * a.Omax + sl];
          sat[r] = a.oatom[int64_t(gs) * a.Omax + sl];
        }
      }
      nxt = lane < LKB_ITEMS && k0 + LKB_ITEMS + lane < n_items ? items[k0 + LKB_ITEMS + lane]
                                                              : -1;
      // the items' records, one per lane (lanes < LKB_ITEMS), into LDS: the
      // pair tests below read them by a per-lane item index, which on private
      // arrays would place them in scratch memory
      if (lane < LKB_ITEMS) {
        const int4 inf2 = my >= 0 ? a.info[my] : make_int4(-1, 0, 0, -1);
        LkbItem it;
        it.g2 = my;
        it.bt2 = inf2.x;
        it.off2 = inf2.z;
        it.no2 = inf2.y & 255;
        it.nc2 = inf2.y >> 16;
        it.np2 = (inf2.y >> 8) & 255;
        it.msep = 6;
        if (my == g1) {
          it.cand = np1 * no1;
        } else if (my >= 0) {
          // ragged: the last bonded partner of the lower block rules out most
          // pairs without a search
          const int lo = g1 < my ? g1 : my, hi = g1 < my ? my : g1;
          const int bmax = lo == g1 ? inf1.w : inf2.w;
          if (a.s.layout == TMOL_LJLK_PADDED || hi <= bmax)
            it.msep = min_block_sep(a.s, g1, my, nc1, it.nc2);
          it.cand = np1 * it.no2 + it.np2 * no1;
        } else {
          it.cand = 0;
        }
        L.nb.it[lane] = it;
      }
      wave_sync();
      int ctot = 0, cbeg[LKB_ITEMS];
#pragma unroll
      for (int h = 0; h < LKB_ITEMS; ++h) {
        cbeg[h] = ctot;
        ctot += L.nb.it[h].cand;
      }
      // stage the neighbor blocks, one per quarter wave
      {
        const int h = quarter;
        const LkbItem it = L.nb.it[h];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const int sl = ql + 16 * r;
          if (stage && sl < it.no2) {
            L.nb.x[h][sl] = sp[r].x;
            L.nb.y[h][sl] = sp[r].y;
            L.nb.z[h][sl] = sp[r].z;
            L.nb.atom[h][sl] = sat[r];
            if (it.msep <= 4) {
              const int la = sat[r] - it.off2;
#pragma unroll
              for (int c = 0; c < MAX_CONN; ++c) {
                int v = 6;
                if (c < it.nc2) {
                  const int ca = a.bt.conn_atoms[it.bt2 * a.bt.C + c];
                  v = a.path_distance[(int64_t(it.bt2) * a.bt.Amax + ca) * a.bt.Amax + la];
                }
                L.nb.pd[h][c][sl] = uint8_t(v);
              }
            }
          }
        }
        if (stage && it.msep <= 4) {
          const int c1 = ql / MAX_CONN, c2 = ql - c1 * MAX_CONN;
          L.nb.cs[h][ql] = uint8_t(c1 < nc1 && c2 < it.nc2 ? conn_sep(a.s, g1, gs, c1, c2) : 6);
        }
      }
      wave_sync();
      // pair tests over the items; passing pairs into the FIFO
      if (CENSUS) {
#pragma unroll
        for (int h = 0; h < LKB_ITEMS; ++h) {
          const int g2 = L.nb.it[h].g2;
          if (g2 == g1)
            cnt[0] += uint64_t(np1) * (no1 - 1);
          else if (g2 > g1)
            cnt[0] += uint64_t(L.nb.it[h].cand);
        }
      }
      for (int base = 0; base < ctot; base += WAVE) {
        const int p = base + lane;
        bool pass = false;
        uint32_t meta = 0;
        int32_t oatom = 0, otslot = 0, oblk = 0;
        if (p < ctot) {
          int h = 0, pp = p;
#pragma unroll
          for (int q = 1; q < LKB_ITEMS; ++q)
            if (p >= cbeg[q]) {
              h = q;
              pp = p - cbeg[q];
            }
          const LkbItem& it = L.nb.it[h];
          const int g2 = it.g2;
          const bool counted = g2 >= g1;
          if (g2 == g1) {
            const int i = pp / no1, j = pp - i * no1;
            if (i != j) {
              const int sep = a.path_distance[(int64_t(bt1) * a.bt.Amax + L.row.loc[i]) *
                                                  a.bt.Amax + L.row.loc[j]];
              if (sep > 3) {
                const Real dx = L.row.pos[i][0] - L.row.pos[j][0],
                           dy = L.row.pos[i][1] - L.row.pos[j][1],
                           dz = L.row.pos[i][2] - L.row.pos[j][2];
                pass = dx * dx + dy * dy + dz * dz < g.max_dis2;
                meta = uint32_t(i) | (1u << 6) | (sep == 4 ? 1u << 7 : 0u) | (1u << 8) |
                       (uint32_t(j) << 9);
                oblk = g1;
              } else if (CENSUS) {
                cnt[1] += 1;
              }
            }
          } else {
            const int no2 = it.no2, nc2 = it.nc2;
            const int nA = np1 * no2;
            const int dir = pp < nA ? 0 : 1;
            int i, j;  // i: slot of the row block, j: slot of the neighbor
            if (dir == 0) {
              i = pp / no2;
              j = pp - i * no2;
            } else {
              const int q = pp - nA;
              j = q / no1;
              i = q - j * no1;
            }
            int sep = it.msep;
            if (sep <= 4) {
              int best = 5;
              for (int c1 = 0; c1 < nc1; ++c1)
                for (int c2 = 0; c2 < nc2; ++c2) {
                  const int v = L.row.pd[c1][i] + L.nb.cs[h][c1 * MAX_CONN + c2] +
                                L.nb.pd[h][c2][j];
                  best = v < best ? v : best;
                }
              sep = best;
            }
            if (sep > 3) {
              const float dx = float(L.row.pos[i][0]) - L.nb.x[h][j];
              const float dy = float(L.row.pos[i][1]) - L.nb.y[h][j];
              const float dz = float(L.row.pos[i][2]) - L.nb.z[h][j];
              // a float pre-test with margin; the exact cutoff is applied in
              // the evaluation's precision below
              pass = dx * dx + dy * dy + dz * dz < max_dis2f * 1.0001f;
              if (pass && !(sizeof(Real) == 4)) {
                const int64_t at = L.nb.atom[h][j];
                const Real ex = L.row.pos[i][0] - a.coords[3 * at],
                           ey = L.row.pos[i][1] - a.coords[3 * at + 1],
                           ez = L.row.pos[i][2] - a.coords[3 * at + 2];
                pass = ex * ex + ey * ey + ez * ez < g.max_dis2;
              } else if (pass) {
                pass = dx * dx + dy * dy + dz * dz < max_dis2f;
              }
              meta = uint32_t(i) | (uint32_t(dir) << 5) | (counted ? 1u << 6 : 0u) |
                     (sep == 4 ? 1u << 7 : 0u);
              oatom = L.nb.atom[h][j];
              otslot = it.bt2 * a.Omax + j;
              oblk = g2;
            } else if (CENSUS && counted) {
              cnt[1] += 1;
            }
          }
        }
        const uint64_t mask = __ballot(pass);
        if (pass) {
          const int e = (tail + lanes_below(mask)) & (LKB_FIFO - 1);
          L.q.atom[e] = oatom;
          L.q.tslot[e] = otslot;
          L.q.blk[e] = oblk;
          L.q.meta[e] = meta;
        }
        tail += __popcll(mask);
        if (tail - head >= WAVE) {
          wave_sync();
          lkb_phase_b<Real, MODE>(L, a, g1, head, WAVE, lane, g, water_dist, w4row, e4, cnt);
          head += WAVE;
        }
      }
      wave_sync();
    }
    while (tail > head) {
      const int n = tail - head < WAVE ? tail - head : WAVE;
      wave_sync();
      lkb_phase_b<Real, MODE>(L, a, g1, head, n, lane, g, water_dist, w4row, e4, cnt);
      head += n;
    }
    wave_sync();
    // the row owns its block's atoms: plain read-modify-write, no atomics
    if (DERIV && lane < no1) {
      const int64_t at = L.row.atom[lane];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const auto v = L.row.acc[k][lane];
        if (v != 0) a.dxyz[3 * at + k] += LdsAcc<Real>::dec(v);
      }
      if (lane < np1)
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const auto v = L.row.accw[k][lane];
          if (v != 0) a.dwat[12 * at + k] += LdsAcc<Real>::dec(v);
        }
    }
    if (!BP && !NOE && !CENSUS) {
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const double v = wave_sum(double(e4[st]));
        if (lane == 0 && v != 0.0) atomicAdd(&a.energies64[int64_t(st) * P + pose], v);
      }
    }
    wave_sync();
  }
  if (CENSUS) {
    // cnt[0] is wave-uniform (added by every lane); the others per lane
    const unsigned long long c1 = wave_sum(cnt[1]), c2 = wave_sum(cnt[2]),
                             c3 = wave_sum(cnt[3]);
    if (lane == 0) {
      atomicAdd(&a.census[0], (unsigned long long)cnt[0]);
      atomicAdd(&a.census[1], c1);
      atomicAdd(&a.census[2], c2);
      atomicAdd(&a.census[3], c3);
    }
  }
}

}  // namespace tmol_ljlk
