/* CPU restatement of the reference lk_ball ops -- TEST INFRASTRUCTURE.
 *
 * The checker, never the product: only tests/, the smoke check and bench.py's
 * cpu_baseline leg load it.  Included by ljlk_oracle.c after
 * ljlk_oracle_impl.h (it reuses tparams, lj_sigma, the LK pieces and the
 * block spheres), once per REAL.
 *
 * Restates:
 *   gen_pose_waters         tmol/score/lk_ball/potentials/gen_pose_waters.impl.hh:28-234
 *     (forward), :236-432 (backward); water.hh:29-98 (donor water),
 *     :323-420 (acceptor water), :422-595 (their derivatives);
 *     tmol/score/common/gen_coord.hh (build_coordinate);
 *     acceptor bases tmol/score/hbond/identification.hh:26-167 over the
 *     all-bonds walk of tmol/score/bonded_atom.hh:157-270;
 *   lk_ball_pose_score      lk_ball_pose_score.impl.hh:358-700 (forward),
 *     the block-pair / whole-pose backward; lk_ball.hh:42-494 (lk_fraction,
 *     lk_bridge_fraction, lk_ball_score), :1095-1272 (atom pair energy and
 *     derivatives), :1275-1361 (polar x occluder enumeration).
 *
 * The derivative of the acceptor water's frame is written here as a
 * vector-Jacobian product by the chain rule through the normalisations and
 * cross products, not as the reference's expanded closed form.
 *
 * Pinned by: tmol's lk_ball goldens (1ubq whole pose, jagged, block pair),
 * the potential-level golden values of tmol/tests/score/lk_ball/potentials/
 * test_compiled_lk_ball.py (water positions, lk_fraction, bridge fraction,
 * pair scores), central finite differences, and the reference's own CPU ops
 * built into oracle/_ref (tests/test_lk_ball_oracle.py).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, RSUF)

#ifndef LKB_MAX_WATER
#define LKB_MAX_WATER 4
#define LKB_TILE 32
typedef struct {
  int32_t n_poses, max_n_blocks, n_block_types, max_n_block_atoms, max_n_tiles,
      max_n_conn, max_n_all_bonds;
  const int32_t* rot_coord_offset;       /* [P*B] */
  const int32_t* block_type;             /* [P*B], -1 = padding */
  const int32_t* irc;                    /* [P, B, C, 2] inter-residue connections */
  const int32_t* min_bondsep;            /* [P, B, B] */
  const int32_t* inter_block_bondsep;    /* [P, B, B, C, C] */
  const int32_t* bt_n_atoms;             /* [T] */
  const int32_t* bt_n_conn;              /* [T] */
  const int32_t* bt_conn_atoms;          /* [T, C] */
  const int32_t* bt_all_bonds;           /* [T, NB, 3] */
  const int32_t* bt_atom_all_bond_ranges;/* [T, Am, 2] */
  const int32_t* tile_n_donH;            /* [T, nt] */
  const int32_t* tile_n_acc;             /* [T, nt] */
  const int32_t* tile_donH_inds;         /* [T, nt, 32] */
  const int32_t* tile_donH_hvy_inds;     /* [T, nt, 32] block-local donor heavy atom */
  const int32_t* tile_which_donH;        /* [T, nt, 32] */
  const int32_t* tile_acc_inds;          /* [T, nt, 32] */
  const int32_t* tile_acc_hyb;           /* [T, nt, 32] */
  const int32_t* tile_acc_nH;            /* [T, nt, 32] */
  const int32_t* is_hydrogen;            /* [T, Am] */
  const int32_t* tile_n_polar;           /* [T, nt] */
  const int32_t* tile_n_occ;             /* [T, nt] */
  const int32_t* tile_pol_occ_inds;      /* [T, nt, 32] */
  const int32_t* path_distance;          /* [T, Am, Am] */
  int32_t n_tors[3];                     /* sp2, sp3, ring torsion counts */
} oracle_lkb_stack;

/* a block-centric atom (bonded_atom.hh BlockCentricAtom) */
typedef struct {
  int block, block_type, atom;
} lkb_bca;

static int lkb_bca_eq(lkb_bca a, lkb_bca b) {
  return a.block == b.block && a.block_type == b.block_type && a.atom == b.atom;
}

/* k-th partner of atom a (RotamerCentricIndexedBonds::BondJIter) */
static lkb_bca lkb_partner(const oracle_lkb_stack* s, int p, lkb_bca a, int bidx) {
  const int32_t* e = s->bt_all_bonds + ((int64_t)a.block_type * s->max_n_all_bonds + bidx) * 3;
  lkb_bca out = {-1, -1, -1};
  if (e[1] >= 0) {
    out.block = a.block;
    out.block_type = a.block_type;
    out.atom = e[1];
    return out;
  }
  int B = s->max_n_blocks, C = s->max_n_conn;
  const int32_t* c = s->irc + (((int64_t)p * B + a.block) * C + e[2]) * 2;
  if (c[1] == -1) return out; /* undefined connection */
  out.block = c[0];
  out.block_type = s->block_type[p * B + c[0]];
  out.atom = s->bt_conn_atoms[out.block_type * C + c[1]];
  return out;
}

static void lkb_range(const oracle_lkb_stack* s, lkb_bca a, int* lo, int* hi) {
  const int32_t* r = s->bt_atom_all_bond_ranges + ((int64_t)a.block_type * s->max_n_block_atoms + a.atom) * 2;
  *lo = r[0];
  *hi = r[1];
}

/* hydrogen flag; an undefined partner (-1) reads as heavy */
static int lkb_is_h(const oracle_lkb_stack* s, lkb_bca a) {
  if (a.atom < 0) return 0;
  return s->is_hydrogen[(int64_t)a.block_type * s->max_n_block_atoms + a.atom] != 0;
}

/* acceptor bases (identification.hh:26-167); returns 0 when no water frame
 * exists (no B, or B0 == B) */
static int lkb_acc_bases(const oracle_lkb_stack* s, int p, lkb_bca A, int hyb, lkb_bca* Bo,
                         lkb_bca* B0o) {
  lkb_bca none = {-1, -1, -1}, B = none, B0 = none;
  int lo, hi;
  lkb_range(s, A, &lo, &hi);
  for (int k = lo; k < hi; ++k) {
    lkb_bca o = lkb_partner(s, p, A, k);
    if (!lkb_is_h(s, o)) {
      B = o;
      break;
    }
  }
  if (B.atom == -1) return 0;
  if (hyb == 1) { /* sp2 */
    int lo2, hi2;
    lkb_range(s, B, &lo2, &hi2);
    for (int k = lo2; k < hi2; ++k) {
      lkb_bca o = lkb_partner(s, p, B, k);
      if (B.block != A.block && o.block == A.block) continue;
      if (!lkb_bca_eq(o, A)) {
        B0 = o;
        break;
      }
    }
    if (B0.atom == -1) return 0;
  } else if (hyb == 2) { /* sp3 */
    for (int k = lo; k < hi; ++k) {
      lkb_bca o = lkb_partner(s, p, A, k);
      if (!lkb_bca_eq(o, B) && lkb_is_h(s, o)) {
        B0 = o;
        break;
      }
    }
    if (B0.atom == -1)
      for (int k = lo; k < hi; ++k) {
        lkb_bca o = lkb_partner(s, p, A, k);
        if (!lkb_bca_eq(o, B)) {
          B0 = o;
          break;
        }
      }
    if (B0.atom == -1) B0 = B;
  } else if (hyb == 3) { /* ring */
    for (int k = lo; k < hi; ++k) {
      lkb_bca o = lkb_partner(s, p, A, k);
      if (!lkb_bca_eq(o, B) && !lkb_is_h(s, o)) {
        B0 = o;
        break;
      }
    }
    if (B0.atom == -1) return 0;
  } else {
    return 0;
  }
  if (lkb_bca_eq(B0, B)) return 0;
  if (B0.atom < 0) return 0; /* an undefined connection: the reference asserts */
  *Bo = B;
  *B0o = B0;
  return 1;
}

static int64_t lkb_atom_index(const oracle_lkb_stack* s, int p, lkb_bca a) {
  return (int64_t)s->rot_coord_offset[p * s->max_n_blocks + a.block] + a.atom;
}
#endif /* LKB_MAX_WATER */

/* ---------- water builders -------------------------------------------- */

/* donor water, water.hh:44-46: D + dist * (H - D) / |H - D| */
static void FN(lkb_don_water)(const REAL* D, const REAL* H, REAL dist, REAL* W) {
  REAL v[3] = {H[0] - D[0], H[1] - D[1], H[2] - D[2]};
  REAL n = SQRT(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; ++k) W[k] = D[k] + dist * (v[k] / n);
}

/* vector-Jacobian product of the donor water: g = dE/dW -> dE/dD, dE/dH
 * (added).  W = D + d u, u = v / |v|, v = H - D:
 * dE/dv = d / |v| (g - u (u . g)); dE/dH = dE/dv; dE/dD = g - dE/dv. */
static void FN(lkb_don_water_vjp)(const REAL* D, const REAL* H, REAL dist, const REAL* g,
                                  REAL* gD, REAL* gH) {
  REAL v[3] = {H[0] - D[0], H[1] - D[1], H[2] - D[2]};
  REAL n = SQRT(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  REAL u[3] = {v[0] / n, v[1] / n, v[2] / n};
  REAL ug = u[0] * g[0] + u[1] * g[1] + u[2] * g[2];
  for (int k = 0; k < 3; ++k) {
    REAL gv = dist / n * (g[k] - u[k] * ug);
    gH[k] += gv;
    gD[k] += g[k] - gv;
  }
}

static void FN(lkb_cross)(const REAL* a, const REAL* b, REAL* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* the frame of build_coordinate (gen_coord.hh): e0 = (A - B) / |A - B|,
 * e2 = e0 x (C - B) normalised (a fixed perpendicular when collinear),
 * e1 = e2 x e0; returns |e0 x (C - B)| before normalisation */
static REAL FN(lkb_frame)(const REAL* A, const REAL* B, const REAL* C, REAL* e0, REAL* e1,
                          REAL* e2) {
  REAL ab[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
  REAL nab = SQRT(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
  for (int k = 0; k < 3; ++k) e0[k] = ab[k] / nab;
  REAL cb[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
  FN(lkb_cross)(e0, cb, e2);
  REAL n2 = SQRT(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
  REAL nraw = n2;
  if (n2 == 0) {
    REAL c1[3] = {1, 0, 0};
    if (e0[0] == 1 || e0[0] == -1) {
      c1[0] = 0;
      c1[1] = 1;
    }
    FN(lkb_cross)(e0, c1, e2);
    n2 = SQRT(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
  }
  for (int k = 0; k < 3; ++k) e2[k] /= n2;
  FN(lkb_cross)(e2, e0, e1);
  return nraw;
}

/* acceptor water, build_coordinate::V(A, B, C, dist, angle, torsion) */
static void FN(lkb_acc_water)(const REAL* A, const REAL* B, const REAL* C, REAL dist,
                              REAL angle, REAL tors, REAL* W) {
  const REAL pi = (REAL)3.14159265358979323846;
  REAL e0[3], e1[3], e2[3];
  FN(lkb_frame)(A, B, C, e0, e1, e2);
  REAL k0 = dist * COS(pi - angle);
  REAL s = dist * SIN(pi - angle);
  REAL k1 = s * COS(tors), k2 = s * SIN(tors);
  for (int k = 0; k < 3; ++k) W[k] = (e0[k] * k0 + e1[k] * k1 + e2[k] * k2) + A[k];
}

/* vector-Jacobian product of build_coordinate: g = dE/dW -> dE/dA, dE/dB,
 * dE/dC (added).  Collinear A, B, C: dW/dA = I, nothing to B and C (the
 * reference's dV convention). */
static void FN(lkb_acc_water_vjp)(const REAL* A, const REAL* B, const REAL* C, REAL dist,
                                  REAL angle, REAL tors, const REAL* g, REAL* gA, REAL* gB,
                                  REAL* gC) {
  const REAL pi = (REAL)3.14159265358979323846;
  REAL e0[3], e1[3], e2[3];
  REAL nraw = FN(lkb_frame)(A, B, C, e0, e1, e2);
  for (int k = 0; k < 3; ++k) gA[k] += g[k];
  if (nraw == 0) return;
  REAL k0 = dist * COS(pi - angle);
  REAL s = dist * SIN(pi - angle);
  REAL k1 = s * COS(tors), k2 = s * SIN(tors);
  REAL g0[3], g1[3], g2[3], t[3];
  for (int k = 0; k < 3; ++k) {
    g0[k] = k0 * g[k];
    g1[k] = k1 * g[k];
    g2[k] = k2 * g[k];
  }
  /* e1 = e2 x e0 */
  FN(lkb_cross)(e0, g1, t);
  for (int k = 0; k < 3; ++k) g2[k] += t[k];
  FN(lkb_cross)(g1, e2, t);
  for (int k = 0; k < 3; ++k) g0[k] += t[k];
  /* e2 = n / |n|, n = e0 x (C - B) */
  REAL d2 = e2[0] * g2[0] + e2[1] * g2[1] + e2[2] * g2[2];
  REAL gn[3];
  for (int k = 0; k < 3; ++k) gn[k] = (g2[k] - e2[k] * d2) / nraw;
  REAL cb[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
  REAL gcb[3];
  FN(lkb_cross)(cb, gn, t);
  for (int k = 0; k < 3; ++k) g0[k] += t[k];
  FN(lkb_cross)(gn, e0, gcb);
  /* e0 = (A - B) / |A - B| */
  REAL ab[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
  REAL nab = SQRT(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
  REAL d0 = e0[0] * g0[0] + e0[1] * g0[1] + e0[2] * g0[2];
  for (int k = 0; k < 3; ++k) {
    REAL gab = (g0[k] - e0[k] * d0) / nab;
    gA[k] += gab;
    gB[k] += -gab - gcb[k];
    gC[k] += gcb[k];
  }
}

/* exported single-water builders (tests): W[3]; with g != NULL the
 * vector-Jacobian product into gP, gG, gGG (donor: gP = dE/dD, gG = dE/dH) */
void FN(oracle_lkb_build_water)(int kind, const REAL* P, const REAL* G, const REAL* GG,
                                REAL dist, REAL angle, REAL tors, REAL* W, const REAL* g,
                                REAL* gP, REAL* gG, REAL* gGG) {
  if (kind == 0) {
    FN(lkb_don_water)(P, G, dist, W);
    if (g) FN(lkb_don_water_vjp)(P, G, dist, g, gP, gG);
    return;
  }
  FN(lkb_acc_water)(P, G, GG, dist, angle, tors, W);
  if (g) FN(lkb_acc_water_vjp)(P, G, GG, dist, angle, tors, g, gP, gG, gGG);
}

/* ---------- potentials ------------------------------------------------- */

/* lk_fraction::V / dV (lk_ball.hh:42-124).  WI [4][3] with NaN rows for
 * absent waters; dWI [4][3], dJ [3] overwritten when non-NULL. */
static REAL FN(lkb_fraction)(const REAL* WI, const REAL* J, REAL lj_radius_j, REAL* dWI,
                             REAL* dJ) {
  const REAL ramp = (REAL)3.709;
  REAL t = (REAL)(1.4 + lj_radius_j);
  REAL d2_low = t * t - ramp;
  if (d2_low < 0) d2_low = 0;
  REAL sum = 0, sJ[3] = {0, 0, 0}, sW[LKB_MAX_WATER][3];
  for (int w = 0; w < LKB_MAX_WATER; ++w) {
    REAL dx = J[0] - WI[3 * w], dy = J[1] - WI[3 * w + 1], dz = J[2] - WI[3 * w + 2];
    REAL d2 = (dx * dx + dy * dy + dz * dz) - d2_low;
    sW[w][0] = sW[w][1] = sW[w][2] = 0;
    if (d2 == d2) {
      REAL e = EXP(-d2);
      sum += e;
      sJ[0] += 2 * e * dx;
      sJ[1] += 2 * e * dy;
      sJ[2] += 2 * e * dz;
      sW[w][0] = -2 * e * dx;
      sW[w][1] = -2 * e * dy;
      sW[w][2] = -2 * e * dz;
    }
  }
  REAL wd = -LOG(sum);
  REAL frac = 0;
  if (wd < 0)
    frac = 1;
  else if (wd < ramp) {
    REAL q = wd / ramp;
    frac = (1 - q * q) * (1 - q * q);
  }
  if (dWI) {
    REAL dfr = 0;
    if (wd > 0 && wd < ramp) dfr = -4 * wd * (ramp * ramp - wd * wd) / ((ramp * ramp) * (ramp * ramp));
    /* divide (not multiply by 1 / sum): sum can be subnormal */
    REAL div = sum != 0 ? sum : 1;
    for (int k = 0; k < 3; ++k) dJ[k] = sJ[k] / div * dfr;
    for (int w = 0; w < LKB_MAX_WATER; ++w)
      for (int k = 0; k < 3; ++k) dWI[3 * w + k] = sW[w][k] / div * dfr;
  }
  return frac;
}

/* lk_bridge_fraction::V / dV (lk_ball.hh:126-277) */
static REAL FN(lkb_bridge)(const REAL* I, const REAL* J, const REAL* WI, const REAL* WJ,
                           REAL water_dist, REAL* dI, REAL* dJ, REAL* dWI, REAL* dWJ) {
  const REAL gap = (REAL)0.5, width = (REAL)2.6, aw = (REAL)(2.8 * (REAL)2.6);
  REAL sum = 0, sWI[LKB_MAX_WATER][3], sWJ[LKB_MAX_WATER][3];
  for (int w = 0; w < LKB_MAX_WATER; ++w)
    for (int k = 0; k < 3; ++k) sWI[w][k] = sWJ[w][k] = 0;
  for (int wi = 0; wi < LKB_MAX_WATER; ++wi)
    for (int wj = 0; wj < LKB_MAX_WATER; ++wj) {
      REAL d[3] = {WI[3 * wi] - WJ[3 * wj], WI[3 * wi + 1] - WJ[3 * wj + 1],
                   WI[3 * wi + 2] - WJ[3 * wj + 2]};
      REAL d2 = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - gap;
      if (d2 == d2) {
        REAL e = EXP(-d2);
        sum += e;
        for (int k = 0; k < 3; ++k) {
          sWI[wi][k] += 2 * e * d[k];
          sWJ[wj][k] -= 2 * e * d[k];
        }
      }
    }
  REAL wd = -LOG(sum);
  REAL ofrac, dof = 0;
  if (wd > width)
    ofrac = 0;
  else if (wd > 0) {
    REAL q = wd / width;
    ofrac = (1 - q * q) * (1 - q * q);
    dof = -4 * wd * (width * width - wd * wd) / ((width * width) * (width * width));
  } else
    ofrac = 1;
  REAL target = (REAL)(8.0 / 3.0 * (water_dist * water_dist));
  REAL ij[3] = {I[0] - J[0], I[1] - J[1], I[2] - J[2]};
  REAL len2 = ij[0] * ij[0] + ij[1] * ij[1] + ij[2] * ij[2];
  REAL bd = len2 - target;
  REAL abd = bd < 0 ? -bd : bd;
  REAL afrac, daf = 0;
  if (abd > aw)
    afrac = 0;
  else {
    REAL q = bd / aw;
    afrac = (1 - q * q) * (1 - q * q);
    if (abd > 0) daf = -4 * bd * (aw * aw - bd * bd) / ((aw * aw) * (aw * aw));
  }
  if (dI) {
    REAL div = sum != 0 ? sum : 1;
    for (int k = 0; k < 3; ++k) {
      dI[k] = ofrac * daf * (2 * ij[k]);
      dJ[k] = ofrac * daf * (-2 * ij[k]);
    }
    for (int w = 0; w < LKB_MAX_WATER; ++w)
      for (int k = 0; k < 3; ++k) {
        dWI[3 * w + k] = afrac * dof * (sWI[w][k] / div);
        dWJ[3 * w + k] = afrac * dof * (sWJ[w][k] / div);
      }
  }
  return ofrac * afrac;
}

/* exported fractions (tests): lk_fraction and lk_bridge_fraction with
 * optional derivatives (NULL to skip) */
REAL FN(oracle_lkb_fraction)(const REAL* WI, const REAL* J, REAL lj_radius_j, REAL* dWI,
                             REAL* dJ) {
  return FN(lkb_fraction)(WI, J, lj_radius_j, dWI, dJ);
}

REAL FN(oracle_lkb_bridge)(const REAL* I, const REAL* J, const REAL* WI, const REAL* WJ,
                           REAL water_dist, REAL* dI, REAL* dJ, REAL* dWI, REAL* dWJ) {
  return FN(lkb_bridge)(I, J, WI, WJ, water_dist, dI, dJ, dWI, dWJ);
}

/* one direction of LK energy, lk_isotropic_pair::V (lk_isotropic.hh:104-160) */
static REAL FN(lkb_iso_V)(REAL dist, int sep, REAL sigma, REAL r_i, REAL dg_i, REAL lam_i,
                          REAL vol_j, REAL max_dis, int is_cc) {
  REAL d_min, cdmin, cdmax;
  FN(lk_knots)(sigma, is_cc, &d_min, &cdmin, &cdmax);
  REAL far_dmax = max_dis, far_dmin = max_dis - (REAL)1.5;
  REAL lk;
  if (dist > far_dmax)
    lk = 0;
  else if (dist > far_dmin) {
    REAL v, dv;
    FN(fdesolv_V_dV)(far_dmin, r_i, dg_i, lam_i, vol_j, &v, &dv);
    lk = FN(interp0)(dist, far_dmin, v, dv, far_dmax);
  } else if (dist > cdmax)
    lk = FN(fdesolv_V)(dist, r_i, dg_i, lam_i, vol_j);
  else if (dist > cdmin) {
    REAL v, dv;
    FN(fdesolv_V_dV)(cdmax, r_i, dg_i, lam_i, vol_j, &v, &dv);
    lk = FN(interp)(dist, cdmin, FN(fdesolv_V)(d_min, r_i, dg_i, lam_i, vol_j), 0, cdmax, v, dv);
  } else
    lk = FN(fdesolv_V)(d_min, r_i, dg_i, lam_i, vol_j);
  return FN(cp_weight)(sep) * lk;
}

/* lk_ball type params (9): lj_radius, lk_dgfree, lk_lambda, lk_volume,
 * is_donor, is_hydroxyl, is_polarh, is_acceptor, is_carbon_lk */
static void FN(lkb_tparams)(const REAL* p, FN(tparams) * t) {
  t->lj_radius = p[0];
  t->lj_wdepth = 0;
  t->lk_dgfree = p[1];
  t->lk_lambda = p[2];
  t->lk_volume = p[3];
  t->is_donor = p[4];
  t->is_hydroxyl = p[5];
  t->is_polarh = p[6];
  t->is_acceptor = p[7];
  t->is_carbon_lk = p[8];
}

/* lk_ball_score::V / dV (lk_ball.hh:335-494) for polar I against occluder J
 * at count-pair separation sep (> 3) and distance below the threshold.
 * gp = lj_hbond_dis, lj_hbond_OH_donor_dis, lj_hbond_hdis, lkb_water_dist,
 * max_dis.  E[4] overwritten; with d != NULL, d = dI[4][3], dJ[4][3],
 * dWI[4][4][3], dWJ[4][4][3] (score type major) overwritten. */
void FN(oracle_lkb_pair)(const REAL* I, const REAL* J, const REAL* WI, const REAL* WJ, int sep,
                         const REAL* pi, const REAL* pj, const REAL* gp, REAL* E, REAL* d) {
  for (int k = 0; k < 4; ++k) E[k] = 0;
  if (d)
    for (int k = 0; k < 4 * 3 * 2 + 4 * 12 * 2; ++k) d[k] = 0;
  if (!(WI[0] == WI[0])) return;
  FN(tparams) ti, tj;
  FN(lkb_tparams)(pi, &ti);
  FN(lkb_tparams)(pj, &tj);
  FN(gparams) g;
  g.max_dis = gp[4];
  g.lj_dlin_sigma_factor = 0;
  g.lj_hbond_dis = gp[0];
  g.lj_hbond_OH_donor_dis = gp[1];
  g.lj_hbond_hdis = gp[2];
  REAL sigma = FN(lj_sigma)(&ti, &tj, &g);
  REAL v[3] = {I[0] - J[0], I[1] - J[1], I[2] - J[2]};
  REAL dist = SQRT(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  int is_cc = ti.is_carbon_lk != 0 && tj.is_carbon_lk != 0;
  int bridge = tj.is_donor != 0 || tj.is_acceptor != 0;
  if (!d) {
    REAL iso = FN(lkb_iso_V)(dist, sep, sigma, ti.lj_radius, ti.lk_dgfree, ti.lk_lambda,
                             tj.lk_volume, gp[4], is_cc);
    REAL fr = FN(lkb_fraction)(WI, J, tj.lj_radius, NULL, NULL);
    REAL ov = bridge ? FN(lkb_bridge)(I, J, WI, WJ, gp[3], NULL, NULL, NULL, NULL) : 0;
    E[0] = iso;
    E[1] = iso * fr;
    E[2] = iso * ov;
    E[3] = ov / 2;
    return;
  }
  REAL iso, diso;
  FN(lk_pair_V_dV)(dist, sep, sigma, ti.lj_radius, ti.lk_dgfree, ti.lk_lambda, tj.lk_volume,
                   gp[4], is_cc, &iso, &diso);
  REAL fdWI[12], fdJ[3];
  REAL fr = FN(lkb_fraction)(WI, J, tj.lj_radius, fdWI, fdJ);
  REAL bI[3] = {0, 0, 0}, bJ[3] = {0, 0, 0}, bWI[12], bWJ[12];
  for (int k = 0; k < 12; ++k) bWI[k] = bWJ[k] = 0;
  REAL ov = bridge ? FN(lkb_bridge)(I, J, WI, WJ, gp[3], bI, bJ, bWI, bWJ) : 0;
  E[0] = iso;
  E[1] = iso * fr;
  E[2] = iso * ov;
  E[3] = ov / 2;
  REAL* dI = d;
  REAL* dJ = d + 12;
  REAL* dWI = d + 24;
  REAL* dWJ = d + 24 + 48;
  for (int k = 0; k < 3; ++k) {
    REAL ddI = diso * (v[k] / dist), ddJ = -ddI;
    dI[0 + k] = ddI;
    dI[3 + k] = ddI * fr;
    dI[6 + k] = ddI * ov + iso * bI[k];
    dI[9 + k] = bI[k] / 2;
    dJ[0 + k] = ddJ;
    dJ[3 + k] = ddJ * fr + iso * fdJ[k];
    dJ[6 + k] = ddJ * ov + iso * bJ[k];
    dJ[9 + k] = bJ[k] / 2;
  }
  for (int k = 0; k < 12; ++k) {
    dWI[12 + k] = iso * fdWI[k];
    dWI[24 + k] = iso * bWI[k];
    dWI[36 + k] = bWI[k] / 2;
    dWJ[24 + k] = iso * bWJ[k];
    dWJ[36 + k] = bWJ[k] / 2;
  }
}

/* ---------- gen_pose_waters -------------------------------------------- */

/* per water: calls f(ctx, p, kind, target, slot, atoms[3], tors index)
 * with kind 0 donor (atoms D, H), 1 sp2 / 2 sp3 / 3 ring (A, B, B0) */
typedef void (*FN(lkb_water_fn))(void* ctx, int kind, int64_t target, int slot,
                                 const int64_t* atoms, int tors_ind);

static void FN(lkb_each_water)(const oracle_lkb_stack* s, void* ctx, FN(lkb_water_fn) f) {
  int B = s->max_n_blocks, nt = s->max_n_tiles;
  for (int p = 0; p < s->n_poses; ++p)
    for (int b = 0; b < B; ++b) {
      int bt = s->block_type[p * B + b];
      if (bt < 0) continue;
      int n = s->bt_n_atoms[bt];
      int64_t off = s->rot_coord_offset[p * B + b];
      for (int t = 0; t * LKB_TILE < n; ++t) {
        int64_t tb = ((int64_t)bt * nt + t) * LKB_TILE;
        for (int k = 0; k < s->tile_n_donH[bt * nt + t]; ++k) {
          int H = t * LKB_TILE + s->tile_donH_inds[tb + k];
          int D = s->tile_donH_hvy_inds[tb + k];
          int64_t atoms[3] = {off + D, off + H, -1};
          f(ctx, 0, off + D, s->tile_which_donH[tb + k], atoms, 0);
        }
        for (int k = 0; k < s->tile_n_acc[bt * nt + t]; ++k) {
          int hyb = s->tile_acc_hyb[tb + k];
          if (hyb < 1 || hyb > 3) continue;
          lkb_bca A = {b, bt, t * LKB_TILE + s->tile_acc_inds[tb + k]}, Bb, B0;
          if (!lkb_acc_bases(s, p, A, hyb, &Bb, &B0)) continue;
          int64_t atoms[3] = {off + A.atom, lkb_atom_index(s, p, Bb), lkb_atom_index(s, p, B0)};
          for (int w = 0; w < s->n_tors[hyb - 1]; ++w)
            f(ctx, hyb, off + A.atom, w + s->tile_acc_nH[tb + k], atoms, w);
        }
      }
    }
}

typedef struct {
  const REAL* coords;
  const REAL* wg; /* lkb_water_dist, angle_sp2, angle_sp3, angle_ring */
  const REAL* tors[3];
  REAL* waters;   /* [NA, 4, 3] */
  const REAL* dW; /* backward: dE/dW */
  REAL* dX;       /* backward: dE/dx (added) */
} FN(lkb_wctx);

static void FN(lkb_frame_atoms)(FN(lkb_wctx) * c, int kind, const int64_t* at, REAL* P, REAL* G,
                                REAL* GG) {
  for (int k = 0; k < 3; ++k) {
    P[k] = c->coords[3 * at[0] + k];
    G[k] = c->coords[3 * at[1] + k];
    GG[k] = c->coords[3 * at[2] + k];
  }
  if (kind == 3)
    for (int k = 0; k < 3; ++k) G[k] = (G[k] + GG[k]) / 2;
}

static void FN(lkb_gen_one)(void* vc, int kind, int64_t target, int slot, const int64_t* at,
                            int ti) {
  FN(lkb_wctx)* c = (FN(lkb_wctx)*)vc;
  REAL* W = c->waters + (target * LKB_MAX_WATER + slot) * 3;
  if (kind == 0) {
    FN(lkb_don_water)(c->coords + 3 * at[0], c->coords + 3 * at[1], c->wg[0], W);
    return;
  }
  REAL P[3], G[3], GG[3];
  FN(lkb_frame_atoms)(c, kind, at, P, G, GG);
  FN(lkb_acc_water)(P, G, GG, c->wg[0], c->wg[kind], c->tors[kind - 1][ti], W);
}

static void FN(lkb_dgen_one)(void* vc, int kind, int64_t target, int slot, const int64_t* at,
                             int ti) {
  FN(lkb_wctx)* c = (FN(lkb_wctx)*)vc;
  const REAL* g = c->dW + (target * LKB_MAX_WATER + slot) * 3;
  if (kind == 0) {
    FN(lkb_don_water_vjp)(c->coords + 3 * at[0], c->coords + 3 * at[1], c->wg[0], g,
                          c->dX + 3 * at[0], c->dX + 3 * at[1]);
    return;
  }
  REAL P[3], G[3], GG[3], gP[3] = {0, 0, 0}, gG[3] = {0, 0, 0}, gGG[3] = {0, 0, 0};
  FN(lkb_frame_atoms)(c, kind, at, P, G, GG);
  FN(lkb_acc_water_vjp)(P, G, GG, c->wg[0], c->wg[kind], c->tors[kind - 1][ti], g, gP, gG, gGG);
  for (int k = 0; k < 3; ++k) {
    c->dX[3 * at[0] + k] += gP[k];
    if (kind == 3) {
      c->dX[3 * at[1] + k] += (REAL)0.5 * gG[k];
      c->dX[3 * at[2] + k] += (REAL)0.5 * gG[k] + gGG[k];
    } else {
      c->dX[3 * at[1] + k] += gG[k];
      c->dX[3 * at[2] + k] += gGG[k];
    }
  }
}

/* waters [NA, 4, 3]: NaN where no water is built */
void FN(oracle_lkb_waters)(const oracle_lkb_stack* s, const REAL* coords, int64_t na,
                           const REAL* wg, const REAL* sp2, const REAL* sp3, const REAL* ring,
                           REAL* waters) {
  for (int64_t i = 0; i < na * LKB_MAX_WATER * 3; ++i) waters[i] = (REAL)NAN;
  FN(lkb_wctx) c = {coords, wg, {sp2, sp3, ring}, waters, NULL, NULL};
  FN(lkb_each_water)(s, &c, FN(lkb_gen_one));
}

/* dE/dx += J^T dE/dW over every built water */
void FN(oracle_lkb_waters_backward)(const oracle_lkb_stack* s, const REAL* coords,
                                    const REAL* wg, const REAL* sp2, const REAL* sp3,
                                    const REAL* ring, const REAL* dW, REAL* dX) {
  FN(lkb_wctx) c = {coords, wg, {sp2, sp3, ring}, NULL, dW, dX};
  FN(lkb_each_water)(s, &c, FN(lkb_dgen_one));
}

/* ---------- lk_ball_pose_score ----------------------------------------- */

/* inter-block separation for lk_ball: the count-pair rule of the pose score
 * (lk_ball_pose_score.impl.hh:338-352, max important separation 4), with
 * lk_ball's path distances and no ligand rule */
static int FN(lkb_inter_sep)(const oracle_lkb_stack* s, int p, int b1, int b2, int bt1, int bt2,
                             int a1, int a2) {
  int B = s->max_n_blocks, C = s->max_n_conn, Am = s->max_n_block_atoms;
  int sep = s->min_bondsep[((int64_t)p * B + b1) * B + b2];
  if (sep <= 4) {
    int best = 5;
    for (int c1 = 0; c1 < s->bt_n_conn[bt1]; ++c1) {
      int ca1 = s->bt_conn_atoms[bt1 * C + c1];
      int pd1 = s->path_distance[((int64_t)bt1 * Am + ca1) * Am + a1];
      for (int c2 = 0; c2 < s->bt_n_conn[bt2]; ++c2) {
        int ca2 = s->bt_conn_atoms[bt2 * C + c2];
        int pd2 = s->path_distance[((int64_t)bt2 * Am + ca2) * Am + a2];
        int cs = s->inter_block_bondsep[((((int64_t)p * B + b1) * B + b2) * C + c1) * C + c2];
        if (pd1 + cs + pd2 < best) best = pd1 + cs + pd2;
      }
    }
    sep = best;
  }
  return sep;
}

/* one (polar, occluder) pair: energies into E4 (added), derivatives with
 * weights w4 into dX / dW (added) */
static void FN(lkb_pair_acc)(const REAL* coords, const REAL* waters, int64_t ip, int64_t jo,
                             int sep, const REAL* pp, const REAL* po, const REAL* gp, REAL* E4,
                             const REAL* w4, REAL* dX, REAL* dW) {
  if (sep <= 3) return;
  const REAL* I = coords + 3 * ip;
  const REAL* J = coords + 3 * jo;
  REAL v[3] = {I[0] - J[0], I[1] - J[1], I[2] - J[2]};
  REAL dist = SQRT(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (dist >= gp[4]) return;
  REAL E[4], d[4 * 3 * 2 + 4 * 12 * 2];
  FN(oracle_lkb_pair)(I, J, waters + 12 * ip, waters + 12 * jo, sep, pp, po, gp, E,
                      w4 ? d : NULL);
  if (E4)
    for (int k = 0; k < 4; ++k) E4[k] += E[k];
  if (!w4) return;
  for (int st = 0; st < 4; ++st) {
    for (int k = 0; k < 3; ++k) {
      dX[3 * ip + k] += w4[st] * d[3 * st + k];
      dX[3 * jo + k] += w4[st] * d[12 + 3 * st + k];
    }
    for (int k = 0; k < 12; ++k) {
      dW[12 * ip + k] += w4[st] * d[24 + 12 * st + k];
      dW[12 * jo + k] += w4[st] * d[72 + 12 * st + k];
    }
  }
}

/* The pose score over neighbor block pairs nbr [P, B, B] (b1 <= b2).
 *   energies: [4, P] or [4, P, B, B] (block_pair), overwritten; may be NULL
 *   weights:  dTdV, same shape as energies, or NULL (no derivatives)
 *   dX [NA, 3], dW [NA, 4, 3]: overwritten with weights */
void FN(oracle_lkb_score)(const oracle_lkb_stack* s, const REAL* coords, const REAL* waters,
                          int64_t na, const REAL* bt_params, const REAL* gp, const int32_t* nbr,
                          int block_pair, REAL* energies, const REAL* weights, REAL* dX,
                          REAL* dW) {
  int P = s->n_poses, B = s->max_n_blocks, nt = s->max_n_tiles, Am = s->max_n_block_atoms;
  int64_t nE = block_pair ? (int64_t)P * B * B : P;
  if (energies)
    for (int64_t i = 0; i < 4 * nE; ++i) energies[i] = 0;
  if (weights) {
    for (int64_t i = 0; i < na * 3; ++i) dX[i] = 0;
    for (int64_t i = 0; i < na * 12; ++i) dW[i] = 0;
  }
  for (int p = 0; p < P; ++p)
    for (int b1 = 0; b1 < B; ++b1)
      for (int b2 = b1; b2 < B; ++b2) {
        if (!nbr[((int64_t)p * B + b1) * B + b2]) continue;
        int bt1 = s->block_type[p * B + b1], bt2 = s->block_type[p * B + b2];
        if (bt1 < 0 || bt2 < 0) continue;
        int64_t cell = block_pair ? ((int64_t)p * B + b1) * B + b2 : p;
        REAL E4[4] = {0, 0, 0, 0}, w4[4];
        if (weights)
          for (int k = 0; k < 4; ++k) w4[k] = weights[k * nE + cell];
        int64_t o1 = s->rot_coord_offset[p * B + b1], o2 = s->rot_coord_offset[p * B + b2];
        int n1 = s->bt_n_atoms[bt1], n2 = s->bt_n_atoms[bt2];
        /* (polar of x) x (occluder of y) for (x, y) = (b1, b2) and (b2, b1);
         * intra-block: every ordered pair of distinct atoms */
        for (int dir = 0; dir < (b1 == b2 ? 1 : 2); ++dir) {
          int bx = dir ? b2 : b1, by = dir ? b1 : b2;
          int btx = dir ? bt2 : bt1, bty = dir ? bt1 : bt2;
          int nx = dir ? n2 : n1, ny = dir ? n1 : n2;
          int64_t ox = dir ? o2 : o1, oy = dir ? o1 : o2;
          for (int tx = 0; tx * LKB_TILE < nx; ++tx)
            for (int kx = 0; kx < s->tile_n_polar[btx * nt + tx]; ++kx) {
              int64_t ix = ((int64_t)btx * nt + tx) * LKB_TILE + kx;
              int ax = tx * LKB_TILE + s->tile_pol_occ_inds[ix];
              for (int ty = 0; ty * LKB_TILE < ny; ++ty)
                for (int ky = 0; ky < s->tile_n_occ[bty * nt + ty]; ++ky) {
                  int64_t iy = ((int64_t)bty * nt + ty) * LKB_TILE + ky;
                  int ay = ty * LKB_TILE + s->tile_pol_occ_inds[iy];
                  int sep;
                  if (bx == by) {
                    if (ax == ay) continue;
                    sep = s->path_distance[((int64_t)btx * Am + ax) * Am + ay];
                  } else {
                    sep = dir ? FN(lkb_inter_sep)(s, p, b1, b2, bt1, bt2, ay, ax)
                              : FN(lkb_inter_sep)(s, p, b1, b2, bt1, bt2, ax, ay);
                  }
                  FN(lkb_pair_acc)(coords, waters, ox + ax, oy + ay, sep, bt_params + 9 * ix,
                                   bt_params + 9 * iy, gp, E4, weights ? w4 : NULL, dX, dW);
                }
            }
        }
        if (energies)
          for (int k = 0; k < 4; ++k) energies[k * nE + cell] += E4[k];
      }
}

#undef FN
#undef CAT
#undef CAT_
