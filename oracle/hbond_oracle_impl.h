/* CPU restatement of the reference hbond ops -- TEST INFRASTRUCTURE.
 *
 * The checker, never the product: only tests/, the smoke check and bench.py's
 * cpu_baseline leg load it.  Included by ljlk_oracle.c after
 * lkball_oracle_impl.h (it reuses the lk_ball stack view, its all-bonds walk
 * and the block spheres / neighbors), once per REAL.
 *
 * Restates:
 *   gen_hbond_bases     tmol/score/hbond/potentials/gen_hbond_bases.impl.hh:200-468
 *     (donor base: the last heavy partner of the polar hydrogen; acceptor
 *     bases B, B0 by tmol/score/hbond/identification.hh:26-198 over the
 *     all-bonds walk of tmol/score/bonded_atom.hh:157-270; a missing base
 *     reads as (0, 0, 0) with index -1, atoms with no role stay NaN / -1);
 *   hbond_pose_scores   hbond_pose_score.impl.hh:47-200 (count pair: an
 *     inter-block pair's separation is looked up with the donor's and the
 *     acceptor's positions in their tile lists as the atom indices, as the
 *     reference does; intra-block pairs use the true atoms' path distance;
 *     pairs below separation 5 score 0), :430-720 (neighbors at 5.5 A, upper
 *     triangle, tiles of 32 atoms), :860-990 (block-pair backward);
 *     hbond.hh:640-913 (H-A cutoff max_ha_dis, the five-atom energy and its
 *     derivatives through derived_atom_inds);
 *   the potential      potentials.hh:36-508 (A-H distance, AHD angle, BAH
 *     angle per hybridization with the sp3 soft-max, the sp2 B0-B-A-H chi
 *     term, the donor x acceptor weight and the [-0.1, 0.1] fade);
 *     tmol/score/common/polynomial.hh (bounded degree-10 polynomials in
 *     double), geom.hh (distance, interior angles, the Blondel-Karplus
 *     dihedral derivatives).
 *
 * Pinned by: tmol's 1ubq hbond baselines (whole pose x10, jagged, block
 * pair), the point scores of tmol/tests/score/hbond/potentials/
 * test_potentials.py (sp2 -2.40, sp3 -2.00, ring -2.17), central finite
 * differences of the potential, and the reference's own CPU ops built into
 * oracle/_ref (tests/test_hbond_oracle.py).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, RSUF)

#ifndef HB_POLY_WIDTH
#define HB_POLY_WIDTH 16 /* 11 coefficients (c10 first), pad, range[2], bound[2] */
#define HB_TILE 32
#define HB_MAX_SEP 4     /* max important bond separation */
typedef struct {
  oracle_lkb_stack base;        /* offsets, types, bonds, tiles, path distances */
  const int32_t* tile_don_type; /* [T, nt, 32] donor type of each tile donor H */
  const int32_t* tile_acc_type; /* [T, nt, 32] acceptor type of each tile acceptor */
  int32_t n_don_types, n_acc_types;
} oracle_hb_stack;

/* base atoms of a donor hydrogen: the last heavy partner
 * (identification.hh:176-193) */
static lkb_bca hb_donor_base(const oracle_lkb_stack* s, int p, lkb_bca H) {
  lkb_bca D = {-1, -1, -1};
  int lo, hi;
  lkb_range(s, H, &lo, &hi);
  for (int k = lo; k < hi; ++k) {
    lkb_bca o = lkb_partner(s, p, H, k);
    if (o.atom >= 0 && !lkb_is_h(s, o)) D = o;
  }
  return D;
}

/* acceptor bases for hbond (identification.hh:26-167): sp2 and ring
 * acceptors without B or B0 get neither; sp3 falls back to B0 = B */
static void hb_acc_bases(const oracle_lkb_stack* s, int p, lkb_bca A, int hyb, lkb_bca* Bo,
                         lkb_bca* B0o) {
  lkb_bca none = {-1, -1, -1}, B = none, B0 = none;
  *Bo = none;
  *B0o = none;
  if (hyb < 1 || hyb > 3) return;
  int lo, hi;
  lkb_range(s, A, &lo, &hi);
  for (int k = lo; k < hi; ++k) {
    lkb_bca o = lkb_partner(s, p, A, k);
    if (!lkb_is_h(s, o)) {
      B = o;
      break;
    }
  }
  if (hyb == 1) {
    if (B.atom == -1) return;
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
    if (B0.atom == -1) return;
  } else if (hyb == 2) {
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
  } else {
    if (B.atom == -1) return;
    for (int k = lo; k < hi; ++k) {
      lkb_bca o = lkb_partner(s, p, A, k);
      if (!lkb_bca_eq(o, B) && !lkb_is_h(s, o)) {
        B0 = o;
        break;
      }
    }
    if (B0.atom == -1) return;
  }
  *Bo = B;
  *B0o = B0;
}
#endif /* HB_POLY_WIDTH */

/* ---------- geometry (geom.hh), REAL ----------------------------------- */

static void FN(hb_sub)(const REAL* a, const REAL* b, REAL* o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static REAL FN(hb_dot)(const REAL* a, const REAL* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static REAL FN(hb_norm)(const REAL* a) { return SQRT(FN(hb_dot)(a, a)); }
static void FN(hb_cross)(const REAL* a, const REAL* b, REAL* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* interior angle between vectors a and b and its gradients
 * (geom.hh interior_angle: 2 atan2(|a x b|, |a||b| + a.b)) */
static REAL FN(hb_interior)(const REAL* a, const REAL* b, REAL* da, REAL* db) {
  REAL cr[3], z[3];
  FN(hb_cross)(a, b, cr);
  REAL n2 = FN(hb_dot)(cr, cr);
  if (n2 > 0) {
    REAL n = SQRT(n2);
    z[0] = cr[0] / n;
    z[1] = cr[1] / n;
    z[2] = cr[2] / n;
  } else {
    z[0] = cr[0];
    z[1] = cr[1];
    z[2] = cr[2];
  }
  REAL an = FN(hb_norm)(a), bn = FN(hb_norm)(b);
  REAL v = 2 * ATAN2(FN(hb_dot)(cr, z), an * bn + FN(hb_dot)(a, b));
  if (da) {
    REAL ua[3] = {a[0] / an, a[1] / an, a[2] / an}, ub[3] = {b[0] / bn, b[1] / bn, b[2] / bn};
    REAL t[3];
    FN(hb_cross)(ua, z, t);
    for (int k = 0; k < 3; ++k) da[k] = t[k] / an;
    FN(hb_cross)(ub, z, t);
    for (int k = 0; k < 3; ++k) db[k] = -t[k] / bn;
  }
  return v;
}

/* angle at B of A-B-C (pt_interior_angle) */
static REAL FN(hb_pt_angle)(const REAL* A, const REAL* B, const REAL* C, REAL* dA, REAL* dB,
                            REAL* dC) {
  REAL ba[3], bc[3];
  FN(hb_sub)(A, B, ba);
  FN(hb_sub)(C, B, bc);
  REAL v = FN(hb_interior)(ba, bc, dA, dC);
  if (dA)
    for (int k = 0; k < 3; ++k) dB[k] = -(dA[k] + dC[k]);
  return v;
}

/* dihedral I-J-K-L (Blondel & Karplus 1996, geom.hh dihedral_angle) */
static REAL FN(hb_dihedral)(const REAL* I, const REAL* J, const REAL* K, const REAL* L, REAL* dI,
                            REAL* dJ, REAL* dK, REAL* dL) {
  REAL F[3], G[3], H[3], A[3], B[3], AxB[3];
  FN(hb_sub)(I, J, F);
  FN(hb_sub)(J, K, G);
  FN(hb_sub)(L, K, H);
  FN(hb_cross)(F, G, A);
  FN(hb_cross)(H, G, B);
  FN(hb_cross)(A, B, AxB);
  REAL sign = FN(hb_dot)(G, AxB) >= 0 ? (REAL)-1.0 : (REAL)1.0;
  /* fmin / fmax, which return the other argument for a NaN */
  REAL c = FMAX((REAL)-1.0, FMIN(FN(hb_dot)(A, B) / (FN(hb_norm)(A) * FN(hb_norm)(B)), (REAL)1.0));
  REAL v = sign * ACOS(c);
  if (dI) {
    REAL gn = FN(hb_norm)(G), aa = FN(hb_dot)(A, A), bb = FN(hb_dot)(B, B);
    REAL fg = FN(hb_dot)(F, G), hg = FN(hb_dot)(H, G);
    for (int k = 0; k < 3; ++k) {
      dI[k] = -(gn / aa) * A[k];
      dJ[k] = gn / aa * A[k] + fg / (aa * gn) * A[k] - (hg / (bb * gn)) * B[k];
      dK[k] = -gn / bb * B[k] - fg / (aa * gn) * A[k] + (hg / (bb * gn)) * B[k];
      dL[k] = gn / bb * B[k];
    }
  }
  return v;
}

/* bounded polynomial in double (polynomial.hh bound_poly<11, double>) */
static REAL FN(hb_poly)(REAL x, const double* P, REAL* dV) {
  double xd = (double)x, v, d;
  if (xd < P[12]) {
    v = P[14];
    d = 0;
  } else if (xd > P[13]) {
    v = P[15];
    d = 0;
  } else {
    v = P[0];
    d = P[0];
    for (int i = 1; i < 10; ++i) {
      v = v * xd + P[i];
      d = d * xd + v;
    }
    v = v * xd + P[10];
  }
  if (dV) *dV = (REAL)d;
  return (REAL)v;
}

/* BAH base form (potentials.hh _BAH_angle_base_form_V_dV): polynomial of
 * cos(angle between H - A and A - B) */
static REAL FN(hb_bah_base)(const REAL* B, const REAL* A, const REAL* H, const double* P,
                            REAL* dB, REAL* dA, REAL* dH) {
  REAL AH[3], BA[3];
  FN(hb_sub)(H, A, AH);
  FN(hb_sub)(A, B, BA);
  REAL an = FN(hb_norm)(AH), bn = FN(hb_norm)(BA);
  REAL abn = FN(hb_norm)(AH) * FN(hb_norm)(BA);
  REAL c = FN(hb_dot)(AH, BA) / abn;
  REAL dp;
  REAL v = FN(hb_poly)(c, P, dB ? &dp : NULL);
  if (dB) {
    for (int k = 0; k < 3; ++k) {
      REAL cA = -c * AH[k] / (an * an) + BA[k] / abn;
      REAL cB = -c * BA[k] / (bn * bn) + AH[k] / abn;
      dB[k] = dp * (-cB);
      dA[k] = dp * (cB - cA);
      dH[k] = dp * cA;
    }
  }
  return v;
}

/* sp2 chi term (potentials.hh sp2chi_energy_V_dV), d = BAH180_rise,
 * m = range_span, l = outer_width; the double literals of the reference
 * lift those sub-expressions to double */
static REAL FN(hb_sp2chi)(REAL ang, REAL chi, REAL d, REAL m, REAL l, REAL* dang, REAL* dchi) {
  const REAL pi = (REAL)3.14159265358979323846;
  REAL H = (REAL)(0.5 * (COS(2 * chi) + 1));
  REAL E, ea = 0, ec = 0;
  if (ang > (double)pi * 2.0 / 3.0) {
    REAL F = (REAL)(d / 2 * COS(3 * (pi - ang)) + d / 2 - 0.5);
    REAL G = (REAL)(d - 0.5);
    E = H * F + (1 - H) * G;
    ea = (REAL)(1.5 * d * SIN(3 * ang) * COS(chi) * COS(chi));
    ec = (REAL)(0.5 * d * (COS(3 * ang) + 1) * SIN(2 * chi));
  } else if (ang >= (double)pi * (2.0 / 3.0 - l)) {
    REAL outer_rise = COS(pi - (pi * 2 / 3 - ang) / l);
    REAL F = (REAL)(m / 2 * outer_rise + m / 2 - 0.5);
    REAL G = (REAL)((m - d) / 2 * outer_rise + (m - d) / 2 + d - 0.5);
    E = H * F + (1 - H) * G;
    ea = (REAL)(0.5 * (-d * SIN(chi) * SIN(chi) + m) * SIN((ang - 2 * pi / 3) / l) / l);
    ec = (REAL)(0.5 * d * (COS((ang - 2 * pi / 3) / l) + 1) * SIN(2 * chi));
  } else {
    E = (REAL)(m - 0.5);
  }
  if (dang) {
    *dang = ea;
    *dchi = ec;
  }
  return E;
}

/* The five-atom hbond energy (potentials.hh hbond_score::V_dV).
 *   P: the pair's three polynomials (AHdist, cosBAH, cosAHD), 16 doubles each
 *   pp: [acceptor hybridization, acceptor weight, donor weight]
 *   gp: [sp2_range_span, sp2_BAH180_rise, sp2_outer_width, sp3_softmax_fade,
 *        threshold_distance, max_ha_dis]
 *   g (nullable): [15] = dD, dH, dA, dB, dB0 */
REAL FN(oracle_hb_score)(const REAL* D, const REAL* H, const REAL* A, const REAL* B,
                         const REAL* B0, const REAL* pp, const double* P, const REAL* gp,
                         REAL* g) {
  REAL E = 0, dD[3] = {0, 0, 0}, dH[3] = {0, 0, 0}, dA[3] = {0, 0, 0}, dB[3] = {0, 0, 0},
       dB0[3] = {0, 0, 0};
  const int deriv = g != NULL;
  /* A-H distance */
  {
    REAL v[3];
    FN(hb_sub)(A, H, v);
    REAL dist = FN(hb_norm)(v), dp;
    E += FN(hb_poly)(dist, P, deriv ? &dp : NULL);
    if (deriv && dist != 0)
      for (int k = 0; k < 3; ++k) {
        dA[k] += dp * (v[k] / dist);
        dH[k] += dp * (-v[k] / dist);
      }
  }
  /* A-H-D angle */
  {
    REAL ga[3], gh[3], gd[3], dp;
    REAL ang = FN(hb_pt_angle)(A, H, D, deriv ? ga : NULL, gh, gd);
    E += FN(hb_poly)(ang, P + 2 * HB_POLY_WIDTH, deriv ? &dp : NULL);
    if (deriv)
      for (int k = 0; k < 3; ++k) {
        dA[k] += dp * ga[k];
        dH[k] += dp * gh[k];
        dD[k] += dp * gd[k];
      }
  }
  /* B-A-H angle by acceptor hybridization */
  int hyb = (int)pp[0];
  const double* Pb = P + HB_POLY_WIDTH;
  if (hyb == 1) {
    REAL gb[3], ga[3], gh[3];
    E += FN(hb_bah_base)(B, A, H, Pb, deriv ? gb : NULL, ga, gh);
    if (deriv)
      for (int k = 0; k < 3; ++k) {
        dB[k] += gb[k];
        dA[k] += ga[k];
        dH[k] += gh[k];
      }
  } else if (hyb == 3) {
    REAL Bm[3] = {(B[0] + B0[0]) / 2, (B[1] + B0[1]) / 2, (B[2] + B0[2]) / 2};
    REAL gb[3], ga[3], gh[3];
    E += FN(hb_bah_base)(Bm, A, H, Pb, deriv ? gb : NULL, ga, gh);
    if (deriv)
      for (int k = 0; k < 3; ++k) {
        dB[k] += gb[k] / 2;
        dB0[k] += gb[k] / 2;
        dA[k] += ga[k];
        dH[k] += gh[k];
      }
  } else if (hyb == 2) {
    REAL gb[3], ga[3], gh[3], gb0[3], ga0[3], gh0[3];
    REAL x = FN(hb_bah_base)(B, A, H, Pb, deriv ? gb : NULL, ga, gh);
    REAL x0 = FN(hb_bah_base)(B0, A, H, Pb, deriv ? gb0 : NULL, ga0, gh0);
    REAL f = gp[3];
    REAL e1 = EXP(x * f), e0 = EXP(x0 * f);
    E += LOG(e1 + e0) / f;
    if (deriv) {
      REAL w1 = EXP(x * f) / (EXP(x * f) + EXP(x0 * f));
      REAL w0 = EXP(x0 * f) / (EXP(x * f) + EXP(x0 * f));
      for (int k = 0; k < 3; ++k) {
        dB[k] += w1 * gb[k];
        dB0[k] += w0 * gb0[k];
        dA[k] += (w1 * ga[k]) + (w0 * ga0[k]);
        dH[k] += (w1 * gh[k]) + (w0 * gh0[k]);
      }
    }
  }
  /* B0-B-A-H chi (sp2 only) */
  if (hyb == 1) {
    REAL bb[3], ba[3], bh[3], cb0[3], cb[3], ca[3], ch[3], da, dc;
    REAL bah = FN(hb_pt_angle)(B, A, H, deriv ? bb : NULL, ba, bh);
    REAL chi = FN(hb_dihedral)(B0, B, A, H, deriv ? cb0 : NULL, cb, ca, ch);
    E += FN(hb_sp2chi)(bah, chi, gp[1], gp[0], gp[2], deriv ? &da : NULL, &dc);
    if (deriv)
      for (int k = 0; k < 3; ++k) {
        dB0[k] += dc * cb0[k];
        dB[k] += dc * cb[k] + da * bb[k];
        dA[k] += dc * ca[k] + da * ba[k];
        dH[k] += dc * ch[k] + da * bh[k];
      }
  }
  /* donor x acceptor weight, computed and rounded as a float; then the
   * truncation above 0.1 and the quadratic fade over [-0.1, 0.1] */
  const float w = (float)(pp[1] * pp[2]);
  E *= w;
  REAL slope = 1;
  int zero = 0;
  if (E > (REAL)0.1) {
    E = 0;
    zero = 1;
  } else if (E > (REAL)-0.1) {
    REAL E0 = E;
    E = (REAL)(-0.025 + 0.5 * E - 2.5 * E * E);
    slope = (REAL)(-5.0 * E0 + 0.5);
  }
  if (deriv) {
    const REAL* src[5] = {dD, dH, dA, dB, dB0};
    for (int j = 0; j < 5; ++j)
      for (int k = 0; k < 3; ++k) {
        REAL v = src[j][k] * (REAL)w;
        g[3 * j + k] = zero ? (REAL)0 : v * slope;
      }
  }
  return E;
}


/* ---------- gen_hbond_bases ------------------------------------------- */

/* derived coords [NA, 3, 3] (D of donor H; B, B0 of acceptors) and their
 * source atoms [NA, 3]; NaN / -1 where an atom has no role, (0, 0, 0) / -1
 * for a base that does not exist */
void FN(oracle_hb_bases)(const oracle_hb_stack* hs, const REAL* coords, int64_t na,
                         REAL* derived, int32_t* inds) {
  const oracle_lkb_stack* s = &hs->base;
  int P = s->n_poses, B = s->max_n_blocks, nt = s->max_n_tiles;
  for (int64_t i = 0; i < na * 9; ++i) derived[i] = (REAL)NAN;
  for (int64_t i = 0; i < na * 3; ++i) inds[i] = -1;
  for (int p = 0; p < P; ++p)
    for (int b = 0; b < B; ++b) {
      int bt = s->block_type[p * B + b];
      if (bt < 0) continue;
      int n = s->bt_n_atoms[bt];
      int64_t off = s->rot_coord_offset[p * B + b];
      for (int t = 0; t * HB_TILE < n; ++t) {
        int64_t ti = (int64_t)bt * nt + t;
        for (int k = 0; k < s->tile_n_donH[ti]; ++k) {
          lkb_bca H = {b, bt, t * HB_TILE + s->tile_donH_inds[ti * HB_TILE + k]};
          lkb_bca D = hb_donor_base(s, p, H);
          int64_t h = off + H.atom;
          for (int c = 0; c < 3; ++c) derived[9 * h + c] = 0;
          if (D.atom >= 0) {
            int64_t d = lkb_atom_index(s, p, D);
            for (int c = 0; c < 3; ++c) derived[9 * h + c] = coords[3 * d + c];
            inds[3 * h] = (int32_t)d;
          }
        }
        for (int k = 0; k < s->tile_n_acc[ti]; ++k) {
          int64_t e = ti * HB_TILE + k;
          lkb_bca A = {b, bt, t * HB_TILE + s->tile_acc_inds[e]};
          lkb_bca Bb, B0;
          hb_acc_bases(s, p, A, s->tile_acc_hyb[e], &Bb, &B0);
          int64_t a = off + A.atom;
          const lkb_bca* bs[2] = {&Bb, &B0};
          for (int j = 0; j < 2; ++j) {
            for (int c = 0; c < 3; ++c) derived[9 * a + 3 * (j + 1) + c] = 0;
            if (bs[j]->atom >= 0) {
              int64_t x = lkb_atom_index(s, p, *bs[j]);
              for (int c = 0; c < 3; ++c) derived[9 * a + 3 * (j + 1) + c] = coords[3 * x + c];
              inds[3 * a + j + 1] = (int32_t)x;
            }
          }
        }
      }
    }
}

/* ---------- hbond_pose_scores ----------------------------------------- */

/* inter-block count-pair separation with the reference's lookup: i1 / i2
 * are the pair's positions in block 1's and block 2's tile lists, read as
 * tile atom indices (hbond_pose_score.impl.hh:47-62, 106-140) */
static int FN(hb_inter_sep)(const oracle_lkb_stack* s, int p, int b1, int b2, int bt1, int bt2,
                        int start1, int start2, int i1, int i2) {
  int B = s->max_n_blocks, C = s->max_n_conn, Am = s->max_n_block_atoms;
  int sep = s->min_bondsep[((int64_t)p * B + b1) * B + b2];
  if (sep > HB_MAX_SEP) return sep;
  int best = HB_MAX_SEP + 1;
  for (int c1 = 0; c1 < s->bt_n_conn[bt1]; ++c1) {
    int ca1 = s->bt_conn_atoms[bt1 * C + c1];
    int pd1 = s->path_distance[((int64_t)bt1 * Am + ca1) * Am + start1 + i1];
    for (int c2 = 0; c2 < s->bt_n_conn[bt2]; ++c2) {
      int ca2 = s->bt_conn_atoms[bt2 * C + c2];
      int pd2 = s->path_distance[((int64_t)bt2 * Am + ca2) * Am + start2 + i2];
      int cs = s->inter_block_bondsep[((((int64_t)p * B + b1) * B + b2) * C + c1) * C + c2];
      if (pd1 + cs + pd2 < best) best = pd1 + cs + pd2;
    }
  }
  return best;
}

/* one donor H x acceptor pair (global atom indices h, a): energy, and with
 * dX the weighted derivatives through the derived atoms */
static REAL FN(hb_pair)(const oracle_hb_stack* hs, const REAL* coords, const REAL* derived,
                        const int32_t* inds, int64_t h, int64_t a, int dt, int at,
                        const REAL* pair_params, const double* polys, const REAL* gp, REAL wt,
                        REAL* dX) {
  const REAL* H = coords + 3 * h;
  const REAL* A = coords + 3 * a;
  REAL v[3];
  FN(hb_sub)(H, A, v);
  REAL dist = FN(hb_norm)(v);
  if (!(dist < gp[5])) return 0;
  int64_t pi = (int64_t)dt * hs->n_acc_types + at;
  REAL g[15];
  REAL E = FN(oracle_hb_score)(derived + 9 * h, H, A, derived + 9 * a + 3, derived + 9 * a + 6,
                               pair_params + 3 * pi, polys + 3 * HB_POLY_WIDTH * pi, gp,
                               dX ? g : NULL);
  if (dX) {
    int64_t at5[5] = {inds[3 * h], h, a, inds[3 * a + 1], inds[3 * a + 2]};
    for (int j = 0; j < 5; ++j) {
      if (at5[j] < 0) continue;
      for (int k = 0; k < 3; ++k)
        if (g[3 * j + k] != 0) dX[3 * at5[j] + k] += wt * g[3 * j + k];
    }
  }
  return E;
}

/* The pose score over neighbor block pairs nbr [P, B, B] (b1 <= b2), tile
 * by tile as the reference's tile_evaluate_rot_pair.
 *   energies: [P] or [P, B, B] (block_pair), overwritten; may be NULL
 *   weights:  dTdV of the same shape, or NULL (no derivatives)
 *   dX [NA, 3]: overwritten with the weighted derivatives */
void FN(oracle_hb_score_pose)(const oracle_hb_stack* hs, const REAL* coords, int64_t na,
                              const REAL* derived, const int32_t* inds, const REAL* pair_params,
                              const double* polys, const REAL* gp, const int32_t* nbr,
                              int block_pair, REAL* energies, const REAL* weights, REAL* dX) {
  const oracle_lkb_stack* s = &hs->base;
  int P = s->n_poses, B = s->max_n_blocks, nt = s->max_n_tiles, Am = s->max_n_block_atoms;
  int64_t nE = block_pair ? (int64_t)P * B * B : P;
  if (energies)
    for (int64_t i = 0; i < nE; ++i) energies[i] = 0;
  if (weights)
    for (int64_t i = 0; i < na * 3; ++i) dX[i] = 0;
  for (int p = 0; p < P; ++p)
    for (int b1 = 0; b1 < B; ++b1)
      for (int b2 = b1; b2 < B; ++b2) {
        if (!nbr[((int64_t)p * B + b1) * B + b2]) continue;
        int bt1 = s->block_type[p * B + b1], bt2 = s->block_type[p * B + b2];
        if (bt1 < 0 || bt2 < 0) continue;
        int64_t cell = block_pair ? ((int64_t)p * B + b1) * B + b2 : p;
        REAL wt = weights ? weights[cell] : 0;
        REAL* d = weights ? dX : NULL;
        int64_t o1 = s->rot_coord_offset[p * B + b1], o2 = s->rot_coord_offset[p * B + b2];
        int n1 = s->bt_n_atoms[bt1], n2 = s->bt_n_atoms[bt2];
        REAL total = 0;
        if (b1 != b2) {
          for (int t1 = 0; t1 * HB_TILE < n1; ++t1)
            for (int t2 = 0; t2 * HB_TILE < n2; ++t2) {
              int64_t ti1 = (int64_t)bt1 * nt + t1, ti2 = (int64_t)bt2 * nt + t2;
              int s1 = t1 * HB_TILE, s2 = t2 * HB_TILE;
              /* block 1 donates, block 2 accepts; then the reverse */
              for (int dir = 0; dir < 2; ++dir) {
                int64_t tid = dir ? ti2 : ti1, tia = dir ? ti1 : ti2;
                int sd = dir ? s2 : s1, sa = dir ? s1 : s2;
                int64_t od = dir ? o2 : o1, oa = dir ? o1 : o2;
                for (int i = 0; i < s->tile_n_donH[tid]; ++i)
                  for (int j = 0; j < s->tile_n_acc[tia]; ++j) {
                    int sep = dir ? FN(hb_inter_sep)(s, p, b1, b2, bt1, bt2, s1, s2, j, i)
                                  : FN(hb_inter_sep)(s, p, b1, b2, bt1, bt2, s1, s2, i, j);
                    if (sep < HB_MAX_SEP + 1) continue;
                    int64_t h = od + sd + s->tile_donH_inds[tid * HB_TILE + i];
                    int64_t a = oa + sa + s->tile_acc_inds[tia * HB_TILE + j];
                    total += FN(hb_pair)(hs, coords, derived, inds, h, a,
                                         hs->tile_don_type[tid * HB_TILE + i],
                                         hs->tile_acc_type[tia * HB_TILE + j], pair_params,
                                         polys, gp, wt, d);
                  }
              }
            }
        } else {
          for (int t1 = 0; t1 * HB_TILE < n1; ++t1)
            for (int t2 = t1; t2 * HB_TILE < n1; ++t2) {
              int64_t ti1 = (int64_t)bt1 * nt + t1, ti2 = (int64_t)bt1 * nt + t2;
              int s1 = t1 * HB_TILE, s2 = t2 * HB_TILE;
              for (int dir = 0; dir < (t1 == t2 ? 1 : 2); ++dir) {
                int64_t tid = dir ? ti2 : ti1, tia = dir ? ti1 : ti2;
                int sd = dir ? s2 : s1, sa = dir ? s1 : s2;
                for (int i = 0; i < s->tile_n_donH[tid]; ++i)
                  for (int j = 0; j < s->tile_n_acc[tia]; ++j) {
                    int ad = sd + s->tile_donH_inds[tid * HB_TILE + i];
                    int aa = sa + s->tile_acc_inds[tia * HB_TILE + j];
                    int sep = s->path_distance[((int64_t)bt1 * Am + ad) * Am + aa];
                    if (sep < HB_MAX_SEP + 1) continue;
                    total += FN(hb_pair)(hs, coords, derived, inds, o1 + ad, o1 + aa,
                                         hs->tile_don_type[tid * HB_TILE + i],
                                         hs->tile_acc_type[tia * HB_TILE + j], pair_params,
                                         polys, gp, wt, d);
                  }
              }
            }
        }
        if (energies) energies[cell] += total;
      }
}

#undef FN
#undef CAT
#undef CAT_
