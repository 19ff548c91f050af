/* CPU restatement of the reference LJ+LK pose-score op -- TEST INFRASTRUCTURE.
 *
 * This file is the checker, never the product: only tests/, the smoke check
 * and bench.py's cpu_baseline leg load it.  It restates the serial CPU path
 * of tmol's ljlk_pose_scores (DeviceOperations<CPU>, one thread per
 * "workgroup", tmol/score/common/device_operations.cpu.impl.hh:44-49) in
 * plain C, instantiated twice (float, double) by ljlk_oracle.c, which
 * defines REAL, RSUF (name suffix), SQRT, EXP, FLOOR before including it.
 *
 * Pinned by: tmol's 1ubq goldens (whole pose, jagged, block pair), the
 * potential invariants of tmol/tests/score/ljlk/potentials, central finite
 * differences, brute-force bond separation on the atom graph, and the
 * reference op itself where oracle/_ref builds (tests/test_oracle*.py).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, RSUF)

/* ---------- pair potentials ------------------------------------------ */

/* connectivity weight, tmol/score/ljlk/potentials/common.hh:13-22 */
static REAL FN(cp_weight)(int sep) {
  if (sep > 4) return (REAL)1.0;
  if (sep == 4) return (REAL)0.2;
  return (REAL)0.0;
}

typedef struct {
  REAL lj_radius, lj_wdepth, lk_dgfree, lk_lambda, lk_volume;
  REAL is_donor, is_hydroxyl, is_polarh, is_acceptor, is_carbon_lk;
} FN(tparams);

typedef struct {
  REAL max_dis, lj_dlin_sigma_factor, lj_hbond_dis, lj_hbond_OH_donor_dis,
      lj_hbond_hdis;
} FN(gparams);

/* hbond-aware sigma, common.hh:24-44 */
static REAL FN(lj_sigma)(const FN(tparams) * i, const FN(tparams) * j,
                         const FN(gparams) * g) {
  if ((i->is_donor != 0 && j->is_acceptor != 0) ||
      (j->is_donor != 0 && i->is_acceptor != 0)) {
    if ((i->is_donor != 0 && i->is_hydroxyl != 0) ||
        (j->is_donor != 0 && j->is_hydroxyl != 0))
      return g->lj_hbond_OH_donor_dis;
    return g->lj_hbond_dis;
  } else if ((i->is_polarh != 0 && j->is_acceptor != 0) ||
             (j->is_polarh != 0 && i->is_acceptor != 0)) {
    return g->lj_hbond_hdis;
  }
  return i->lj_radius + j->lj_radius;
}

/* 12-6 LJ, lj.hh:27-47 */
static REAL FN(vdw_V)(REAL dist, REAL sigma, REAL eps) {
  REAL sd = sigma / dist;
  REAL sd2 = sd * sd;
  REAL sd6 = sd2 * sd2 * sd2;
  REAL sd12 = sd6 * sd6;
  return eps * (sd12 - 2 * sd6);
}
static void FN(vdw_V_dV)(REAL dist, REAL sigma, REAL eps, REAL* V, REAL* dV) {
  REAL sd = sigma / dist;
  REAL sd2 = sd * sd;
  REAL sd6 = sd2 * sd2 * sd2;
  REAL sd12 = sd6 * sd6;
  *V = eps * (sd12 - (REAL)2.0 * sd6);
  *dV = eps * (((REAL)-12.0 * sd12 / dist) - ((REAL)-12.0 * sd6 / dist));
}

/* cubic Hermite helpers, tmol/score/common/cubic_hermite_polynomial.hh */
static REAL FN(interp)(REAL x, REAL x0, REAL p0, REAL dpdx0, REAL x1, REAL p1,
                       REAL dpdx1) {
  REAL t = (x - x0) / (x1 - x0);
  REAL dp0 = dpdx0 * (x1 - x0);
  REAL dp1 = dpdx1 * (x1 - x0);
  return p0 + t * (dp0 + t * (-2 * dp0 - dp1 - 3 * p0 + 3 * p1 +
                              t * (dp0 + dp1 + 2 * p0 - 2 * p1)));
}
static void FN(interp_V_dV)(REAL x, REAL x0, REAL p0, REAL dpdx0, REAL x1,
                            REAL p1, REAL dpdx1, REAL* V, REAL* dV) {
  REAL t = (x - x0) / (x1 - x0);
  REAL dp0 = dpdx0 * (x1 - x0);
  REAL dp1 = dpdx1 * (x1 - x0);
  *V = p0 + t * (dp0 + t * (-2 * dp0 - dp1 - 3 * p0 + 3 * p1 +
                            t * (dp0 + dp1 + 2 * p0 - 2 * p1)));
  REAL dp = dp0 + t * (-4 * dp0 - 2 * dp1 - 6 * p0 + 6 * p1 +
                       t * (3 * dp0 + 3 * dp1 + 6 * p0 - 6 * p1));
  *dV = dp / (x1 - x0);
}
static REAL FN(interp0)(REAL x, REAL x0, REAL p0, REAL dpdx0, REAL x1) {
  REAL t = (x - x0) / (x1 - x0);
  REAL dp0 = dpdx0 * (x1 - x0);
  return p0 + t * (dp0 + t * (-2 * dp0 - 3 * p0 + t * (dp0 + 2 * p0)));
}
static void FN(interp0_V_dV)(REAL x, REAL x0, REAL p0, REAL dpdx0, REAL x1,
                             REAL* V, REAL* dV) {
  REAL t = (x - x0) / (x1 - x0);
  REAL dp0 = dpdx0 * (x1 - x0);
  *V = p0 + t * (dp0 + t * (-2 * dp0 - 3 * p0 + t * (dp0 + 2 * p0)));
  *dV = (dp0 + t * (-4 * dp0 - 6 * p0 + t * (3 * dp0 + 6 * p0))) / (x1 - x0);
}

/* LJ energy only, lj.hh:65-113 */
static void FN(lj_V)(REAL dist, int sep, const FN(tparams) * i,
                     const FN(tparams) * j, const FN(gparams) * g, REAL* atr,
                     REAL* rep) {
  REAL cpoly_dmax = g->max_dis;
  REAL spline_start = g->max_dis - (REAL)1.5;
  REAL weight, Vatr, Vrep;
  if (dist > cpoly_dmax) {
    Vatr = 0;
    Vrep = 0;
    weight = 0;
  } else {
    REAL sigma = FN(lj_sigma)(i, j, g);
    REAL eps = SQRT(i->lj_wdepth * j->lj_wdepth);
    REAL d_lin = sigma * g->lj_dlin_sigma_factor;
    REAL cpoly_dmin =
        sigma > spline_start
            ? (sigma > cpoly_dmax - (REAL)0.1 ? cpoly_dmax - (REAL)0.1 : sigma)
            : spline_start;
    weight = FN(cp_weight)(sep);
    if (dist > cpoly_dmin) {
      REAL v, dv;
      FN(vdw_V_dV)(cpoly_dmin, sigma, eps, &v, &dv);
      Vatr = FN(interp0)(dist, cpoly_dmin, v, dv, cpoly_dmax);
    } else if (dist > d_lin) {
      Vatr = FN(vdw_V)(dist, sigma, eps);
    } else {
      REAL v, dv;
      FN(vdw_V_dV)(d_lin, sigma, eps, &v, &dv);
      Vatr = v + dv * (dist - d_lin);
    }
    if (dist < sigma) {
      Vrep = Vatr + eps;
      Vatr = -eps;
    } else {
      Vrep = 0;
    }
  }
  *atr = weight * Vatr;
  *rep = weight * Vrep;
}

/* LJ energy + derivative, lj.hh:115-177 */
static void FN(lj_V_dV)(REAL dist, int sep, const FN(tparams) * i,
                        const FN(tparams) * j, const FN(gparams) * g,
                        REAL* atr, REAL* rep, REAL* datr, REAL* drep) {
  REAL sigma = FN(lj_sigma)(i, j, g);
  REAL weight = FN(cp_weight)(sep);
  REAL eps = SQRT(i->lj_wdepth * j->lj_wdepth);
  REAL d_lin = sigma * g->lj_dlin_sigma_factor;
  REAL cpoly_dmax = g->max_dis;
  REAL cpoly_dmin = g->max_dis - (REAL)1.5;
  if (sigma > cpoly_dmin) cpoly_dmin = sigma;
  if (cpoly_dmin > cpoly_dmax - (REAL)0.1) cpoly_dmin = cpoly_dmax - (REAL)0.1;
  REAL lj, dlj;
  if (dist < d_lin) {
    REAL v, dv;
    FN(vdw_V_dV)(d_lin, sigma, eps, &v, &dv);
    lj = v + dv * (dist - d_lin);
    dlj = dv;
  } else if (dist < cpoly_dmin) {
    FN(vdw_V_dV)(dist, sigma, eps, &lj, &dlj);
  } else if (dist < cpoly_dmax) {
    REAL v, dv;
    FN(vdw_V_dV)(cpoly_dmin, sigma, eps, &v, &dv);
    FN(interp0_V_dV)(dist, cpoly_dmin, v, dv, cpoly_dmax, &lj, &dlj);
  } else {
    lj = 0;
    dlj = 0;
  }
  REAL Va, Vr, dVa, dVr;
  if (dist < sigma) {
    Va = -eps;
    Vr = lj + eps;
    dVa = 0;
    dVr = dlj;
  } else {
    Va = lj;
    Vr = 0;
    dVa = dlj;
    dVr = 0;
  }
  *atr = weight * Va;
  *rep = weight * Vr;
  *datr = weight * dVa;
  *drep = weight * dVr;
}

/* f_desolv, lk_isotropic.hh:22-93 */
static REAL FN(fdesolv_V)(REAL dist, REAL lj_radius_i, REAL lk_dgfree_i,
                          REAL lk_lambda_i, REAL lk_volume_j) {
  const REAL pi_pow1p5 = (REAL)5.56832799683f;
  return -lk_volume_j * lk_dgfree_i / (2 * pi_pow1p5 * lk_lambda_i) /
         (dist * dist) *
         EXP(-(dist - lj_radius_i) * (dist - lj_radius_i) /
             (lk_lambda_i * lk_lambda_i));
}
static void FN(fdesolv_V_dV)(REAL dist, REAL lj_radius_i, REAL lk_dgfree_i,
                             REAL lk_lambda_i, REAL lk_volume_j, REAL* V,
                             REAL* dV) {
  const REAL pi_pow1p5 = (REAL)5.56832799683f;
  REAL exp_val = EXP(-(dist - lj_radius_i) * (dist - lj_radius_i) /
                     (lk_lambda_i * lk_lambda_i));
  *V = -lk_volume_j * lk_dgfree_i / (2 * pi_pow1p5 * lk_lambda_i) /
       (dist * dist) * exp_val;
  *dV = -lk_volume_j * lk_dgfree_i / (2 * pi_pow1p5 * lk_lambda_i) * exp_val *
        ((-2 / (dist * dist * dist)) +
         (1 / (dist * dist) * -(2 * dist - 2 * lj_radius_i) /
          (lk_lambda_i * lk_lambda_i)));
}

/* knot placement shared by both LK variants, lk_isotropic.hh:116-125 */
static void FN(lk_knots)(REAL sigma, int is_cc, REAL* d_min, REAL* close_dmin,
                         REAL* close_dmax) {
  REAL dm = (REAL)(sigma * .89); /* double literal, as in lk_isotropic.hh:116 */
  if (is_cc && dm < (REAL)4.2) dm = (REAL)4.2;
  REAL n = FLOOR((REAL)20 * dm * dm);
  REAL lo = n - 29;
  if (lo < 0) lo = 0;
  REAL hi = n + 21;
  if (hi > (REAL)405) hi = (REAL)405;
  *d_min = dm;
  *close_dmin = SQRT(lo / 20);
  *close_dmax = SQRT(hi / 20);
}

/* one direction of LK with derivative, lk_isotropic.hh:167-235 */
static void FN(lk_pair_V_dV)(REAL dist, int sep, REAL sigma, REAL r_i,
                             REAL dg_i, REAL lam_i, REAL vol_j, REAL max_dis,
                             int is_cc, REAL* V, REAL* dV) {
  REAL d_min, cdmin, cdmax;
  FN(lk_knots)(sigma, is_cc, &d_min, &cdmin, &cdmax);
  REAL far_dmax = max_dis;
  REAL far_dmin = max_dis - (REAL)1.5;
  REAL weight = FN(cp_weight)(sep);
  REAL lk, dlk;
  if (dist < cdmin) {
    lk = FN(fdesolv_V)(d_min, r_i, dg_i, lam_i, vol_j);
    dlk = 0;
  } else if (dist < cdmax) {
    REAL v, dv;
    FN(fdesolv_V_dV)(cdmax, r_i, dg_i, lam_i, vol_j, &v, &dv);
    FN(interp_V_dV)(dist, cdmin, FN(fdesolv_V)(d_min, r_i, dg_i, lam_i, vol_j),
                    0, cdmax, v, dv, &lk, &dlk);
  } else if (dist < far_dmin) {
    FN(fdesolv_V_dV)(dist, r_i, dg_i, lam_i, vol_j, &lk, &dlk);
  } else if (dist < far_dmax) {
    REAL v, dv;
    FN(fdesolv_V_dV)(far_dmin, r_i, dg_i, lam_i, vol_j, &v, &dv);
    FN(interp0_V_dV)(dist, far_dmin, v, dv, far_dmax, &lk, &dlk);
  } else {
    lk = 0;
    dlk = 0;
  }
  *V = weight * lk;
  *dV = weight * dlk;
}

/* LK energy + derivative, both directions, lk_isotropic.hh:328-360 */
static void FN(lk_V_dV)(REAL dist, int sep, const FN(tparams) * i,
                        const FN(tparams) * j, const FN(gparams) * g, REAL* V,
                        REAL* dV) {
  REAL sigma = FN(lj_sigma)(i, j, g);
  int is_cc = (i->is_carbon_lk != 0) && (j->is_carbon_lk != 0);
  REAL v1, d1, v2, d2;
  FN(lk_pair_V_dV)(dist, sep, sigma, i->lj_radius, i->lk_dgfree, i->lk_lambda,
                   j->lk_volume, g->max_dis, is_cc, &v1, &d1);
  FN(lk_pair_V_dV)(dist, sep, sigma, j->lj_radius, j->lk_dgfree, j->lk_lambda,
                   i->lk_volume, g->max_dis, is_cc, &v2, &d2);
  *V = v1 + v2;
  *dV = d1 + d2;
}

/* LK energy only, lk_isotropic.hh:247-326 */
static REAL FN(lk_V)(REAL dist, int sep, const FN(tparams) * i,
                     const FN(tparams) * j, const FN(gparams) * g) {
  REAL sigma = FN(lj_sigma)(i, j, g);
  int is_cc = (i->is_carbon_lk != 0) && (j->is_carbon_lk != 0);
  REAL d_min, cdmin, cdmax;
  FN(lk_knots)(sigma, is_cc, &d_min, &cdmin, &cdmax);
  REAL far_dmax = g->max_dis;
  REAL far_dmin = g->max_dis - (REAL)1.5;
  REAL weight = FN(cp_weight)(sep);
  REAL lk;
  if (dist > far_dmax) {
    lk = 0;
  } else if (dist > far_dmin) {
    REAL v, dv;
    FN(fdesolv_V_dV)(far_dmin, i->lj_radius, i->lk_dgfree, i->lk_lambda,
                     j->lk_volume, &v, &dv);
    lk = FN(interp0)(dist, far_dmin, v, dv, far_dmax);
    FN(fdesolv_V_dV)(far_dmin, j->lj_radius, j->lk_dgfree, j->lk_lambda,
                     i->lk_volume, &v, &dv);
    lk += FN(interp0)(dist, far_dmin, v, dv, far_dmax);
  } else if (dist > cdmax) {
    lk = FN(fdesolv_V)(dist, i->lj_radius, i->lk_dgfree, i->lk_lambda,
                       j->lk_volume) +
         FN(fdesolv_V)(dist, j->lj_radius, j->lk_dgfree, j->lk_lambda,
                       i->lk_volume);
  } else if (dist > cdmin) {
    REAL v, dv;
    FN(fdesolv_V_dV)(cdmax, i->lj_radius, i->lk_dgfree, i->lk_lambda,
                     j->lk_volume, &v, &dv);
    lk = FN(interp)(dist, cdmin,
                    FN(fdesolv_V)(d_min, i->lj_radius, i->lk_dgfree,
                                  i->lk_lambda, j->lk_volume),
                    0, cdmax, v, dv);
    FN(fdesolv_V_dV)(cdmax, j->lj_radius, j->lk_dgfree, j->lk_lambda,
                     i->lk_volume, &v, &dv);
    lk += FN(interp)(dist, cdmin,
                     FN(fdesolv_V)(d_min, j->lj_radius, j->lk_dgfree,
                                   j->lk_lambda, i->lk_volume),
                     0, cdmax, v, dv);
  } else {
    lk = FN(fdesolv_V)(d_min, i->lj_radius, i->lk_dgfree, i->lk_lambda,
                       j->lk_volume) +
         FN(fdesolv_V)(d_min, j->lj_radius, j->lk_dgfree, j->lk_lambda,
                       i->lk_volume);
  }
  return weight * lk;
}

/* exported single-pair evaluators (potential invariant tests) */
void FN(oracle_lj_pair)(REAL dist, int sep, const REAL* pi, const REAL* pj,
                        const REAL* gp, int with_deriv, REAL* out4) {
  const FN(tparams)* i = (const FN(tparams)*)pi;
  const FN(tparams)* j = (const FN(tparams)*)pj;
  const FN(gparams)* g = (const FN(gparams)*)gp;
  if (with_deriv) {
    FN(lj_V_dV)(dist, sep, i, j, g, &out4[0], &out4[1], &out4[2], &out4[3]);
  } else {
    FN(lj_V)(dist, sep, i, j, g, &out4[0], &out4[1]);
    out4[2] = out4[3] = 0;
  }
}

void FN(oracle_lk_pair)(REAL dist, int sep, const REAL* pi, const REAL* pj,
                        const REAL* gp, int with_deriv, REAL* out2) {
  const FN(tparams)* i = (const FN(tparams)*)pi;
  const FN(tparams)* j = (const FN(tparams)*)pj;
  const FN(gparams)* g = (const FN(gparams)*)gp;
  if (with_deriv) {
    FN(lk_V_dV)(dist, sep, i, j, g, &out2[0], &out2[1]);
  } else {
    out2[0] = FN(lk_V)(dist, sep, i, j, g);
    out2[1] = 0;
  }
}

/* ---------- neighbor detection ---------------------------------------- */

/* bounding sphere per block, sphere_overlap.impl.hh:109-196 on the CPU:
 * sequential coordinate sum in atom order, divided by n; radius is the
 * square root of the largest squared distance to the centre.  The expression
 * order ((dx*dx + dy*dy) + dz*dz) is kept and the file is compiled without
 * contraction (-ffp-contract=off). */
void FN(oracle_block_spheres)(const REAL* coords, int n_blocks,
                              const int32_t* block_atom_offset,
                              const int32_t* block_type,
                              const int32_t* bt_n_atoms, REAL* spheres) {
  for (int b = 0; b < n_blocks; ++b) {
    REAL* s = spheres + 4 * (int64_t)b;
    s[0] = s[1] = s[2] = s[3] = 0;
    int bt = block_type[b];
    if (bt < 0) continue;
    int n = bt_n_atoms[bt];
    const REAL* c = coords + 3 * (int64_t)block_atom_offset[b];
    REAL cx = 0, cy = 0, cz = 0;
    for (int a = 0; a < n; ++a) {
      cx += c[3 * a + 0];
      cy += c[3 * a + 1];
      cz += c[3 * a + 2];
    }
    cx /= n;
    cy /= n;
    cz /= n;
    REAL d2max = 0;
    for (int a = 0; a < n; ++a) {
      REAL dx = c[3 * a + 0] - cx, dy = c[3 * a + 1] - cy,
           dz = c[3 * a + 2] - cz;
      REAL d2 = dx * dx + dy * dy + dz * dz;
      if (d2 > d2max) d2max = d2;
    }
    s[0] = cx;
    s[1] = cy;
    s[2] = cz;
    s[3] = SQRT(d2max);
  }
}

/* dense neighbor flags per pose, sphere_overlap.impl.hh:287-351:
 * neighbors[p][b1][b2] = 1 iff b1 <= b2, both real, d2 < ((r1+r2)+reach)^2 */
void FN(oracle_block_neighbors)(int n_poses, int max_n_blocks,
                                const int32_t* block_type, const REAL* spheres,
                                REAL reach, int32_t* neighbors) {
  int B = max_n_blocks;
  for (int p = 0; p < n_poses; ++p)
    for (int b1 = 0; b1 < B; ++b1)
      for (int b2 = 0; b2 < B; ++b2) {
        int32_t* out = neighbors + ((int64_t)p * B + b1) * B + b2;
        *out = 0;
        if (b1 > b2) continue;
        if (block_type[p * B + b1] < 0 || block_type[p * B + b2] < 0) continue;
        const REAL* s1 = spheres + 4 * ((int64_t)p * B + b1);
        const REAL* s2 = spheres + 4 * ((int64_t)p * B + b2);
        REAL dx = s1[0] - s2[0], dy = s1[1] - s2[1], dz = s1[2] - s2[2];
        REAL d2 = dx * dx + dy * dy + dz * dz;
        REAL thr = s1[3] + s2[3] + reach;
        if (d2 < thr * thr) *out = 1;
      }
}

/* ---------- block-pair scoring ---------------------------------------- */

static void FN(load_params)(const REAL* type_params, int at, FN(tparams) * out) {
  const REAL* t = type_params + 10 * (int64_t)at;
  out->lj_radius = t[0];
  out->lj_wdepth = t[1];
  out->lk_dgfree = t[2];
  out->lk_lambda = t[3];
  out->lk_volume = t[4];
  out->is_donor = t[5];
  out->is_hydroxyl = t[6];
  out->is_polarh = t[7];
  out->is_acceptor = t[8];
  out->is_carbon_lk = t[9];
}

/* inter-block count-pair separation (impl.hh:47-67, count_pair.hh:162-207):
 * if min_sep <= 4 the minimum over connection pairs of
 * pd1[c1][a1] + conn_seps[c1][c2] + pd2[c2][a2], starting at 5; both blocks
 * ligand fragments -> 3/4 become 5. */
static int FN(inter_sep)(const oracle_stack* st, int p, int b1, int b2,
                         int bt1, int bt2, int a1, int a2) {
  int B = st->max_n_blocks, C = st->max_n_conn, Am = st->max_n_block_atoms;
  int sep = st->min_bondsep[((int64_t)p * B + b1) * B + b2];
  if (sep <= 4) {
    int best = 5;
    int nc1 = st->bt_n_conn[bt1], nc2 = st->bt_n_conn[bt2];
    for (int c1 = 0; c1 < nc1; ++c1) {
      int ca1 = st->bt_conn_atoms[bt1 * C + c1];
      int pd1 = st->bt_path_distance[((int64_t)bt1 * Am + ca1) * Am + a1];
      for (int c2 = 0; c2 < nc2; ++c2) {
        int ca2 = st->bt_conn_atoms[bt2 * C + c2];
        int pd2 = st->bt_path_distance[((int64_t)bt2 * Am + ca2) * Am + a2];
        int cs = st->inter_block_bondsep[((((int64_t)p * B + b1) * B + b2) * C +
                                          c1) * C + c2];
        int s = pd1 + cs + pd2;
        if (s < best) best = s;
      }
    }
    sep = best;
  }
  if (st->bt_is_ligand[bt1] && st->bt_is_ligand[bt2] && (sep == 3 || sep == 4))
    sep = 5;
  return sep;
}


static void FN(add_grad)(REAL* dvdx, int64_t atom, int term, int64_t na,
                         REAL g, const REAL* u) {
  for (int k = 0; k < 3; ++k) {
    REAL v = g * u[k];
    if (v != 0) dvdx[((int64_t)term * na + atom) * 3 + k] += v;
  }
}

/* one atom pair: LJ or LK, energy (+derivative) */
static void FN(eval_pair)(const REAL* coords, int64_t na, int64_t i1,
                          int64_t i2, int sep, const FN(tparams) * t1,
                          const FN(tparams) * t2, const FN(gparams) * g,
                          int lk, int require_grad, REAL* e, REAL* dvdx) {
  const REAL* x1 = coords + 3 * i1;
  const REAL* x2 = coords + 3 * i2;
  REAL d[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
  REAL dist = SQRT(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (!lk) {
    if (require_grad) {
      REAL a, r, da, dr;
      FN(lj_V_dV)(dist, sep, t1, t2, g, &a, &r, &da, &dr);
      e[0] = a;
      e[1] = r;
      if (dvdx) {
        REAL u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
        if (dist != 0)
          for (int k = 0; k < 3; ++k) {
            u1[k] = d[k] / dist;
            u2[k] = -d[k] / dist;
          }
        /* impl order (ljlk.hh:629-664): atom 1 atr/rep, then atom 2 */
        for (int k = 0; k < 3; ++k) {
          REAL va = da * u1[k], vr = dr * u1[k];
          if (va != 0) dvdx[((int64_t)0 * na + i1) * 3 + k] += va;
          if (vr != 0) dvdx[((int64_t)1 * na + i1) * 3 + k] += vr;
        }
        for (int k = 0; k < 3; ++k) {
          REAL va = da * u2[k], vr = dr * u2[k];
          if (va != 0) dvdx[((int64_t)0 * na + i2) * 3 + k] += va;
          if (vr != 0) dvdx[((int64_t)1 * na + i2) * 3 + k] += vr;
        }
      }
    } else {
      FN(lj_V)(dist, sep, t1, t2, g, &e[0], &e[1]);
    }
  } else {
    if (require_grad) {
      REAL v, dv;
      FN(lk_V_dV)(dist, sep, t1, t2, g, &v, &dv);
      e[0] = v;
      if (dvdx) {
        REAL u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
        if (dist != 0)
          for (int k = 0; k < 3; ++k) {
            u1[k] = d[k] / dist;
            u2[k] = -d[k] / dist;
          }
        FN(add_grad)(dvdx, i1, 2, na, dv, u1);
        FN(add_grad)(dvdx, i2, 2, na, dv, u2);
      }
    } else {
      e[0] = FN(lk_V)(dist, sep, t1, t2, g);
    }
  }
}

/* Score every neighbor block pair of a padded stack, in the reference CPU
 * order: pose-major, upper triangle (b1 <= b2) row by row, tiles of 32 atoms,
 * pairs row-major inside a tile pair (impl.hh:633-1158,
 * tile_atom_pair_evaluation.hh:88-173, 377-505).  Energies accumulate in
 * REAL: pair -> tile-pair partial -> block-pair total -> pose total.
 *   out:       [3, P] (whole pose) or [3, P, B, B] (block_pair != 0)
 *   dvdx:      [3, NA, 3] or NULL, filled only when require_grad
 *   neighbors: [P, B, B] dense flags from oracle_block_neighbors
 */
void FN(oracle_pose_scores)(const oracle_stack* st, const REAL* coords,
                            int64_t n_atoms, const REAL* type_params,
                            const REAL* global_params, const int32_t* neighbors,
                            int block_pair, int require_grad, REAL* out,
                            REAL* dvdx) {
  const int P = st->n_poses, B = st->max_n_blocks;
  const int Am = st->max_n_block_atoms, nt = st->max_n_tiles;
  const FN(gparams)* g = (const FN(gparams)*)global_params;
  const int TILE = 32;
  if (block_pair)
    for (int64_t k = 0; k < 3LL * P * B * B; ++k) out[k] = 0;
  else
    for (int64_t k = 0; k < 3LL * P; ++k) out[k] = 0;
  if (dvdx)
    for (int64_t k = 0; k < 9LL * n_atoms; ++k) dvdx[k] = 0;
  REAL* gradp = require_grad ? dvdx : 0;

  for (int p = 0; p < P; ++p) {
    for (int b1 = 0; b1 < B; ++b1) {
      for (int b2 = b1; b2 < B; ++b2) {
        if (!neighbors[((int64_t)p * B + b1) * B + b2]) continue;
        int bt1 = st->block_type[p * B + b1], bt2 = st->block_type[p * B + b2];
        if (bt1 < 0 || bt2 < 0) continue;
        int64_t o1 = st->rot_coord_offset[p * B + b1];
        int64_t o2 = st->rot_coord_offset[p * B + b2];
        int n1 = st->bt_n_atoms[bt1], n2 = st->bt_n_atoms[bt2];
        REAL tot_atr = 0, tot_rep = 0, tot_lk = 0;
        if (b1 != b2) {
          int nti1 = (n1 - 1) / TILE + 1, nti2 = (n2 - 1) / TILE + 1;
          for (int ti = 0; ti < nti1; ++ti) {
            int s1 = TILE * ti;
            int m1 = n1 - s1 < TILE ? n1 - s1 : TILE;
            for (int tj = 0; tj < nti2; ++tj) {
              int s2 = TILE * tj;
              int m2 = n2 - s2 < TILE ? n2 - s2 : TILE;
              REAL pa = 0, pr = 0, pl = 0;
              for (int k = 0; k < m1 * m2; ++k) {
                int a1 = s1 + k / m2, a2 = s2 + k % m2;
                FN(tparams) t1, t2;
                FN(load_params)(type_params, st->bt_atom_types[bt1 * Am + a1], &t1);
                FN(load_params)(type_params, st->bt_atom_types[bt2 * Am + a2], &t2);
                int sep = FN(inter_sep)(st, p, b1, b2, bt1, bt2, a1, a2);
                REAL e[2];
                FN(eval_pair)(coords, n_atoms, o1 + a1, o2 + a2, sep, &t1, &t2,
                              g, 0, require_grad, e, gradp);
                pa += e[0];
                pr += e[1];
              }
              int h1 = st->bt_n_heavy_in_tile[bt1 * nt + ti];
              int h2 = st->bt_n_heavy_in_tile[bt2 * nt + tj];
              for (int k = 0; k < h1 * h2; ++k) {
                int a1 = s1 + st->bt_heavy_in_tile[bt1 * TILE * nt + TILE * ti + k / h2];
                int a2 = s2 + st->bt_heavy_in_tile[bt2 * TILE * nt + TILE * tj + k % h2];
                FN(tparams) t1, t2;
                FN(load_params)(type_params, st->bt_atom_types[bt1 * Am + a1], &t1);
                FN(load_params)(type_params, st->bt_atom_types[bt2 * Am + a2], &t2);
                int sep = FN(inter_sep)(st, p, b1, b2, bt1, bt2, a1, a2);
                REAL e[1];
                FN(eval_pair)(coords, n_atoms, o1 + a1, o2 + a2, sep, &t1, &t2,
                              g, 1, require_grad, e, gradp);
                pl += e[0];
              }
              tot_atr += pa;
              tot_rep += pr;
              tot_lk += pl;
            }
          }
        } else {
          int nti = (n1 - 1) / TILE + 1;
          const int32_t* pd = st->bt_path_distance + (int64_t)bt1 * Am * Am;
          for (int ti = 0; ti < nti; ++ti) {
            int s1 = TILE * ti;
            int m1 = n1 - s1 < TILE ? n1 - s1 : TILE;
            for (int tj = ti; tj < nti; ++tj) {
              int s2 = TILE * tj;
              int m2 = n1 - s2 < TILE ? n1 - s2 : TILE;
              REAL pa = 0, pr = 0, pl = 0;
              for (int k = 0; k < m1 * m2; ++k) {
                int a1 = s1 + k / m2, a2 = s2 + k % m2;
                if (a1 >= a2) continue;
                FN(tparams) t1, t2;
                FN(load_params)(type_params, st->bt_atom_types[bt1 * Am + a1], &t1);
                FN(load_params)(type_params, st->bt_atom_types[bt1 * Am + a2], &t2);
                int sep = pd[a1 * Am + a2];
                REAL e[2];
                FN(eval_pair)(coords, n_atoms, o1 + a1, o1 + a2, sep, &t1, &t2,
                              g, 0, require_grad, e, gradp);
                pa += e[0];
                pr += e[1];
              }
              int h1 = st->bt_n_heavy_in_tile[bt1 * nt + ti];
              int h2 = st->bt_n_heavy_in_tile[bt1 * nt + tj];
              for (int k = 0; k < h1 * h2; ++k) {
                int i1 = k / h2, i2 = k % h2;
                /* positions compared, as in IntraResBlockEvaluation */
                if (s1 + i1 >= s2 + i2) continue;
                int a1 = s1 + st->bt_heavy_in_tile[bt1 * TILE * nt + TILE * ti + i1];
                int a2 = s2 + st->bt_heavy_in_tile[bt1 * TILE * nt + TILE * tj + i2];
                FN(tparams) t1, t2;
                FN(load_params)(type_params, st->bt_atom_types[bt1 * Am + a1], &t1);
                FN(load_params)(type_params, st->bt_atom_types[bt1 * Am + a2], &t2);
                int sep = pd[a1 * Am + a2];
                REAL e[1];
                FN(eval_pair)(coords, n_atoms, o1 + a1, o1 + a2, sep, &t1, &t2,
                              g, 1, require_grad, e, gradp);
                pl += e[0];
              }
              tot_atr += pa;
              tot_rep += pr;
              tot_lk += pl;
            }
          }
        }
        if (block_pair) {
          int64_t base = ((int64_t)p * B + b1) * B + b2;
          int64_t PBB = (int64_t)P * B * B;
          out[base] = tot_atr;
          out[PBB + base] = tot_rep;
          out[2 * PBB + base] = tot_lk;
        } else {
          out[p] += tot_atr;
          out[P + p] += tot_rep;
          out[2 * P + p] += tot_lk;
        }
      }
    }
  }
}

#undef FN
#undef CAT
#undef CAT_
