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


/* Where gradients go: per-term planes dvdx [3, NA, 3] (w == NULL, the
 * whole-pose forward), or one plane [NA, 3] with term t scaled by w[t] (the
 * dTdV-weighted backward of the block-pair and rotamer-pair dispatches,
 * lj_atom_derivs / lk_atom_derivs in ljlk.hh:531-600, 686-738). */
typedef struct {
  REAL* g;
  int64_t na;
  const REAL* w;
} FN(gsink);

static void FN(sink_add)(const FN(gsink) * s, int term, int64_t atom, REAL dv,
                         const REAL* u) {
  for (int k = 0; k < 3; ++k) {
    REAL v = s->w ? (s->w[term] * dv) * u[k] : dv * u[k];
    if (v == 0) continue;
    if (s->w)
      s->g[atom * 3 + k] += v;
    else
      s->g[((int64_t)term * s->na + atom) * 3 + k] += v;
  }
}

/* one atom pair: LJ or LK, energy (+derivative into sink) */
static void FN(eval_pair)(const REAL* coords, int64_t i1, int64_t i2, int sep,
                          const FN(tparams) * t1, const FN(tparams) * t2,
                          const FN(gparams) * g, int lk, int require_grad,
                          REAL* e, const FN(gsink) * sink) {
  const REAL* x1 = coords + 3 * i1;
  const REAL* x2 = coords + 3 * i2;
  REAL d[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
  REAL dist = SQRT(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  REAL u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
  if (require_grad && sink && dist != 0)
    for (int k = 0; k < 3; ++k) {
      u1[k] = d[k] / dist;
      u2[k] = -d[k] / dist;
    }
  if (!lk) {
    if (require_grad) {
      REAL a, r, da, dr;
      FN(lj_V_dV)(dist, sep, t1, t2, g, &a, &r, &da, &dr);
      e[0] = a;
      e[1] = r;
      if (sink) {
        /* impl order (ljlk.hh:629-664): atom 1 atr/rep, then atom 2 */
        FN(sink_add)(sink, 0, i1, da, u1);
        FN(sink_add)(sink, 1, i1, dr, u1);
        FN(sink_add)(sink, 0, i2, da, u2);
        FN(sink_add)(sink, 1, i2, dr, u2);
      }
    } else {
      FN(lj_V)(dist, sep, t1, t2, g, &e[0], &e[1]);
    }
  } else {
    if (require_grad) {
      REAL v, dv;
      FN(lk_V_dV)(dist, sep, t1, t2, g, &v, &dv);
      e[0] = v;
      if (sink) {
        FN(sink_add)(sink, 2, i1, dv, u1);
        FN(sink_add)(sink, 2, i2, dv, u2);
      }
    } else {
      e[0] = FN(lk_V)(dist, sep, t1, t2, g);
    }
  }
}

/* One block pair (b1 != b2: inter-block; b1 == b2: intra-block) with the
 * given block types and first-atom offsets, in the reference CPU order:
 * tiles of 32 atoms, pairs row-major inside a tile pair
 * (tile_atom_pair_evaluation.hh:88-173, 377-505).  Energies accumulate in
 * REAL: pair -> tile-pair partial -> block-pair total (tot[3]). */
static void FN(score_block_pair)(const oracle_stack* st, const REAL* coords,
                                 const REAL* type_params, const FN(gparams) * g,
                                 int p, int b1, int b2, int bt1, int bt2,
                                 int64_t o1, int64_t o2, int require_grad,
                                 const FN(gsink) * sink, REAL* tot) {
  const int Am = st->max_n_block_atoms, nt = st->max_n_tiles;
  const int TILE = 32;
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
          FN(eval_pair)(coords, o1 + a1, o2 + a2, sep, &t1, &t2, g, 0,
                        require_grad, e, sink);
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
          FN(eval_pair)(coords, o1 + a1, o2 + a2, sep, &t1, &t2, g, 1,
                        require_grad, e, sink);
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
          FN(eval_pair)(coords, o1 + a1, o1 + a2, sep, &t1, &t2, g, 0,
                        require_grad, e, sink);
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
          FN(eval_pair)(coords, o1 + a1, o1 + a2, sep, &t1, &t2, g, 1,
                        require_grad, e, sink);
          pl += e[0];
        }
        tot_atr += pa;
        tot_rep += pr;
        tot_lk += pl;
      }
    }
  }
  tot[0] = tot_atr;
  tot[1] = tot_rep;
  tot[2] = tot_lk;
}

/* Score every neighbor block pair of a padded stack, in the reference CPU
 * order: pose-major, upper triangle (b1 <= b2) row by row
 * (impl.hh:633-1158).  Block-pair totals are added to the pose total.
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
  const FN(gparams)* g = (const FN(gparams)*)global_params;
  if (block_pair)
    for (int64_t k = 0; k < 3LL * P * B * B; ++k) out[k] = 0;
  else
    for (int64_t k = 0; k < 3LL * P; ++k) out[k] = 0;
  if (dvdx)
    for (int64_t k = 0; k < 9LL * n_atoms; ++k) dvdx[k] = 0;
  FN(gsink) sink = {dvdx, n_atoms, 0};
  const FN(gsink)* sp = (require_grad && dvdx) ? &sink : 0;

  for (int p = 0; p < P; ++p) {
    for (int b1 = 0; b1 < B; ++b1) {
      for (int b2 = b1; b2 < B; ++b2) {
        if (!neighbors[((int64_t)p * B + b1) * B + b2]) continue;
        int bt1 = st->block_type[p * B + b1], bt2 = st->block_type[p * B + b2];
        if (bt1 < 0 || bt2 < 0) continue;
        int64_t o1 = st->rot_coord_offset[p * B + b1];
        int64_t o2 = st->rot_coord_offset[p * B + b2];
        REAL tot[3];
        FN(score_block_pair)(st, coords, type_params, g, p, b1, b2, bt1, bt2,
                             o1, o2, require_grad, sp, tot);
        if (block_pair) {
          int64_t base = ((int64_t)p * B + b1) * B + b2;
          int64_t PBB = (int64_t)P * B * B;
          out[base] = tot[0];
          out[PBB + base] = tot[1];
          out[2 * PBB + base] = tot[2];
        } else {
          out[p] += tot[0];
          out[P + p] += tot[1];
          out[2 * P + p] += tot[2];
        }
      }
    }
  }
}

/* ---------- rotamer-pair scoring (packing) ---------------------------- */

/* Block spheres enclosing every rotamer of a block, from the rotamer
 * spheres (sphere_overlap.impl.hh:455-514): centre = sequential mean of the
 * rotamer centres, radius = max over rotamers of |c_r - c| + r_r.  Blocks
 * without rotamers keep zeros.  Rotamer spheres themselves are
 * oracle_block_spheres over rotamers (compute_rot_spheres, :23-107, nt = 1). */
void FN(oracle_rot_block_spheres)(int n_poses, int max_n_blocks,
                                  const REAL* rot_spheres,
                                  const int32_t* n_rots_for_block,
                                  const int32_t* rot_offset_for_block,
                                  REAL* block_spheres) {
  for (int64_t i = 0; i < (int64_t)n_poses * max_n_blocks; ++i) {
    REAL* s = block_spheres + 4 * i;
    s[0] = s[1] = s[2] = s[3] = 0;
    int n = n_rots_for_block[i];
    int r0 = rot_offset_for_block[i];
    if (n <= 0 || r0 < 0) continue;
    REAL cx = 0, cy = 0, cz = 0;
    for (int r = 0; r < n; ++r) {
      cx += rot_spheres[4 * ((int64_t)r0 + r) + 0];
      cy += rot_spheres[4 * ((int64_t)r0 + r) + 1];
      cz += rot_spheres[4 * ((int64_t)r0 + r) + 2];
    }
    cx /= n;
    cy /= n;
    cz /= n;
    REAL rmax = 0;
    for (int r = 0; r < n; ++r) {
      const REAL* q = rot_spheres + 4 * ((int64_t)r0 + r);
      REAL dx = q[0] - cx, dy = q[1] - cy, dz = q[2] - cz;
      REAL d = SQRT(dx * dx + dy * dy + dz * dz) + q[3];
      if (d > rmax) rmax = d;
    }
    s[0] = cx;
    s[1] = cy;
    s[2] = cz;
    s[3] = rmax;
  }
}

/* Score (dTdV == NULL: energies out[3, n]) or differentiate (dTdV [3, n]:
 * grad [NA, 3] += sum_t dTdV[t, i] dV_t(pair i)/dx) every rotamer pair of a
 * dispatch list indices [3, n] = (pose, rot1, rot2)
 * (ljlk_pose_score.impl.hh:1479-2126).  A pair of rotamers of one block is
 * the rotamer's intra-block energy.  Energies use the energy-only potentials
 * (lj_atom_energy / lk_atom_energy), as the reference forward does. */
void FN(oracle_rotamer_scores)(const oracle_stack* st, const REAL* coords,
                               int64_t n_atoms, const REAL* type_params,
                               const REAL* global_params, int64_t n_pairs,
                               const int32_t* indices,
                               const int32_t* rot_coord_offset,
                               const int32_t* block_ind_for_rot,
                               const int32_t* block_type_ind_for_rot,
                               const REAL* dTdV, REAL* out, REAL* grad) {
  const FN(gparams)* g = (const FN(gparams)*)global_params;
  if (out)
    for (int64_t k = 0; k < 3 * n_pairs; ++k) out[k] = 0;
  if (grad)
    for (int64_t k = 0; k < 3 * n_atoms; ++k) grad[k] = 0;
  for (int64_t i = 0; i < n_pairs; ++i) {
    int p = indices[i], r1 = indices[n_pairs + i], r2 = indices[2 * n_pairs + i];
    if (p < 0 || r1 < 0 || r2 < 0) continue;
    int bt1 = block_type_ind_for_rot[r1], bt2 = block_type_ind_for_rot[r2];
    if (bt1 < 0 || bt2 < 0) continue;
    int b1 = block_ind_for_rot[r1], b2 = block_ind_for_rot[r2];
    REAL tot[3];
    if (dTdV) {
      REAL w[3] = {dTdV[i], dTdV[n_pairs + i], dTdV[2 * n_pairs + i]};
      FN(gsink) sink = {grad, n_atoms, w};
      FN(score_block_pair)(st, coords, type_params, g, p, b1, b2, bt1, bt2,
                           rot_coord_offset[r1], rot_coord_offset[r2], 1,
                           &sink, tot);
    } else {
      FN(score_block_pair)(st, coords, type_params, g, p, b1, b2, bt1, bt2,
                           rot_coord_offset[r1], rot_coord_offset[r2], 0, 0,
                           tot);
      out[i] = tot[0];
      out[n_pairs + i] = tot[1];
      out[2 * n_pairs + i] = tot[2];
    }
  }
}

#undef FN
#undef CAT
#undef CAT_
