/* CPU restatement of the reference elec pose-score op -- TEST INFRASTRUCTURE.
 *
 * The checker, never the product (same rules as ljlk_oracle_impl.h, which
 * must be included first: it provides the Hermite helpers and the count-pair
 * routine inter_sep).  Instantiated for float and double by ljlk_oracle.c.
 *
 * Follows tmol/score/elec/potentials/potentials.hh:31-201 (the potential,
 * including its float/double promotions: eps() of max_dis and min_dis runs in
 * float whatever Real is, and the 0.5 literal lifts part of eps to double)
 * and elec_pose_score.impl.hh:383-856 with elec.hh:406-548 (serial CPU
 * order: tiles of 32 atoms, pairs row-major, intra-block pairs a1 < a2;
 * count pair through the representative-atom path distances).
 *
 * Pinned by: tmol's elec term baselines for 1ubq (whole pose x 10, jagged,
 * block pair, the 3-residue smoke value), potential invariants (continuity
 * at the region joins, the Coulomb closed form, the derivative against
 * central differences), and the reference's own CPU op where oracle/_ref
 * builds (tests/test_elec_oracle.py).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, RSUF)

/* sigmoidal distance-dependent dielectric, potentials.hh:43-53; D, D0, S are
 * float parameters, the 0.5 literal is double */
static REAL FN(elec_eps)(REAL dist, float D, float D0, float S) {
  double a = 0.5 * (D - D0);
  REAL poly = 2 + 2 * dist * S + dist * dist * S * S;
  return (REAL)(D - a * poly * EXP(-dist * S));
}

static REAL FN(elec_deps)(REAL dist, float D, float D0, float S) {
  return (REAL)(0.5 * (D - D0) * dist * dist * S * S * S * EXP(-dist * S));
}

/* elec_delec_ddist (potentials.hh:55-138) and elec (:140-201) in one body;
 * dE == NULL selects the energy-only form.  Parameters are floats, as in the
 * reference's signature. */
static REAL FN(elec_pair)(REAL dist, REAL e_i, REAL e_j, REAL bonded_path_length,
                          float D, float D0, float S, float min_dis, float max_dis,
                          REAL* dE) {
  REAL low_poly_start = (REAL)(min_dis - 0.25);
  REAL low_poly_end = (REAL)(min_dis + 0.25);
  REAL hi_poly_start = (REAL)(max_dis - 1.0);
  REAL hi_poly_end = max_dis;
  REAL weight = bonded_path_length > 4 ? (REAL)1.0
                : bonded_path_length == 4 ? (REAL)0.2 : (REAL)0.0;
  REAL C1 = (REAL)322.0637;
  REAL C2 = C1 / (max_dis * elec_eps_float(max_dis, D, D0, S));
  REAL eiej = e_i * e_j;
  REAL E = 0, dV = 0;
  if (dE) *dE = 0;
  if (eiej == 0) return 0;
  if (dist < low_poly_start) {
    REAL min_dis_score = C1 / (min_dis * elec_eps_float(min_dis, D, D0, S)) - C2;
    E = eiej * min_dis_score;
    dV = 0;
  } else if (dist < low_poly_end) {
    REAL min_dis_score = C1 / (min_dis * elec_eps_float(min_dis, D, D0, S)) - C2;
    REAL eps_elec = FN(elec_eps)(low_poly_end, D, D0, S);
    REAL deps = FN(elec_deps)(low_poly_end, D, D0, S);
    REAL dmax = eiej * (C1 / (low_poly_end * eps_elec) - C2);
    REAL dmax_d = -C1 * eiej * (eps_elec + low_poly_end * deps) /
                  (low_poly_end * low_poly_end * eps_elec * eps_elec);
    if (dE)
      FN(interp_V_dV)(dist, low_poly_start, eiej * min_dis_score, (REAL)0.0, low_poly_end,
                      dmax, dmax_d, &E, &dV);
    else
      E = FN(interp)(dist, low_poly_start, eiej * min_dis_score, (REAL)0.0, low_poly_end, dmax,
                     dmax_d);
  } else if (dist < hi_poly_start) {
    REAL eps_elec = FN(elec_eps)(dist, D, D0, S);
    REAL deps = FN(elec_deps)(dist, D, D0, S);
    E = eiej * (C1 / (dist * eps_elec) - C2);
    dV = -C1 * eiej * (eps_elec + dist * deps) / (dist * dist * eps_elec * eps_elec);
  } else if (dist < hi_poly_end) {
    REAL eps_elec = FN(elec_eps)(hi_poly_start, D, D0, S);
    REAL deps = FN(elec_deps)(hi_poly_start, D, D0, S);
    REAL dmin = eiej * (C1 / (hi_poly_start * eps_elec) - C2);
    REAL dmin_d = -C1 * eiej * (eps_elec + hi_poly_start * deps) /
                  (hi_poly_start * hi_poly_start * eps_elec * eps_elec);
    if (dE)
      FN(interp0_V_dV)(dist, hi_poly_start, dmin, dmin_d, hi_poly_end, &E, &dV);
    else
      E = FN(interp0)(dist, hi_poly_start, dmin, dmin_d, hi_poly_end);
  }
  if (dE) *dE = weight * dV;
  return weight * E;
}

/* gp = {D, D0, S, min_dis, max_dis} (ElecGlobalParams, params.hh:12-18) */
void FN(oracle_elec_pair)(REAL dist, REAL q1, REAL q2, int sep, const REAL* gp,
                          int with_deriv, REAL* out) {
  REAL d = 0;
  out[0] = FN(elec_pair)(dist, q1, q2, (REAL)sep, (float)gp[0], (float)gp[1], (float)gp[2],
                         (float)gp[3], (float)gp[4], with_deriv ? &d : 0);
  out[1] = d;
}

/* one atom pair into the block-pair partial (elec.hh:406-548): energy, and
 * w * dV/dx of both atoms into dvdx (w = 1 for the pose op, dTdV of the
 * rotamer pair for the rotamer op, elec.hh:466, 478), zero components
 * skipped */
static REAL FN(elec_eval_pair)(const REAL* coords, int64_t i1, int64_t i2, int sep,
                               REAL q1, REAL q2, const REAL* gp, int require_grad,
                               REAL w, REAL* dvdx) {
  const REAL* x1 = coords + 3 * i1;
  const REAL* x2 = coords + 3 * i2;
  REAL d[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
  REAL dist = SQRT(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  float D = (float)gp[0], D0 = (float)gp[1], S = (float)gp[2];
  float mn = (float)gp[3], mx = (float)gp[4];
  if (!require_grad) return FN(elec_pair)(dist, q1, q2, (REAL)sep, D, D0, S, mn, mx, 0);
  REAL dV = 0;
  REAL V = FN(elec_pair)(dist, q1, q2, (REAL)sep, D, D0, S, mn, mx, &dV);
  if (dvdx) {
    REAL wdV = w * dV;
    for (int k = 0; k < 3; ++k) {
      REAL u = dist != 0 ? d[k] / dist : 0;
      REAL v = wdV * u;
      if (v != 0) dvdx[i1 * 3 + k] += v;
    }
    for (int k = 0; k < 3; ++k) {
      REAL u = dist != 0 ? -d[k] / dist : 0;
      REAL v = wdV * u;
      if (v != 0) dvdx[i2 * 3 + k] += v;
    }
  }
  return V;
}

/* Elec of one block pair (b1 <= b2 of pose p; types bt1, bt2; first atoms
 * o1, o2), reference CPU order: tiles of 32 atoms, pairs row-major inside a
 * tile pair, one partial per tile pair; b1 == b2 takes the pairs a1 < a2
 * with the intra-block representative separations. */
static REAL FN(elec_block_pair)(const oracle_stack* st, const REAL* coords, const REAL* charges,
                                const int32_t* intra_pd, const REAL* gp, int p, int b1, int b2,
                                int bt1, int bt2, int64_t o1, int64_t o2, int require_grad,
                                REAL w, REAL* dvdx) {
  const int Am = st->max_n_block_atoms;
  const int TILE = 32;
  int n1 = st->bt_n_atoms[bt1], n2 = st->bt_n_atoms[bt2];
  const REAL* q1 = charges + (int64_t)bt1 * Am;
  const REAL* q2 = charges + (int64_t)bt2 * Am;
  REAL tot = 0;
  int nti1 = (n1 - 1) / TILE + 1, nti2 = (n2 - 1) / TILE + 1;
  for (int ti = 0; ti < nti1; ++ti) {
    int s1 = TILE * ti;
    int m1 = n1 - s1 < TILE ? n1 - s1 : TILE;
    for (int tj = (b1 == b2 ? ti : 0); tj < nti2; ++tj) {
      int s2 = TILE * tj;
      int m2 = n2 - s2 < TILE ? n2 - s2 : TILE;
      REAL part = 0;
      for (int k = 0; k < m1 * m2; ++k) {
        int a1 = s1 + k / m2, a2 = s2 + k % m2;
        int sep;
        if (b1 == b2) {
          if (a1 >= a2) continue;
          sep = intra_pd[((int64_t)bt1 * Am + a1) * Am + a2];
        } else {
          sep = FN(inter_sep)(st, p, b1, b2, bt1, bt2, a1, a2);
        }
        part += FN(elec_eval_pair)(coords, o1 + a1, o2 + a2, sep, q1[a1], q2[a2], gp,
                                   require_grad, w, dvdx);
      }
      tot += part;
    }
  }
  return tot;
}

/* Elec over every neighbor block pair of a padded stack, reference CPU
 * order (pose-major upper triangle; tiles of 32; pairs row-major).
 *   st:       bt_path_distance holds the inter-block representative path
 *             distances [T, Am, Am] (entry [c_atom][a] = path_dist[c_atom][rep(a)])
 *   intra_pd: [T, Am, Am] path_dist[rep(a1)][rep(a2)]
 *   charges:  [T, Am]
 *   out:      [P] or [P, B, B] (block_pair)
 *   dvdx:     [NA, 3] or NULL */
void FN(oracle_elec_scores)(const oracle_stack* st, const REAL* coords, int64_t n_atoms,
                            const REAL* charges, const int32_t* intra_pd, const REAL* gp,
                            const int32_t* neighbors, int block_pair, int require_grad,
                            REAL* out, REAL* dvdx) {
  const int P = st->n_poses, B = st->max_n_blocks;
  if (block_pair)
    for (int64_t k = 0; k < (int64_t)P * B * B; ++k) out[k] = 0;
  else
    for (int p = 0; p < P; ++p) out[p] = 0;
  if (dvdx)
    for (int64_t k = 0; k < 3 * n_atoms; ++k) dvdx[k] = 0;
  for (int p = 0; p < P; ++p)
    for (int b1 = 0; b1 < B; ++b1)
      for (int b2 = b1; b2 < B; ++b2) {
        if (!neighbors[((int64_t)p * B + b1) * B + b2]) continue;
        int bt1 = st->block_type[p * B + b1], bt2 = st->block_type[p * B + b2];
        if (bt1 < 0 || bt2 < 0) continue;
        int64_t o1 = st->rot_coord_offset[p * B + b1];
        int64_t o2 = st->rot_coord_offset[p * B + b2];
        REAL tot = FN(elec_block_pair)(st, coords, charges, intra_pd, gp, p, b1, b2, bt1, bt2,
                                       o1, o2, require_grad, (REAL)1, dvdx);
        if (block_pair)
          out[((int64_t)p * B + b1) * B + b2] = tot;
        else
          out[p] += tot;
      }
}

/* Elec of every rotamer pair of a dispatch list indices [3, n] = (pose,
 * rot1, rot2) (ElecRotamerScoreDispatch, elec_pose_score.impl.hh:1132-1436
 * forward, :1437-1700 backward; tile_evaluate_rot_pair,
 * tile_atom_pair_evaluation.hh:377-500): dTdV == NULL gives the energies
 * out[n]; dTdV [n] gives grad [NA, 3] = sum_i dTdV[i] dV(pair i)/dx.  The
 * two rotamers of a pair take their own block types and first atoms; the
 * count pair of different blocks uses the pose's inter-block separations,
 * and a pair of rotamers of one block (a rotamer with itself) is that
 * rotamer's intra-block energy.  Pairs with a rotamer of no block type score
 * zero. */
void FN(oracle_elec_rotamer_scores)(const oracle_stack* st, const REAL* coords, int64_t n_atoms,
                                    const REAL* charges, const int32_t* intra_pd, const REAL* gp,
                                    int64_t n_pairs, const int32_t* indices,
                                    const int32_t* rot_coord_offset,
                                    const int32_t* block_ind_for_rot,
                                    const int32_t* block_type_ind_for_rot, const REAL* dTdV,
                                    REAL* out, REAL* grad) {
  if (out)
    for (int64_t k = 0; k < n_pairs; ++k) out[k] = 0;
  if (grad)
    for (int64_t k = 0; k < 3 * n_atoms; ++k) grad[k] = 0;
  for (int64_t i = 0; i < n_pairs; ++i) {
    int p = indices[i], r1 = indices[n_pairs + i], r2 = indices[2 * n_pairs + i];
    if (p < 0 || r1 < 0 || r2 < 0) continue;
    int bt1 = block_type_ind_for_rot[r1], bt2 = block_type_ind_for_rot[r2];
    if (bt1 < 0 || bt2 < 0) continue;
    int b1 = block_ind_for_rot[r1], b2 = block_ind_for_rot[r2];
    if (dTdV) {
      FN(elec_block_pair)(st, coords, charges, intra_pd, gp, p, b1, b2, bt1, bt2,
                          rot_coord_offset[r1], rot_coord_offset[r2], 1, dTdV[i], grad);
    } else {
      out[i] = FN(elec_block_pair)(st, coords, charges, intra_pd, gp, p, b1, b2, bt1, bt2,
                                   rot_coord_offset[r1], rot_coord_offset[r2], 0, (REAL)1, 0);
    }
  }
}

#undef FN
#undef CAT
#undef CAT_
