// LJ + LK pair potentials as device functions, written for wave64 VALU code.
//
// Same functions as the reference (tmol/score/ljlk/potentials/lj.hh:21-177,
// lk_isotropic.hh:22-360, common.hh:13-44, tmol/score/common/
// cubic_hermite_polynomial.hh), arranged so that each region of a piecewise
// potential evaluates the underlying closed form once at a clamped abscissa
// instead of branching per region:
//
//   LJ:  x = clamp(d, d_lin, cpoly_dmin); one 12-6 evaluation V(x), V'(x);
//        linear extension below d_lin, Hermite-to-zero above cpoly_dmin.
//   LK:  one f_desolv evaluation at clamp(d, close_dmax, far_dmin) per
//        direction, a second one at d_min only where d < close_dmax.
//
// Float builds use the hardware rcp / sqrt / exp (1-2 ulp); the double
// instantiation (the oracle run) uses correctly rounded division, sqrt and
// exp.  Results agree with the reference's formulas to rounding.
#pragma once

#include <hip/hip_runtime.h>

namespace tmol_ljlk {

template <typename Real>
struct GlobalParams {
  Real max_dis;               // cpoly_dmax (lj.hh:125), cpoly_far_dmax (lk)
  Real dlin_factor;           // lj_dlin_sigma_factor
  Real hbond_dis;             // lj_hbond_dis
  Real hbond_oh_dis;          // lj_hbond_OH_donor_dis
  Real hbond_hdis;            // lj_hbond_hdis
  Real spline_start;          // max_dis - 1.5
  Real cpoly_cap;             // max_dis - 0.1
  Real max_dis2;              // max_dis^2 (pair culling)
};

// per-atom flag bits
enum : int {
  F_DONOR = 1,
  F_OH_DONOR = 2,  // donor and hydroxyl
  F_POLARH = 4,
  F_ACCEPTOR = 8,
  F_CARBON_LK = 16,
};

template <typename Real>
__device__ __forceinline__ Real fast_rcp(Real x);
template <>
__device__ __forceinline__ float fast_rcp<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}
template <>
__device__ __forceinline__ double fast_rcp<double>(double x) {
  return 1.0 / x;
}

template <typename Real>
__device__ __forceinline__ Real fast_sqrt(Real x);
template <>
__device__ __forceinline__ float fast_sqrt<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}
template <>
__device__ __forceinline__ double fast_sqrt<double>(double x) {
  return __builtin_sqrt(x);
}

// 1 / sqrt(x): hardware v_rsq_f32 for float, correctly rounded ops for double
template <typename Real>
__device__ __forceinline__ Real fast_rsqrt(Real x);
template <>
__device__ __forceinline__ float fast_rsqrt<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}
template <>
__device__ __forceinline__ double fast_rsqrt<double>(double x) {
  return 1.0 / __builtin_sqrt(x);
}

template <typename Real>
__device__ __forceinline__ Real fast_exp(Real x);
template <>
__device__ __forceinline__ float fast_exp<float>(float x) {
  return __expf(x);
}
template <>
__device__ __forceinline__ double fast_exp<double>(double x) {
  return exp(x);
}

// connectivity weight: >4 bonds -> 1, ==4 -> 0.2, else 0 (common.hh:13-22)
template <typename Real>
__device__ __forceinline__ Real cp_weight(int sep) {
  return sep > 4 ? Real(1) : (sep == 4 ? Real(0.2) : Real(0));
}

// hbond-aware sigma (common.hh:24-44): donor-acceptor pairs use the hbond
// distance (the hydroxyl-donor one if either atom is a hydroxyl donor),
// polar-hydrogen-acceptor pairs the hbond-H distance, others r_i + r_j.
// With the flag bits D = 1, OH = 2, PH = 4, A = 8: bit 0 of
// (Ai & (Dj | PHj)) | (Aj & (Di | PHi)) is the donor-acceptor test and bit 2
// the polar-H-acceptor test, each the OR of both orientations.
template <typename Real>
__device__ __forceinline__ Real lj_sigma(int fi, int fj, Real ri, Real rj,
                                         const GlobalParams<Real>& g) {
  const int both = ((fi >> 3) & 1 ? fj : 0) | ((fj >> 3) & 1 ? fi : 0);
  const bool da = both & F_DONOR;
  const bool ha = both & F_POLARH;
  const bool oh = (fi | fj) & F_OH_DONOR;
  Real s = ri + rj;
  s = ha ? g.hbond_hdis : s;
  s = da ? (oh ? g.hbond_oh_dis : g.hbond_dis) : s;
  return s;
}

// LJ for one pair: returns weighted (atr, rep) and d(atr+rep)/dd.
// Only one of atr/rep carries the derivative: rep inside sigma, atr outside.
template <typename Real, bool DERIV>
__device__ __forceinline__ void lj_pair(Real dist, Real sigma, Real eps,
                                        Real weight,
                                        const GlobalParams<Real>& g,
                                        Real& atr, Real& rep, Real& dlj) {
  Real d_lin = sigma * g.dlin_factor;
  Real cpoly_dmin = sigma > g.spline_start ? sigma : g.spline_start;
  cpoly_dmin = cpoly_dmin > g.cpoly_cap ? g.cpoly_cap : cpoly_dmin;
  // evaluation point of the 12-6 form
  Real x = dist < d_lin ? d_lin : (dist < cpoly_dmin ? dist : cpoly_dmin);
  Real inv_x = fast_rcp(x);
  Real sd = sigma * inv_x;
  Real sd2 = sd * sd;
  Real sd6 = sd2 * sd2 * sd2;
  Real sd12 = sd6 * sd6;
  Real V = eps * (sd12 - Real(2) * sd6);
  Real dV = eps * Real(12) * (sd6 - sd12) * inv_x;
  Real lj, dl;
  if (dist < d_lin) {
    lj = V + dV * (dist - d_lin);
    dl = dV;
  } else if (dist < cpoly_dmin) {
    lj = V;
    dl = dV;
  } else {
    // Hermite to (0, 0) at max_dis (interpolate_to_zero_V_dV)
    Real w = g.max_dis - cpoly_dmin;
    Real inv_w = fast_rcp(w);
    Real t = (dist - cpoly_dmin) * inv_w;
    Real dp0 = dV * w;
    lj = V + t * (dp0 + t * (Real(-2) * dp0 - Real(3) * V + t * (dp0 + Real(2) * V)));
    dl = (dp0 + t * (Real(-4) * dp0 - Real(6) * V + t * (Real(3) * dp0 + Real(6) * V))) * inv_w;
    bool out = dist >= g.max_dis;
    lj = out ? Real(0) : lj;
    dl = out ? Real(0) : dl;
  }
  bool inside = dist < sigma;
  atr = weight * (inside ? -eps : lj);
  rep = weight * (inside ? lj + eps : Real(0));
  dlj = DERIV ? weight * dl : Real(0);
}

// LK knots on the 1/20 A^2 etable grid (lk_isotropic.hh:116-125)
template <typename Real>
__device__ __forceinline__ void lk_knots(Real sigma, bool is_cc, Real& d_min,
                                         Real& close_dmin, Real& close_dmax) {
  Real dm = Real(double(sigma) * 0.89);
  if (is_cc) dm = dm > Real(4.2) ? dm : Real(4.2);
  Real n = floor(Real(20) * dm * dm);
  Real lo = n - Real(29);
  lo = lo > Real(0) ? lo : Real(0);
  Real hi = n + Real(21);
  hi = hi < Real(405) ? hi : Real(405);
  d_min = dm;
  close_dmin = fast_sqrt(lo * Real(0.05));
  close_dmax = fast_sqrt(hi * Real(0.05));
}

// f_desolv and its derivative with c = -vol_j * dg_i / (2 pi^1.5 lambda_i)
// precomputed per atom as -vol_j * K_i, K_i = dg_i / (2 pi^1.5 lambda_i)
template <typename Real>
__device__ __forceinline__ void fdesolv(Real x, Real c, Real r_i,
                                        Real inv_lam2_i, Real& f, Real& df) {
  Real inv_x = fast_rcp(x);
  Real dr = x - r_i;
  Real e = fast_exp(-dr * dr * inv_lam2_i);
  f = c * inv_x * inv_x * e;
  df = f * (Real(-2) * inv_x - Real(2) * dr * inv_lam2_i);
}

// one direction of isotropic LK (lk_isotropic.hh:167-235), unweighted
template <typename Real, bool DERIV>
__device__ __forceinline__ void lk_dir(Real dist, Real d_min, Real cdmin,
                                       Real cdmax, Real c, Real r_i,
                                       Real inv_lam2_i,
                                       const GlobalParams<Real>& g, Real& V,
                                       Real& dV) {
  Real x = dist < cdmax ? cdmax : (dist < g.spline_start ? dist : g.spline_start);
  Real f, df;
  fdesolv(x, c, r_i, inv_lam2_i, f, df);
  if (dist >= cdmax) {
    if (dist < g.spline_start) {
      V = f;
      dV = df;
    } else {
      // Hermite to zero over [max_dis - 1.5, max_dis]
      Real w = g.max_dis - g.spline_start;
      Real inv_w = fast_rcp(w);
      Real t = (dist - g.spline_start) * inv_w;
      Real dp0 = df * w;
      V = f + t * (dp0 + t * (Real(-2) * dp0 - Real(3) * f + t * (dp0 + Real(2) * f)));
      dV = (dp0 + t * (Real(-4) * dp0 - Real(6) * f + t * (Real(3) * dp0 + Real(6) * f))) * inv_w;
      bool out = dist >= g.max_dis;
      V = out ? Real(0) : V;
      dV = out ? Real(0) : dV;
    }
  } else {
    Real f0, df0;
    fdesolv(d_min, c, r_i, inv_lam2_i, f0, df0);
    if (dist < cdmin) {
      V = f0;
      dV = Real(0);
    } else {
      // Hermite from (cdmin, f(d_min), 0) to (cdmax, f, df)
      Real w = cdmax - cdmin;
      Real inv_w = fast_rcp(w);
      Real t = (dist - cdmin) * inv_w;
      Real dp1 = df * w;
      Real p0 = f0, p1 = f;
      V = p0 + t * (t * (-dp1 - Real(3) * p0 + Real(3) * p1 + t * (dp1 + Real(2) * p0 - Real(2) * p1)));
      dV = (t * (Real(-2) * dp1 - Real(6) * p0 + Real(6) * p1 + t * (Real(3) * dp1 + Real(6) * p0 - Real(6) * p1))) * inv_w;
    }
  }
  if (!DERIV) dV = Real(0);
}

// ---- elec (tmol/score/elec/potentials/potentials.hh)

// Per-launch constants of the potential (computed on the host in the
// reference's precision: eps() at max_dis and min_dis in float, the rest in
// Real).  Hermite pieces take the charge product out as a common factor.
template <typename Real>
struct ElecConst {
  Real D, S, hdd;       // hdd = 0.5 (D - D0)
  Real C1, C2;          // 322.0637, C1 / (max_dis eps(max_dis))
  Real lps, lpe;        // min_dis -+ 0.25
  Real hps, hpe;        // max_dis - 1, max_dis
  Real mds;             // C1 / (min_dis eps(min_dis)) - C2 (flat part)
  Real e_lo, de_lo;     // Coulomb value and slope at lpe (per unit q1 q2)
  Real e_hi, de_hi;     // ... at hps
  Real lw, hw;          // lpe - lps, hpe - hps
  Real inv_lw, inv_hw;
  Real cull2;           // (max_dis + 1e-3)^2: phase A keeps pairs below it
};

// V and dV/dd per unit charge product (potentials.hh:55-138).  Branch-free
// over the wave: the Coulomb form is evaluated at clamp(d, lpe, hps), and one
// Hermite piece, its ends chosen by region, covers both fades (below lps it
// sits at t = 0: the flat part, slope 0).
template <typename Real, bool DERIV>
__device__ __forceinline__ void elec_unit(Real dist, const ElecConst<Real>& k, Real& V,
                                          Real& dV) {
  const Real x = dist < k.lpe ? k.lpe : (dist < k.hps ? dist : k.hps);
  const Real ex = fast_exp(-x * k.S);
  const Real xs = x * k.S;
  const Real eps = k.D - k.hdd * (Real(2) + Real(2) * xs + xs * xs) * ex;
  const Real r = fast_rcp(eps);
  const Real ix = fast_rcp(x);
  const Real vc = k.C1 * ix * r - k.C2;
  Real dc = Real(0);
  if (DERIV) {
    const Real deps = k.hdd * xs * xs * k.S * ex;  // 0.5 (D - D0) x^2 S^3 e^{-xS}
    dc = -k.C1 * (eps + x * deps) * ix * ix * r * r;
  }
  const bool lo = dist < k.lpe;
  const bool hi = dist >= k.hps;
  const Real x0 = lo ? k.lps : k.hps;
  const Real w = lo ? k.lw : k.hw;
  const Real iw = lo ? k.inv_lw : k.inv_hw;
  const Real p0 = lo ? k.mds : k.e_hi;
  const Real p1 = lo ? k.e_lo : Real(0);
  const Real dp0 = lo ? Real(0) : k.de_hi * w;
  const Real dp1 = lo ? k.de_lo * w : Real(0);
  Real t = (dist - x0) * iw;
  t = t > Real(0) ? t : Real(0);
  const Real c2 = Real(-2) * dp0 - dp1 - Real(3) * p0 + Real(3) * p1;
  const Real c3 = dp0 + dp1 + Real(2) * p0 - Real(2) * p1;
  const Real vh = p0 + t * (dp0 + t * (c2 + t * c3));
  const Real dh = (dp0 + t * (Real(2) * c2 + Real(3) * t * c3)) * iw;
  const bool fade = lo || hi;
  V = fade ? vh : vc;
  dV = DERIV ? (fade ? dh : dc) : Real(0);
  const bool out = dist >= k.hpe;
  V = out ? Real(0) : V;
  dV = out ? Real(0) : dV;
}

}  // namespace tmol_ljlk
