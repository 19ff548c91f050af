// Rotamer-pair LJ+LK scoring for packing (reference: LJLKRotamerScoreDispatch,
// ljlk_pose_score.impl.hh:1479-2126; sphere_overlap.impl.hh:23-107, 455-610).
//
//   spheres   one bounding sphere per rotamer (spheres_kernel run over
//             rotamers), then one per block enclosing all of its rotamers
//   rows      one wave per block b1 of a pose: the neighbor test against
//             every b2 >= b1 (bit-identical to the reference), and the row's
//             number of rotamer pairs and of work units; an exclusive scan
//             gives each row its first dispatch entry and first unit
//   fill      the same test again; each neighbor cell (b1, b2) writes its
//             dispatch entries (pose, rot1, rot2) in the reference's order
//             (cells row-major, rot1-major inside a cell, self pairs only on
//             the diagonal) and its work units: up to UNIT_R1 rotamers of b1
//             against every rotamer of b2
//   score     persistent waves take units from per-XCD queues.  The lanes of
//             a wave hold the atoms of a pack of consecutive b2 rotamers (as
//             many as fit in 64 lanes); for each b1 rotamer of the unit its
//             atoms are staged in LDS and read as wave-uniform broadcasts, so
//             each lane evaluates (atom a1, its own atom) and a whole wave
//             covers 64 atom pairs of up to a few rotamer pairs at once
//             (block types of at most 64 atoms; larger ones are rejected).  An
//             a1 is skipped when no lane is within max_dis (wave-uniform
//             branch).  Each rotamer pair's energies are a segmented lane
//             reduction; the backward scales each pair by its dTdV.
#pragma once

#include "ljlk_score_rows.hh"

namespace tmol_ljlk {

constexpr int UNIT_R1 = 4;  // b1 rotamers per work unit

struct RotView {
  int P, B, R, C;
  const int32_t* rot_coord_offset;      // [R]
  const int32_t* block_ind_for_rot;     // [R]
  const int32_t* block_type_for_rot;    // [R]
  const int32_t* n_rots_for_block;      // [P * B]
  const int32_t* rot_offset_for_block;  // [P * B]
  const int32_t* first_rot_block_type;  // [P * B]
  const int32_t* min_bondsep;           // [P, B, B]
  const int32_t* inter_block_bondsep;   // [P, B, B, C, C]
};

// One work unit: rotamers [r1_first, r1_first + r1_count) of block b1 against
// every rotamer of block b2 (b1 == b2: the self pairs of those rotamers).
struct alignas(16) RotUnit {
  int32_t g1;        // pose * B + b1
  int32_t b2;
  int32_t r1_first;  // global rotamer index
  int32_t r1_count;  // < 0: -count of a cell whose rotamer offsets are negative
                     // (its entries get zero energies)
  int64_t entry;     // dispatch entry of (r1_first, first rotamer of b2)
  int64_t pad;
};
static_assert(sizeof(RotUnit) == 32, "RotUnit is 32 bytes");

// ------------------------------------------------------------------ spheres

// Block spheres from rotamer spheres, bit-identical to the reference
// (compute_block_spheres_from_rot_spheres, sphere_overlap.impl.hh:455-514):
// centre = sequential mean of the rotamer centres, radius = max over the
// rotamers of sqrt((dx*dx + dy*dy) + dz*dz) + r_rot; no contraction.
template <typename Real>
__global__ void rot_block_spheres_kernel(RotView v, const Real* __restrict__ rot_spheres,
                                         Real* __restrict__ block_spheres) {
#pragma clang fp contract(off)
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(v.P) * v.B) return;
  Real* s = block_spheres + i * 4;
  const int n = v.n_rots_for_block[i], r0 = v.rot_offset_for_block[i];
  if (n <= 0 || r0 < 0) {
    s[0] = s[1] = s[2] = s[3] = Real(0);
    return;
  }
  const Real* q = rot_spheres + int64_t(r0) * 4;
  Real cx = 0, cy = 0, cz = 0;
  for (int r = 0; r < n; ++r) {
    cx = cx + q[4 * r + 0];
    cy = cy + q[4 * r + 1];
    cz = cz + q[4 * r + 2];
  }
  const Real rn = Real(n);
  cx = cx / rn;
  cy = cy / rn;
  cz = cz / rn;
  Real rmax = 0;
  for (int r = 0; r < n; ++r) {
    const Real dx = q[4 * r + 0] - cx;
    const Real dy = q[4 * r + 1] - cy;
    const Real dz = q[4 * r + 2] - cz;
    const Real dxx = dx * dx;
    const Real dyy = dy * dy;
    const Real dzz = dz * dz;
    const Real d = sqrt((dxx + dyy) + dzz) + q[4 * r + 3];
    rmax = d > rmax ? d : rmax;
  }
  s[0] = cx;
  s[1] = cy;
  s[2] = cz;
  s[3] = rmax;
}

// neighbor test of cell (g1, g2), detect_block_neighbors
// (sphere_overlap.impl.hh:287-351): both blocks have rotamers and
// ((dx*dx + dy*dy) + dz*dz) < ((r1 + r2) + reach)^2
template <typename Real>
__device__ __forceinline__ bool rot_cell_hit(const RotView& v, const Real* __restrict__ bsph,
                                             Real s0, Real s1, Real s2, Real s3, int64_t g2,
                                             Real reach) {
#pragma clang fp contract(off)
  if (v.first_rot_block_type[g2] < 0) return false;
  const Real dx = s0 - bsph[g2 * 4 + 0];
  const Real dy = s1 - bsph[g2 * 4 + 1];
  const Real dz = s2 - bsph[g2 * 4 + 2];
  const Real dxx = dx * dx;
  const Real dyy = dy * dy;
  const Real dzz = dz * dz;
  const Real d2 = (dxx + dyy) + dzz;
  const Real thr = (s3 + bsph[g2 * 4 + 3]) + reach;
  return d2 < thr * thr;
}

// ------------------------------------------------------------------ dispatch

// One wave per row g1 = pose * B + b1: rotamer pairs and work units of the
// row's neighbor cells.
template <typename Real>
__global__ __launch_bounds__(256) void rot_rows_count_kernel(RotView v,
                                                             const Real* __restrict__ bsph,
                                                             Real reach,
                                                             int64_t* __restrict__ row_pairs,
                                                             int64_t* __restrict__ row_units) {
  const int lane = lane_id();
  const int64_t g1 = int64_t(blockIdx.x) * WAVES_PER_WG + threadIdx.x / WAVE;
  if (g1 >= int64_t(v.P) * v.B) return;
  const int p = int(g1 / v.B), b1 = int(g1 - int64_t(p) * v.B);
  int64_t pairs = 0, units = 0;
  if (v.first_rot_block_type[g1] >= 0) {
    const Real s0 = bsph[g1 * 4 + 0], s1 = bsph[g1 * 4 + 1], s2 = bsph[g1 * 4 + 2],
               s3 = bsph[g1 * 4 + 3];
    const int64_t nr1 = v.n_rots_for_block[g1];
    const int64_t nu = (nr1 + UNIT_R1 - 1) / UNIT_R1;
    for (int b2 = b1 + lane; b2 < v.B; b2 += WAVE) {
      const int64_t g2 = int64_t(p) * v.B + b2;
      if (rot_cell_hit(v, bsph, s0, s1, s2, s3, g2, reach)) {
        pairs += b2 == b1 ? nr1 : nr1 * int64_t(v.n_rots_for_block[g2]);
        units += nu;
      }
    }
  }
  pairs = wave_sum(pairs);
  units = wave_sum(units);
  if (lane == 0) {
    row_pairs[g1] = pairs;
    row_units[g1] = units;
  }
}

// Exclusive scans of two int64 arrays [n] into [n + 1] (last = total), one
// workgroup: each thread sums a contiguous chunk, the chunk sums are scanned
// in LDS, then each thread writes its chunk's prefixes.
__global__ __launch_bounds__(1024) void scan2_kernel(const int64_t* __restrict__ a,
                                                     const int64_t* __restrict__ b, int64_t n,
                                                     int64_t* __restrict__ oa,
                                                     int64_t* __restrict__ ob) {
  __shared__ int64_t sa[1024], sb[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t lo = min(n, int64_t(t) * per), hi = min(n, lo + per);
  int64_t xa = 0, xb = 0;
  for (int64_t i = lo; i < hi; ++i) {
    xa += a[i];
    xb += b[i];
  }
  sa[t] = xa;
  sb[t] = xb;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int64_t va = t >= o ? sa[t - o] : 0, vb = t >= o ? sb[t - o] : 0;
    __syncthreads();
    sa[t] += va;
    sb[t] += vb;
    __syncthreads();
  }
  int64_t ra = sa[t] - xa, rb = sb[t] - xb;
  for (int64_t i = lo; i < hi; ++i) {
    oa[i] = ra;
    ob[i] = rb;
    ra += a[i];
    rb += b[i];
  }
  if (t == 1023) {
    oa[n] = sa[1023];
    ob[n] = sb[1023];
  }
}

// One wave per row: the dispatch entries [3, n_pairs] and work units of the
// row's neighbor cells, at the row's scanned offsets.
template <typename Real>
__global__ __launch_bounds__(256) void rot_rows_fill_kernel(
    RotView v, const Real* __restrict__ bsph, Real reach, const int64_t* __restrict__ pair_off,
    const int64_t* __restrict__ unit_off, int32_t* __restrict__ indices, int64_t n_pairs,
    RotUnit* __restrict__ units) {
  const int lane = lane_id();
  const int64_t g1 = int64_t(blockIdx.x) * WAVES_PER_WG + threadIdx.x / WAVE;
  if (g1 >= int64_t(v.P) * v.B) return;
  if (v.first_rot_block_type[g1] < 0) return;
  const int p = int(g1 / v.B), b1 = int(g1 - int64_t(p) * v.B);
  const Real s0 = bsph[g1 * 4 + 0], s1 = bsph[g1 * 4 + 1], s2 = bsph[g1 * 4 + 2],
             s3 = bsph[g1 * 4 + 3];
  const int nr1 = v.n_rots_for_block[g1], off1 = v.rot_offset_for_block[g1];
  const int nu = (nr1 + UNIT_R1 - 1) / UNIT_R1;
  int64_t e = pair_off[g1], u = unit_off[g1];
  for (int c0 = b1; c0 < v.B; c0 += WAVE) {
    const int b2l = c0 + lane;
    const int64_t g2l = int64_t(p) * v.B + b2l;
    const bool hit = b2l < v.B && rot_cell_hit(v, bsph, s0, s1, s2, s3, g2l, reach);
    const int nr2l = hit ? v.n_rots_for_block[g2l] : 0;
    const int off2l = hit ? v.rot_offset_for_block[g2l] : 0;
    uint64_t m = __ballot(hit);
    while (m) {
      const int j = __builtin_ctzll(m);
      m &= m - 1;
      const int b2 = c0 + j;
      const int nr2 = lane_val(nr2l, j), off2 = lane_val(off2l, j);
      const bool diag = b2 == b1;
      const int64_t cnt = diag ? int64_t(nr1) : int64_t(nr1) * nr2;
      const bool ok = off1 >= 0 && off2 >= 0;
      for (int64_t k = lane; k < cnt; k += WAVE) {
        int32_t v0 = -1, v1 = -1, v2 = -1;
        if (ok) {
          v0 = p;
          if (diag) {
            v1 = off1 + int(k);
            v2 = v1;
          } else {
            const int q = int(k) / nr2;
            v1 = off1 + q;
            v2 = off2 + (int(k) - q * nr2);
          }
        }
        indices[e + k] = v0;
        indices[n_pairs + e + k] = v1;
        indices[2 * n_pairs + e + k] = v2;
      }
      for (int k = lane; k < nu; k += WAVE) {
        RotUnit U;
        U.g1 = int(g1);
        U.b2 = b2;
        U.r1_first = off1 + k * UNIT_R1;
        U.r1_count = (ok ? 1 : -1) * min(UNIT_R1, nr1 - k * UNIT_R1);
        U.entry = e + (diag ? int64_t(k) * UNIT_R1 : int64_t(k) * UNIT_R1 * nr2);
        U.pad = 0;
        units[u + k] = U;
      }
      e += cnt;
      u += nu;
    }
  }
}

// ------------------------------------------------------------------ scoring

template <typename Real>
struct RotArgs {
  RotView v;
  BTView bt;
  const Real* coords;
  const Real* records;
  const uint8_t* conn_pd;
  const Real* global_params;
  const Real* rot_spheres;  // [R, 4] centre, radius (from the neighbor pass)
  const RotUnit* units;
  int64_t n_units;
  int64_t n_pairs;
  uint32_t* cursor;   // [8 * QSTRIDE] queue cursors
  Real* energies;     // [3, n_pairs]
  const Real* dTdV;   // [3, n_pairs] (backward)
  Real* grad;         // [NA, 3] (backward)
  unsigned long long* census;  // [16] with M_CENSUS
};

// Branch census of one enumerated atom pair (the reference enumerates every
// atom pair of a rotamer pair for LJ and every heavy pair for LK): weight-0,
// beyond max_dis, or the potential region the distance falls in.
template <typename Real>
__device__ __forceinline__ void census_pair(uint32_t* cnt, int f1, Real rad1,
                                            const AtomRegs<Real>& R2, Real d2, Real wcp,
                                            const GlobalParams<Real>& gp) {
  const bool lk = (f1 & F_HEAVY) && (R2.flags & F_HEAVY);
  cnt[C_LJ_PAIRS] += 1;
  cnt[C_LK_PAIRS] += lk ? 1 : 0;
  if (wcp == Real(0)) {
    cnt[C_LJ_CP0] += 1;
    cnt[C_LK_CP0] += lk ? 1 : 0;
    return;
  }
  if (!(d2 < gp.max_dis2)) {
    cnt[C_LJ_FAR] += 1;
    cnt[C_LK_FAR] += lk ? 1 : 0;
    return;
  }
  const Real dist = fast_sqrt(d2);
  const Real sigma = lj_sigma(f1, R2.flags, rad1, R2.r0.x, gp);
  const Real d_lin = sigma * gp.dlin_factor;
  Real cmin = sigma > gp.spline_start ? sigma : gp.spline_start;
  cmin = cmin > gp.cpoly_cap ? gp.cpoly_cap : cmin;
  cnt[dist < d_lin ? C_LJ_LIN : (dist < cmin ? C_LJ_MID : C_LJ_SPL)] += 1;
  if (lk) {
    const bool cc = (f1 & F_CARBON_LK) && (R2.flags & F_CARBON_LK);
    Real d_min, cdmin, cdmax;
    lk_knots(sigma, cc, d_min, cdmin, cdmax);
    cnt[dist < cdmin ? C_LK_CONST
                     : (dist < cdmax ? C_LK_CLOSE
                                     : (dist < gp.spline_start ? C_LK_MID : C_LK_SPL))] += 1;
  }
}

// wave-private LDS: the staged atoms of the current b1 rotamer (64 at a time)
template <typename Real>
struct RotLDS {
  alignas(4 * sizeof(Real)) Real pos[WAVE][4];  // x, y, z, lj_radius
  alignas(4 * sizeof(Real)) Real prm[WAVE][4];  // sqrt(lj_wdepth), K, 1/lambda^2, lk_volume
  int flg[WAVE];
  uint32_t pdw[WAVE];                 // packed path distances to the connection atoms
  uint8_t cs[MAX_CONN * MAX_CONN];    // connection-pair bond separations of the cell
  uint8_t mark[WAVE];                 // lane starts a rotamer of the pack
};

template <typename Real>
__device__ __forceinline__ void load_rot_atom(AtomRegs<Real>& R, const RotArgs<Real>& A, int t,
                                              int o, int a) {
  using V = typename V4<Real>::T;
  const Real* xyz = A.coords + (int64_t(o) + a) * 3;
  R.x = xyz[0];
  R.y = xyz[1];
  R.z = xyz[2];
  const Real* rec = A.records + (int64_t(t) * A.bt.Amax + a) * REC;
  R.r0 = *(const V*)rec;
  const V r1 = *(const V*)(rec + 4);
  R.vol = r1.x;
  R.flags = int(r1.y);
  R.pd = *(const uint32_t*)(A.conn_pd + (int64_t(t) * A.bt.Amax + a) * MAX_CONN);
}

// stage atoms [c0, c0 + n) of a rotamer (type t, first atom o) into LDS
// (n <= 64)
template <typename Real>
__device__ __forceinline__ void stage_rot_chunk(RotLDS<Real>& L, const RotArgs<Real>& A, int t,
                                                int o, int c0, int n, int lane) {
  using V = typename V4<Real>::T;
  wave_sync();
  if (lane < n) {
    AtomRegs<Real> R;
    load_rot_atom(R, A, t, o, c0 + lane);
    V p, q;
    p.x = R.x;
    p.y = R.y;
    p.z = R.z;
    p.w = R.r0.x;
    q.x = R.r0.y;
    q.y = R.r0.z;
    q.z = R.r0.w;
    q.w = R.vol;
    *(V*)L.pos[lane] = p;
    *(V*)L.prm[lane] = q;
    L.flg[lane] = R.flags;
    L.pdw[lane] = R.pd;
  }
  wave_sync();
}

// Sum of the lanes of each segment [seg_start, ...] of contiguous lanes,
// left in each segment's last lane (inclusive segmented scan).
template <typename T>
__device__ __forceinline__ T segment_scan(T x, int seg_start, int lane) {
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const T t = __shfl_up(x, o, WAVE);
    if (lane - o >= seg_start) x += t;
  }
  return x;
}

// LJ + LK of one atom pair (a1 staged in LDS slot j, a2 this lane's atom),
// both already known to be within max_dis with weight wcp != 0 (`in`).
// Adds energies (forward) or w-scaled gradients (backward: returns the
// coefficient gs of (x1 - x2) for atom 1; atom 2 gets the negative).
template <typename Real, bool BWD>
__device__ __forceinline__ Real rot_pair(const RotLDS<Real>& L, int j, const AtomRegs<Real>& R2,
                                         Real dx, Real dy, Real dz, Real d2, Real wcp, bool in,
                                         bool lk_any, const GlobalParams<Real>& gp, Real& ea,
                                         Real& er, Real& el, Real w0, Real w1, Real w2) {
  using V = typename V4<Real>::T;
  const V p1 = *(const V*)L.pos[j];
  const V q1 = *(const V*)L.prm[j];
  const int f1 = L.flg[j];
  const Real inv_d = fast_rsqrt(d2);
  const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);  // coincident atoms: d = 0
  const Real sigma = lj_sigma(f1, R2.flags, p1.w, R2.r0.x, gp);
  const Real eps = q1.x * R2.r0.y;
  Real atr, rep, dlj;
  lj_pair<Real, BWD>(dist, sigma, eps, wcp, gp, atr, rep, dlj);
  Real gs = Real(0);
  if (in) {
    if (!BWD) {
      ea += atr;
      er += rep;
    } else if (dist > Real(0)) {
      gs = (dist < sigma ? w1 : w0) * dlj * inv_d;
    }
  }
  if (lk_any) {
    const bool heavy = in && (f1 & F_HEAVY) && (R2.flags & F_HEAVY);
    const bool cc = (f1 & F_CARBON_LK) && (R2.flags & F_CARBON_LK);
    Real d_min, cdmin, cdmax;
    lk_knots(sigma, cc, d_min, cdmin, cdmax);
    Real va, dva, vb, dvb;
    lk_dir<Real, BWD>(dist, d_min, cdmin, cdmax, -R2.vol * q1.y, p1.w, q1.z, gp, va, dva);
    lk_dir<Real, BWD>(dist, d_min, cdmin, cdmax, -q1.w * R2.r0.z, R2.r0.x, R2.r0.w, gp, vb, dvb);
    if (heavy) {
      if (!BWD)
        el += wcp * (va + vb);
      else if (dist > Real(0))
        gs += w2 * (wcp * (dva + dvb)) * inv_d;
    }
  }
  return gs;
}

// a1's gradient: sum over the wave, one atomic per component
template <typename Real>
__device__ __forceinline__ void flush_a1(const RotArgs<Real>& A, int64_t atom, Real gx, Real gy,
                                         Real gz, int lane) {
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  if (lane < 3) {
    const Real v = lane == 0 ? gx : (lane == 1 ? gy : gz);
    if (v != Real(0)) unsafeAtomicAdd(&A.grad[atom * 3 + lane], v);
  }
}

// zero energies of entries [e0, e0 + n) (pairs that are not scored)
template <typename Real>
__device__ __forceinline__ void zero_entries(const RotArgs<Real>& A, int64_t e0, int64_t n,
                                             int lane) {
  for (int64_t k = lane; k < n; k += WAVE) {
    A.energies[e0 + k] = Real(0);
    A.energies[A.n_pairs + e0 + k] = Real(0);
    A.energies[2 * A.n_pairs + e0 + k] = Real(0);
  }
}

// inter-block unit: rotamers [r1_first, +r1_count) of g1 x every rotamer of g2
template <typename Real, int MODE>
__device__ void rot_inter(const RotArgs<Real>& A, RotLDS<Real>& L, const GlobalParams<Real>& gp,
                          int g1, int g2, int b2, int r1_first, int r1_count, int64_t entry,
                          int lane, uint32_t* cnt) {
  constexpr bool BWD = (MODE & M_WGRAD) != 0;
  constexpr bool CENSUS = (MODE & M_CENSUS) != 0;
  const RotView& v = A.v;
  const int nr2 = uni(v.n_rots_for_block[g2]), off2 = uni(v.rot_offset_for_block[g2]);
  const int min_sep = uni(v.min_bondsep[int64_t(g1) * v.B + b2]);
  if (CENSUS && lane == 0) cnt[C_ITEMS] += uint32_t(r1_count) * uint32_t(nr2);
  const bool cp = min_sep <= 4;
  if (cp) {
    wave_sync();
    if (lane < MAX_CONN * MAX_CONN) {
      const int c1 = lane / MAX_CONN, c2 = lane - c1 * MAX_CONN;
      L.cs[lane] = (c1 < v.C && c2 < v.C)
                       ? (uint8_t)v.inter_block_bondsep[((int64_t(g1) * v.B + b2) * v.C + c1) *
                                                            v.C + c2]
                       : (uint8_t)6;
    }
    wave_sync();
  }
  const int64_t n = A.n_pairs;
  for (int ps = 0; ps < nr2;) {
    // pack: consecutive rotamers ps.. whose atoms fit in 64 lanes (every
    // block type has at most 64 atoms, so a pack holds at least one)
    const int rl = ps + lane;
    int nat = 0;
    if (rl < nr2) {
      const int t = v.block_type_for_rot[off2 + rl];
      nat = t >= 0 ? A.bt.n_atoms[t] : 0;
    }
    int incl = nat;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int t = __shfl_up(incl, o, WAVE);
      if (lane >= o) incl += t;
    }
    const int k = __popcll(__ballot(rl < nr2 && incl <= WAVE));
    const int excl = incl - nat;
    // a rotamer with atoms marks its first lane with its own index + 1
    wave_sync();
    L.mark[lane] = 0;
    wave_sync();
    if (lane < k && nat > 0) L.mark[excl] = (uint8_t)(lane + 1);
    wave_sync();
    const uint64_t starts = __ballot(L.mark[lane] != 0);
    const int total = lane_val(incl, k - 1);
    // this lane's rotamer (slot): the one whose first lane is the nearest
    // start at or below this lane
    const uint64_t below = starts & ((2ull << lane) - 1ull);
    const int start_pos = below ? 63 - __builtin_clzll(below) : -1;
    const int slot = start_pos >= 0 ? int(L.mark[start_pos]) - 1 : -1;
    const int slot_c = slot < 0 ? 0 : slot;
    const int seg_start = __shfl(excl, slot_c, WAVE);
    const int seg_nat = __shfl(nat, slot_c, WAVE);
    const int a2 = lane - seg_start;
    const bool last = lane == seg_start + seg_nat - 1;
    bool active = lane < total && slot >= 0;
    const int r2 = off2 + ps + slot_c;
    const int t2 = active ? v.block_type_for_rot[r2] : 0;
    const int o2 = active ? v.rot_coord_offset[r2] : 0;
    active = active && t2 >= 0 && a2 < seg_nat;
    const int nc2 = active ? A.bt.n_conn[t2] : 0;
    const bool lig2 = active && A.bt.is_ligand[t2];
    AtomRegs<Real> R2{};
    if (active) load_rot_atom(R2, A, t2, o2, a2);
    Real g2x = 0, g2y = 0, g2z = 0;  // this lane's atom gradient (backward)
    for (int i = 0; i < r1_count; ++i) {
      const int r1 = r1_first + i;
      const int t1 = uni(v.block_type_for_rot[r1]);
      if (t1 < 0) {
        if (!BWD && !CENSUS) zero_entries(A, entry + int64_t(i) * nr2 + ps, k, lane);
        continue;
      }
      // rotamers of the pack without atoms (no block type) score zero
      if (!BWD && !CENSUS && lane < k && nat == 0) {
        const int64_t e0 = entry + int64_t(i) * nr2 + ps + lane;
        A.energies[e0] = Real(0);
        A.energies[n + e0] = Real(0);
        A.energies[2 * n + e0] = Real(0);
      }
      const int o1 = uni(v.rot_coord_offset[r1]);
      const int n1 = uni(A.bt.n_atoms[t1]);
      const int nc1 = uni(A.bt.n_conn[t1]);
      const bool lig = lig2 && A.bt.is_ligand[t1];
      const int64_t e = entry + int64_t(i) * nr2 + ps + slot_c;
      Real w0 = 0, w1 = 0, w2 = 0;
      if (BWD && active) {
        w0 = A.dTdV[e];
        w1 = A.dTdV[n + e];
        w2 = A.dTdV[2 * n + e];
      }
      Real ea = 0, er = 0, el = 0;
      stage_rot_chunk(L, A, t1, o1, 0, n1, lane);
      for (int j = 0; j < n1; ++j) {
        using V = typename V4<Real>::T;
        const V p1 = *(const V*)L.pos[j];
        const Real dx = p1.x - R2.x, dy = p1.y - R2.y, dz = p1.z - R2.z;
        const Real d2 = dx * dx + dy * dy + dz * dz;
        bool in = active && d2 < gp.max_dis2;
        if (!CENSUS && __ballot(in) == 0) continue;
        Real wcp = Real(1);
        if (cp) {
          const int sep = packed_pair_sep(L.pdw[j], R2.pd, L.cs, min_sep, nc1, nc2, lig);
          wcp = cp_weight<Real>(sep);
          in = in && wcp != Real(0);
        }
        if (CENSUS) {
          if (active) census_pair(cnt, L.flg[j], p1.w, R2, d2, wcp, gp);
          continue;
        }
        const bool lk_any = (L.flg[j] & F_HEAVY) && __ballot(in && (R2.flags & F_HEAVY)) != 0;
        const Real gs = rot_pair<Real, BWD>(L, j, R2, dx, dy, dz, d2, wcp, in, lk_any, gp, ea, er,
                                            el, w0, w1, w2);
        if (BWD) {
          g2x -= gs * dx;
          g2y -= gs * dy;
          g2z -= gs * dz;
          if (__ballot(gs != Real(0)) != 0)
            flush_a1(A, int64_t(o1) + j, gs * dx, gs * dy, gs * dz, lane);
        }
      }
      if (!BWD && !CENSUS) {
        ea = segment_scan(ea, seg_start, lane);
        er = segment_scan(er, seg_start, lane);
        el = segment_scan(el, seg_start, lane);
        if (active && last) {
          A.energies[e] = ea;
          A.energies[n + e] = er;
          A.energies[2 * n + e] = el;
        }
      }
    }
    if (BWD && active) {
      const int64_t at = (int64_t(o2) + a2) * 3;
      if (g2x != Real(0)) unsafeAtomicAdd(&A.grad[at + 0], g2x);
      if (g2y != Real(0)) unsafeAtomicAdd(&A.grad[at + 1], g2y);
      if (g2z != Real(0)) unsafeAtomicAdd(&A.grad[at + 2], g2z);
    }
    ps += k;
  }
}

// intra-block unit: each rotamer r of [r1_first, +r1_count) with itself
// (atom pairs a1 < a2, count pair from the block type's path distances);
// one lane per atom (at most 64)
template <typename Real, int MODE>
__device__ void rot_intra(const RotArgs<Real>& A, RotLDS<Real>& L, const GlobalParams<Real>& gp,
                          int r1_first, int r1_count, int64_t entry, int lane, uint32_t* cnt) {
  constexpr bool BWD = (MODE & M_WGRAD) != 0;
  constexpr bool CENSUS = (MODE & M_CENSUS) != 0;
  if (CENSUS && lane == 0) {
    cnt[C_ITEMS] += uint32_t(r1_count);
    cnt[C_INTRA_ITEMS] += uint32_t(r1_count);
  }
  const RotView& v = A.v;
  const int64_t n = A.n_pairs;
  for (int i = 0; i < r1_count; ++i) {
    const int r = r1_first + i;
    const int t = uni(v.block_type_for_rot[r]);
    if (t < 0) {
      if (!BWD && !CENSUS) zero_entries(A, entry + i, 1, lane);
      continue;
    }
    const int o = uni(v.rot_coord_offset[r]);
    const int na = uni(A.bt.n_atoms[t]);
    const int64_t e = entry + i;
    Real w0 = 0, w1 = 0, w2 = 0;
    if (BWD) {
      w0 = A.dTdV[e];
      w1 = A.dTdV[n + e];
      w2 = A.dTdV[2 * n + e];
    }
    const int32_t* pd = A.bt.path_distance + int64_t(t) * A.bt.Amax * A.bt.Amax;
    Real ea = 0, er = 0, el = 0;
    const int a2 = lane;
    const bool active = a2 < na;
    AtomRegs<Real> R2{};
    if (active) load_rot_atom(R2, A, t, o, a2);
    Real g2x = 0, g2y = 0, g2z = 0;
    stage_rot_chunk(L, A, t, o, 0, na, lane);
    // pairs a1 < a2: a1 runs over all but the last atom
    for (int j = 0; j + 1 < na; ++j) {
      using V = typename V4<Real>::T;
      const V p1 = *(const V*)L.pos[j];
      const Real dx = p1.x - R2.x, dy = p1.y - R2.y, dz = p1.z - R2.z;
      const Real d2 = dx * dx + dy * dy + dz * dz;
      bool in = active && j < a2 && d2 < gp.max_dis2;
      if (!CENSUS && __ballot(in) == 0) continue;
      Real wcp = Real(0);
      if (active && j < a2) wcp = cp_weight<Real>(pd[int64_t(j) * A.bt.Amax + a2]);
      in = in && wcp != Real(0);
      if (CENSUS) {
        if (active && j < a2) census_pair(cnt, L.flg[j], p1.w, R2, d2, wcp, gp);
        continue;
      }
      const bool lk_any = (L.flg[j] & F_HEAVY) && __ballot(in && (R2.flags & F_HEAVY)) != 0;
      const Real gs = rot_pair<Real, BWD>(L, j, R2, dx, dy, dz, d2, wcp, in, lk_any, gp, ea, er,
                                          el, w0, w1, w2);
      if (BWD) {
        g2x -= gs * dx;
        g2y -= gs * dy;
        g2z -= gs * dz;
        if (__ballot(gs != Real(0)) != 0)
          flush_a1(A, int64_t(o) + j, gs * dx, gs * dy, gs * dz, lane);
      }
    }
    if (BWD && active) {
      const int64_t at = (int64_t(o) + a2) * 3;
      if (g2x != Real(0)) unsafeAtomicAdd(&A.grad[at + 0], g2x);
      if (g2y != Real(0)) unsafeAtomicAdd(&A.grad[at + 1], g2y);
      if (g2z != Real(0)) unsafeAtomicAdd(&A.grad[at + 2], g2z);
    }
    if (!BWD && !CENSUS) {
      ea = wave_sum(ea);
      er = wave_sum(er);
      el = wave_sum(el);
      if (lane == 0) {
        A.energies[e] = ea;
        A.energies[n + e] = er;
        A.energies[2 * n + e] = el;
      }
    }
  }
}

// Persistent rotamer-pair kernel: 4 independent waves per workgroup, units
// from the per-XCD queues.  MODE 0: energies [3, n_pairs] (every entry
// written, zero for pairs that are not scored);
// M_WGRAD | M_NOE: grad [NA, 3] += sum_t dTdV[t, i] dV_t(pair i)/dx;
// M_CENSUS: counts of enumerated atom pairs per potential region.
// Float gradient builds are held to 8 waves per SIMD (64 VGPRs, a few
// spilled): the backward is latency-bound and runs 1.3x faster that way on
// the 16 x 300-residue repacking set (18.4 -> 14.2 ms).
template <typename Real, int MODE>
__global__ __launch_bounds__(256, (sizeof(Real) == 4 && (MODE & M_WGRAD)) ? 8 : 1) void
rotamer_score_kernel(RotArgs<Real> A) {
  __shared__ RotLDS<Real> lds[WAVES_PER_WG];
  const int lane = lane_id();
  RotLDS<Real>& L = lds[threadIdx.x / WAVE];
  const GlobalParams<Real> gp = load_global_params(A.global_params);
  uint32_t cnt[C_N];
#pragma unroll
  for (int k = 0; k < C_N; ++k) cnt[k] = 0;
  RowQueue queue(A.cursor, A.n_units);
  for (int64_t u = queue.next(); u >= 0; u = queue.next()) {
    const RotUnit* up = A.units + u;
    const int g1 = uni(up->g1), b2 = uni(up->b2);
    const int r1_first = uni(up->r1_first), r1_count = uni(up->r1_count);
    const int64_t entry = uni64(up->entry);
    const int p = g1 / A.v.B, b1 = g1 - p * A.v.B;
    if (r1_count <= 0) {
      if (MODE == 0 && r1_count < 0)
        zero_entries(A, entry,
                     int64_t(-r1_count) * (b1 == b2 ? 1 : uni(A.v.n_rots_for_block[p * A.v.B + b2])),
                     lane);
      continue;
    }
    if (b1 == b2)
      rot_intra<Real, MODE>(A, L, gp, r1_first, r1_count, entry, lane, cnt);
    else
      rot_inter<Real, MODE>(A, L, gp, g1, p * A.v.B + b2, b2, r1_first, r1_count, entry, lane,
                            cnt);
  }
  if (MODE & M_CENSUS) {
#pragma unroll
    for (int k = 0; k < C_N; ++k) {
      const unsigned long long v = wave_sum((unsigned long long)cnt[k]);
      if (lane == 0 && v) atomicAdd(&A.census[k], v);
    }
  }
}


// ------------------------------------------------------------------ energies, culled

// Energy forward with culling and compaction.  Per b1 rotamer r1 of a unit:
//   cull      every atom of b2's rotamers (the candidates, taken 64 at a time
//             in rotamer order) is tested against r1's sphere: an atom
//             farther than R1 + max_dis from r1's centre reaches no atom of
//             r1 (triangle inequality; 1e-3 A margin for rounding).
//             Survivors are compacted, in order, into a 64-lane chunk.
//   phase A   for a full (or the last) chunk: each lane holds one surviving
//             atom, a1 runs over r1's atoms (LDS broadcasts); pairs within
//             max_dis and of non-zero count-pair weight are ballot-compacted
//             into an LJ FIFO, heavy-heavy ones also into an LK FIFO.
//   phase B   whenever a FIFO holds 64 entries, and at the end of a chunk:
//             one entry per lane, every lane busy.  Energies go to the
//             chunk lane's 64-bit fixed-point LDS accumulator (integer LDS
//             atomics; exact, order-independent).
//   fold      at the end of a chunk, each lane's sums are added per rotamer
//             (segments of contiguous lanes, segmented scan) into LDS sums
//             of the rotamers of the current window of 64 b2 rotamers.
//   write     after the window, one store per rotamer pair and term (zero
//             for pairs with no surviving atoms).

constexpr int RFIFO = 2 * WAVE;  // FIFO capacity: < 64 left over + 64 appended

template <typename Real>
struct RotLDS2 {
  RotLDS<Real> r1;  // staged atoms of the b1 rotamer, cell separations, marks
  alignas(4 * sizeof(Real)) Real pos2[WAVE][4];  // chunk: x, y, z, lj_radius
  alignas(4 * sizeof(Real)) Real prm2[WAVE][4];  // sqrt(lj_wdepth), K, 1/lambda^2, lk_volume
  int meta2[WAVE];       // flags | n_conn << 8 | ligand << 11 | slot << 16
  uint32_t pd2[WAVE];    // packed path distances to the connection atoms
  uint16_t lj[RFIFO];    // entry: a1 | lane << 6 | weight-0.2 << 12
  uint16_t lk[RFIFO];
  typename LdsAcc<Real>::T lacc[3][WAVE];  // per chunk lane: atr, rep, lk
  Real sum[3][WAVE];     // per rotamer of the window
};

template <typename Real>
__device__ __forceinline__ void rot_fifo_lj(RotLDS2<Real>& L, int consume,
                                            const GlobalParams<Real>& gp, int lane) {
  using V = typename V4<Real>::T;
  using Acc = LdsAcc<Real>;
  if (lane < consume) {
    const int e = L.lj[lane];
    const int a1 = e & 63, l2 = (e >> 6) & 63;
    const Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
    const V p1 = *(const V*)L.r1.pos[a1], p2 = *(const V*)L.pos2[l2];
    const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
    const Real d2 = dx * dx + dy * dy + dz * dz;
    const Real inv_d = fast_rsqrt(d2);
    const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);
    const Real sigma = lj_sigma(L.r1.flg[a1], L.meta2[l2] & 255, p1.w, p2.w, gp);
    const Real eps = L.r1.prm[a1][0] * L.prm2[l2][0];
    Real atr, rep, dlj;
    lj_pair<Real, false>(dist, sigma, eps, wcp, gp, atr, rep, dlj);
    atomicAdd(&L.lacc[0][l2], Acc::enc(atr));
    atomicAdd(&L.lacc[1][l2], Acc::enc(rep));
  }
}

template <typename Real>
__device__ __forceinline__ void rot_fifo_lk(RotLDS2<Real>& L, int consume,
                                            const GlobalParams<Real>& gp, int lane) {
  using V = typename V4<Real>::T;
  using Acc = LdsAcc<Real>;
  if (lane < consume) {
    const int e = L.lk[lane];
    const int a1 = e & 63, l2 = (e >> 6) & 63;
    const Real wcp = (e >> 12) & 1 ? Real(0.2) : Real(1);
    const V p1 = *(const V*)L.r1.pos[a1], p2 = *(const V*)L.pos2[l2];
    const V q1 = *(const V*)L.r1.prm[a1], q2 = *(const V*)L.prm2[l2];
    const Real dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
    const Real d2 = dx * dx + dy * dy + dz * dz;
    const Real inv_d = fast_rsqrt(d2);
    const Real dist = d2 > Real(0) ? d2 * inv_d : Real(0);
    const int f1 = L.r1.flg[a1], f2 = L.meta2[l2] & 255;
    const Real sigma = lj_sigma(f1, f2, p1.w, p2.w, gp);
    const bool cc = (f1 & F_CARBON_LK) && (f2 & F_CARBON_LK);
    Real d_min, cdmin, cdmax;
    lk_knots(sigma, cc, d_min, cdmin, cdmax);
    Real va, dva, vb, dvb;
    lk_dir<Real, false>(dist, d_min, cdmin, cdmax, -q2.w * q1.y, p1.w, q1.z, gp, va, dva);
    lk_dir<Real, false>(dist, d_min, cdmin, cdmax, -q1.w * q2.y, p2.w, q2.z, gp, vb, dvb);
    atomicAdd(&L.lacc[2][l2], Acc::enc(wcp * (va + vb)));
  }
}

// consume the first `c` entries of a FIFO of n and move the rest to the front
__device__ __forceinline__ void rot_fifo_shift(uint16_t* f, int c, int n, int lane) {
  const int rem = n - c;
  if (rem <= 0) return;
  uint16_t v = 0;
  if (lane < rem) v = f[c + lane];
  wave_sync();
  if (lane < rem) f[lane] = v;
  wave_sync();
}

// phase A + B over the chunk's nb lanes against r1 (n1 atoms staged), then
// fold the lanes' sums into the window's per-rotamer sums
template <typename Real>
__device__ void rot_chunk(RotLDS2<Real>& L, const GlobalParams<Real>& gp, int nb, int n1,
                          bool cp, int min_sep, int nc1, bool lig1, const AtomRegs<Real>& R1,
                          int lane) {
  using V = typename V4<Real>::T;
  using Acc = LdsAcc<Real>;
  wave_sync();
  const bool valid = lane < nb;
  V p2 = {};
  int meta = 0;
  uint32_t pd2 = 0;
  if (valid) {
    p2 = *(const V*)L.pos2[lane];
    meta = L.meta2[lane];
    pd2 = L.pd2[lane];
  }
  const bool heavy2 = meta & F_HEAVY;
  const int nc2 = (meta >> 8) & 7;
  const bool lig = lig1 && ((meta >> 11) & 1);
  L.lacc[0][lane] = 0;
  L.lacc[1][lane] = 0;
  L.lacc[2][lane] = 0;
  int nlj = 0, nlk = 0;  // wave-uniform FIFO fill
  // a1's data come from the registers of the lane that staged it
  // (wave-uniform readlane), so the test chain has no LDS latency
  for (int j = 0; j < n1; ++j) {
    const Real dx = lane_val(R1.x, j) - p2.x, dy = lane_val(R1.y, j) - p2.y,
               dz = lane_val(R1.z, j) - p2.z;
    const Real d2 = dx * dx + dy * dy + dz * dz;
    bool in = valid && d2 < gp.max_dis2;
    bool w02 = false;
    if (cp) {
      const int sep = packed_pair_sep((uint32_t)lane_val((int)R1.pd, j), pd2, L.r1.cs, min_sep,
                                      nc1, nc2, lig);
      in = in && sep >= 4;
      w02 = sep == 4;
    }
    const uint64_t m = __ballot(in);
    if (m == 0) continue;
    const uint16_t ent = (uint16_t)(j | (lane << 6) | (w02 ? (1 << 12) : 0));
    if (in) L.lj[nlj + lanes_below(m)] = ent;
    nlj += __popcll(m);
    if (lane_val(R1.flags, j) & F_HEAVY) {
      const uint64_t mh = __ballot(in && heavy2);
      if (in && heavy2) L.lk[nlk + lanes_below(mh)] = ent;
      nlk += __popcll(mh);
    }
    if (nlj >= WAVE) {
      wave_sync();
      rot_fifo_lj(L, WAVE, gp, lane);
      rot_fifo_shift(L.lj, WAVE, nlj, lane);
      nlj -= WAVE;
    }
    if (nlk >= WAVE) {
      wave_sync();
      rot_fifo_lk(L, WAVE, gp, lane);
      rot_fifo_shift(L.lk, WAVE, nlk, lane);
      nlk -= WAVE;
    }
  }
  wave_sync();
  if (nlj > 0) rot_fifo_lj(L, nlj, gp, lane);
  if (nlk > 0) rot_fifo_lk(L, nlk, gp, lane);
  wave_sync();
  // fold: lanes of one rotamer (slot) are contiguous
  const int slot = meta >> 16;
  const int prev = lane > 0 ? L.meta2[lane - 1] >> 16 : -1;
  const uint64_t starts = __ballot(valid && (lane == 0 || slot != prev));
  const uint64_t below = starts & ((2ull << lane) - 1ull);
  const int seg_start = below ? 63 - __builtin_clzll(below) : 0;
  const bool last = valid && (lane == nb - 1 || (L.meta2[lane + 1 < WAVE ? lane + 1 : lane] >> 16) != slot);
  Real e0 = valid ? Acc::dec(L.lacc[0][lane]) : Real(0);
  Real e1 = valid ? Acc::dec(L.lacc[1][lane]) : Real(0);
  Real e2 = valid ? Acc::dec(L.lacc[2][lane]) : Real(0);
  e0 = segment_scan(e0, seg_start, lane);
  e1 = segment_scan(e1, seg_start, lane);
  e2 = segment_scan(e2, seg_start, lane);
  if (last) {
    L.sum[0][slot] += e0;
    L.sum[1][slot] += e1;
    L.sum[2][slot] += e2;
  }
  wave_sync();
}

// inter-block unit, energies only
template <typename Real>
__device__ void rot_inter_energy(const RotArgs<Real>& A, RotLDS2<Real>& L,
                                 const GlobalParams<Real>& gp, int g1, int g2, int b2,
                                 int r1_first, int r1_count, int64_t entry, int lane) {
  using V = typename V4<Real>::T;
  const RotView& v = A.v;
  const int nr2 = uni(v.n_rots_for_block[g2]), off2 = uni(v.rot_offset_for_block[g2]);
  const int min_sep = uni(v.min_bondsep[int64_t(g1) * v.B + b2]);
  const bool cp = min_sep <= 4;
  if (cp) {
    wave_sync();
    if (lane < MAX_CONN * MAX_CONN) {
      const int c1 = lane / MAX_CONN, c2 = lane - c1 * MAX_CONN;
      L.r1.cs[lane] = (c1 < v.C && c2 < v.C)
                          ? (uint8_t)v.inter_block_bondsep[((int64_t(g1) * v.B + b2) * v.C + c1) *
                                                               v.C + c2]
                          : (uint8_t)6;
    }
    wave_sync();
  }
  const int64_t n = A.n_pairs;
  for (int w0 = 0; w0 < nr2; w0 += WAVE) {
    const int wn = min(WAVE, nr2 - w0);
    // the window's rotamers (lane = rotamer): type, first atom, atom count,
    // first candidate; candidates fetch theirs by lane shuffle
    int nat = 0, t2l = -1, o2l = 0;
    if (lane < wn) {
      t2l = v.block_type_for_rot[off2 + w0 + lane];
      o2l = v.rot_coord_offset[off2 + w0 + lane];
      nat = t2l >= 0 ? A.bt.n_atoms[t2l] : 0;
    }
    int incl = nat;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int t = __shfl_up(incl, o, WAVE);
      if (lane >= o) incl += t;
    }
    const int excl = incl - nat;
    const int total = lane_val(incl, wn - 1);
    for (int i = 0; i < r1_count; ++i) {
      const int r1 = r1_first + i;
      const int t1 = uni(v.block_type_for_rot[r1]);
      if (t1 < 0) {
        zero_entries(A, entry + int64_t(i) * nr2 + w0, wn, lane);
        continue;
      }
      const int o1 = uni(v.rot_coord_offset[r1]);
      const int n1 = uni(A.bt.n_atoms[t1]);
      const int nc1 = uni(A.bt.n_conn[t1]);
      const bool lig1 = A.bt.is_ligand[t1];
      const Real* sph = A.rot_spheres + int64_t(r1) * 4;
      const Real cx = sph[0], cy = sph[1], cz = sph[2];
      const Real reach = sph[3] + gp.max_dis + Real(1e-3);
      const Real reach2 = reach * reach;
      AtomRegs<Real> R1{};
      wave_sync();
      if (lane < n1) {
        load_rot_atom(R1, A, t1, o1, lane);
        V p, q;
        p.x = R1.x;
        p.y = R1.y;
        p.z = R1.z;
        p.w = R1.r0.x;
        q.x = R1.r0.y;
        q.y = R1.r0.z;
        q.z = R1.r0.w;
        q.w = R1.vol;
        *(V*)L.r1.pos[lane] = p;
        *(V*)L.r1.prm[lane] = q;
        L.r1.flg[lane] = R1.flags;
      }
      L.sum[0][lane] = 0;
      L.sum[1][lane] = 0;
      L.sum[2][lane] = 0;
      int nb = 0;  // surviving candidates in the chunk (wave-uniform)
      for (int c0 = 0; c0 < total; c0 += WAVE) {
        // candidate c = c0 + lane: its rotamer (slot) and atom
        const int c = c0 + lane;
        const uint64_t before = __ballot(lane < wn && excl <= c0 && nat > 0);
        const int carry = before ? 63 - __builtin_clzll(before) : 0;
        wave_sync();
        L.r1.mark[lane] = 0;
        wave_sync();
        if (lane < wn && nat > 0 && excl > c0 && excl < c0 + WAVE) L.r1.mark[excl - c0] = (uint8_t)(lane + 1);
        wave_sync();
        const uint64_t st = __ballot(L.r1.mark[lane] != 0);
        const uint64_t bl = st & ((2ull << lane) - 1ull);
        const int slot = bl ? int(L.r1.mark[63 - __builtin_clzll(bl)]) - 1 : carry;
        const bool cand = c < total;
        const int a2 = c - __shfl(excl, slot, WAVE);
        const int t2 = __shfl(t2l, slot, WAVE);
        const int o2 = __shfl(o2l, slot, WAVE);
        Real x = 0, y = 0, z = 0;
        bool keep = false;
        if (cand) {
          const Real* xyz = A.coords + (int64_t(o2) + a2) * 3;
          x = xyz[0];
          y = xyz[1];
          z = xyz[2];
          const Real ex = x - cx, ey = y - cy, ez = z - cz;
          keep = ex * ex + ey * ey + ez * ez < reach2;
        }
        const uint64_t mk = __ballot(keep);
        const int nk = __popcll(mk);
        if (nk == 0) continue;
        if (nb + nk > WAVE) {
          rot_chunk(L, gp, nb, n1, cp, min_sep, nc1, lig1, R1, lane);
          nb = 0;
        }
        if (keep) {
          const int pos = nb + lanes_below(mk);
          const Real* rec = A.records + (int64_t(t2) * A.bt.Amax + a2) * REC;
          const V r0 = *(const V*)rec;
          const V r1v = *(const V*)(rec + 4);
          V p, q;
          p.x = x;
          p.y = y;
          p.z = z;
          p.w = r0.x;
          q.x = r0.y;
          q.y = r0.z;
          q.z = r0.w;
          q.w = r1v.x;
          *(V*)L.pos2[pos] = p;
          *(V*)L.prm2[pos] = q;
          L.meta2[pos] = int(r1v.y) | (A.bt.n_conn[t2] << 8) | ((A.bt.is_ligand[t2] ? 1 : 0) << 11) |
                         (slot << 16);
          L.pd2[pos] = *(const uint32_t*)(A.conn_pd + (int64_t(t2) * A.bt.Amax + a2) * MAX_CONN);
        }
        nb += nk;
      }
      if (nb > 0) rot_chunk(L, gp, nb, n1, cp, min_sep, nc1, lig1, R1, lane);
      wave_sync();
      if (lane < wn) {
        const int64_t e = entry + int64_t(i) * nr2 + w0 + lane;
        A.energies[e] = L.sum[0][lane];
        A.energies[n + e] = L.sum[1][lane];
        A.energies[2 * n + e] = L.sum[2][lane];
      }
    }
  }
}

// Energy kernel: inter-block units culled and compacted, self pairs as in
// rotamer_score_kernel.
template <typename Real>
__global__ __launch_bounds__(256) void rotamer_energy_kernel(RotArgs<Real> A) {
  __shared__ RotLDS2<Real> lds[WAVES_PER_WG];
  const int lane = lane_id();
  RotLDS2<Real>& L = lds[threadIdx.x / WAVE];
  const GlobalParams<Real> gp = load_global_params(A.global_params);
  uint32_t cnt[1];
  RowQueue queue(A.cursor, A.n_units);
  for (int64_t u = queue.next(); u >= 0; u = queue.next()) {
    const RotUnit* up = A.units + u;
    const int g1 = uni(up->g1), b2 = uni(up->b2);
    const int r1_first = uni(up->r1_first), r1_count = uni(up->r1_count);
    const int64_t entry = uni64(up->entry);
    const int p = g1 / A.v.B, b1 = g1 - p * A.v.B;
    if (r1_count <= 0) {
      if (r1_count < 0)
        zero_entries(A, entry,
                     int64_t(-r1_count) * (b1 == b2 ? 1 : uni(A.v.n_rots_for_block[p * A.v.B + b2])),
                     lane);
      continue;
    }
    if (b1 == b2)
      rot_intra<Real, 0>(A, L.r1, gp, r1_first, r1_count, entry, lane, cnt);
    else
      rot_inter_energy<Real>(A, L, gp, g1, p * A.v.B + b2, b2, r1_first, r1_count, entry, lane);
  }
}

}  // namespace tmol_ljlk
