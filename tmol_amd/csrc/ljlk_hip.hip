// MI355X (gfx950) LJ+LK pose scoring: host side and the C ABI declared in
// include/tmol_ljlk.h.
//
// Device work per call (reference: ljlk_pose_score.impl.hh:1123-1155):
//   1. prepare + spheres   per block-type atom records; one bounding sphere
//                          per block (bit-identical to the reference CPU op)
//   2. neighbors           64x64 block tiles, wave-ballot compaction into a
//                          row-major work list (row g1: its g2 >= g1 neighbors)
//   3. score               persistent waves walk rows of the work list
//                          (ljlk_score_rows.hh; ljlk_score_general.hh for
//                          blocks over 32 atoms and for the branch census)
//   4. finalize            FP64 per-pose totals -> output dtype
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "../../include/tmol_ljlk.h"
#include "ljlk_common.hh"
#include "ljlk_neighbors.hh"
#include "ljlk_score_general.hh"
#include "ljlk_score_rows.hh"
#include "ljlk_score_stream.hh"
#include "ljlk_prune.hh"
#include "ljlk_rotamer.hh"
#include "ljlk_elec_rotamer.hh"
#include "ljlk_topology.hh"
#include "ljlk_elec.hh"
#include "lbfgs.hh"
#include "lkball.hh"
#include "hbond.hh"
#include "leaf_atoms.hh"

#include <cmath>

namespace tmol_ljlk {

// ------------------------------------------------------------------ host side

static int check(hipError_t e) { return e == hipSuccess ? TMOL_LJLK_OK : TMOL_LJLK_EHIP; }

static int validate(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt, int dtype) {
  if (!s || !bt) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (bt->dtype != dtype) return TMOL_LJLK_EINVAL;
  if (bt->max_n_conn > MAX_CONN || bt->max_n_conn < 0) return TMOL_LJLK_ENOTSUP;
  if (bt->n_block_types <= 0 || bt->max_n_atoms <= 0) return TMOL_LJLK_EINVAL;
  if (bt->max_n_tiles != (bt->max_n_atoms - 1) / TILE + 1) return TMOL_LJLK_EINVAL;
  if (s->n_poses <= 0 || s->n_blocks_total <= 0) return TMOL_LJLK_EINVAL;
  if (s->layout == TMOL_LJLK_PADDED) {
    if (int64_t(s->n_poses) * s->max_n_blocks != s->n_blocks_total) return TMOL_LJLK_EINVAL;
    if (!s->min_bondsep || !s->inter_block_bondsep) return TMOL_LJLK_EINVAL;
  } else if (s->layout == TMOL_LJLK_RAGGED) {
    if (!s->block_pose || !s->bonded_start || !s->bonded) return TMOL_LJLK_EINVAL;
  } else {
    return TMOL_LJLK_EINVAL;
  }
  if (!s->block_atom_offset || !s->block_type) return TMOL_LJLK_EINVAL;
  if (s->n_atoms_total >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  return TMOL_LJLK_OK;
}

static unsigned grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

template <typename Real>
static void fill(Real* p, int64_t n, hipStream_t stream, Real v = Real(0)) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel<Real>, dim3(grid_for(n)), dim3(256), 0, stream, p, n, v);
}

template <typename Real>
static int run_neighbors(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                         const void* coords, double max_dis, const int32_t* tiles,
                         int64_t n_tiles, void* ws, int64_t row_cap, int32_t* dense,
                         hipStream_t stream, const WsLayout& Lw) {
  StackView s = make_stack_view(st, btp->max_n_conn);
  BTView bt = make_bt_view(btp);
  char* w = (char*)ws;
  uint32_t* counters = (uint32_t*)(w + Lw.counters);
  Real* spheres = (Real*)(w + Lw.spheres);
  Real* records = (Real*)(w + Lw.records);
  uint8_t* conn_pd = (uint8_t*)(w + Lw.conn_pd);
  uint32_t* row_count = (uint32_t*)(w + Lw.row_count);
  int32_t* row_items = (int32_t*)(w + Lw.row_items);
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, stream, counters, 16,
                     (uint32_t)0);
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(grid_for(s.G)), dim3(256), 0, stream,
                     row_count, int64_t(s.G), (uint32_t)0);
  int nrec = bt.T * bt.Amax;
  hipLaunchKernelGGL(prepare_records_kernel<Real>, dim3(grid_for(nrec)), dim3(256), 0, stream, bt,
                     (const Real*)btp->type_params, btp->n_atom_types, records, conn_pd);
  hipLaunchKernelGGL(prepare_intra_kernel, dim3(bt.T), dim3(64), 0, stream, bt,
                     (int32_t*)(w + Lw.intra),
                     (uint16_t*)(w + Lw.intra + 4 * int64_t(bt.T)));
  hipLaunchKernelGGL(spheres_kernel<Real>, dim3(grid_for(s.G)), dim3(256), 0, stream, s, bt,
                     (const Real*)coords, spheres);
  if (dense) {
    int64_t n = int64_t(s.P) * s.B * s.B;
    hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream, dense, n, 0);
  }
  if (n_tiles > 0)
    hipLaunchKernelGGL(neighbor_rows_kernel<Real>, dim3((unsigned)n_tiles), dim3(256), 0, stream, s,
                       (const int4*)tiles, n_tiles, spheres, (Real)max_dis, row_count, row_items,
                       row_cap, counters, dense);
  return check(hipGetLastError());
}

template <typename Real>
static ScoreArgs<Real> make_args(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                                 const void* coords, void* ws, int64_t row_cap,
                                 const WsLayout& Lw) {
  ScoreArgs<Real> a{};
  a.s = make_stack_view(st, btp->max_n_conn);
  a.bt = make_bt_view(btp);
  char* w = (char*)ws;
  a.w.row_count = (const uint32_t*)(w + Lw.row_count);
  a.w.row_items = (const int32_t*)(w + Lw.row_items);
  a.w.row_cap = row_cap;
  a.coords = (const Real*)coords;
  a.spheres = (const Real*)(w + Lw.spheres);
  a.records = (const Real*)(w + Lw.records);
  a.conn_pd = (const uint8_t*)(w + Lw.conn_pd);
  a.intra_count = (const int32_t*)(w + Lw.intra);
  a.intra_pairs = (const uint16_t*)(w + Lw.intra + 4 * int64_t(btp->n_block_types));
  a.global_params = (const Real*)btp->global_params;
  a.counters = (const uint32_t*)(w + Lw.counters);
  a.row_cursor = (uint32_t*)(w + Lw.cursors);
  a.energies64 = (double*)(w + Lw.energies);
  a.n_atoms = st->n_atoms_total;
  return a;
}

// workgroups of `kernel` (256 threads) resident on the whole device at once,
// cached per (kernel, device)
struct ResidencyCache {
  const void* kernel[64];
  int dev[64];
  int n[64];
  int used;
};

template <typename K>
static int resident_workgroups(K kernel) {
  static ResidencyCache cache{};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) dev = 0;
  const void* key = (const void*)kernel;
  for (int i = 0; i < cache.used; ++i)
    if (cache.kernel[i] == key && cache.dev[i] == dev) return cache.n[i];
  int cu = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess)
    per_cu = 0;
  const int n = std::max(1, cu) * std::max(1, per_cu);
  if (cache.used < 64) {
    cache.kernel[cache.used] = key;
    cache.dev[cache.used] = dev;
    cache.n[cache.used] = n;
    ++cache.used;
  }
  return n;
}

// n_wg <= 0: one workgroup per resident slot (the row queue balances the load)
template <typename Real, int MODE>
static void launch_score(const ScoreArgs<Real>& a, bool general, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.row_cursor,
                     8 * QSTRIDE, 0u);
  if (general) {
    auto k = score_general_kernel<Real, MODE>;
    if (n_wg <= 0) n_wg = resident_workgroups(k);
    hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
  } else {
    auto k = score_rows_kernel<Real, MODE>;
    if (n_wg <= 0) n_wg = resident_workgroups(k);
    hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
  }
}

// rows kernel only (modes the general kernel does not implement)
template <typename Real, int MODE>
static void launch_rows(const ScoreArgs<Real>& a, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.row_cursor,
                     8 * QSTRIDE, 0u);
  auto k = score_rows_kernel<Real, MODE>;
  if (n_wg <= 0) n_wg = resident_workgroups(k);
  hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
}

// the rows kernel compiled for `occ` waves per SIMD (0: the mode's default;
// float only -- other dtypes take the default)
template <typename Real, int MODE>
static void launch_rows_occ(const ScoreArgs<Real>& a, int occ, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.row_cursor,
                     8 * QSTRIDE, 0u);
  auto go = [&](auto k) {
    if (n_wg <= 0) n_wg = resident_workgroups(k);
    hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
  };
  if constexpr (std::is_same<Real, float>::value) {
    switch (occ) {
      case 3: go(score_rows_kernel<Real, MODE, 3>); return;
      case 4: go(score_rows_kernel<Real, MODE, 4>); return;
      case 5: go(score_rows_kernel<Real, MODE, 5>); return;
      default: break;
    }
  }
  go(score_rows_kernel<Real, MODE>);
}

// the stream kernel (whole-pose energies / WGRAD, blocks of <= 32 atoms),
// compiled for `occ` waves per SIMD (float: 3, 4 or 5; 0 = 4)
template <typename Real, int MODE>
static void launch_stream(const ScoreArgs<Real>& a, int occ, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.row_cursor,
                     8 * QSTRIDE, 0u);
  auto go = [&](auto k) {
    if (n_wg <= 0) n_wg = resident_workgroups(k);
    hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
  };
  if constexpr (std::is_same<Real, float>::value) {
    switch (occ) {
      case 3: go(score_stream_kernel<Real, MODE, 3>); return;
      case 5: go(score_stream_kernel<Real, MODE, 5>); return;
      default: break;
    }
  }
  go(score_stream_kernel<Real, MODE>);
}

// ek != nullptr: elec is scored in the same pass (M_ELEC; tables prepared in
// the workspace by tmol_ljlk_elec_prepare), energies and weights [4, P]
template <typename Real>
static int run_score(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                     const void* coords, void* ws, int64_t row_cap, int flags,
                     const void* weights, void* energies, void* dvdx, void* grad,
                     int n_wave_groups, hipStream_t stream, const WsLayout& Lw,
                     const tmol_elec_tables* e = nullptr, const ElecConst<Real>* ek = nullptr) {
  ScoreArgs<Real> a = make_args<Real>(st, btp, coords, ws, row_cap, Lw);
  const int n_terms = ek ? 4 : 3;
  if (ek) {
    a.e_charge = (const Real*)e->partial_charge;
    a.e_pd = (const uint32_t*)((char*)ws + Lw.e_pd);
    a.e_intra_count = (const int32_t*)((char*)ws + Lw.e_intra);
    a.e_intra_pairs = (const uint16_t*)((char*)ws + Lw.e_intra + 4 * int64_t(btp->n_block_types));
    a.ek = *ek;
  }
  a.weights = (const Real*)weights;
  a.energies_bp = (Real*)energies;
  a.dvdx = (Real*)dvdx;
  a.grad = (Real*)grad;
  const bool bp = flags & TMOL_LJLK_BLOCK_PAIR;
  const bool noe = flags & TMOL_LJLK_NO_ENERGY;
  const bool acc = flags & TMOL_LJLK_ACCUMULATE;
  const bool det = flags & TMOL_LJLK_DETERMINISTIC;
  const bool general = (flags & TMOL_LJLK_GENERAL_KERNEL) || btp->max_n_atoms > TILE;
  const int P = st->n_poses;
  const int64_t NA3 = st->n_atoms_total * 3;
  if (det) {
    // float32 whole-pose DVDX / WGRAD (stream or row kernel; not the general one)
    if (!std::is_same<Real, float>::value || general || bp || ek) return TMOL_LJLK_ENOTSUP;
    if (!(flags & (TMOL_LJLK_DVDX | TMOL_LJLK_WGRAD))) return TMOL_LJLK_ENOTSUP;
    const int planes = (flags & TMOL_LJLK_DVDX) ? 3 : 1;
    a.fixed = (unsigned long long*)((char*)ws + Lw.fixed);
    fill<unsigned long long>(a.fixed, planes * NA3, stream);
  }
  if (!noe) {
    if (bp)
      fill<Real>((Real*)energies,
                 3 * (st->layout == TMOL_LJLK_PADDED
                          ? int64_t(P) * st->max_n_blocks * st->max_n_blocks
                          : int64_t(st->n_blocks_total) * row_cap),
                 stream);
    else
      fill<double>(a.energies64, int64_t(n_terms) * P, stream);
  }
  if ((flags & TMOL_LJLK_DVDX) && !acc && !det) fill<Real>((Real*)dvdx, 3 * NA3, stream);
  if ((flags & TMOL_LJLK_WGRAD) && !acc && !det) fill<Real>((Real*)grad, NA3, stream);
  const int n_wg = n_wave_groups;
  int mode = flags & (TMOL_LJLK_DVDX | TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR | TMOL_LJLK_NO_ENERGY);
  if (ek) mode |= M_ELEC;
  // launch variants (autotuner): waves per SIMD and rows per queue grab
  const int occ = (flags >> 16) & 15;
  a.row_grab = (flags >> 20) & 15;
  // whole-pose energies and the weighted gradient stream each row's atoms
  const bool stream_k = !general && !ek && !(flags & TMOL_LJLK_ROW_KERNEL) &&
                        (mode == 0 || mode == M_WGRAD || mode == (M_WGRAD | M_NOE));
  if (det) {
    if constexpr (std::is_same<Real, float>::value) {
      if (stream_k) {
        if (mode == M_WGRAD)
          launch_stream<Real, M_WGRAD | M_DET>(a, occ, n_wg, stream);
        else
          launch_stream<Real, M_WGRAD | M_NOE | M_DET>(a, occ, n_wg, stream);
      } else {
        switch (mode) {
          case M_DVDX: launch_rows<Real, M_DVDX | M_DET>(a, n_wg, stream); break;
          case M_WGRAD: launch_rows<Real, M_WGRAD | M_DET>(a, n_wg, stream); break;
          case M_WGRAD | M_NOE: launch_rows<Real, M_WGRAD | M_NOE | M_DET>(a, n_wg, stream); break;
          default: return TMOL_LJLK_ENOTSUP;
        }
      }
      const int64_t n = (mode == M_DVDX ? 3 : 1) * NA3;
      hipLaunchKernelGGL(fixed_to_real_kernel<Real>, dim3(grid_for(n)), dim3(256), 0, stream,
                         (const unsigned long long*)a.fixed,
                         (Real*)(mode == M_DVDX ? dvdx : grad), n, acc);
      if (!noe)
        hipLaunchKernelGGL(finalize_fixed_kernel<Real>, dim3(grid_for(3 * P)), dim3(256), 0,
                           stream, (const unsigned long long*)a.energies64, (Real*)energies,
                           3 * P);
      return check(hipGetLastError());
    }
    return TMOL_LJLK_ENOTSUP;
  }
  if (stream_k) {
    switch (mode) {
      case 0: launch_stream<Real, 0>(a, occ, n_wg, stream); break;
      case M_WGRAD: launch_stream<Real, M_WGRAD>(a, occ, n_wg, stream); break;
      default: launch_stream<Real, M_WGRAD | M_NOE>(a, occ, n_wg, stream); break;
    }
    mode = -1;  // launched
  } else if (occ && !general && !ek && (mode == 0 || mode == M_WGRAD)) {
    if (mode == 0)
      launch_rows_occ<Real, 0>(a, occ, n_wg, stream);
    else
      launch_rows_occ<Real, M_WGRAD>(a, occ, n_wg, stream);
    mode = -1;  // launched
  }
  switch (mode) {
    case -1: break;
    case M_ELEC: launch_rows<Real, M_ELEC>(a, n_wave_groups, stream); break;
    case M_ELEC | M_WGRAD: launch_rows<Real, M_ELEC | M_WGRAD>(a, n_wave_groups, stream); break;
    case M_ELEC | M_WGRAD | M_NOE:
      launch_rows<Real, M_ELEC | M_WGRAD | M_NOE>(a, n_wave_groups, stream);
      break;
    case 0: launch_score<Real, 0>(a, general, n_wg, stream); break;
    case M_DVDX: launch_score<Real, M_DVDX>(a, general, n_wg, stream); break;
    case M_WGRAD: launch_score<Real, M_WGRAD>(a, general, n_wg, stream); break;
    case M_WGRAD | M_NOE: launch_score<Real, M_WGRAD | M_NOE>(a, general, n_wg, stream); break;
    case M_BP: launch_score<Real, M_BP>(a, general, n_wg, stream); break;
    case M_BP | M_WGRAD | M_NOE:
      launch_score<Real, M_BP | M_WGRAD | M_NOE>(a, general, n_wg, stream);
      break;
    case M_BP | M_WGRAD: launch_score<Real, M_BP | M_WGRAD>(a, general, n_wg, stream); break;
    default: return TMOL_LJLK_ENOTSUP;
  }
  if (!noe && !bp)
    hipLaunchKernelGGL(finalize_kernel<Real>, dim3(grid_for(n_terms * P)), dim3(256), 0, stream,
                       (const double*)a.energies64, (Real*)energies, n_terms * P);
  return check(hipGetLastError());
}

// ------------------------------------------------------------------ no-allocation entry point

// Device scratch of tmol_ljlk_score_noalloc: one array per GPU, held from the
// library's load until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): the
// device address is looked up once per device with hipGetSymbolAddress and
// that address is used everywhere.
static __device__ __align__(256) unsigned char scratch[TMOL_LJLK_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void* noalloc_ptr[NOALLOC_MAX_DEVICES];
static std::mutex noalloc_mu;

// the current device's scratch address (looked up on the first call on that
// device, with that device current)
static int noalloc_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return TMOL_LJLK_EHIP;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return TMOL_LJLK_ENOTSUP;
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return TMOL_LJLK_EHIP;
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return TMOL_LJLK_OK;
}

// ------------------------------------------------------------------ rotamers

struct RotWsLayout {
  int64_t cursors, rot_spheres, block_spheres, records, conn_pd;
  int64_t row_pairs, row_units, pair_off, unit_off;
  int64_t e_pd, e_intra;  // elec: packed representative path distances, intra pair tables
  int64_t total;
};

static RotWsLayout rot_ws_layout(const tmol_ljlk_rotamer_set* r, const tmol_ljlk_block_types* bt,
                                 int dtype) {
  const int64_t rs = dtype == TMOL_LJLK_F64 ? 8 : 4;
  const int64_t G = int64_t(r->n_poses) * r->max_n_blocks;
  RotWsLayout L;
  int64_t o = 0;
  L.cursors = o;
  o = align256(o + 8 * 256);
  L.rot_spheres = o;
  o = align256(o + rs * 4 * int64_t(r->n_rots));
  L.block_spheres = o;
  o = align256(o + rs * 4 * G);
  L.records = o;
  o = align256(o + rs * REC * int64_t(bt->n_block_types) * bt->max_n_atoms);
  L.conn_pd = o;
  o = align256(o + int64_t(bt->n_block_types) * MAX_CONN * bt->max_n_atoms);
  L.row_pairs = o;
  o = align256(o + 8 * G);
  L.row_units = o;
  o = align256(o + 8 * G);
  L.pair_off = o;
  o = align256(o + 8 * (G + 1));
  L.unit_off = o;
  o = align256(o + 8 * (G + 1));
  L.e_pd = o;
  o = align256(o + 4 * int64_t(bt->n_block_types) * bt->max_n_atoms);
  L.e_intra = o;
  o = align256(o + 4 * int64_t(bt->n_block_types) + 2 * int64_t(bt->n_block_types) * INTRA_MAX);
  L.total = o;
  return L;
}

static int validate_rot(const tmol_ljlk_rotamer_set* r, const tmol_ljlk_block_types* bt,
                        int dtype) {
  if (!r || !bt) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (bt->dtype != dtype) return TMOL_LJLK_EINVAL;
  if (bt->max_n_conn > MAX_CONN || bt->max_n_conn < 0) return TMOL_LJLK_ENOTSUP;
  if (bt->n_block_types <= 0 || bt->max_n_atoms <= 0) return TMOL_LJLK_EINVAL;
  if (bt->max_n_tiles != (bt->max_n_atoms - 1) / TILE + 1) return TMOL_LJLK_EINVAL;
  if (bt->max_n_atoms > WAVE) return TMOL_LJLK_ENOTSUP;  // one lane per atom of a rotamer
  if (r->n_poses <= 0 || r->max_n_blocks <= 0 || r->n_rots < 0) return TMOL_LJLK_EINVAL;
  if (int64_t(r->n_poses) * r->max_n_blocks >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  if (r->n_atoms_total >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  if (!r->rot_coord_offset || !r->block_ind_for_rot || !r->block_type_ind_for_rot ||
      !r->n_rots_for_block || !r->rot_offset_for_block || !r->first_rot_block_type ||
      !r->min_bondsep || !r->inter_block_bondsep)
    return TMOL_LJLK_EINVAL;
  return TMOL_LJLK_OK;
}

static RotView make_rot_view(const tmol_ljlk_rotamer_set* r, int C) {
  RotView v;
  v.P = r->n_poses;
  v.B = r->max_n_blocks;
  v.R = r->n_rots;
  v.C = C;
  v.rot_coord_offset = r->rot_coord_offset;
  v.block_ind_for_rot = r->block_ind_for_rot;
  v.block_type_for_rot = r->block_type_ind_for_rot;
  v.n_rots_for_block = r->n_rots_for_block;
  v.rot_offset_for_block = r->rot_offset_for_block;
  v.first_rot_block_type = r->first_rot_block_type;
  v.min_bondsep = r->min_bondsep;
  v.inter_block_bondsep = r->inter_block_bondsep;
  return v;
}

// Rotamer spheres, block spheres, neighbor rows and their scans, after the
// term's per-block-type tables: LJ+LK atom records (elec == nullptr) or
// elec's packed path distances and intra pair tables.
template <typename Real>
static int run_rot_neighbors(const tmol_ljlk_rotamer_set* r, const tmol_ljlk_block_types* btp,
                             const void* coords, double max_dis, void* ws, hipStream_t stream,
                             const RotWsLayout& Lw, const tmol_elec_tables* elec = nullptr) {
  RotView v = make_rot_view(r, btp->max_n_conn);
  BTView bt = make_bt_view(btp);
  char* w = (char*)ws;
  Real* rsph = (Real*)(w + Lw.rot_spheres);
  Real* bsph = (Real*)(w + Lw.block_spheres);
  const int64_t G = int64_t(v.P) * v.B;
  const int nrec = bt.T * bt.Amax;
  if (elec)
    hipLaunchKernelGGL(elec_prepare_kernel<Real>, dim3(bt.T), dim3(64), 0, stream, bt,
                       (const Real*)elec->partial_charge, elec->inter_repr_path_distance,
                       elec->intra_repr_path_distance, (uint32_t*)(w + Lw.e_pd),
                       (int32_t*)(w + Lw.e_intra),
                       (uint16_t*)(w + Lw.e_intra + 4 * int64_t(bt.T)));
  else
    hipLaunchKernelGGL(prepare_records_kernel<Real>, dim3(grid_for(nrec)), dim3(256), 0, stream,
                       bt, (const Real*)btp->type_params, btp->n_atom_types,
                       (Real*)(w + Lw.records), (uint8_t*)(w + Lw.conn_pd));
  if (v.R > 0) {
    // rotamer spheres: the block-sphere pass with rotamers as blocks
    StackView s{};
    s.G = v.R;
    s.block_atom_offset = v.rot_coord_offset;
    s.block_type = v.block_type_for_rot;
    hipLaunchKernelGGL(spheres_kernel<Real>, dim3((unsigned)((v.R + 255) / 256)), dim3(256), 0,
                       stream, s, bt, (const Real*)coords, rsph);
  }
  hipLaunchKernelGGL(rot_block_spheres_kernel<Real>, dim3((unsigned)((G + 255) / 256)),
                     dim3(256), 0, stream, v, (const Real*)rsph, bsph);
  int64_t* rp = (int64_t*)(w + Lw.row_pairs);
  int64_t* ru = (int64_t*)(w + Lw.row_units);
  hipLaunchKernelGGL(rot_rows_count_kernel<Real>,
                     dim3((unsigned)((G + WAVES_PER_WG - 1) / WAVES_PER_WG)), dim3(256), 0,
                     stream, v, (const Real*)bsph, (Real)max_dis, rp, ru);
  hipLaunchKernelGGL(scan2_kernel, dim3(1), dim3(1024), 0, stream, (const int64_t*)rp,
                     (const int64_t*)ru, G, (int64_t*)(w + Lw.pair_off),
                     (int64_t*)(w + Lw.unit_off));
  return check(hipGetLastError());
}

template <typename Real>
static RotArgs<Real> make_rot_args(const tmol_ljlk_rotamer_set* r,
                                   const tmol_ljlk_block_types* btp, const void* coords, void* ws,
                                   const RotWsLayout& Lw) {
  RotArgs<Real> a{};
  a.v = make_rot_view(r, btp->max_n_conn);
  a.bt = make_bt_view(btp);
  char* w = (char*)ws;
  a.coords = (const Real*)coords;
  a.records = (const Real*)(w + Lw.records);
  a.conn_pd = (const uint8_t*)(w + Lw.conn_pd);
  a.global_params = (const Real*)btp->global_params;
  a.rot_spheres = (const Real*)(w + Lw.rot_spheres);
  a.cursor = (uint32_t*)(w + Lw.cursors);
  return a;
}

template <typename Real, int MODE>
static void launch_rot_score(const RotArgs<Real>& a, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.cursor,
                     int64_t(8 * QSTRIDE), 0u);
  if (MODE == 0) {
    auto k = rotamer_energy_kernel<Real>;
    if (n_wg <= 0) n_wg = resident_workgroups(k);
    hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
  } else {
    auto k = rotamer_score_kernel<Real, MODE>;
    if (n_wg <= 0) n_wg = resident_workgroups(k);
    hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
  }
}

}  // namespace tmol_ljlk

using namespace tmol_ljlk;

// ------------------------------------------------------------------ elec

// eps() of the elec potential with the reference's argument types
// (potentials.hh:43-53): Real dist, float D, D0, S; the 0.5 literal is
// double, so part of the product is formed in double whatever Real is
template <typename Real>
static Real elec_eps_host(Real dist, float D, float D0, float S) {
  const double a = 0.5 * (D - D0);
  const Real poly = 2 + 2 * dist * S + dist * dist * S * S;
  return (Real)(D - a * poly * std::exp(-dist * S));
}

template <typename Real>
static Real elec_deps_host(Real dist, float D, float D0, float S) {
  return (Real)(0.5 * (D - D0) * dist * dist * S * S * S * std::exp(-dist * S));
}

// per-launch constants (potentials.hh:55-138): region bounds, the flat
// value, and the Coulomb value and slope at the two fade joins, per unit
// charge product; eps at max_dis and min_dis in float as in the reference
template <typename Real>
static ElecConst<Real> elec_constants(const float* g) {
  const float D = g[0], D0 = g[1], S = g[2], mn = g[3], mx = g[4];
  ElecConst<Real> k;
  k.D = D;
  k.S = S;
  k.hdd = (Real)(0.5 * (D - D0));
  k.C1 = (Real)322.0637;
  k.C2 = k.C1 / (mx * elec_eps_host<float>(mx, D, D0, S));
  k.lps = (Real)(mn - 0.25);
  k.lpe = (Real)(mn + 0.25);
  k.hps = (Real)(mx - 1.0);
  k.hpe = mx;
  k.mds = k.C1 / (mn * elec_eps_host<float>(mn, D, D0, S)) - k.C2;
  const Real el = elec_eps_host<Real>(k.lpe, D, D0, S), dl = elec_deps_host<Real>(k.lpe, D, D0, S);
  k.e_lo = k.C1 / (k.lpe * el) - k.C2;
  k.de_lo = -k.C1 * (el + k.lpe * dl) / (k.lpe * k.lpe * el * el);
  const Real eh = elec_eps_host<Real>(k.hps, D, D0, S), dh = elec_deps_host<Real>(k.hps, D, D0, S);
  k.e_hi = k.C1 / (k.hps * eh) - k.C2;
  k.de_hi = -k.C1 * (eh + k.hps * dh) / (k.hps * k.hps * eh * eh);
  k.lw = k.lpe - k.lps;
  k.hw = k.hpe - k.hps;
  k.inv_lw = Real(1) / k.lw;
  k.inv_hw = Real(1) / k.hw;
  const Real reach = (Real)mx + Real(1e-3);
  k.cull2 = reach * reach;
  return k;
}

// elec's block-type tables and global parameters (shared by the pose and
// rotamer entry points)
static int validate_elec_tables(const tmol_ljlk_block_types* bt, const tmol_elec_tables* e) {
  if (!e || !e->partial_charge || !e->inter_repr_path_distance || !e->intra_repr_path_distance)
    return TMOL_LJLK_EINVAL;
  if (!bt->n_atoms || !bt->n_conn || !bt->is_ligand || (bt->max_n_conn > 0 && !bt->conn_atoms))
    return TMOL_LJLK_EINVAL;
  if (bt->max_n_atoms > TILE) return TMOL_LJLK_ENOTSUP;
  const float* g = e->global_params;
  for (int i = 0; i < 5; ++i)
    if (!std::isfinite(g[i])) return TMOL_LJLK_EINVAL;
  // region bounds must be ordered: min_dis - 0.25 < min_dis + 0.25 <= max_dis - 1 < max_dis
  if (!(g[3] > 0.25f && double(g[3]) + 0.25 <= double(g[4]) - 1.0)) return TMOL_LJLK_EINVAL;
  return TMOL_LJLK_OK;
}

static int validate_elec(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                         const tmol_elec_tables* e, int dtype) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  return validate_elec_tables(bt, e);
}

static int validate_elec_rot(const tmol_ljlk_rotamer_set* r, const tmol_ljlk_block_types* bt,
                             const tmol_elec_tables* e, int dtype) {
  int v = validate_rot(r, bt, dtype);
  if (v) return v;
  return validate_elec_tables(bt, e);
}

template <typename Real, bool BWD>
static int run_elec_rot_score(const tmol_ljlk_rotamer_set* r, const tmol_ljlk_block_types* btp,
                              const tmol_elec_tables* e, const void* coords, void* ws,
                              const RotWsLayout& Lw, const void* units, int64_t n_units,
                              int64_t n_pairs, const void* dTdV, void* energies, void* grad,
                              bool accumulate, int n_wg, hipStream_t stream) {
  ElecRotArgs<Real> a{};
  a.v = make_rot_view(r, btp->max_n_conn);
  a.bt = make_bt_view(btp);
  char* w = (char*)ws;
  a.coords = (const Real*)coords;
  a.rot_spheres = (const Real*)(w + Lw.rot_spheres);
  a.charge = (const Real*)e->partial_charge;
  a.epd = (const uint32_t*)(w + Lw.e_pd);
  a.intra_count = (const int32_t*)(w + Lw.e_intra);
  a.intra_pairs = (const uint16_t*)(w + Lw.e_intra + 4 * int64_t(btp->n_block_types));
  a.k = elec_constants<Real>(e->global_params);
  a.units = (const RotUnit*)units;
  a.n_units = n_units;
  a.n_pairs = n_pairs;
  a.cursor = (uint32_t*)(w + Lw.cursors);
  a.energies = (Real*)energies;
  a.dTdV = (const Real*)dTdV;
  a.grad = (Real*)grad;
  if (BWD && !accumulate) fill<Real>((Real*)grad, r->n_atoms_total * 3, stream);
  if (n_units > 0) {
    hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.cursor,
                       int64_t(8 * QSTRIDE), 0u);
    auto k = elec_rotamer_kernel<Real, BWD>;
    if (n_wg <= 0) n_wg = resident_workgroups(k);
    hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
  }
  return check(hipGetLastError());
}

template <typename Real>
static int run_elec_neighbors(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                              const tmol_elec_tables* e, const void* coords, double max_dis,
                              const int32_t* tiles, int64_t n_tiles, void* ws, int64_t row_cap,
                              int32_t* dense, hipStream_t stream, const WsLayout& Lw) {
  StackView s = make_stack_view(st, btp->max_n_conn);
  BTView bt = make_bt_view(btp);
  char* w = (char*)ws;
  uint32_t* counters = (uint32_t*)(w + Lw.counters);
  Real* spheres = (Real*)(w + Lw.spheres);
  uint32_t* row_count = (uint32_t*)(w + Lw.row_count);
  int32_t* row_items = (int32_t*)(w + Lw.row_items);
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, stream, counters, 16,
                     (uint32_t)0);
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(grid_for(s.G)), dim3(256), 0, stream,
                     row_count, int64_t(s.G), (uint32_t)0);
  hipLaunchKernelGGL(elec_prepare_kernel<Real>, dim3(bt.T), dim3(64), 0, stream, bt,
                     (const Real*)e->partial_charge, e->inter_repr_path_distance,
                     e->intra_repr_path_distance, (uint32_t*)(w + Lw.conn_pd),
                     (int32_t*)(w + Lw.intra), (uint16_t*)(w + Lw.intra + 4 * int64_t(bt.T)));
  hipLaunchKernelGGL(spheres_kernel<Real>, dim3(grid_for(s.G)), dim3(256), 0, stream, s, bt,
                     (const Real*)coords, spheres);
  if (dense) {
    int64_t n = int64_t(s.P) * s.B * s.B;
    hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream, dense, n, 0);
  }
  if (n_tiles > 0)
    hipLaunchKernelGGL(neighbor_rows_kernel<Real>, dim3((unsigned)n_tiles), dim3(256), 0, stream, s,
                       (const int4*)tiles, n_tiles, spheres, (Real)max_dis, row_count, row_items,
                       row_cap, counters, dense);
  return check(hipGetLastError());
}

template <typename Real, int MODE>
static void launch_elec(const ElecArgs<Real>& a, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.row_cursor,
                     8 * QSTRIDE, 0u);
  auto k = elec_rows_kernel<Real, MODE>;
  if (n_wg <= 0) n_wg = resident_workgroups(k);
  hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
}

template <typename Real>
static int run_elec_score(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                          const tmol_elec_tables* e, const void* coords, void* ws,
                          int64_t row_cap, int flags, const void* weights, void* energies,
                          void* dvdx, void* grad, int n_wg, hipStream_t stream,
                          const WsLayout& Lw, unsigned long long* census = nullptr) {
  ElecArgs<Real> a{};
  a.s = make_stack_view(st, btp->max_n_conn);
  a.bt = make_bt_view(btp);
  char* w = (char*)ws;
  a.w.row_count = (const uint32_t*)(w + Lw.row_count);
  a.w.row_items = (const int32_t*)(w + Lw.row_items);
  a.w.row_cap = row_cap;
  a.coords = (const Real*)coords;
  a.spheres = (const Real*)(w + Lw.spheres);
  a.charge = (const Real*)e->partial_charge;
  a.epd = (const uint32_t*)(w + Lw.conn_pd);
  a.intra_count = (const int32_t*)(w + Lw.intra);
  a.intra_pairs = (const uint16_t*)(w + Lw.intra + 4 * int64_t(btp->n_block_types));
  a.k = elec_constants<Real>(e->global_params);
  a.row_cursor = (uint32_t*)(w + Lw.cursors);
  a.weights = (const Real*)weights;
  a.energies64 = (double*)(w + Lw.energies);
  a.energies_bp = (Real*)energies;
  a.dvdx = (Real*)dvdx;
  a.grad = (Real*)grad;
  a.n_atoms = st->n_atoms_total;
  a.census = census;
  if (census) {
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, stream, census, 16,
                       0ull);
    launch_elec<Real, M_CENSUS>(a, n_wg, stream);
    return check(hipGetLastError());
  }
  const bool bp = flags & TMOL_LJLK_BLOCK_PAIR;
  const bool noe = flags & TMOL_LJLK_NO_ENERGY;
  const bool acc = flags & TMOL_LJLK_ACCUMULATE;
  const int P = st->n_poses;
  const int64_t NA3 = st->n_atoms_total * 3;
  if (!noe) {
    if (bp)
      fill<Real>((Real*)energies,
                 st->layout == TMOL_LJLK_PADDED
                     ? int64_t(P) * st->max_n_blocks * st->max_n_blocks
                     : int64_t(st->n_blocks_total) * row_cap,
                 stream);
    else
      fill<double>(a.energies64, P, stream);
  }
  if ((flags & TMOL_LJLK_DVDX) && !acc) fill<Real>((Real*)dvdx, NA3, stream);
  if ((flags & TMOL_LJLK_WGRAD) && !acc) fill<Real>((Real*)grad, NA3, stream);
  const int mode = flags & (TMOL_LJLK_DVDX | TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR |
                            TMOL_LJLK_NO_ENERGY);
  switch (mode) {
    case 0: launch_elec<Real, 0>(a, n_wg, stream); break;
    case M_DVDX: launch_elec<Real, M_DVDX>(a, n_wg, stream); break;
    case M_WGRAD: launch_elec<Real, M_WGRAD>(a, n_wg, stream); break;
    case M_WGRAD | M_NOE: launch_elec<Real, M_WGRAD | M_NOE>(a, n_wg, stream); break;
    case M_BP: launch_elec<Real, M_BP>(a, n_wg, stream); break;
    case M_BP | M_WGRAD | M_NOE: launch_elec<Real, M_BP | M_WGRAD | M_NOE>(a, n_wg, stream); break;
    case M_BP | M_WGRAD: launch_elec<Real, M_BP | M_WGRAD>(a, n_wg, stream); break;
    default: return TMOL_LJLK_ENOTSUP;
  }
  if (!noe && !bp)
    hipLaunchKernelGGL(finalize_kernel<Real>, dim3(grid_for(P)), dim3(256), 0, stream,
                       (const double*)a.energies64, (Real*)energies, P);
  return check(hipGetLastError());
}

// ------------------------------------------------------------------ lk_ball

// scoring = false: the waters and their backward, which need no count-pair
// tables
static int validate_lkb(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                        const tmol_lkball_tables* k, int dtype, bool scoring) {
  if (!s || !bt || !k) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (bt->max_n_conn > MAX_CONN || bt->max_n_conn < 0) return TMOL_LJLK_ENOTSUP;
  if (bt->n_block_types <= 0 || bt->max_n_atoms <= 0) return TMOL_LJLK_EINVAL;
  if (!bt->n_atoms || !bt->n_conn || (bt->max_n_conn > 0 && !bt->conn_atoms))
    return TMOL_LJLK_EINVAL;
  if (k->max_n_occ < 0 || k->max_n_recipes < 0) return TMOL_LJLK_EINVAL;
  if (k->max_n_occ > LKB_OMAX || k->max_n_polar > LKB_PMAX || k->max_n_polar < 0)
    return TMOL_LJLK_ENOTSUP;
  if (!k->n_occ || !k->n_polar || !k->occ_atoms || !k->occ_params || !k->n_recipes ||
      (k->max_n_recipes > 0 && !k->recipes) || !k->path_distance || !k->global_params ||
      !k->watergen_params || !k->water_tors)
    return TMOL_LJLK_EINVAL;
  if (s->n_poses <= 0 || s->n_blocks_total <= 0) return TMOL_LJLK_EINVAL;
  if (s->layout == TMOL_LJLK_PADDED) {
    if (int64_t(s->n_poses) * s->max_n_blocks != s->n_blocks_total) return TMOL_LJLK_EINVAL;
    if (scoring && (!s->min_bondsep || !s->inter_block_bondsep)) return TMOL_LJLK_EINVAL;
  } else if (s->layout == TMOL_LJLK_RAGGED) {
    if (!s->block_pose || (scoring && (!s->bonded_start || !s->bonded))) return TMOL_LJLK_EINVAL;
  } else {
    return TMOL_LJLK_EINVAL;
  }
  if (!s->block_atom_offset || !s->block_type) return TMOL_LJLK_EINVAL;
  if (s->n_atoms_total >= (int64_t(1) << 31) / 12) return TMOL_LJLK_ENOTSUP;
  return TMOL_LJLK_OK;
}

template <typename Real>
static int run_lkb_waters(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                          const tmol_lkball_tables* k, const void* coords, void* waters,
                          bool backward, const void* dwat, void* dxyz, hipStream_t stream) {
  StackView s = make_stack_view(st, btp->max_n_conn);
  const int R = k->max_n_recipes;
  if (!backward) {
    const int64_t n = st->n_atoms_total * 12;
    hipLaunchKernelGGL(fill_kernel<Real>, dim3(grid_for(n)), dim3(256), 0, stream, (Real*)waters,
                       n, Real(NAN));
  }
  if (R == 0) return check(hipGetLastError());
  const int64_t n = int64_t(s.G) * R;
  if (!backward)
    hipLaunchKernelGGL(lkb_waters_kernel<Real>, dim3(grid_for(n)), dim3(256), 0, stream, s,
                       k->n_recipes, k->recipes, R, (const Real*)coords,
                       (const Real*)k->watergen_params, (const Real*)k->water_tors,
                       (Real*)waters);
  else
    hipLaunchKernelGGL(lkb_waters_backward_kernel<Real>, dim3(grid_for(n)), dim3(256), 0, stream,
                       s, k->n_recipes, k->recipes, R, (const Real*)coords,
                       (const Real*)k->watergen_params, (const Real*)k->water_tors,
                       (const Real*)dwat, (Real*)dxyz);
  return check(hipGetLastError());
}

// spheres + work list without the LJ+LK atom records; FULL: rows list
// both triangles (lk_ball), otherwise g2 >= g1 (hbond)
template <typename Real, bool FULL = true>
static int run_lkb_neighbors(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                             const void* coords, double max_dis, const int32_t* tiles,
                             int64_t n_tiles, void* ws, int64_t row_cap, int32_t* dense,
                             hipStream_t stream, const WsLayout& Lw) {
  StackView s = make_stack_view(st, btp->max_n_conn);
  BTView bt = make_bt_view(btp);
  char* w = (char*)ws;
  uint32_t* counters = (uint32_t*)(w + Lw.counters);
  Real* spheres = (Real*)(w + Lw.spheres);
  uint32_t* row_count = (uint32_t*)(w + Lw.row_count);
  int32_t* row_items = (int32_t*)(w + Lw.row_items);
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, stream, counters, 16,
                     (uint32_t)0);
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(grid_for(s.G)), dim3(256), 0, stream,
                     row_count, int64_t(s.G), (uint32_t)0);
  hipLaunchKernelGGL(spheres_kernel<Real>, dim3(grid_for(s.G)), dim3(256), 0, stream, s, bt,
                     (const Real*)coords, spheres);
  if (dense) {
    int64_t n = int64_t(s.P) * s.B * s.B;
    hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream, dense, n, 0);
  }
  if (n_tiles > 0) {
    auto k = neighbor_rows_kernel<Real, FULL>;
    hipLaunchKernelGGL(k, dim3((unsigned)n_tiles), dim3(256), 0, stream, s, (const int4*)tiles,
                       n_tiles, spheres, (Real)max_dis, row_count, row_items, row_cap, counters,
                       dense, st->pose_block_start);
  }
  return check(hipGetLastError());
}

template <typename Real, int MODE>
static void launch_lkb(const LkbArgs<Real>& a, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.row_cursor,
                     8 * QSTRIDE, 0u);
  auto k = lkb_gather_kernel<Real, MODE>;
  if (n_wg <= 0) n_wg = resident_workgroups(k);
  hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
}

template <typename Real>
static int run_lkb_score(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                         const tmol_lkball_tables* k, const void* coords, const void* waters,
                         void* ws, int64_t row_cap, int flags, const void* weights,
                         void* energies, void* dxyz, void* dwat, int n_wg, hipStream_t stream,
                         const WsLayout& Lw, unsigned long long* census = nullptr) {
  LkbArgs<Real> a{};
  a.s = make_stack_view(st, btp->max_n_conn);
  a.bt = make_bt_view(btp);
  char* w = (char*)ws;
  a.w.row_count = (const uint32_t*)(w + Lw.row_count);
  a.w.row_items = (const int32_t*)(w + Lw.row_items);
  a.w.row_cap = row_cap;
  a.coords = (const Real*)coords;
  a.waters = (const Real*)waters;
  a.n_occ = k->n_occ;
  a.n_pol = k->n_polar;
  a.occ_atoms = k->occ_atoms;
  a.occ_params = (const Real*)k->occ_params;
  a.path_distance = k->path_distance;
  a.Omax = k->max_n_occ;
  a.gp = (const Real*)k->global_params;
  a.row_cursor = (uint32_t*)(w + Lw.cursors);
  a.weights = (const Real*)weights;
  a.energies64 = (double*)(w + Lw.energies);
  a.energies_bp = (Real*)energies;
  a.dxyz = (Real*)dxyz;
  a.dwat = (Real*)dwat;
  a.census = census;
  const LkbWs Lk = lkb_ws_layout(Lw, st->n_blocks_total, k->max_n_occ);
  a.info = (const int4*)(w + Lk.info);
  a.opos = (const float4*)(w + Lk.opos);
  a.oatom = (const int32_t*)(w + Lk.oatom);
  {
    const int64_t n = int64_t(st->n_blocks_total) * k->max_n_occ;
    hipLaunchKernelGGL(lkb_prep_kernel<Real>, dim3(grid_for(n)), dim3(256), 0, stream, a.s,
                       k->n_occ, k->n_polar, btp->n_conn, k->occ_atoms, k->max_n_occ,
                       (const Real*)coords, (int4*)(w + Lk.info), (float4*)(w + Lk.opos),
                       (int32_t*)(w + Lk.oatom));
  }
  if (census) {
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, stream, census, 4,
                       0ull);
    launch_lkb<Real, M_CENSUS>(a, n_wg, stream);
    return check(hipGetLastError());
  }
  const bool bp = flags & TMOL_LJLK_BLOCK_PAIR;
  const bool noe = flags & TMOL_LJLK_NO_ENERGY;
  const bool acc = flags & TMOL_LJLK_ACCUMULATE;
  const int P = st->n_poses;
  if (!noe) {
    if (bp)
      fill<Real>((Real*)energies,
                 4 * (st->layout == TMOL_LJLK_PADDED
                          ? int64_t(P) * st->max_n_blocks * st->max_n_blocks
                          : int64_t(st->n_blocks_total) * row_cap),
                 stream);
    else
      fill<double>(a.energies64, 4 * int64_t(P), stream);
  }
  if ((flags & TMOL_LJLK_WGRAD) && !acc) {
    fill<Real>((Real*)dxyz, st->n_atoms_total * 3, stream);
    fill<Real>((Real*)dwat, st->n_atoms_total * 12, stream);
  }
  const int mode = flags & (TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR | TMOL_LJLK_NO_ENERGY);
  switch (mode) {
    case 0: launch_lkb<Real, 0>(a, n_wg, stream); break;
    case M_WGRAD: launch_lkb<Real, M_WGRAD>(a, n_wg, stream); break;
    case M_WGRAD | M_NOE: launch_lkb<Real, M_WGRAD | M_NOE>(a, n_wg, stream); break;
    case M_BP: launch_lkb<Real, M_BP>(a, n_wg, stream); break;
    case M_BP | M_WGRAD: launch_lkb<Real, M_BP | M_WGRAD>(a, n_wg, stream); break;
    case M_BP | M_WGRAD | M_NOE: launch_lkb<Real, M_BP | M_WGRAD | M_NOE>(a, n_wg, stream); break;
    default: return TMOL_LJLK_ENOTSUP;
  }
  if (!noe && !bp)
    hipLaunchKernelGGL(finalize_kernel<Real>, dim3(grid_for(4 * int64_t(P))), dim3(256), 0,
                       stream, (const double*)a.energies64, (Real*)energies, 4 * P);
  return check(hipGetLastError());
}

// ------------------------------------------------------------------ hbond

static int validate_hb(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                       const tmol_hbond_tables* h, int dtype, bool scoring) {
  if (!s || !bt || !h) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (bt->max_n_conn > MAX_CONN || bt->max_n_conn < 0) return TMOL_LJLK_ENOTSUP;
  if (bt->n_block_types <= 0 || bt->max_n_atoms <= 0) return TMOL_LJLK_EINVAL;
  if (bt->max_n_atoms > TILE) return TMOL_LJLK_ENOTSUP;
  if (!bt->n_atoms || !bt->n_conn || (bt->max_n_conn > 0 && !bt->conn_atoms))
    return TMOL_LJLK_EINVAL;
  if (h->max_n_don < 0 || h->max_n_acc < 0 || h->n_don_types <= 0 || h->n_acc_types <= 0)
    return TMOL_LJLK_EINVAL;
  if (h->max_n_don > TILE || h->max_n_acc > TILE) return TMOL_LJLK_ENOTSUP;
  if (h->n_don_types >= (1 << 15) || h->n_acc_types >= (1 << 15)) return TMOL_LJLK_ENOTSUP;
  if (!h->n_don || !h->n_acc || (h->max_n_don > 0 && !h->don) ||
      (h->max_n_acc > 0 && !h->acc) || !h->base_src || !h->path_distance || !h->pair_params ||
      !h->pair_polys || !h->global_params)
    return TMOL_LJLK_EINVAL;
  if (s->n_poses <= 0 || s->n_blocks_total <= 0) return TMOL_LJLK_EINVAL;
  if (s->layout == TMOL_LJLK_PADDED) {
    if (int64_t(s->n_poses) * s->max_n_blocks != s->n_blocks_total) return TMOL_LJLK_EINVAL;
    if (scoring && (!s->min_bondsep || !s->inter_block_bondsep)) return TMOL_LJLK_EINVAL;
  } else if (s->layout == TMOL_LJLK_RAGGED) {
    if (!s->block_pose || (scoring && (!s->bonded_start || !s->bonded))) return TMOL_LJLK_EINVAL;
  } else {
    return TMOL_LJLK_EINVAL;
  }
  if (!s->block_atom_offset || !s->block_type) return TMOL_LJLK_EINVAL;
  if (s->n_atoms_total >= (int64_t(1) << 31) / 9) return TMOL_LJLK_ENOTSUP;
  return TMOL_LJLK_OK;
}

template <typename Real>
static int run_hb_bases(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                        const tmol_hbond_tables* h, const void* coords, void* derived,
                        int32_t* inds, hipStream_t stream) {
  StackView s = make_stack_view(st, btp->max_n_conn);
  const int64_t na = st->n_atoms_total;
  fill<Real>((Real*)derived, 9 * na, stream, Real(NAN));
  hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(grid_for(3 * na)), dim3(256), 0, stream, inds,
                     3 * na, -1);
  const int64_t n = int64_t(s.G) * btp->max_n_atoms;
  hipLaunchKernelGGL(hb_bases_kernel<Real>, dim3(grid_for(n)), dim3(256), 0, stream, s,
                     btp->n_atoms, h->base_src, btp->max_n_atoms, (const Real*)coords,
                     (Real*)derived, inds);
  return check(hipGetLastError());
}

template <typename Real, int MODE>
static void launch_hb(const HbArgs<Real>& a, int n_wg, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(2), dim3(256), 0, stream, a.row_cursor,
                     8 * QSTRIDE, 0u);
  auto k = hb_score_kernel<Real, MODE>;
  if (n_wg <= 0) n_wg = resident_workgroups(k);
  hipLaunchKernelGGL(k, dim3(n_wg), dim3(256), 0, stream, a);
}

template <typename Real>
static int run_hb_score(const tmol_ljlk_stack* st, const tmol_ljlk_block_types* btp,
                        const tmol_hbond_tables* h, const void* coords, const void* derived,
                        const int32_t* dinds, void* ws, int64_t row_cap, int flags,
                        const void* weights, void* energies, void* dxyz, int n_wg,
                        hipStream_t stream, const WsLayout& Lw,
                        unsigned long long* census = nullptr) {
  HbArgs<Real> a{};
  a.s = make_stack_view(st, btp->max_n_conn);
  char* w = (char*)ws;
  a.w.row_count = (const uint32_t*)(w + Lw.row_count);
  a.w.row_items = (const int32_t*)(w + Lw.row_items);
  a.w.row_cap = row_cap;
  a.coords = (const Real*)coords;
  a.derived = (const Real*)derived;
  a.dinds = dinds;
  a.n_conn = btp->n_conn;
  a.conn_atoms = btp->conn_atoms;
  a.n_don = h->n_don;
  a.n_acc = h->n_acc;
  a.don = h->don;
  a.acc = h->acc;
  a.Dmax = h->max_n_don;
  a.Xmax = h->max_n_acc;
  a.Am = btp->max_n_atoms;
  a.path_distance = h->path_distance;
  a.pair_params = (const Real*)h->pair_params;
  a.polys = h->pair_polys;
  a.n_acc_types = h->n_acc_types;
  a.gp = (const Real*)h->global_params;
  a.spheres = (const Real*)(w + Lw.spheres);
  a.row_cursor = (uint32_t*)(w + Lw.cursors);
  a.weights = (const Real*)weights;
  a.energies64 = (double*)(w + Lw.energies);
  a.energies_bp = (Real*)energies;
  a.dxyz = (Real*)dxyz;
  a.census = census;
  if (census) {
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, stream, census, 3,
                       0ull);
    launch_hb<Real, M_CENSUS>(a, n_wg, stream);
    return check(hipGetLastError());
  }
  const bool bp = flags & TMOL_LJLK_BLOCK_PAIR;
  const bool noe = flags & TMOL_LJLK_NO_ENERGY;
  const bool acc = flags & TMOL_LJLK_ACCUMULATE;
  const int P = st->n_poses;
  if (!noe) {
    if (bp)
      fill<Real>((Real*)energies, int64_t(P) * st->max_n_blocks * st->max_n_blocks, stream);
    else
      fill<double>(a.energies64, int64_t(P), stream);
  }
  if ((flags & TMOL_LJLK_WGRAD) && !acc) fill<Real>((Real*)dxyz, st->n_atoms_total * 3, stream);
  const int mode = flags & (TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR | TMOL_LJLK_NO_ENERGY);
  switch (mode) {
    case 0: launch_hb<Real, 0>(a, n_wg, stream); break;
    case M_WGRAD: launch_hb<Real, M_WGRAD>(a, n_wg, stream); break;
    case M_WGRAD | M_NOE: launch_hb<Real, M_WGRAD | M_NOE>(a, n_wg, stream); break;
    case M_BP: launch_hb<Real, M_BP>(a, n_wg, stream); break;
    case M_BP | M_WGRAD: launch_hb<Real, M_BP | M_WGRAD>(a, n_wg, stream); break;
    case M_BP | M_WGRAD | M_NOE: launch_hb<Real, M_BP | M_WGRAD | M_NOE>(a, n_wg, stream); break;
    default: return TMOL_LJLK_ENOTSUP;
  }
  if (!noe && !bp)
    hipLaunchKernelGGL(finalize_kernel<Real>, dim3(grid_for(int64_t(P))), dim3(256), 0, stream,
                       (const double*)a.energies64, (Real*)energies, P);
  return check(hipGetLastError());
}

extern "C" {

int64_t tmol_ljlk_count_neighbor_tiles(const int32_t* pose_block_start_host, int32_t n_poses) {
  int64_t n = 0;
  for (int p = 0; p < n_poses; ++p) {
    int64_t nb = pose_block_start_host[p + 1] - pose_block_start_host[p];
    int64_t t = (nb + NBR_TILE - 1) / NBR_TILE;
    n += t * (t + 1) / 2;
  }
  return n;
}

int tmol_ljlk_fill_neighbor_tiles(const int32_t* pbs, int32_t n_poses, int32_t* tiles) {
  if (!pbs || !tiles || n_poses < 0) return TMOL_LJLK_EINVAL;
  int64_t k = 0;
  for (int p = 0; p < n_poses; ++p) {
    int start = pbs[p], end = pbs[p + 1];
    int t = (end - start + NBR_TILE - 1) / NBR_TILE;
    for (int i = 0; i < t; ++i)
      for (int j = i; j < t; ++j) {
        tiles[4 * k + 0] = start + i * NBR_TILE;
        tiles[4 * k + 1] = start + j * NBR_TILE;
        tiles[4 * k + 2] = end;
        tiles[4 * k + 3] = p;
        ++k;
      }
  }
  return TMOL_LJLK_OK;
}

int64_t tmol_ljlk_workspace_bytes(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                                  int64_t row_capacity, int32_t dtype) {
  if (!s || !bt || row_capacity < 0) return -1;
  return ws_layout(s, bt, row_capacity, dtype).total;
}

int64_t tmol_ljlk_workspace_bytes_flags(const tmol_ljlk_stack* s,
                                        const tmol_ljlk_block_types* bt, int64_t row_capacity,
                                        int32_t dtype, int32_t flags) {
  if (!s || !bt || row_capacity < 0) return -1;
  const int planes = (flags & TMOL_LJLK_DETERMINISTIC) ? ((flags & TMOL_LJLK_DVDX) ? 3 : 1) : 0;
  return ws_layout(s, bt, row_capacity, dtype, planes).total;
}

int tmol_ljlk_work_list(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                        void* workspace, int64_t row_capacity, int32_t dtype,
                        uint32_t** row_count_out, int32_t** row_items_out) {
  if (!s || !bt || !workspace || !row_count_out || !row_items_out) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  *row_count_out = (uint32_t*)((char*)workspace + L.row_count);
  *row_items_out = (int32_t*)((char*)workspace + L.row_items);
  return TMOL_LJLK_OK;
}

int tmol_ljlk_spheres(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                      void* workspace, int64_t row_capacity, int32_t dtype, void** spheres_out) {
  if (!s || !bt || !workspace || !spheres_out) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  *spheres_out = (char*)workspace + L.spheres;
  return TMOL_LJLK_OK;
}

int tmol_ljlk_neighbors(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                        const void* coords, int32_t dtype, double max_dis, const int32_t* tiles,
                        int64_t n_tiles, void* workspace, int64_t ws_bytes,
                        int64_t row_capacity, int32_t* dense, void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!coords || !workspace || (n_tiles > 0 && !tiles) || row_capacity <= 0)
    return TMOL_LJLK_EINVAL;
  if (n_tiles >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  if (dense && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_neighbors<float>(s, bt, coords, max_dis, tiles, n_tiles, workspace, row_capacity,
                                dense, st, L);
  return run_neighbors<double>(s, bt, coords, max_dis, tiles, n_tiles, workspace, row_capacity,
                               dense, st, L);
}

int tmol_ljlk_score(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                    const void* coords, int32_t dtype, void* workspace, int64_t ws_bytes,
                    int64_t row_capacity, int32_t flags, const void* weights, void* energies,
                    void* dvdx, void* grad, int32_t n_wave_groups, void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!coords || !workspace || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  const int det_planes =
      (flags & TMOL_LJLK_DETERMINISTIC) ? ((flags & TMOL_LJLK_DVDX) ? 3 : 1) : 0;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype, det_planes);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  if (!(flags & TMOL_LJLK_NO_ENERGY) && !energies) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_DVDX) && !dvdx) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_WGRAD) && (!grad || !weights)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_DVDX) && (flags & TMOL_LJLK_WGRAD)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_BLOCK_PAIR) && (flags & TMOL_LJLK_DVDX)) return TMOL_LJLK_ENOTSUP;
  if (n_wave_groups < 0 || n_wave_groups > (1 << 20)) return TMOL_LJLK_EINVAL;
  {
    const int occ = (flags >> 16) & 15, grab = (flags >> 20) & 15;
    if ((occ != 0 && (occ < 3 || occ > 5)) || grab > 8 || (flags >> 24) != 0)
      return TMOL_LJLK_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_score<float>(s, bt, coords, workspace, row_capacity, flags, weights, energies,
                            dvdx, grad, n_wave_groups, st, L);
  return run_score<double>(s, bt, coords, workspace, row_capacity, flags, weights, energies,
                           dvdx, grad, n_wave_groups, st, L);
}

int64_t tmol_ljlk_noalloc_scratch_bytes(void) { return TMOL_LJLK_NOALLOC_SCRATCH_BYTES; }

int tmol_ljlk_score_noalloc(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                            const void* coords, int32_t dtype, double max_dis,
                            const int32_t* tiles, int64_t n_tiles, int64_t row_capacity,
                            int32_t flags, const void* weights, void* energies, void* dvdx,
                            void* grad, int32_t* dense_neighbors, void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!coords || (n_tiles > 0 && !tiles) || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  // a row holds the neighbors g2 >= g1 of one pose: with a capacity of the
  // longest pose the work list cannot overflow, and no host read-back of the
  // overflow counter is needed
  if (row_capacity < s->max_n_blocks) return TMOL_LJLK_EINVAL;
  if (flags & (TMOL_LJLK_GENERAL_KERNEL | (0xff << 16))) return TMOL_LJLK_EINVAL;
  const int64_t need = ws_layout(s, bt, row_capacity, dtype).total;
  if (need > TMOL_LJLK_NOALLOC_SCRATCH_BYTES) return TMOL_LJLK_ESCRATCH;
  void* ws = nullptr;
  v = noalloc_scratch(&ws);
  if (v) return v;
  v = tmol_ljlk_neighbors(s, bt, coords, dtype, max_dis, tiles, n_tiles, ws, need, row_capacity,
                          dense_neighbors, stream);
  if (v) return v;
  return tmol_ljlk_score(s, bt, coords, dtype, ws, need, row_capacity, flags, weights, energies,
                         dvdx, grad, 0, stream);
}

int tmol_ljlk_census(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                     const void* coords, int32_t dtype, void* workspace, int64_t ws_bytes,
                     int64_t row_capacity, uint64_t* census, void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!coords || !workspace || !census || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32) return TMOL_LJLK_ENOTSUP;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, st,
                     (unsigned long long*)census, 16, 0ull);
  ScoreArgs<float> a = make_args<float>(s, bt, coords, workspace, row_capacity, L);
  a.census = (unsigned long long*)census;
  launch_score<float, M_CENSUS>(a, true, 0, st);
  return check(hipGetLastError());
}

int tmol_ljlk_backward_combine(const void* dvdx, const void* grad_scores,
                               const int32_t* atom_pose, int64_t n_atoms, int32_t n_poses,
                               int32_t dtype, void* grad, void* stream) {
  if (!dvdx || !grad_scores || !atom_pose || !grad || n_atoms < 0 || n_poses <= 0)
    return TMOL_LJLK_EINVAL;
  int64_t n = n_atoms * 3;
  if (n == 0) return TMOL_LJLK_OK;
  unsigned nb = (unsigned)((n + 255) / 256);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    hipLaunchKernelGGL(combine_kernel<float>, dim3(nb), dim3(256), 0, st, (const float*)dvdx,
                       (const float*)grad_scores, atom_pose, n_atoms, n_poses, (float*)grad);
  else if (dtype == TMOL_LJLK_F64)
    hipLaunchKernelGGL(combine_kernel<double>, dim3(nb), dim3(256), 0, st, (const double*)dvdx,
                       (const double*)grad_scores, atom_pose, n_atoms, n_poses, (double*)grad);
  else
    return TMOL_LJLK_EINVAL;
  return check(hipGetLastError());
}

int tmol_ljlk_read_counters(const void* workspace, int64_t* out, void* stream) {
  if (!workspace || !out) return TMOL_LJLK_EINVAL;
  uint32_t c[2];
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(c, workspace, sizeof(c), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return TMOL_LJLK_EHIP;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return TMOL_LJLK_EHIP;
  out[0] = c[0];
  out[1] = c[1];
  return TMOL_LJLK_OK;
}

int tmol_ljlk_debug_pair_separations(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                                     const void* workspace, int64_t row_capacity, int32_t dtype,
                                     const int32_t* items, int64_t n_items, uint8_t* out,
                                     void* stream) {
  int v = validate(s, bt, dtype);
  if (v) return v;
  if (!workspace || !items || !out || n_items < 0) return TMOL_LJLK_EINVAL;
  if (n_items == 0) return TMOL_LJLK_OK;
  if (n_items >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  const uint8_t* conn_pd = (const uint8_t*)((const char*)workspace + L.conn_pd);
  hipLaunchKernelGGL(debug_sep_kernel, dim3((unsigned)n_items), dim3(64), 0, (hipStream_t)stream,
                     make_stack_view(s, bt->max_n_conn), make_bt_view(bt), conn_pd, items,
                     n_items, out);
  return check(hipGetLastError());
}

int tmol_ljlk_device_cu_count(int32_t* cu) {
  if (!cu) return TMOL_LJLK_EINVAL;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TMOL_LJLK_EHIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return TMOL_LJLK_EHIP;
  *cu = prop.multiProcessorCount;
  return TMOL_LJLK_OK;
}


// ------------------------------------------------------------------ pruner

int tmol_ljlk_prune_block_torsions(int32_t n_blocks, const int32_t* torsion_atoms,
                                   const void* coords, int32_t dtype, float* sincos,
                                   void* stream) {
  if (n_blocks < 0 || !torsion_atoms || !coords || !sincos) return TMOL_LJLK_EINVAL;
  if (n_blocks == 0) return TMOL_LJLK_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    hipLaunchKernelGGL(prune_block_torsions_kernel<float>, dim3(grid_for(n_blocks)), dim3(256), 0,
                       st, n_blocks, torsion_atoms, (const float*)coords, sincos);
  else if (dtype == TMOL_LJLK_F64)
    hipLaunchKernelGGL(prune_block_torsions_kernel<double>, dim3(grid_for(n_blocks)), dim3(256), 0,
                       st, n_blocks, torsion_atoms, (const double*)coords, sincos);
  else
    return TMOL_LJLK_EINVAL;
  return check(hipGetLastError());
}

static PruneBlocks prune_blocks(const tmol_ljlk_prune_blocks* b) {
  PruneBlocks p;
  p.sphere = b->sphere;
  p.sincos = b->sincos;
  p.emb_row = b->emb_row;
  p.seqpos = b->seqpos;
  p.emb = b->emb;
  return p;
}

int tmol_ljlk_prune_pair_features(int64_t n_pairs, const int32_t* pairs,
                                  const tmol_ljlk_prune_blocks* blocks, uint16_t* features,
                                  void* stream) {
  if (n_pairs < 0 || !pairs || !blocks || !features) return TMOL_LJLK_EINVAL;
  if (n_pairs == 0) return TMOL_LJLK_OK;
  const int64_t n = n_pairs * PRUNE_K1;
  hipLaunchKernelGGL(prune_pair_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_pairs, pairs, prune_blocks(blocks), features);
  return check(hipGetLastError());
}

int tmol_ljlk_prune_gemm(int32_t epilogue, int32_t a_km, int32_t b_kn, const uint16_t* A,
                         const uint16_t* B, int32_t M, int32_t N, int32_t K, int32_t lda,
                         int32_t ldb, int32_t ldc, int32_t split_k, float* C, const float* bias,
                         float* Z, uint16_t* H, float* D, void* stream) {
  if (!A || !B || M <= 0 || N <= 0 || K <= 0 || split_k <= 0) return TMOL_LJLK_EINVAL;
  if (epilogue == EPI_F32 && !C) return TMOL_LJLK_EINVAL;
  if (epilogue == EPI_BIAS_SILU && (!bias || !Z || !H)) return TMOL_LJLK_EINVAL;
  if (epilogue == EPI_DSILU && (!Z || !H || !D)) return TMOL_LJLK_EINVAL;
  if (epilogue != EPI_F32 && split_k != 1) return TMOL_LJLK_EINVAL;
  GemmArgs g{A, B, M, N, K, lda, ldb, ldc, 0, C, bias, Z, H, D};
  const int kc = ((K + split_k - 1) / split_k + GK - 1) / GK * GK;
  g.k_chunk = kc;
  const int nz = (K + kc - 1) / kc;
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, nz);
  hipStream_t st = (hipStream_t)stream;
#define GEMM_CASE(AK, BK, E) \
  if (a_km == AK && b_kn == BK && epilogue == E) { \
    hipLaunchKernelGGL((prune_gemm_kernel<AK, BK, E>), grid, dim3(256), 0, st, g); \
    return check(hipGetLastError()); \
  }
  GEMM_CASE(0, 0, EPI_BIAS_SILU)
  GEMM_CASE(0, 0, EPI_F32)
  GEMM_CASE(0, 1, EPI_DSILU)
  GEMM_CASE(0, 1, EPI_F32)
  GEMM_CASE(1, 1, EPI_F32)
#undef GEMM_CASE
  return TMOL_LJLK_ENOTSUP;
}

int tmol_ljlk_prune_head(int32_t n, const uint16_t* H3, const float* Z3, const float* w4,
                         const float* b4, const float* labels, const float* sample_weights,
                         float pos_weight, int32_t train, float* logits, float* loss, float* dZ3,
                         uint16_t* dZ3_bf16, float* dw4, float* db4, void* stream) {
  if (n <= 0 || !H3 || !w4 || !b4 || !logits) return TMOL_LJLK_EINVAL;
  if (train && (!Z3 || !labels || !loss || !dZ3 || !dZ3_bf16 || !dw4 || !db4))
    return TMOL_LJLK_EINVAL;
  hipLaunchKernelGGL(prune_head_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     n, H3, Z3, w4, b4, labels, sample_weights, pos_weight, train, logits, loss,
                     dZ3, dZ3_bf16, dw4, db4);
  return check(hipGetLastError());
}

int tmol_ljlk_prune_colsum(int32_t n, int32_t ncol, int32_t ld, const float* D, float* out,
                           void* stream) {
  if (n <= 0 || ncol <= 0 || ld < ncol || !D || !out) return TMOL_LJLK_EINVAL;
  const int ychunks = std::min(64, (n + 255) / 256);
  hipLaunchKernelGGL(prune_colsum_kernel, dim3(ncol, ychunks), dim3(256), 0, (hipStream_t)stream,
                     n, ncol, ld, D, out);
  return check(hipGetLastError());
}

int tmol_ljlk_prune_emb_grad(int64_t n_pairs, const int32_t* pairs, const int32_t* emb_row,
                             const float* dX, float* demb, void* stream) {
  if (n_pairs < 0 || !pairs || !emb_row || !dX || !demb) return TMOL_LJLK_EINVAL;
  if (n_pairs == 0) return TMOL_LJLK_OK;
  const int64_t n = n_pairs * 2 * PRUNE_EMB;
  hipLaunchKernelGGL(prune_emb_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_pairs, pairs, emb_row, dX, demb);
  return check(hipGetLastError());
}


int64_t tmol_ljlk_prune_scratch_bytes(int32_t n_blocks) {
  if (n_blocks < 0) return -1;
  return 8 * (int64_t(n_blocks + 1023) / 1024 + 1);
}

int tmol_ljlk_prune_flatten(int32_t n_blocks, const uint32_t* row_count, const int32_t* row_items,
                            int64_t row_capacity, int64_t* scratch, int64_t* row_offset,
                            int32_t* pairs, void* stream) {
  if (n_blocks <= 0 || !row_count || !row_items || row_capacity <= 0 || !scratch || !row_offset ||
      !pairs)
    return TMOL_LJLK_EINVAL;
  const int64_t nb = (int64_t(n_blocks) + 1023) / 1024;
  if (nb > (int64_t(1) << 20)) return TMOL_LJLK_ENOTSUP;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(prune_count_kernel, dim3((unsigned)nb), dim3(1024), 0, st, int64_t(n_blocks),
                     row_count, row_capacity, scratch);
  hipLaunchKernelGGL(prune_scan_blocks_kernel, dim3(1), dim3(1024), 0, st, nb, scratch);
  hipLaunchKernelGGL(prune_flatten_kernel, dim3((unsigned)nb), dim3(1024), 0, st, int64_t(n_blocks),
                     row_count, row_items, row_capacity, (const int64_t*)scratch, row_offset, pairs);
  return check(hipGetLastError());
}

int tmol_ljlk_prune_infer(const int64_t* n_pairs, const int32_t* pairs,
                          const tmol_ljlk_prune_blocks* blocks, const uint16_t* W1,
                          const uint16_t* W2, const uint16_t* W3, const float* b1, const float* b2,
                          const float* b3, const float* w4, const float* b4, float logit_cut,
                          float min_gap, uint8_t* keep, float* logits, int32_t n_workgroups,
                          void* stream) {
  if (!n_pairs || !pairs || !blocks || !W1 || !W2 || !W3 || !b1 || !b2 || !b3 || !w4 || !b4 || !keep)
    return TMOL_LJLK_EINVAL;
  PruneNet net{W1, W2, W3, b1, b2, b3, w4, b4};
  int nwg = n_workgroups;
  if (nwg <= 0) nwg = resident_workgroups(prune_infer_kernel);
  hipLaunchKernelGGL(prune_infer_kernel, dim3(nwg), dim3(PI_WAVES * 64), 0, (hipStream_t)stream,
                     n_pairs, pairs, prune_blocks(blocks), net, logit_cut, min_gap, keep, logits);
  return check(hipGetLastError());
}

int tmol_ljlk_prune_compact(int32_t n_blocks, uint32_t* row_count, int32_t* row_items,
                            int64_t row_capacity, const int64_t* row_offset, const uint8_t* keep,
                            void* stream) {
  if (n_blocks <= 0 || !row_count || !row_items || !row_offset || !keep) return TMOL_LJLK_EINVAL;
  hipLaunchKernelGGL(prune_compact_kernel, dim3(grid_for(n_blocks)), dim3(256), 0,
                     (hipStream_t)stream, int64_t(n_blocks), row_count, row_items, row_capacity,
                     row_offset, keep);  // one wave per 64 rows: n_blocks / 64 waves
  return check(hipGetLastError());
}

int64_t tmol_ljlk_rotamer_workspace_bytes(const tmol_ljlk_rotamer_set* rs,
                                          const tmol_ljlk_block_types* bt, int32_t dtype) {
  if (validate_rot(rs, bt, dtype)) return -1;
  return rot_ws_layout(rs, bt, dtype).total;
}

int tmol_ljlk_rotamer_neighbors(const tmol_ljlk_rotamer_set* rs,
                                const tmol_ljlk_block_types* bt, const void* coords,
                                int32_t dtype, double max_dis, void* workspace,
                                int64_t ws_bytes, void* stream) {
  int v = validate_rot(rs, bt, dtype);
  if (v) return v;
  if (!coords || !workspace) return TMOL_LJLK_EINVAL;
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_rot_neighbors<float>(rs, bt, coords, max_dis, workspace, st, L);
  return run_rot_neighbors<double>(rs, bt, coords, max_dis, workspace, st, L);
}

int tmol_ljlk_rotamer_totals(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                             int32_t dtype, const void* workspace, int64_t* totals_host,
                             void* stream) {
  int v = validate_rot(rs, bt, dtype);
  if (v) return v;
  if (!workspace || !totals_host) return TMOL_LJLK_EINVAL;
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  const int64_t G = int64_t(rs->n_poses) * rs->max_n_blocks;
  hipStream_t st = (hipStream_t)stream;
  const char* w = (const char*)workspace;
  hipError_t e = hipMemcpyAsync(&totals_host[0], w + L.pair_off + 8 * G, 8,
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&totals_host[1], w + L.unit_off + 8 * G, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return check(e);
}

int tmol_ljlk_rotamer_dispatch(const tmol_ljlk_rotamer_set* rs,
                               const tmol_ljlk_block_types* bt, int32_t dtype, double max_dis,
                               void* workspace, int64_t ws_bytes, int32_t* indices,
                               int64_t n_pairs, void* units, int64_t n_units, void* stream) {
  int v = validate_rot(rs, bt, dtype);
  if (v) return v;
  if (!workspace || n_pairs < 0 || n_units < 0) return TMOL_LJLK_EINVAL;
  if ((n_pairs > 0 && !indices) || (n_units > 0 && !units)) return TMOL_LJLK_EINVAL;
  if (n_pairs >= (int64_t(1) << 31)) return TMOL_LJLK_EOVERFLOW;  // int32 dispatch indices
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  const int64_t G = int64_t(rs->n_poses) * rs->max_n_blocks;
  hipStream_t st = (hipStream_t)stream;
  char* w = (char*)workspace;
  RotView rv = make_rot_view(rs, bt->max_n_conn);
  const dim3 grid((unsigned)((G + WAVES_PER_WG - 1) / WAVES_PER_WG));
  const int64_t* po = (const int64_t*)(w + L.pair_off);
  const int64_t* uo = (const int64_t*)(w + L.unit_off);
  if (dtype == TMOL_LJLK_F32)
    hipLaunchKernelGGL(rot_rows_fill_kernel<float>, grid, dim3(256), 0, st, rv,
                       (const float*)(w + L.block_spheres), (float)max_dis, po, uo, indices,
                       n_pairs, (RotUnit*)units);
  else
    hipLaunchKernelGGL(rot_rows_fill_kernel<double>, grid, dim3(256), 0, st, rv,
                       (const double*)(w + L.block_spheres), max_dis, po, uo, indices, n_pairs,
                       (RotUnit*)units);
  return check(hipGetLastError());
}

int tmol_ljlk_rotamer_score(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                            const void* coords, int32_t dtype, void* workspace,
                            int64_t ws_bytes, const void* units, int64_t n_units,
                            int64_t n_pairs, int32_t flags, const void* dTdV, void* energies,
                            void* grad, int32_t n_workgroups, void* stream) {
  int v = validate_rot(rs, bt, dtype);
  if (v) return v;
  if (!coords || !workspace || n_units < 0 || n_pairs < 0 || (n_units > 0 && !units))
    return TMOL_LJLK_EINVAL;
  if (n_workgroups < 0 || n_workgroups > (1 << 20)) return TMOL_LJLK_EINVAL;
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  const int mode = flags & ~TMOL_LJLK_ACCUMULATE;
  const bool bwd = mode == (TMOL_LJLK_WGRAD | TMOL_LJLK_NO_ENERGY);
  if (mode != 0 && !bwd) return TMOL_LJLK_ENOTSUP;
  if (n_pairs >= (int64_t(1) << 31)) return TMOL_LJLK_EOVERFLOW;
  if (!bwd && n_pairs > 0 && !energies) return TMOL_LJLK_EINVAL;
  if (bwd && (!grad || (n_pairs > 0 && !dTdV))) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const int64_t NA3 = rs->n_atoms_total * 3;
  if (dtype == TMOL_LJLK_F32) {
    RotArgs<float> a = make_rot_args<float>(rs, bt, coords, workspace, L);
    a.units = (const RotUnit*)units;
    a.n_units = n_units;
    a.n_pairs = n_pairs;
    a.energies = (float*)energies;
    a.dTdV = (const float*)dTdV;
    a.grad = (float*)grad;
    if (bwd) {
      if (!(flags & TMOL_LJLK_ACCUMULATE)) fill<float>((float*)grad, NA3, st);
      if (n_units > 0) launch_rot_score<float, M_WGRAD | M_NOE>(a, n_workgroups, st);
    } else {
      if (n_units > 0) launch_rot_score<float, 0>(a, n_workgroups, st);
    }
  } else {
    RotArgs<double> a = make_rot_args<double>(rs, bt, coords, workspace, L);
    a.units = (const RotUnit*)units;
    a.n_units = n_units;
    a.n_pairs = n_pairs;
    a.energies = (double*)energies;
    a.dTdV = (const double*)dTdV;
    a.grad = (double*)grad;
    if (bwd) {
      if (!(flags & TMOL_LJLK_ACCUMULATE)) fill<double>((double*)grad, NA3, st);
      if (n_units > 0) launch_rot_score<double, M_WGRAD | M_NOE>(a, n_workgroups, st);
    } else {
      if (n_units > 0) launch_rot_score<double, 0>(a, n_workgroups, st);
    }
  }
  return check(hipGetLastError());
}

int tmol_ljlk_rotamer_census(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                             const void* coords, int32_t dtype, void* workspace,
                             int64_t ws_bytes, const void* units, int64_t n_units,
                             int64_t n_pairs, uint64_t* census, void* stream) {
  int v = validate_rot(rs, bt, dtype);
  if (v) return v;
  if (!coords || !workspace || !census || n_units < 0 || (n_units > 0 && !units))
    return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32) return TMOL_LJLK_ENOTSUP;
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, st,
                     (unsigned long long*)census, int64_t(16), 0ull);
  RotArgs<float> a = make_rot_args<float>(rs, bt, coords, workspace, L);
  a.units = (const RotUnit*)units;
  a.n_units = n_units;
  a.n_pairs = n_pairs;
  a.census = (unsigned long long*)census;
  if (n_units > 0) launch_rot_score<float, M_CENSUS>(a, 0, st);
  return check(hipGetLastError());
}

int tmol_ljlk_rotamer_spheres(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                              int32_t dtype, void* workspace, void** rot_spheres,
                              void** block_spheres) {
  int v = validate_rot(rs, bt, dtype);
  if (v) return v;
  if (!workspace || !rot_spheres || !block_spheres) return TMOL_LJLK_EINVAL;
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  *rot_spheres = (char*)workspace + L.rot_spheres;
  *block_spheres = (char*)workspace + L.block_spheres;
  return TMOL_LJLK_OK;
}

int tmol_ljlk_inter_block_bondsep(const tmol_ljlk_block_types* bt, int32_t n_poses,
                                  int32_t max_n_blocks, const int32_t* block_type,
                                  const int32_t* inter_residue_connections,
                                  int32_t* inter_block_bondsep, int32_t* min_bondsep,
                                  uint32_t* overflow, void* stream) {
  if (!bt || !block_type || !inter_residue_connections || !inter_block_bondsep)
    return TMOL_LJLK_EINVAL;
  if (n_poses <= 0 || max_n_blocks <= 0) return TMOL_LJLK_EINVAL;
  if (bt->max_n_conn <= 0 || bt->max_n_conn > MAX_CONN) return TMOL_LJLK_ENOTSUP;
  if (int64_t(max_n_blocks) >= (int64_t(1) << 27)) return TMOL_LJLK_ENOTSUP;
  TopoView v{n_poses, max_n_blocks, bt->max_n_conn, block_type, inter_residue_connections};
  BTView b = make_bt_view(bt);
  hipStream_t st = (hipStream_t)stream;
  const int64_t PB = int64_t(n_poses) * max_n_blocks;
  const int64_t n_ibs = PB * max_n_blocks * v.C * v.C;
  hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(grid_for(n_ibs)), dim3(256), 0, st,
                     inter_block_bondsep, n_ibs, int32_t(BS_CUT));
  if (overflow)
    hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, st, overflow, int64_t(1), 0u);
  hipLaunchKernelGGL(bondsep_rows_kernel, dim3(grid_for(PB * v.C)), dim3(256), 0, st, v, b,
                     inter_block_bondsep, (int32_t*)overflow);
  if (min_bondsep)
    hipLaunchKernelGGL(bondsep_min_kernel, dim3(grid_for(PB * max_n_blocks)), dim3(256), 0, st, v,
                       (const int32_t*)inter_block_bondsep, min_bondsep);
  return check(hipGetLastError());
}

int tmol_elec_neighbors(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                        const tmol_elec_tables* e, const void* coords, int32_t dtype,
                        double max_dis, const int32_t* tiles, int64_t n_tiles, void* workspace,
                        int64_t ws_bytes, int64_t row_capacity, int32_t* dense, void* stream) {
  int v = validate_elec(s, bt, e, dtype);
  if (v) return v;
  if (!coords || !workspace || (n_tiles > 0 && !tiles) || row_capacity <= 0)
    return TMOL_LJLK_EINVAL;
  if (!std::isfinite(max_dis) || max_dis < 0) return TMOL_LJLK_EINVAL;
  if (n_tiles >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  if (dense && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_elec_neighbors<float>(s, bt, e, coords, max_dis, tiles, n_tiles, workspace,
                                     row_capacity, dense, st, L);
  return run_elec_neighbors<double>(s, bt, e, coords, max_dis, tiles, n_tiles, workspace,
                                    row_capacity, dense, st, L);
}

int tmol_elec_score(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                    const tmol_elec_tables* e, const void* coords, int32_t dtype,
                    void* workspace, int64_t ws_bytes, int64_t row_capacity, int32_t flags,
                    const void* weights, void* energies, void* dvdx, void* grad,
                    int32_t n_wave_groups, void* stream) {
  int v = validate_elec(s, bt, e, dtype);
  if (v) return v;
  if (!coords || !workspace || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  if (flags & ~(TMOL_LJLK_DVDX | TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR | TMOL_LJLK_NO_ENERGY |
                TMOL_LJLK_ACCUMULATE))
    return TMOL_LJLK_EINVAL;
  if (!(flags & TMOL_LJLK_NO_ENERGY) && !energies) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_DVDX) && !dvdx) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_WGRAD) && (!grad || !weights)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_DVDX) && (flags & TMOL_LJLK_WGRAD)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_BLOCK_PAIR) && (flags & TMOL_LJLK_DVDX)) return TMOL_LJLK_ENOTSUP;
  if (n_wave_groups < 0 || n_wave_groups > (1 << 20)) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_elec_score<float>(s, bt, e, coords, workspace, row_capacity, flags, weights,
                                 energies, dvdx, grad, n_wave_groups, st, L);
  return run_elec_score<double>(s, bt, e, coords, workspace, row_capacity, flags, weights,
                                energies, dvdx, grad, n_wave_groups, st, L);
}

int tmol_ljlk_elec_prepare(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                           const tmol_elec_tables* e, void* workspace, int64_t ws_bytes,
                           int64_t row_capacity, int32_t dtype, void* stream) {
  int v = validate_elec(s, bt, e, dtype);
  if (v) return v;
  if (!workspace || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  BTView b = make_bt_view(bt);
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    hipLaunchKernelGGL(elec_prepare_kernel<float>, dim3(b.T), dim3(64), 0, st, b,
                       (const float*)e->partial_charge, e->inter_repr_path_distance,
                       e->intra_repr_path_distance, (uint32_t*)(w + L.e_pd),
                       (int32_t*)(w + L.e_intra), (uint16_t*)(w + L.e_intra + 4 * int64_t(b.T)));
  else
    hipLaunchKernelGGL(elec_prepare_kernel<double>, dim3(b.T), dim3(64), 0, st, b,
                       (const double*)e->partial_charge, e->inter_repr_path_distance,
                       e->intra_repr_path_distance, (uint32_t*)(w + L.e_pd),
                       (int32_t*)(w + L.e_intra), (uint16_t*)(w + L.e_intra + 4 * int64_t(b.T)));
  return check(hipGetLastError());
}

int tmol_ljlk_elec_score(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                         const tmol_elec_tables* e, const void* coords, int32_t dtype,
                         void* workspace, int64_t ws_bytes, int64_t row_capacity, int32_t flags,
                         const void* weights, void* energies, void* grad, int32_t n_wave_groups,
                         void* stream) {
  int v = validate_elec(s, bt, e, dtype);
  if (v) return v;
  if (!coords || !workspace || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  if (flags & ~(TMOL_LJLK_WGRAD | TMOL_LJLK_NO_ENERGY | TMOL_LJLK_ACCUMULATE))
    return TMOL_LJLK_ENOTSUP;
  if (!(flags & TMOL_LJLK_NO_ENERGY) && !energies) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_NO_ENERGY) && !(flags & TMOL_LJLK_WGRAD)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_WGRAD) && (!grad || !weights)) return TMOL_LJLK_EINVAL;
  if (!bt->type_params || !bt->global_params || !bt->path_distance) return TMOL_LJLK_EINVAL;
  if (n_wave_groups < 0 || n_wave_groups > (1 << 20)) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32) {
    const ElecConst<float> k = elec_constants<float>(e->global_params);
    return run_score<float>(s, bt, coords, workspace, row_capacity, flags, weights, energies,
                            nullptr, grad, n_wave_groups, st, L, e, &k);
  }
  const ElecConst<double> k = elec_constants<double>(e->global_params);
  return run_score<double>(s, bt, coords, workspace, row_capacity, flags, weights, energies,
                           nullptr, grad, n_wave_groups, st, L, e, &k);
}

int tmol_elec_census(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                     const tmol_elec_tables* e, const void* coords, int32_t dtype,
                     void* workspace, int64_t ws_bytes, int64_t row_capacity, uint64_t* census,
                     void* stream) {
  int v = validate_elec(s, bt, e, dtype);
  if (v) return v;
  if (!coords || !workspace || !census || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32) return TMOL_LJLK_ENOTSUP;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  return run_elec_score<float>(s, bt, e, coords, workspace, row_capacity, 0, nullptr, nullptr,
                               nullptr, nullptr, 0, (hipStream_t)stream, L,
                               (unsigned long long*)census);
}

int tmol_elec_rotamer_neighbors(const tmol_ljlk_rotamer_set* rs,
                                const tmol_ljlk_block_types* bt, const tmol_elec_tables* e,
                                const void* coords, int32_t dtype, double max_dis,
                                void* workspace, int64_t ws_bytes, void* stream) {
  int v = validate_elec_rot(rs, bt, e, dtype);
  if (v) return v;
  if (!coords || !workspace) return TMOL_LJLK_EINVAL;
  if (!std::isfinite(max_dis) || max_dis < 0) return TMOL_LJLK_EINVAL;
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_rot_neighbors<float>(rs, bt, coords, max_dis, workspace, st, L, e);
  return run_rot_neighbors<double>(rs, bt, coords, max_dis, workspace, st, L, e);
}

int tmol_elec_rotamer_score(const tmol_ljlk_rotamer_set* rs, const tmol_ljlk_block_types* bt,
                            const tmol_elec_tables* e, const void* coords, int32_t dtype,
                            void* workspace, int64_t ws_bytes, const void* units,
                            int64_t n_units, int64_t n_pairs, int32_t flags, const void* dTdV,
                            void* energies, void* grad, int32_t n_workgroups, void* stream) {
  int v = validate_elec_rot(rs, bt, e, dtype);
  if (v) return v;
  if (!coords || !workspace || n_units < 0 || n_pairs < 0 || (n_units > 0 && !units))
    return TMOL_LJLK_EINVAL;
  if (n_workgroups < 0 || n_workgroups > (1 << 20)) return TMOL_LJLK_EINVAL;
  RotWsLayout L = rot_ws_layout(rs, bt, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  const int mode = flags & ~TMOL_LJLK_ACCUMULATE;
  const bool bwd = mode == (TMOL_LJLK_WGRAD | TMOL_LJLK_NO_ENERGY);
  if (mode != 0 && !bwd) return TMOL_LJLK_ENOTSUP;
  if (n_pairs >= (int64_t(1) << 31)) return TMOL_LJLK_EOVERFLOW;
  if (!bwd && n_pairs > 0 && !energies) return TMOL_LJLK_EINVAL;
  if (bwd && (!grad || (n_pairs > 0 && !dTdV))) return TMOL_LJLK_EINVAL;
  const bool acc = (flags & TMOL_LJLK_ACCUMULATE) != 0;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32) {
    if (bwd)
      return run_elec_rot_score<float, true>(rs, bt, e, coords, workspace, L, units, n_units,
                                             n_pairs, dTdV, energies, grad, acc, n_workgroups,
                                             st);
    return run_elec_rot_score<float, false>(rs, bt, e, coords, workspace, L, units, n_units,
                                            n_pairs, dTdV, energies, grad, acc, n_workgroups, st);
  }
  if (bwd)
    return run_elec_rot_score<double, true>(rs, bt, e, coords, workspace, L, units, n_units,
                                            n_pairs, dTdV, energies, grad, acc, n_workgroups, st);
  return run_elec_rot_score<double, false>(rs, bt, e, coords, workspace, L, units, n_units,
                                           n_pairs, dTdV, energies, grad, acc, n_workgroups, st);
}

// ------------------------------------------------------------------ L-BFGS history products

int64_t tmol_lkball_workspace_bytes(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                                    const tmol_lkball_tables* k, int64_t row_capacity,
                                    int32_t dtype) {
  if (!s || !bt || !k || row_capacity < 0 || k->max_n_occ < 0) return -1;
  return lkb_ws_layout(ws_layout(s, bt, row_capacity, dtype), s->n_blocks_total, k->max_n_occ)
      .total;
}

int tmol_lkball_waters(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                       const tmol_lkball_tables* k, const void* coords, int32_t dtype,
                       void* waters, void* stream) {
  int v = validate_lkb(s, bt, k, dtype, false);
  if (v) return v;
  if (!coords || !waters) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_lkb_waters<float>(s, bt, k, coords, waters, false, nullptr, nullptr, st);
  return run_lkb_waters<double>(s, bt, k, coords, waters, false, nullptr, nullptr, st);
}

int tmol_lkball_waters_backward(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                                const tmol_lkball_tables* k, const void* coords, int32_t dtype,
                                const void* dwaters, void* dxyz, void* stream) {
  int v = validate_lkb(s, bt, k, dtype, false);
  if (v) return v;
  if (!coords || !dwaters || !dxyz) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_lkb_waters<float>(s, bt, k, coords, nullptr, true, dwaters, dxyz, st);
  return run_lkb_waters<double>(s, bt, k, coords, nullptr, true, dwaters, dxyz, st);
}

int tmol_lkball_neighbors(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                          const void* coords, int32_t dtype, double max_dis,
                          const int32_t* tiles, int64_t n_tiles, void* workspace,
                          int64_t ws_bytes, int64_t row_capacity, int32_t* dense, void* stream) {
  if (!s || !bt) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (s->n_poses <= 0 || s->n_blocks_total <= 0 || !bt->n_atoms) return TMOL_LJLK_EINVAL;
  if (!coords || !workspace || (n_tiles > 0 && !tiles) || row_capacity <= 0)
    return TMOL_LJLK_EINVAL;
  if (!std::isfinite(max_dis) || max_dis < 0) return TMOL_LJLK_EINVAL;
  if (n_tiles >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  if (dense && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_EINVAL;
  if (s->layout == TMOL_LJLK_RAGGED && !s->pose_block_start) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_lkb_neighbors<float>(s, bt, coords, max_dis, tiles, n_tiles, workspace,
                                    row_capacity, dense, st, L);
  return run_lkb_neighbors<double>(s, bt, coords, max_dis, tiles, n_tiles, workspace,
                                   row_capacity, dense, st, L);
}

int tmol_lkball_score(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                      const tmol_lkball_tables* k, const void* coords, const void* waters,
                      int32_t dtype, void* workspace, int64_t ws_bytes, int64_t row_capacity,
                      int32_t flags, const void* weights, void* energies, void* dxyz,
                      void* dwaters, int32_t n_wave_groups, void* stream) {
  int v = validate_lkb(s, bt, k, dtype, true);
  if (v) return v;
  if (!coords || !waters || !workspace || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < lkb_ws_layout(L, s->n_blocks_total, k->max_n_occ).total) return TMOL_LJLK_EINVAL;
  if (flags & ~(TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR | TMOL_LJLK_NO_ENERGY |
                TMOL_LJLK_ACCUMULATE))
    return TMOL_LJLK_EINVAL;
  if (!(flags & TMOL_LJLK_NO_ENERGY) && !energies) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_NO_ENERGY) && !(flags & TMOL_LJLK_WGRAD)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_WGRAD) && (!dxyz || !dwaters || !weights)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_BLOCK_PAIR) && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_ENOTSUP;
  if (n_wave_groups < 0 || n_wave_groups > (1 << 20)) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_lkb_score<float>(s, bt, k, coords, waters, workspace, row_capacity, flags,
                                weights, energies, dxyz, dwaters, n_wave_groups, st, L);
  return run_lkb_score<double>(s, bt, k, coords, waters, workspace, row_capacity, flags,
                               weights, energies, dxyz, dwaters, n_wave_groups, st, L);
}

int tmol_lkball_census(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                       const tmol_lkball_tables* k, const void* coords, const void* waters,
                       int32_t dtype, void* workspace, int64_t ws_bytes, int64_t row_capacity,
                       unsigned long long* census, void* stream) {
  int v = validate_lkb(s, bt, k, dtype, true);
  if (v) return v;
  if (!coords || !waters || !workspace || !census || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < lkb_ws_layout(L, s->n_blocks_total, k->max_n_occ).total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_lkb_score<float>(s, bt, k, coords, waters, workspace, row_capacity, 0, nullptr,
                                nullptr, nullptr, nullptr, 0, st, L, census);
  return run_lkb_score<double>(s, bt, k, coords, waters, workspace, row_capacity, 0, nullptr,
                               nullptr, nullptr, nullptr, 0, st, L, census);
}

int tmol_hbond_bases(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                     const tmol_hbond_tables* h, const void* coords, int32_t dtype,
                     void* derived, int32_t* inds, void* stream) {
  int v = validate_hb(s, bt, h, dtype, false);
  if (v) return v;
  if (!coords || !derived || !inds) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32) return run_hb_bases<float>(s, bt, h, coords, derived, inds, st);
  return run_hb_bases<double>(s, bt, h, coords, derived, inds, st);
}

int tmol_hbond_neighbors(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                         const void* coords, int32_t dtype, double max_dis,
                         const int32_t* tiles, int64_t n_tiles, void* workspace,
                         int64_t ws_bytes, int64_t row_capacity, int32_t* dense, void* stream) {
  if (!s || !bt) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (s->n_poses <= 0 || s->n_blocks_total <= 0 || !bt->n_atoms) return TMOL_LJLK_EINVAL;
  if (!coords || !workspace || (n_tiles > 0 && !tiles) || row_capacity <= 0)
    return TMOL_LJLK_EINVAL;
  if (!std::isfinite(max_dis) || max_dis < 0) return TMOL_LJLK_EINVAL;
  if (n_tiles >= (int64_t(1) << 31)) return TMOL_LJLK_ENOTSUP;
  if (dense && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_lkb_neighbors<float, false>(s, bt, coords, max_dis, tiles, n_tiles, workspace,
                                           row_capacity, dense, st, L);
  return run_lkb_neighbors<double, false>(s, bt, coords, max_dis, tiles, n_tiles, workspace,
                                          row_capacity, dense, st, L);
}

int tmol_hbond_score(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                     const tmol_hbond_tables* h, const void* coords, const void* derived,
                     const int32_t* dinds, int32_t dtype, void* workspace, int64_t ws_bytes,
                     int64_t row_capacity, int32_t flags, const void* weights, void* energies,
                     void* dxyz, int32_t n_wave_groups, void* stream) {
  int v = validate_hb(s, bt, h, dtype, true);
  if (v) return v;
  if (!coords || !derived || !dinds || !workspace || row_capacity <= 0) return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  if (flags & ~(TMOL_LJLK_WGRAD | TMOL_LJLK_BLOCK_PAIR | TMOL_LJLK_NO_ENERGY |
                TMOL_LJLK_ACCUMULATE))
    return TMOL_LJLK_EINVAL;
  if (!(flags & TMOL_LJLK_NO_ENERGY) && !energies) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_NO_ENERGY) && !(flags & TMOL_LJLK_WGRAD)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_WGRAD) && (!dxyz || !weights)) return TMOL_LJLK_EINVAL;
  if ((flags & TMOL_LJLK_BLOCK_PAIR) && s->layout != TMOL_LJLK_PADDED) return TMOL_LJLK_ENOTSUP;
  if (n_wave_groups < 0 || n_wave_groups > (1 << 20)) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_hb_score<float>(s, bt, h, coords, derived, dinds, workspace, row_capacity, flags,
                               weights, energies, dxyz, n_wave_groups, st, L);
  return run_hb_score<double>(s, bt, h, coords, derived, dinds, workspace, row_capacity, flags,
                              weights, energies, dxyz, n_wave_groups, st, L);
}

int tmol_hbond_census(const tmol_ljlk_stack* s, const tmol_ljlk_block_types* bt,
                      const tmol_hbond_tables* h, const void* coords, const void* derived,
                      const int32_t* dinds, int32_t dtype, void* workspace, int64_t ws_bytes,
                      int64_t row_capacity, unsigned long long* census, void* stream) {
  int v = validate_hb(s, bt, h, dtype, true);
  if (v) return v;
  if (!coords || !derived || !dinds || !workspace || !census || row_capacity <= 0)
    return TMOL_LJLK_EINVAL;
  WsLayout L = ws_layout(s, bt, row_capacity, dtype);
  if (ws_bytes < L.total) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TMOL_LJLK_F32)
    return run_hb_score<float>(s, bt, h, coords, derived, dinds, workspace, row_capacity, 0,
                               nullptr, nullptr, nullptr, 0, st, L, census);
  return run_hb_score<double>(s, bt, h, coords, derived, dinds, workspace, row_capacity, 0,
                              nullptr, nullptr, nullptr, 0, st, L, census);
}

static int leaf_args(const tmol_leaf_atoms_args* x, int dtype, LeafArgs& a) {
  if (!x) return TMOL_LJLK_EINVAL;
  if (dtype != TMOL_LJLK_F32 && dtype != TMOL_LJLK_F64) return TMOL_LJLK_EINVAL;
  if (x->n_poses <= 0 || x->max_n_blocks <= 0 || x->max_n_block_atoms <= 0 ||
      x->max_n_pose_atoms <= 0 || x->max_n_conn < 0)
    return TMOL_LJLK_EINVAL;
  if (!x->block_atom_missing || !x->pose_atom_missing || !x->block_coord_offset ||
      !x->block_type || (x->max_n_conn > 0 && !x->inter_block_connections) ||
      !x->bt_n_atoms || !x->bt_atom_downstream_of_conn || !x->bt_atom_ancestors ||
      !x->bt_atom_icoors || !x->bt_atom_ancestors_backup || !x->bt_atom_icoors_backup)
    return TMOL_LJLK_EINVAL;
  if (int64_t(x->n_poses) * x->max_n_pose_atoms >= (int64_t(1) << 31) / 3 ||
      int64_t(x->n_poses) * x->max_n_blocks * x->max_n_block_atoms >= (int64_t(1) << 31) ||
      x->max_n_block_atoms > LEAF_MAX_ATOMS)
    return TMOL_LJLK_ENOTSUP;
  a.P = x->n_poses;
  a.NA = x->max_n_pose_atoms;
  a.B = x->max_n_blocks;
  a.Amax = x->max_n_block_atoms;
  a.C = x->max_n_conn;
  a.block_missing = x->block_atom_missing;
  a.pose_missing = x->pose_atom_missing;
  a.block_offset = x->block_coord_offset;
  a.block_type = x->block_type;
  a.irc = x->inter_block_connections;
  a.n_atoms = x->bt_n_atoms;
  a.downstream = x->bt_atom_downstream_of_conn;
  a.anc = x->bt_atom_ancestors;
  a.anc_backup = x->bt_atom_ancestors_backup;
  a.geom = x->bt_atom_icoors;
  a.geom_backup = x->bt_atom_icoors_backup;
  return TMOL_LJLK_OK;
}

// one-wave workgroups over the (pose, block) grid, LEAF_BPW blocks per wave
static unsigned leaf_grid(const LeafArgs& a) {
  const int64_t waves = (int64_t(a.P) * a.B + LEAF_BPW - 1) / LEAF_BPW;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(waves, 65536));
}

int tmol_leaf_atoms_build(const tmol_leaf_atoms_args* x, const void* orig, int32_t dtype,
                          void* out, void* stream) {
  LeafArgs a;
  int v = leaf_args(x, dtype, a);
  if (v) return v;
  if (!orig || !out) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  // atoms of the pose layout beyond any block stay zero, as in the reference
  if (dtype == TMOL_LJLK_F32) {
    fill<float>((float*)out, 3 * int64_t(a.P) * a.NA, st);
    hipLaunchKernelGGL(leaf_atoms_kernel<float>, dim3(leaf_grid(a)), dim3(LEAF_WAVE), 0, st, a,
                       (const float*)orig, (float*)out);
  } else {
    fill<double>((double*)out, 3 * int64_t(a.P) * a.NA, st);
    hipLaunchKernelGGL(leaf_atoms_kernel<double>, dim3(leaf_grid(a)), dim3(LEAF_WAVE), 0, st,
                       a, (const double*)orig, (double*)out);
  }
  return check(hipGetLastError());
}

int tmol_leaf_atoms_backward(const tmol_leaf_atoms_args* x, const void* orig, int32_t dtype,
                             const void* dnew, void* dorig, void* stream) {
  LeafArgs a;
  int v = leaf_args(x, dtype, a);
  if (v) return v;
  if (!orig || !dnew || !dorig) return TMOL_LJLK_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds_entries = 3 * size_t(a.Amax <= LEAF_WAVE ? LEAF_WAVE : a.Amax);
  if (dtype == TMOL_LJLK_F32) {
    fill<float>((float*)dorig, 3 * int64_t(a.P) * a.NA, st);
    hipLaunchKernelGGL(leaf_atoms_backward_local_kernel<float>, dim3(leaf_grid(a)),
                       dim3(LEAF_WAVE), lds_entries * sizeof(float), st, a, (const float*)orig,
                       (const float*)dnew, (float*)dorig);
    hipLaunchKernelGGL(leaf_atoms_backward_remote_kernel<float>, dim3(leaf_grid(a)),
                       dim3(LEAF_WAVE), 0, st, a, (const float*)orig, (const float*)dnew,
                       (float*)dorig);
  } else {
    fill<double>((double*)dorig, 3 * int64_t(a.P) * a.NA, st);
    hipLaunchKernelGGL(leaf_atoms_backward_local_kernel<double>, dim3(leaf_grid(a)),
                       dim3(LEAF_WAVE), lds_entries * sizeof(double), st, a,
                       (const double*)orig, (const double*)dnew, (double*)dorig);
    hipLaunchKernelGGL(leaf_atoms_backward_remote_kernel<double>, dim3(leaf_grid(a)),
                       dim3(LEAF_WAVE), 0, st, a, (const double*)orig, (const double*)dnew,
                       (double*)dorig);
  }
  return check(hipGetLastError());
}

int tmol_lbfgs_dots(int32_t n_segments, int32_t history, int64_t seg_len, int32_t n_used,
                    const double* H, const double* vecs, int32_t n_vecs, double* out,
                    void* stream) {
  if (n_segments <= 0 || history <= 0 || seg_len <= 0 || n_used < 0 || n_used > history ||
      !H || !vecs || !out || n_vecs < 1 || n_vecs > LBFGS_K)
    return TMOL_LJLK_EINVAL;
  if (n_used == 0) return TMOL_LJLK_OK;
  hipLaunchKernelGGL(lbfgs_dots_kernel, dim3(unsigned(n_segments) * unsigned(n_used)), dim3(256),
                     0, (hipStream_t)stream, H, history, seg_len, n_used, n_segments, vecs, n_vecs,
                     out);
  return check(hipGetLastError());
}

int tmol_lbfgs_combine(int32_t n_segments, int32_t history, int64_t seg_len, int32_t n_used,
                       const double* g, const double* H1, const double* c1, const double* H2,
                       const double* c2, double* out, void* stream) {
  if (n_segments <= 0 || history <= 0 || seg_len <= 0 || n_used < 0 || n_used > history || !g ||
      !H1 || !c1 || !H2 || !c2 || !out)
    return TMOL_LJLK_EINVAL;
  hipLaunchKernelGGL(lbfgs_combine_kernel, dim3(unsigned((seg_len + 255) / 256), n_segments),
                     dim3(256), 0, (hipStream_t)stream, g, H1, c1, H2, c2, history, seg_len,
                     n_used, n_segments, out);
  return check(hipGetLastError());
}

}  // extern "C"
