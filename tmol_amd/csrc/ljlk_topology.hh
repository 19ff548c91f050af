// Pose topology on the device: inter_block_bondsep [P, B, B, C, C] and
// min_block_bondsep [P, B, B] from the block types and the inter-residue
// connections (reference: tmol/pose/pose_stack_builder.py:459-599,
// 1230-1349 with tmol/pose/compiled/apsp.*; tmol/score/
// bond_dependent_term.py:91-100).
//
// The graph has one node per connection point of a real block.  Two points
// of one block are joined with the path distance between their atoms, two
// points joined by an inter-residue connection with weight 1; separations
// are shortest-path lengths capped at 6.  The reference builds the dense
// point-by-point matrix of each pose and runs all-pairs shortest paths on
// it.  Here one lane per connection point runs a bounded search from its
// point: with a cap of 6 only a handful of points are within reach, so
// each lane keeps them in a short private list, relaxes their edges until
// nothing changes, and writes its row of the output; a second pass takes
// the minimum over connection pairs.  The inter-residue connections are
// read in both directions as given (the reference's builder stores both).
#pragma once

#include "ljlk_common.hh"

namespace tmol_ljlk {

constexpr int BS_CUT = 6;      // MAX_SIG_BOND_SEPARATION
constexpr int BS_REACH = 48;   // points within reach of one point (list size)

struct TopoView {
  int P, B, C;
  const int32_t* block_type;  // [P * B]
  const int32_t* irc;         // [P, B, C, 2]
};

__device__ __forceinline__ int topo_n_conn(const TopoView& v, const BTView& bt, int p, int b) {
  const int t = v.block_type[int64_t(p) * v.B + b];
  return t >= 0 ? bt.n_conn[t] : 0;
}

// one lane per (pose, block, connection)
__global__ __launch_bounds__(256) void bondsep_rows_kernel(TopoView v, BTView bt,
                                                           int32_t* __restrict__ ibs,
                                                           int32_t* __restrict__ overflow) {
  const int64_t n = int64_t(v.P) * v.B * v.C;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int c = int(i % v.C);
    const int64_t pb = i / v.C;
    const int b = int(pb % v.B), p = int(pb / v.B);
    if (c >= topo_n_conn(v, bt, p, b)) continue;
    // reached points: (block << 3 | connection), distance
    int node[BS_REACH];
    int dist[BS_REACH];
    int cnt = 1;
    node[0] = (b << 3) | c;
    dist[0] = 0;
    bool changed = true, over = false;
    while (changed) {
      changed = false;
      for (int k = 0; k < cnt; ++k) {
        const int u = node[k], du = dist[k];
        const int ub = u >> 3, uc = u & 7;
        const int ut = v.block_type[int64_t(p) * v.B + ub];
        if (ut < 0) continue;  // a connection to a block that is not there
        const int unc = bt.n_conn[ut];
        if (uc >= unc) continue;
        // neighbors: the other points of u's block, and u's partner
        for (int e = 0; e <= unc; ++e) {
          int w, tgt;
          if (e < unc) {
            if (e == uc) continue;
            const int a1 = bt.conn_atoms[ut * bt.C + uc], a2 = bt.conn_atoms[ut * bt.C + e];
            w = bt.path_distance[(int64_t(ut) * bt.Amax + a1) * bt.Amax + a2];
            tgt = (ub << 3) | e;
          } else {
            const int32_t* ir = v.irc + ((int64_t(p) * v.B + ub) * v.C + uc) * 2;
            if (ir[0] < 0 || ir[0] >= v.B || ir[1] < 0 || ir[1] >= v.C) continue;
            w = 1;
            tgt = (ir[0] << 3) | ir[1];
          }
          const int nd = du + w;
          if (nd >= BS_CUT) continue;
          int f = -1;
          for (int q = 0; q < cnt; ++q)
            if (node[q] == tgt) f = q;
          if (f >= 0) {
            if (nd < dist[f]) {
              dist[f] = nd;
              changed = true;
            }
          } else if (cnt < BS_REACH) {
            node[cnt] = tgt;
            dist[cnt] = nd;
            ++cnt;
            changed = true;
          } else {
            over = true;
          }
        }
      }
    }
    if (over && overflow) atomicOr((uint32_t*)overflow, 1u);
    for (int k = 0; k < cnt; ++k) {
      const int tb = node[k] >> 3, tc = node[k] & 7;
      ibs[((((int64_t(p) * v.B + b) * v.B + tb) * v.C + c) * v.C) + tc] = dist[k];
    }
  }
}

// min_bondsep[p, b1, b2] = min over the [C, C] entries
__global__ __launch_bounds__(256) void bondsep_min_kernel(TopoView v, const int32_t* __restrict__ ibs,
                                                          int32_t* __restrict__ min_sep) {
  const int64_t n = int64_t(v.P) * v.B * v.B;
  const int cc = v.C * v.C;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int m = BS_CUT;
    for (int k = 0; k < cc; ++k) {
      const int x = ibs[i * cc + k];
      m = x < m ? x : m;
    }
    min_sep[i] = m;
  }
}

}  // namespace tmol_ljlk
