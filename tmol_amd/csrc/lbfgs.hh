// Batched products with the L-BFGS history of a segmented minimization, in
// float64: every segment (pose) p keeps m history vectors of length S
// ([P, m, S], rows contiguous).  The compact inverse-Hessian product needs,
// per iteration, the dot products of the history with a few vectors and one
// linear combination of the history rows; both are streaming passes over the
// history (memory bound), done here in one pass each.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmol_ljlk {

constexpr int LBFGS_K = 2;  // vectors per pass

// out[k, p, i] = sum_s H[p, i, s] * v[k, p, s] for i < n, k < nv: one
// workgroup per (p, i); threads stride over s, then a workgroup reduction
__global__ __launch_bounds__(256) void lbfgs_dots_kernel(const double* __restrict__ H, int m,
                                                         int64_t S, int n, int P,
                                                         const double* __restrict__ v, int nv,
                                                         double* __restrict__ out) {
  const int p = blockIdx.x / n, i = blockIdx.x - p * n;
  const double* h = H + (int64_t(p) * m + i) * S;
  double acc[LBFGS_K] = {0.0, 0.0};
  for (int64_t s = threadIdx.x; s < S; s += blockDim.x) {
    const double x = h[s];
#pragma unroll
    for (int k = 0; k < LBFGS_K; ++k)
      if (k < nv) acc[k] += x * v[(int64_t(k) * P + p) * S + s];
  }
  __shared__ double part[LBFGS_K][4];
#pragma unroll
  for (int k = 0; k < LBFGS_K; ++k) {
    double a = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) part[k][threadIdx.x >> 6] = a;
  }
  __syncthreads();
  if (threadIdx.x < nv)
    out[(int64_t(threadIdx.x) * P + p) * n + i] =
        (part[threadIdx.x][0] + part[threadIdx.x][1]) + (part[threadIdx.x][2] + part[threadIdx.x][3]);
}

// out[p, s] = g[p, s] + sum_{i < n} (c1[p, i] H1[p, i, s] + c2[p, i] H2[p, i, s])
__global__ __launch_bounds__(256) void lbfgs_combine_kernel(
    const double* __restrict__ g, const double* __restrict__ H1, const double* __restrict__ c1,
    const double* __restrict__ H2, const double* __restrict__ c2, int m, int64_t S, int n, int P,
    double* __restrict__ out) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (s >= S) return;
  const double* h1 = H1 + int64_t(p) * m * S + s;
  const double* h2 = H2 + int64_t(p) * m * S + s;
  double a = g[int64_t(p) * S + s];
  for (int i = 0; i < n; ++i)
    a += c1[int64_t(p) * n + i] * h1[int64_t(i) * S] + c2[int64_t(p) * n + i] * h2[int64_t(i) * S];
  out[int64_t(p) * S + s] = a;
}

}  // namespace tmol_ljlk
