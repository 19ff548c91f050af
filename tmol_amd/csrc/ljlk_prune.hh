// Learned block-pair pruner: a 40 -> 128 -> 128 -> 128 -> 1 SiLU MLP that
// predicts P(|E_pair| < 1e-3 kcal/mol) for a neighbor block pair, trained on
// pair energies from the exact kernel.  BF16 operands on the matrix cores
// (v_mfma_f32_32x32x8bf16_1k), FP32 accumulation.
//
// Inputs of a block pair (g1, g2) (PRUNE_IN = 40, padded to PRUNE_K1 = 48):
//   [0]      centroid distance / 10 A
//   [1, 2]   bounding-sphere radii / 10 A
//   [3, 11)  learned 8-dim embedding of g1's residue type
//   [11, 19) embedding of g2's residue type
//   [19, 29) sin, cos of g1's phi, psi, omega, chi1, chi2 (0, 0 if absent)
//   [29, 39) the same for g2
//   [39]     min(|i - j|, 32) / 32, sequence separation within the pose
//
// Kernels: block torsion features; pair features; a tiled BF16 GEMM with
// training epilogues (bias + SiLU, SiLU backward, FP32 / split-K output);
// the output head with the logistic loss and its backward; bias and
// embedding gradients; the fused inference kernel that scores every
// work-list pair and compacts the kept ones.
#pragma once

#include "ljlk_common.hh"

namespace tmol_ljlk {

constexpr int PRUNE_IN = 40;
constexpr int PRUNE_K1 = 48;
constexpr int PRUNE_H = 128;
constexpr int PRUNE_EMB = 8;
constexpr int PRUNE_TORS = 5;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16 bits, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t to_bf16(float x) {
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}
// two floats -> packed bf16 pair (lo, hi)
__device__ __forceinline__ uint32_t to_bf16x2(float lo, float hi) {
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  bf16x2_t v;
  v[0] = (__bf16)lo;
  v[1] = (__bf16)hi;
  return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// logistic function with the hardware exp and reciprocal (inputs beyond
// +-88 saturate to exactly 0 / 1)
__device__ __forceinline__ float sigmoidf_(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu_(float z) { return z * sigmoidf_(z); }
__device__ __forceinline__ float dsilu_(float z) {
  const float s = sigmoidf_(z);
  return s * (1.0f + z * (1.0f - s));
}

// D[32 x 32] += A[32 x 8] B[8 x 32]; lane l holds A[l % 32][4 (l / 32) + q],
// B[4 (l / 32) + q][l % 32] and D[8 r + 4 (l / 32) + q][l % 32] = acc[4 r + q]
__device__ __forceinline__ f32x16 mfma_32x32x8(bf16x4 a, bf16x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ features

// sin, cos of 5 torsions per block from global atom indices [G, 5, 4]
// (-1: torsion absent)
template <typename Real>
__global__ void prune_block_torsions_kernel(int G, const int32_t* __restrict__ tors_atoms,
                                            const Real* __restrict__ coords,
                                            float* __restrict__ sincos) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  for (int t = 0; t < PRUNE_TORS; ++t) {
    const int32_t* a = tors_atoms + (int64_t(g) * PRUNE_TORS + t) * 4;
    float s = 0.f, c = 0.f;
    if (a[0] >= 0 && a[1] >= 0 && a[2] >= 0 && a[3] >= 0) {
      float p[4][3];
      for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) p[i][k] = float(coords[int64_t(a[i]) * 3 + k]);
      float b1[3], b2[3], b3[3];
      for (int k = 0; k < 3; ++k) {
        b1[k] = p[1][k] - p[0][k];
        b2[k] = p[2][k] - p[1][k];
        b3[k] = p[3][k] - p[2][k];
      }
      // n1 = b1 x b2, n2 = b2 x b3; angle = atan2(|b2| b1 . n2, n1 . n2)
      const float n1[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2],
                           b1[0] * b2[1] - b1[1] * b2[0]};
      const float n2[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2],
                           b2[0] * b3[1] - b2[1] * b3[0]};
      const float lb2 = sqrtf(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
      const float y = lb2 * (b1[0] * n2[0] + b1[1] * n2[1] + b1[2] * n2[2]);
      const float x = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
      const float r = sqrtf(x * x + y * y);
      if (r > 0.f) {
        s = y / r;
        c = x / r;
      }
    }
    sincos[int64_t(g) * 2 * PRUNE_TORS + 2 * t] = s;
    sincos[int64_t(g) * 2 * PRUNE_TORS + 2 * t + 1] = c;
  }
}

struct PruneBlocks {
  const float* sphere;      // [G, 4] (copy of the neighbor pass's spheres, float)
  const float* sincos;      // [G, 10]
  const int32_t* emb_row;   // [G] embedding row of the block's residue type
  const int32_t* seqpos;    // [G] position within the pose
  const float* emb;         // [n_emb, 8]
};

// one feature of a block pair
__device__ __forceinline__ float pair_feature(const PruneBlocks& B, int g1, int g2, int f) {
  if (f == 0) {
    const float dx = B.sphere[4 * g1] - B.sphere[4 * g2];
    const float dy = B.sphere[4 * g1 + 1] - B.sphere[4 * g2 + 1];
    const float dz = B.sphere[4 * g1 + 2] - B.sphere[4 * g2 + 2];
    return sqrtf(dx * dx + dy * dy + dz * dz) * 0.1f;
  }
  if (f == 1) return B.sphere[4 * g1 + 3] * 0.1f;
  if (f == 2) return B.sphere[4 * g2 + 3] * 0.1f;
  if (f < 11) return B.emb[B.emb_row[g1] * PRUNE_EMB + (f - 3)];
  if (f < 19) return B.emb[B.emb_row[g2] * PRUNE_EMB + (f - 11)];
  if (f < 29) return B.sincos[g1 * 2 * PRUNE_TORS + (f - 19)];
  if (f < 39) return B.sincos[g2 * 2 * PRUNE_TORS + (f - 29)];
  if (f == 39) {
    int d = B.seqpos[g1] - B.seqpos[g2];
    d = d < 0 ? -d : d;
    return float(d < 32 ? d : 32) * (1.0f / 32.0f);
  }
  return 0.f;
}

// X [n, 48] bf16 for an explicit pair list [n, 2]
__global__ void prune_pair_features_kernel(int64_t n, const int32_t* __restrict__ pairs,
                                           PruneBlocks B, uint16_t* __restrict__ X) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * PRUNE_K1) return;
  const int64_t r = i / PRUNE_K1;
  const int f = int(i - r * PRUNE_K1);
  const int g1 = pairs[2 * r], g2 = pairs[2 * r + 1];
  X[i] = to_bf16(pair_feature(B, g1, g2, f));
}

// ------------------------------------------------------------------ GEMM

// C[m, n] = sum_k A(m, k) B(k, n), bf16 in, fp32 accumulate.
//   A_KM: A stored [K][M] (else [M][K]);  B_KN: B stored [K][N] (else [N][K])
// 64 x 64 tile per workgroup (4 waves of 32 x 32), K in steps of 32 through
// LDS; blockIdx.z splits K (atomic fp32 output).
enum : int { EPI_F32 = 0, EPI_BIAS_SILU = 1, EPI_DSILU = 2 };

struct GemmArgs {
  const uint16_t* A;
  const uint16_t* B;
  int M, N, K, lda, ldb, ldc;
  int k_chunk;          // K range per blockIdx.z
  float* C;             // EPI_F32 output (atomic when split)
  const float* bias;    // EPI_BIAS_SILU
  float* Z;             // EPI_BIAS_SILU: pre-activation out; EPI_DSILU: pre-activation in
  uint16_t* H;          // bf16 out: SiLU(z) or the back-propagated value
  float* D;             // EPI_DSILU: fp32 back-propagated value
};

constexpr int GT = 64, GK = 32, GKP = GK + 4;

template <bool A_KM, bool B_KN, int EPI>
__global__ __launch_bounds__(256) void prune_gemm_kernel(GemmArgs g) {
  __shared__ uint16_t As[GT][GKP];
  __shared__ uint16_t Bs[GT][GKP];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int kb = blockIdx.z * g.k_chunk;
  const int ke = min(g.K, kb + g.k_chunk);
  f32x16 acc;
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  for (int k0 = kb; k0 < ke; k0 += GK) {
    // ---- stage A, B tiles (zero outside the matrix)
    for (int e = tid; e < GT * GK; e += 256) {
      int r, k;
      if (A_KM) {
        k = e / GT;
        r = e - k * GT;
      } else {
        r = e / GK;
        k = e - r * GK;
      }
      const int gm = m0 + r, gk = k0 + k;
      uint16_t v = 0;
      if (gm < g.M && gk < ke) v = A_KM ? g.A[int64_t(gk) * g.lda + gm] : g.A[int64_t(gm) * g.lda + gk];
      As[r][k] = v;
      int c;
      if (B_KN) {
        k = e / GT;
        c = e - k * GT;
      } else {
        c = e / GK;
        k = e - c * GK;
      }
      const int gn = n0 + c;
      const int gk2 = k0 + k;
      uint16_t u = 0;
      if (gn < g.N && gk2 < ke) u = B_KN ? g.B[int64_t(gk2) * g.ldb + gn] : g.B[int64_t(gn) * g.ldb + gk2];
      Bs[c][k] = u;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < GK; ks += 8) {
      bf16x4 a, b;
      const int kk = ks + 4 * (lane >> 5);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = (short)As[32 * wm + (lane & 31)][kk + q];
        b[q] = (short)Bs[32 * wn + (lane & 31)][kk + q];
      }
      acc = mfma_32x32x8(a, b, acc);
    }
    __syncthreads();
  }
  // ---- epilogue
  const int n = n0 + 32 * wn + (lane & 31);
  if (n >= g.N) return;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m = m0 + 32 * wm + 8 * r + 4 * (lane >> 5) + q;
      if (m >= g.M) continue;
      const int64_t o = int64_t(m) * g.ldc + n;
      const float v = acc[4 * r + q];
      if (EPI == EPI_F32) {
        if (gridDim.z > 1)
          unsafeAtomicAdd(&g.C[o], v);
        else
          g.C[o] = v;
      } else if (EPI == EPI_BIAS_SILU) {
        const float z = v + g.bias[n];
        g.Z[o] = z;
        g.H[o] = to_bf16(silu_(z));
      } else {
        const float d = v * dsilu_(g.Z[o]);
        g.D[o] = d;
        g.H[o] = to_bf16(d);
      }
    }
}

// ------------------------------------------------------------------ head

// Output layer and logistic loss for a batch:
//   logit = H3 . w4 + b4;  p = sigmoid(logit);  label y in {0, 1}; sample
//   weight s_b (1 when sample_w is null)
//   loss_b = s_b [pos_w y log(1 + e^-l) + (1 - y) log(1 + e^l)]  (summed into *loss)
//   with train: dlogit = s_b (pos_w y (p - 1) + (1 - y) p) / n;
//   dZ3 = dlogit w4 silu'(Z3) (fp32 + bf16); dw4 += dlogit H3; db4 += dlogit
__global__ __launch_bounds__(256) void prune_head_kernel(
    int n, const uint16_t* __restrict__ H3, const float* __restrict__ Z3,
    const float* __restrict__ w4, const float* __restrict__ b4, const float* __restrict__ labels,
    const float* __restrict__ sample_w, float pos_w, int train, float* __restrict__ logits,
    float* __restrict__ loss,
    float* __restrict__ dZ3, uint16_t* __restrict__ dZ3b, float* __restrict__ dw4,
    float* __restrict__ db4) {
  const int lane = lane_id();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = b < n;
  float l = 0.f;
  if (ok) {
    l = b4[0];
    for (int i = 0; i < PRUNE_H; ++i) l += from_bf16(H3[int64_t(b) * PRUNE_H + i]) * w4[i];
    logits[b] = l;
  }
  if (!train) return;
  float dl = 0.f, ls = 0.f;
  if (ok) {
    const float y = labels[b];
    const float sp_neg = fmaxf(-l, 0.f) + log1pf(__expf(-fabsf(l)));  // log(1 + e^-l)
    const float sp_pos = sp_neg + l;                                   // log(1 + e^l)
    const float sw = sample_w ? sample_w[b] : 1.f;
    ls = sw * (pos_w * y * sp_neg + (1.f - y) * sp_pos);
    const float p = sigmoidf_(l);
    dl = sw * (pos_w * y * (p - 1.f) + (1.f - y) * p) / float(n);
    for (int i = 0; i < PRUNE_H; ++i) {
      const int64_t o = int64_t(b) * PRUNE_H + i;
      const float d = dl * w4[i] * dsilu_(Z3[o]);
      dZ3[o] = d;
      dZ3b[o] = to_bf16(d);
    }
  }
  const float lsum = wave_sum(ls);
  const float dsum = wave_sum(dl);
  if (lane == 0) {
    unsafeAtomicAdd(loss, lsum);
    unsafeAtomicAdd(db4, dsum);
  }
  for (int i = 0; i < PRUNE_H; ++i) {
    const float v = ok ? dl * from_bf16(H3[int64_t(b) * PRUNE_H + i]) : 0.f;
    const float s = wave_sum(v);
    if (lane == 0) unsafeAtomicAdd(&dw4[i], s);
  }
}

// db[c] += sum_r D[r, c]  (D [n, ld], columns [0, ncol))
__global__ __launch_bounds__(256) void prune_colsum_kernel(int n, int ncol, int ld,
                                                           const float* __restrict__ D,
                                                           float* __restrict__ db) {
  const int c = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x + blockIdx.y * 256; r < n; r += 256 * gridDim.y) s += D[int64_t(r) * ld + c];
  s = wave_sum(s);
  __shared__ float part[4];
  if (lane_id() == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0 && c < ncol) unsafeAtomicAdd(&db[c], part[0] + part[1] + part[2] + part[3]);
}

// embedding gradient from the input-layer gradient dX [n, 48] (fp32)
__global__ void prune_emb_grad_kernel(int64_t n, const int32_t* __restrict__ pairs,
                                      const int32_t* __restrict__ emb_row,
                                      const float* __restrict__ dX, float* __restrict__ demb) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * 2 * PRUNE_EMB) return;
  const int64_t r = i / (2 * PRUNE_EMB);
  const int j = int(i - r * 2 * PRUNE_EMB);
  const int side = j / PRUNE_EMB, e = j - side * PRUNE_EMB;
  const int g = pairs[2 * r + side];
  const float v = dX[r * PRUNE_K1 + 3 + j];
  if (v != 0.f) unsafeAtomicAdd(&demb[emb_row[g] * PRUNE_EMB + e], v);
}

}  // namespace tmol_ljlk

namespace tmol_ljlk {

// ------------------------------------------------------------------ inference

struct PruneNet {
  const uint16_t* W1;  // [128][48] bf16 (out, in)
  const uint16_t* W2;  // [128][128]
  const uint16_t* W3;  // [128][128]
  const float* b1;
  const float* b2;
  const float* b3;
  const float* w4;     // [128]
  const float* b4;     // [1]
};

constexpr int PI_WAVES = 8;          // waves per inference workgroup
constexpr int PS1 = PRUNE_K1 + 4;    // LDS row stride (bf16) of W1
constexpr int PSH = PRUNE_H + 4;     // LDS row stride of W2, W3 and activations

struct PruneLDS {
  uint16_t w1[PRUNE_H][PS1];
  uint16_t w2[PRUNE_H][PSH];
  uint16_t w3[PRUNE_H][PSH];
  float b[3][PRUNE_H];
  float w4[PRUNE_H];
  float b4;
  uint16_t act[PI_WAVES][32][PSH];  // per wave: features, then hidden activations
};

// one hidden layer for the wave's 32 pairs: act[:, :K] -> act[:, :128].
// The product is computed transposed, D[out][pair] = W[out, :] . act[pair, :]
// (weights as the A operand), so that each lane ends up holding 4
// consecutive outputs of one pair and stores them with one 8-byte write.
template <int K, int WS>
__device__ __forceinline__ void prune_layer(uint16_t (*act)[PSH], const uint16_t (*W)[WS],
                                            const float* bias, int lane) {
  f32x16 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[c][i] = 0.f;
  const int r = lane & 31, kh = 4 * (lane >> 5);
#pragma unroll
  for (int ks = 0; ks < K; ks += 8) {
    const bf16x4 x = *(const bf16x4*)&act[r][ks + kh];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bf16x4 w = *(const bf16x4*)&W[32 * c + r][ks + kh];
      acc[c] = mfma_32x32x8(w, x, acc[c]);
    }
  }
  wave_sync();
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 32 * c + 8 * i + kh;  // outputs o .. o + 3 of pair r
      const float4 bb = *(const float4*)&bias[o];
      const float z0 = acc[c][4 * i] + bb.x, z1 = acc[c][4 * i + 1] + bb.y;
      const float z2 = acc[c][4 * i + 2] + bb.z, z3 = acc[c][4 * i + 3] + bb.w;
      uint2 packed;
      packed.x = to_bf16x2(silu_(z0), silu_(z1));
      packed.y = to_bf16x2(silu_(z2), silu_(z3));
      *(uint2*)&act[r][o] = packed;
    }
  wave_sync();
}

// Scores the pairs of a flat list [n, 2] (n read from device memory, so the
// launch shape does not depend on it) with the MLP; keep[i] = 0 when the pair
// is predicted negligible with probability above the threshold
// (logit > logit_cut), is not a block with itself, and its sphere gap
// |c1 - c2| - r1 - r2 -- a lower bound on every atom-pair distance of the two
// blocks -- is at least min_gap (pairs whose atoms may come closer are always
// scored: a clash there is worth far more than the bound).
// Pairs the rule keeps whatever the network says (a block with itself, a gap
// below min_gap) are marked kept without running the network: each wave
// walks chunks of 64 pairs, tests them one per lane, and queues only the
// others; the network runs on 32 queued pairs at a time.  With logits
// (analysis) every pair goes through the network and gets its logit.
__global__ __launch_bounds__(PI_WAVES * 64) void prune_infer_kernel(
    const int64_t* __restrict__ n_pairs, const int32_t* __restrict__ pairs, PruneBlocks B,
    PruneNet net, float logit_cut, float min_gap, uint8_t* __restrict__ keep,
    float* __restrict__ logits) {
  __shared__ PruneLDS L;
  __shared__ int32_t queue_all[PI_WAVES][96];
  const int64_t n = *n_pairs;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < PRUNE_H * PRUNE_K1; i += blockDim.x) L.w1[i / PRUNE_K1][i % PRUNE_K1] = net.W1[i];
  for (int i = tid; i < PRUNE_H * PRUNE_H; i += blockDim.x) {
    L.w2[i / PRUNE_H][i % PRUNE_H] = net.W2[i];
    L.w3[i / PRUNE_H][i % PRUNE_H] = net.W3[i];
  }
  for (int i = tid; i < PRUNE_H; i += blockDim.x) {
    L.b[0][i] = net.b1[i];
    L.b[1][i] = net.b2[i];
    L.b[2][i] = net.b3[i];
    L.w4[i] = net.w4[i];
  }
  if (tid == 0) L.b4 = net.b4[0];
  __syncthreads();
  uint16_t (*act)[PSH] = L.act[wave];
  int32_t* queue = queue_all[wave];
  const bool all = logits != nullptr;
  const int64_t n_chunks = (n + 63) / 64;
  const int r = lane & 31, half = lane >> 5;
  int nq = 0;  // queued pairs (wave-uniform)
  for (int64_t c = int64_t(blockIdx.x) * PI_WAVES + wave;; c += int64_t(gridDim.x) * PI_WAVES) {
    const bool more = c < n_chunks;
    if (more) {
      // ---- which pairs of this chunk need the network
      const int64_t i = c * 64 + lane;
      bool cand = false;
      if (i < n) {
        const int a = pairs[2 * i], b = pairs[2 * i + 1];
        const float4 s1 = *(const float4*)(B.sphere + 4 * int64_t(a));
        const float4 s2 = *(const float4*)(B.sphere + 4 * int64_t(b));
        const float dx = s1.x - s2.x, dy = s1.y - s2.y, dz = s1.z - s2.z;
        const float gap = sqrtf(dx * dx + dy * dy + dz * dz) - s1.w - s2.w;
        cand = all || (a != b && !(gap < min_gap));
        if (!cand) keep[i] = 1;
      }
      const uint64_t m = __ballot(cand);
      if (cand) queue[nq + lanes_below(m)] = int32_t(i);
      nq += __popcll(m);
      wave_sync();
    }
    while (nq >= 32 || (!more && nq > 0)) {
      // ---- the network on queued pairs 0 .. take - 1
      const int take = nq < 32 ? nq : 32;
      const bool valid = r < take;
      const int64_t i = valid ? queue[r] : 0;
      float gap = 0.f;  // sphere gap of row r (half-0 lanes)
      int g1 = 0, g2 = 0;
      if (valid) {
        g1 = pairs[2 * i];
        g2 = pairs[2 * i + 1];
      }
      // features: lane r writes the columns of block g1 of row r, lane r + 32
      // those of g2 (same code on both halves, only the offsets differ); the
      // centroid distance and sequence separation use the partner lane's block
      {
        const int gb = half ? g2 : g1;
        const float4 sph = *(const float4*)(B.sphere + 4 * int64_t(gb));
        const float4 e0 = *(const float4*)(B.emb + B.emb_row[gb] * PRUNE_EMB);
        const float4 e1 = *(const float4*)(B.emb + B.emb_row[gb] * PRUNE_EMB + 4);
        const float* sc = B.sincos + int64_t(gb) * 2 * PRUNE_TORS;
        float t[2 * PRUNE_TORS];
#pragma unroll
        for (int k = 0; k < 2 * PRUNE_TORS; k += 2) {
          const float2 v = *(const float2*)(sc + k);
          t[k] = v.x;
          t[k + 1] = v.y;
        }
        const int sp = B.seqpos[gb];
        const float ox = __shfl_xor(sph.x, 32, WAVE), oy = __shfl_xor(sph.y, 32, WAVE);
        const float oz = __shfl_xor(sph.z, 32, WAVE);
        const float ow = __shfl_xor(sph.w, 32, WAVE);
        const int osp = __shfl_xor(sp, 32, WAVE);
        const float dx = sph.x - ox, dy = sph.y - oy, dz = sph.z - oz;
        gap = sqrtf(dx * dx + dy * dy + dz * dz) - sph.w - ow;
        const float ev[PRUNE_EMB] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
        const uint16_t z = 0;
        // columns: 0 distance, 1 + half radius, 3 + 8 half embedding,
        // 19 + 10 half torsions, 39 sequence separation, 40..47 zero
        if (half == 0) {
          act[r][0] = valid ? to_bf16(sqrtf(dx * dx + dy * dy + dz * dz) * 0.1f) : z;
          int d = sp - osp;
          d = d < 0 ? -d : d;
          act[r][39] = valid ? to_bf16(float(d < 32 ? d : 32) * (1.0f / 32.0f)) : z;
        } else {
#pragma unroll
          for (int k = PRUNE_IN; k < PRUNE_K1; ++k) act[r][k] = z;
        }
        act[r][1 + half] = valid ? to_bf16(sph.w * 0.1f) : z;
#pragma unroll
        for (int k = 0; k < PRUNE_EMB; ++k) act[r][3 + 8 * half + k] = valid ? to_bf16(ev[k]) : z;
#pragma unroll
        for (int k = 0; k < 2 * PRUNE_TORS; ++k) act[r][19 + 10 * half + k] = valid ? to_bf16(t[k]) : z;
      }
      wave_sync();
      prune_layer<PRUNE_K1, PS1>(act, L.w1, L.b[0], lane);
      prune_layer<PRUNE_H, PSH>(act, L.w2, L.b[1], lane);
      prune_layer<PRUNE_H, PSH>(act, L.w3, L.b[2], lane);
      // head: lanes r and r + 32 each sum half of row r
      float sum = 0.f;
      for (int k = 64 * half; k < 64 * half + 64; k += 2) {
        const uint32_t two = *(const uint32_t*)&act[r][k];
        sum += from_bf16(uint16_t(two & 0xffffu)) * L.w4[k] + from_bf16(uint16_t(two >> 16)) * L.w4[k + 1];
      }
      sum += __shfl_xor(sum, 32, WAVE);
      if (half == 0 && valid) {
        const float logit = sum + L.b4;
        keep[i] = (g1 == g2 || !(logit > logit_cut) || gap < min_gap) ? 1 : 0;
        if (logits) logits[i] = logit;
      }
      // ---- drop the taken pairs from the queue (at most 63 remain)
      const int rem = nq - take;
      int32_t v = 0;
      if (lane < rem) v = queue[take + lane];
      wave_sync();
      if (lane < rem) queue[lane] = v;
      nq = rem;
      wave_sync();
    }
    if (!more) break;
  }
}

// ------------------------------------------------------------------ work-list flattening

// per 1024-row block: sum of min(row_count, cap)
__global__ __launch_bounds__(1024) void prune_count_kernel(int64_t G, const uint32_t* __restrict__ row_count,
                                                            int64_t cap, int64_t* __restrict__ block_sums) {
  const int64_t g = int64_t(blockIdx.x) * 1024 + threadIdx.x;
  int64_t c = g < G ? min<int64_t>(row_count[g], cap) : 0;
  __shared__ int64_t part[16];
  c = wave_sum(c);
  if (lane_id() == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int w = 0; w < 16; ++w) s += part[w];
    block_sums[blockIdx.x] = s;
  }
}

// exclusive scan of block_sums in place (single workgroup, sequential per
// thread over chunks), block_sums[nb] = total
__global__ __launch_bounds__(1024) void prune_scan_blocks_kernel(int64_t nb, int64_t* __restrict__ block_sums) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (nb + 1023) / 1024;
  const int64_t b0 = t * per, b1 = min(nb, b0 + per);
  int64_t s = 0;
  for (int64_t b = b0; b < b1; ++b) s += block_sums[b];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int k = 0; k < 1024; ++k) {
      const int64_t v = part[k];
      part[k] = run;
      run += v;
    }
    block_sums[nb] = run;
  }
  __syncthreads();
  int64_t run = part[t];
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t v = block_sums[b];
    block_sums[b] = run;
    run += v;
  }
}

// row offsets (exclusive) and the flat pair list: lane j of a wave owns row
// base + j for the scan; the wave then writes each of its 64 rows' pairs with
// all lanes (coalesced)
__global__ __launch_bounds__(1024) void prune_flatten_kernel(int64_t G, const uint32_t* __restrict__ row_count,
                                                              const int32_t* __restrict__ row_items, int64_t cap,
                                                              const int64_t* __restrict__ block_sums,
                                                              int64_t* __restrict__ row_offset,
                                                              int32_t* __restrict__ pairs) {
  __shared__ int64_t part[16];
  const int lane = lane_id();
  const int64_t g = int64_t(blockIdx.x) * 1024 + threadIdx.x;
  const int64_t c = g < G ? min<int64_t>(row_count[g], cap) : 0;
  // inclusive scan within the wave, then across the 16 waves
  int64_t x = c;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o, WAVE);
    if (lane >= o) x += y;
  }
  if (lane == 63) part[threadIdx.x >> 6] = x;
  __syncthreads();
  int64_t before = block_sums[blockIdx.x];
  for (int w = 0; w < int(threadIdx.x >> 6); ++w) before += part[w];
  const int64_t start = before + x - c;
  if (g < G) {
    row_offset[g] = start;
    if (g == G - 1) row_offset[G] = start + c;
  }
  const int64_t base = g - lane;
  for (int j = 0; j < 64; ++j) {
    const int64_t gj = base + j;
    if (gj >= G) break;
    const int64_t cj = __shfl(c, j, WAVE), sj = __shfl(start, j, WAVE);
    for (int64_t k = lane; k < cj; k += 64) {
      pairs[2 * (sj + k)] = int32_t(gj);
      pairs[2 * (sj + k) + 1] = row_items[gj * cap + k];
    }
  }
}

// keep only flagged pairs in each row of the work list (in place, order
// kept): one wave per 64 rows, ballot compaction of each row's pairs
__global__ __launch_bounds__(256) void prune_compact_kernel(int64_t G, uint32_t* __restrict__ row_count,
                                                            int32_t* __restrict__ row_items, int64_t cap,
                                                            const int64_t* __restrict__ row_offset,
                                                            const uint8_t* __restrict__ keep) {
  const int lane = lane_id();
  const int64_t base = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) - lane;
  for (int j = 0; j < 64; ++j) {
    const int64_t g = base + j;
    if (g >= G) break;
    const int64_t c = min<int64_t>(row_count[g], cap);
    const int64_t o = row_offset[g];
    int64_t w = 0;
    for (int64_t k0 = 0; k0 < c; k0 += 64) {
      const int64_t k = k0 + lane;
      int32_t item = 0;
      bool kept = false;
      if (k < c) {
        item = row_items[g * cap + k];
        kept = keep[o + k] != 0;
      }
      const uint64_t m = __ballot(kept);
      wave_sync();
      if (kept) row_items[g * cap + w + lanes_below(m)] = item;
      w += __popcll(m);
    }
    if (lane == 0) row_count[g] = uint32_t(w);
  }
}

}  // namespace tmol_ljlk
