// The learned step-size controller's network: layout, controller features
// and the BF16 matrix-core inference of a tile of 16 systems.
//
// Network: 10 -> 64 -> 64 -> 1, tanh hidden layers, linear output = the log
// of the proposed step ratio dt_new / dt.  Training keeps FP32 master
// weights in the flat layout
//   W1[64][10] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3     (kParams)
// and publishes an MlpInfer copy: weights in bfloat16, padded to the 16 x 16
// x 16 matrix-core shape, biases and the input normalisation in FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "internal.hpp"

namespace odehip {

constexpr int kMlpIn = ODEHIP_MLP_IN;
constexpr int kMlpHid = ODEHIP_MLP_HIDDEN;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHid * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHid;
constexpr int kMlpParams = kMlpOffB3 + 1;
static_assert(kMlpParams == ODEHIP_MLP_PARAMS, "parameter count");

// Step ratio limits of the learned proposal: [0.2, 5].
constexpr double kMlpLogRatioMin = -1.6094379124341003;
constexpr double kMlpLogRatioMax = 1.6094379124341003;

// Inference copy (device).  Rows are output units, columns inputs, so a
// 16 x 16 B operand (k, n) reads W[n][k .. k+3] contiguously.
struct MlpInfer {
  float mean[16];
  float inv_std[16];
  uint16_t w1[64][16];  // K padded 10 -> 16 with zeros
  float b1[64];
  uint16_t w2[64][64];
  float b2[64];
  uint16_t w3[16][64];  // row 0 = w3, rows 1..15 zero (N padded to 16)
  float b3;
  float pad[3];
};

__host__ __device__ inline uint16_t float_to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// Raw controller features of one system after an attempt: log error norms of
// the last three attempts, log(dt / (tf - t0)) of the last, the log ratios of
// the two before to it, the three accept flags, log N.  Non-finite entries
// (history not filled yet) are 0.
__device__ inline void mlp_features(const SolverState* st, int n_particles,
                                    float* raw) {
  double f[kMlpIn];
  for (int h = 0; h < 3; ++h) {
    double e = st->hist_err[h];
    e = e < 1.0e-12 ? 1.0e-12 : (e > 1.0e12 ? 1.0e12 : e);
    f[h] = std::log(e);
  }
  const double d0 = st->hist_dt[0];
  f[3] = std::log(d0 / (st->tf - st->t0));
  f[4] = std::log(st->hist_dt[1] / d0);
  f[5] = std::log(st->hist_dt[2] / d0);
  for (int h = 0; h < 3; ++h) f[6 + h] = st->hist_acc[h];
  f[9] = std::log(static_cast<double>(n_particles));
  for (int i = 0; i < kMlpIn; ++i)
    raw[i] = std::isfinite(f[i]) ? static_cast<float>(f[i]) : 0.0f;
}

typedef short mlp_short4 __attribute__((ext_vector_type(4)));
typedef float mlp_float4 __attribute__((ext_vector_type(4)));

// LDS scratch of one wave's tile.
struct MlpTileLds {
  float x[16][16];
  float h[16][65];
  float y[16];
};

// Network output for the 16 rows of t.x (normalised features, written by the
// caller; columns 10..15 zero).  Must be called by all 64 lanes of a wave,
// in a workgroup of exactly one wave (the barriers are workgroup barriers).
// v_mfma_f32_16x16x16_bf16 operand layout: lane l holds A[l % 16][4 (l / 16)
// + i], B[4 (l / 16) + i][l % 16] and D[4 (l / 16) + i][l % 16], i = 0..3.
__device__ inline void mlp_tile_forward(const MlpInfer* M, MlpTileLds& t) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  const int kq = lane >> 4;
  // layer 1: [16 x 16] x [16 x 64]
  mlp_short4 a;
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = static_cast<short>(float_to_bf16(t.x[r][4 * kq + i]));
  mlp_float4 z[4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    mlp_short4 b;
#pragma unroll
    for (int i = 0; i < 4; ++i) b[i] = static_cast<short>(M->w1[16 * nb + r][4 * kq + i]);
    const mlp_float4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    z[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      t.h[4 * kq + i][16 * nb + r] = tanhf(z[nb][i] + M->b1[16 * nb + r]);
  __syncthreads();
  // layer 2: [16 x 64] x [64 x 64]
  mlp_short4 a2[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      a2[ks][i] = static_cast<short>(float_to_bf16(t.h[r][16 * ks + 4 * kq + i]));
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    mlp_float4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      mlp_short4 b;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        b[i] = static_cast<short>(M->w2[16 * nb + r][16 * ks + 4 * kq + i]);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[ks], b, c, 0, 0, 0);
    }
    z[nb] = c;
  }
  __syncthreads();
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      t.h[4 * kq + i][16 * nb + r] = tanhf(z[nb][i] + M->b2[16 * nb + r]);
  __syncthreads();
  // layer 3: [16 x 64] x [64 x 16] (only column 0 is the output)
  mlp_float4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    mlp_short4 a3, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a3[i] = static_cast<short>(float_to_bf16(t.h[r][16 * ks + 4 * kq + i]));
      b[i] = static_cast<short>(M->w3[r][16 * ks + 4 * kq + i]);
    }
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, b, c, 0, 0, 0);
  }
  if (r == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) t.y[4 * kq + i] = c[i] + M->b3;
  }
  __syncthreads();
}

// The published inference copy of a network handle (mlp_kernels.hip).
const MlpInfer* mlp_infer_ptr(const odehip_mlp* m);
// Device the network was created on (its buffers live there).
int mlp_device(const odehip_mlp* m);

// Stores the normalised features of row `row` (0..15) of a tile.
__device__ inline void mlp_tile_set_row(const MlpInfer* M, MlpTileLds& t,
                                        int row, const float* raw) {
#pragma unroll
  for (int i = 0; i < 16; ++i)
    t.x[row][i] = i < kMlpIn ? (raw[i] - M->mean[i]) * M->inv_std[i] : 0.0f;
}

}  // namespace odehip
