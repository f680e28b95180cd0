// The learned controller's network on gfx950: FP32 training kernels
// (forward + hand-written backward + Adam) and BF16 matrix-core inference,
// with the odehip_mlp_* C ABI (declarations in odehip.h).
//
// Training: a mini-batch is split over workgroups of 256 threads; each
// workgroup streams its samples through LDS in chunks of 64, runs the
// forward and backward pass for the chunk as small LDS-resident matrix
// products, and keeps its share of the weight gradient in registers; the
// per-workgroup partial gradients are then summed in a fixed order by the
// Adam kernel (the run is deterministic).  Samples are visited in the order
// of an affine permutation i -> (a i + c) mod n per epoch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "context.hpp"
#include "mlp.hpp"

struct odehip_mlp {
  int device = 0;
  float* w = nullptr;  // FP32 master weights
  float* m = nullptr;  // Adam moments
  float* v = nullptr;
  float* mean = nullptr;  // input normalisation (device, kMlpIn each)
  float* inv_std = nullptr;
  odehip::MlpInfer* infer = nullptr;  // published inference copy
  float* partial = nullptr;           // [kMaxGradBlocks][kMlpParams]
  float* loss_part = nullptr;         // [kMaxGradBlocks]
  float* grad = nullptr;              // [kMlpParams]
  double* loss_acc = nullptr;         // device scalar
  long step = 0;
};

namespace odehip {

const MlpInfer* mlp_infer_ptr(const odehip_mlp* m) { return m ? m->infer : nullptr; }
int mlp_device(const odehip_mlp* m) { return m ? m->device : -1; }

namespace {

constexpr int kGradThreads = 256;
constexpr int kChunk = 64;
constexpr int kMaxGradBlocks = 256;

struct GradArgs {
  const float* w;
  const float* mean;
  const float* inv_std;
  const float* x;  // [n][kMlpIn] raw features
  const float* t;  // [n] targets
  long n;
  unsigned long long perm_a, perm_c;  // sample of position i: (a i + c) mod n
  long first;                         // first position of this batch
  int count;                          // samples in this batch
  float inv_count;
  float* partial;
  float* loss_part;
};

__global__ __launch_bounds__(kGradThreads) void k_mlp_grad(GradArgs g) {
  __shared__ float sw[kMlpParams + 3];
  __shared__ float xs[kChunk][kMlpIn + 1];
  __shared__ float h1[kChunk][kMlpHid + 1];
  __shared__ float h2[kChunk][kMlpHid + 1];
  __shared__ float d2[kChunk][kMlpHid + 1];
  __shared__ float d1[kChunk][kMlpHid + 1];
  __shared__ float dy[kChunk];
  __shared__ float tgt[kChunk];
  __shared__ int vld[kChunk];
  __shared__ float red[kGradThreads];
  const int tid = threadIdx.x;
  for (int p = tid; p < kMlpParams; p += kGradThreads) sw[p] = g.w[p];
  float aw2[16], aw1[3], ab = 0.0f, loss = 0.0f;
#pragma unroll
  for (int q = 0; q < 16; ++q) aw2[q] = 0.0f;
#pragma unroll
  for (int q = 0; q < 3; ++q) aw1[q] = 0.0f;
  const int per = (g.count + gridDim.x - 1) / gridDim.x;
  const int begin = blockIdx.x * per;
  const int end = min(g.count, begin + per);
  for (int c0 = begin; c0 < end; c0 += kChunk) {
    __syncthreads();
    for (int e = tid; e < kChunk * kMlpIn; e += kGradThreads) {
      const int s = e / kMlpIn, i = e % kMlpIn;
      const int b = c0 + s;
      float val = 0.0f;
      if (b < end) {
        const unsigned long long pos = static_cast<unsigned long long>(g.first + b);
        const long idx = static_cast<long>((g.perm_a * pos + g.perm_c) % g.n);
        val = (g.x[idx * kMlpIn + i] - g.mean[i]) * g.inv_std[i];
      }
      xs[s][i] = val;
    }
    if (tid < kChunk) {
      const int b = c0 + tid;
      vld[tid] = b < end;
      float tv = 0.0f;
      if (b < end) {
        const unsigned long long pos = static_cast<unsigned long long>(g.first + b);
        tv = g.t[static_cast<long>((g.perm_a * pos + g.perm_c) % g.n)];
      }
      tgt[tid] = tv;
    }
    __syncthreads();
    // Thread tid owns unit u = tid % 64 of the samples s = tid / 64 + 4 q,
    // q = 0..15, in the forward and backward passes, and the weight entries
    // W2[tid / 64 + 4 q][tid % 64] of the gradient.
    const int u = tid & 63;
    const int s0 = tid >> 6;
    {
      float z[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) z[q] = sw[kMlpOffB1 + u];
      for (int i = 0; i < kMlpIn; ++i) {
        const float wv = sw[kMlpOffW1 + u * kMlpIn + i];
#pragma unroll
        for (int q = 0; q < 16; ++q) z[q] += wv * xs[s0 + 4 * q][i];
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) h1[s0 + 4 * q][u] = tanhf(z[q]);
    }
    __syncthreads();
    {
      float z[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) z[q] = sw[kMlpOffB2 + u];
      for (int i = 0; i < kMlpHid; ++i) {
        const float wv = sw[kMlpOffW2 + u * kMlpHid + i];
#pragma unroll
        for (int q = 0; q < 16; ++q) z[q] += wv * h1[s0 + 4 * q][i];
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) h2[s0 + 4 * q][u] = tanhf(z[q]);
    }
    __syncthreads();
    if (tid < kChunk) {
      float y = sw[kMlpOffB3];
      for (int j = 0; j < kMlpHid; ++j) y += sw[kMlpOffW3 + j] * h2[tid][j];
      const float r = vld[tid] ? y - tgt[tid] : 0.0f;
      dy[tid] = r * g.inv_count;
      loss += 0.5f * r * r;
    }
    __syncthreads();
    // backward
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int s = s0 + 4 * q;
      const float hv = h2[s][u];
      d2[s][u] = dy[s] * sw[kMlpOffW3 + u] * (1.0f - hv * hv);
    }
    __syncthreads();
    {
      float acc[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
      for (int j = 0; j < kMlpHid; ++j) {
        const float wv = sw[kMlpOffW2 + j * kMlpHid + u];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] += d2[s0 + 4 * q][j] * wv;
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const float hv = h1[s0 + 4 * q][u];
        d1[s0 + 4 * q][u] = acc[q] * (1.0f - hv * hv);
      }
    }
    __syncthreads();
    // this chunk's weight gradient, accumulated in registers
    for (int s = 0; s < kChunk; ++s) {
      const float hv = h1[s][u];
#pragma unroll
      for (int q = 0; q < 16; ++q) aw2[q] += d2[s][s0 + 4 * q] * hv;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int e = tid + kGradThreads * q;
      if (e < kMlpHid * kMlpIn) {
        const int j = e / kMlpIn, i = e % kMlpIn;
        float acc = 0.0f;
        for (int s = 0; s < kChunk; ++s) acc += d1[s][j] * xs[s][i];
        aw1[q] += acc;
      }
    }
    {
      float acc = 0.0f;
      if (tid < 64) {
        for (int s = 0; s < kChunk; ++s) acc += d1[s][tid];
      } else if (tid < 128) {
        for (int s = 0; s < kChunk; ++s) acc += d2[s][tid - 64];
      } else if (tid < 192) {
        for (int s = 0; s < kChunk; ++s) acc += dy[s] * h2[s][tid - 128];
      } else if (tid == 192) {
        for (int s = 0; s < kChunk; ++s) acc += dy[s];
      }
      ab += acc;
    }
  }
  float* P = g.partial + static_cast<size_t>(blockIdx.x) * kMlpParams;
  const int u_out = tid & 63, s0_out = tid >> 6;
#pragma unroll
  for (int q = 0; q < 16; ++q) P[kMlpOffW2 + (s0_out + 4 * q) * kMlpHid + u_out] = aw2[q];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int e = tid + kGradThreads * q;
    if (e < kMlpHid * kMlpIn) P[kMlpOffW1 + e] = aw1[q];
  }
  if (tid < 64)
    P[kMlpOffB1 + tid] = ab;
  else if (tid < 128)
    P[kMlpOffB2 + tid - 64] = ab;
  else if (tid < 192)
    P[kMlpOffW3 + tid - 128] = ab;
  else if (tid == 192)
    P[kMlpOffB3] = ab;
  __syncthreads();
  red[tid] = loss;
  __syncthreads();
  if (tid == 0) {
    float s = 0.0f;
    for (int k = 0; k < kChunk; ++k) s += red[k];
    g.loss_part[blockIdx.x] = s;
  }
}

// Sums the partial gradients in workgroup order; optionally one Adam step.
__global__ void k_mlp_adam(const float* partial, int nblk, float* w, float* m,
                           float* v, float* grad, float lr, float b1, float b2,
                           float eps, float bc1, float bc2, int apply,
                           const float* loss_part, double* loss_acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += loss_part[b];
    *loss_acc += s;
  }
  if (p >= kMlpParams) return;
  float gsum = 0.0f;
  for (int b = 0; b < nblk; ++b) gsum += partial[static_cast<size_t>(b) * kMlpParams + p];
  grad[p] = gsum;
  if (!apply) return;
  const float mm = b1 * m[p] + (1.0f - b1) * gsum;
  const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
  m[p] = mm;
  v[p] = vv;
  w[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

__global__ void k_mlp_publish(const float* w, const float* mean,
                              const float* inv_std, MlpInfer* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 16) {
    out->mean[t] = t < kMlpIn ? mean[t] : 0.0f;
    out->inv_std[t] = t < kMlpIn ? inv_std[t] : 0.0f;
  }
  if (t < 64 * 16) {
    const int j = t / 16, i = t % 16;
    out->w1[j][i] = i < kMlpIn ? float_to_bf16(w[kMlpOffW1 + j * kMlpIn + i]) : 0;
  }
  if (t < 64 * 64) {
    const int j = t / 64, i = t % 64;
    out->w2[j][i] = float_to_bf16(w[kMlpOffW2 + j * kMlpHid + i]);
    // w3: 16 x 64, row 0 the weights
    if (j < 16) out->w3[j][i] = j == 0 ? float_to_bf16(w[kMlpOffW3 + i]) : 0;
  }
  if (t < 64) {
    out->b1[t] = w[kMlpOffB1 + t];
    out->b2[t] = w[kMlpOffB2 + t];
  }
  if (t == 0) {
    out->b3 = w[kMlpOffB3];
    out->pad[0] = out->pad[1] = out->pad[2] = 0.0f;
  }
}

// One wave per 16 samples.
__global__ __launch_bounds__(64) void k_mlp_infer(const MlpInfer* M,
                                                  const float* x, long n,
                                                  float* y) {
  __shared__ MlpTileLds t;
  const int lane = threadIdx.x;
  const long row = static_cast<long>(blockIdx.x) * 16 + lane;
  if (lane < 16) {
    float raw[kMlpIn];
    for (int i = 0; i < kMlpIn; ++i) raw[i] = row < n ? x[row * kMlpIn + i] : 0.0f;
    mlp_tile_set_row(M, t, lane, raw);
  }
  __syncthreads();
  mlp_tile_forward(M, t);
  if (lane < 16 && row < n) y[row] = t.y[lane];
}

unsigned long long gcd_u(unsigned long long a, unsigned long long b) {
  while (b) {
    const unsigned long long t = a % b;
    a = b;
    b = t;
  }
  return a;
}

int publish(odehip_context* ctx, odehip_mlp* m, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_publish, dim3(16), dim3(256), 0, s, m->w, m->mean, m->inv_std,
                     m->infer);
  HIP_TRY(ctx, hipGetLastError());
  return ODEHIP_OK;
}

// Gradient of one batch (positions [first, first + count) of the
// permutation) into m->grad; with apply, one Adam step.
int batch_step(odehip_context* ctx, odehip_mlp* m, const float* x, const float* t, long n,
               unsigned long long a, unsigned long long c, long first, int count,
               bool apply, float lr, float b1, float b2, float eps, hipStream_t s) {
  const int nblk = std::min(kMaxGradBlocks, std::max(1, (count + 255) / 256));
  GradArgs g{};
  g.w = m->w;
  g.mean = m->mean;
  g.inv_std = m->inv_std;
  g.x = x;
  g.t = t;
  g.n = n;
  g.perm_a = a;
  g.perm_c = c;
  g.first = first;
  g.count = count;
  g.inv_count = 1.0f / static_cast<float>(count);
  g.partial = m->partial;
  g.loss_part = m->loss_part;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(kGradThreads), 0, s, g);
  HIP_TRY(ctx, hipGetLastError());
  float bc1 = 1.0f, bc2 = 1.0f;
  if (apply) {
    m->step += 1;
    bc1 = 1.0f - std::pow(b1, static_cast<float>(m->step));
    bc2 = 1.0f - std::pow(b2, static_cast<float>(m->step));
  }
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s,
                     m->partial, nblk, m->w, m->m, m->v, m->grad, lr, b1, b2, eps, bc1, bc2,
                     apply ? 1 : 0, m->loss_part, m->loss_acc);
  HIP_TRY(ctx, hipGetLastError());
  return ODEHIP_OK;
}

}  // namespace
}  // namespace odehip

using namespace odehip;

extern "C" {

int odehip_mlp_create(odehip_context* ctx, const float* params, uint64_t seed,
                      const float* mean, const float* stdv, odehip_mlp** out) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!out || !mean || !stdv) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  float inv[kMlpIn];
  for (int i = 0; i < kMlpIn; ++i) {
    if (!(stdv[i] > 0.0f) || !std::isfinite(mean[i]))
      return fail(ctx, ODEHIP_ERR_INVALID, "std must be positive and mean finite");
    inv[i] = 1.0f / stdv[i];
  }
  std::vector<float> p(kMlpParams, 0.0f);
  if (params) {
    std::memcpy(p.data(), params, sizeof(float) * kMlpParams);
  } else {
    // Glorot-uniform weights, zero biases
    std::mt19937_64 gen(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
      const double lim = std::sqrt(6.0 / (fan_in + fan_out));
      std::uniform_real_distribution<double> d(-lim, lim);
      for (int k = 0; k < count; ++k) p[off + k] = static_cast<float>(d(gen));
    };
    fill(kMlpOffW1, kMlpIn, kMlpHid, kMlpHid * kMlpIn);
    fill(kMlpOffW2, kMlpHid, kMlpHid, kMlpHid * kMlpHid);
    fill(kMlpOffW3, kMlpHid, 1, kMlpHid);
  }
  auto* m = new (std::nothrow) odehip_mlp();
  if (!m) return fail(ctx, ODEHIP_ERR_NOMEM, "mlp");
  m->device = ctx->device;
  auto cleanup = [&](int rc) {
    odehip_mlp_destroy(m);
    return rc;
  };
  hipError_t e;
  if ((e = hipMalloc(&m->w, sizeof(float) * kMlpParams)) != hipSuccess ||
      (e = hipMalloc(&m->m, sizeof(float) * kMlpParams)) != hipSuccess ||
      (e = hipMalloc(&m->v, sizeof(float) * kMlpParams)) != hipSuccess ||
      (e = hipMalloc(&m->grad, sizeof(float) * kMlpParams)) != hipSuccess ||
      (e = hipMalloc(&m->mean, sizeof(float) * 16)) != hipSuccess ||
      (e = hipMalloc(&m->inv_std, sizeof(float) * 16)) != hipSuccess ||
      (e = hipMalloc(&m->infer, sizeof(MlpInfer))) != hipSuccess ||
      (e = hipMalloc(&m->partial, sizeof(float) * kMlpParams * kMaxGradBlocks)) != hipSuccess ||
      (e = hipMalloc(&m->loss_part, sizeof(float) * kMaxGradBlocks)) != hipSuccess ||
      (e = hipMalloc(&m->loss_acc, sizeof(double))) != hipSuccess)
    return cleanup(hip_fail(ctx, e, "mlp allocation"));
  std::vector<float> zeros(kMlpParams, 0.0f);
  if ((e = hipMemcpy(m->w, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(m->m, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(m->v, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(m->mean, mean, sizeof(float) * kMlpIn, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(m->inv_std, inv, sizeof(float) * kMlpIn, hipMemcpyHostToDevice)))
    return cleanup(hip_fail(ctx, e, "mlp upload"));
  hipStream_t s = begin_work(ctx);
  int rc = publish(ctx, m, s);
  if (rc) return cleanup(rc);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return cleanup(hip_fail(ctx, e, "mlp init"));
  end_work(ctx);
  *out = m;
  return ODEHIP_OK;
}

int odehip_mlp_destroy(odehip_mlp* m) {
  if (!m) return ODEHIP_OK;
  (void)hipSetDevice(m->device);
  (void)hipDeviceSynchronize();
  for (void* p : {static_cast<void*>(m->w), static_cast<void*>(m->m), static_cast<void*>(m->v),
                  static_cast<void*>(m->grad), static_cast<void*>(m->mean),
                  static_cast<void*>(m->inv_std), static_cast<void*>(m->infer),
                  static_cast<void*>(m->partial), static_cast<void*>(m->loss_part),
                  static_cast<void*>(m->loss_acc)})
    if (p) (void)hipFree(p);
  delete m;
  return ODEHIP_OK;
}

int odehip_mlp_get_params(odehip_context* ctx, const odehip_mlp* m,
                          float* params_out) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!m || !params_out) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (m->device != ctx->device)
    return fail(ctx, ODEHIP_ERR_INVALID, "the network lives on another device");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(params_out, m->w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return ODEHIP_OK;
}

int odehip_mlp_loss_grad(odehip_context* ctx, odehip_mlp* m, const float* x,
                         const float* t, long n, float* grad_out,
                         double* loss_out) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!m || !x || !t || !grad_out) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (m->device != ctx->device)
    return fail(ctx, ODEHIP_ERR_INVALID, "the network lives on another device");
  if (n <= 0 || n > (1L << 30)) return fail(ctx, ODEHIP_ERR_INVALID, "bad sample count");
  hipStream_t s = begin_work(ctx);
  HIP_TRY(ctx, hipMemsetAsync(m->loss_acc, 0, sizeof(double), s));
  int rc = batch_step(ctx, m, x, t, n, 1, 0, 0, static_cast<int>(n), false, 0.f, 0.f, 0.f,
                      0.f, s);
  if (rc) return rc;
  double loss = 0.0;
  HIP_TRY(ctx, hipMemcpyAsync(grad_out, m->grad, sizeof(float) * kMlpParams,
                              hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipMemcpyAsync(&loss, m->loss_acc, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  end_work(ctx);
  if (loss_out) *loss_out = loss / static_cast<double>(n);
  return ODEHIP_OK;
}

int odehip_mlp_train(odehip_context* ctx, odehip_mlp* m, const float* x,
                     const float* t, long n, const odehip_mlp_train_options* o,
                     double* epoch_loss) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!m || !x || !t || !o) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (m->device != ctx->device)
    return fail(ctx, ODEHIP_ERR_INVALID, "the network lives on another device");
  if (n <= 0 || n > (1L << 34)) return fail(ctx, ODEHIP_ERR_INVALID, "bad sample count");
  if (o->epochs < 0 || o->batch <= 0 || !(o->lr > 0.0f))
    return fail(ctx, ODEHIP_ERR_INVALID, "bad training options");
  hipStream_t s = begin_work(ctx);
  std::mt19937_64 gen(o->seed);
  const float b1 = o->beta1 > 0.0f ? o->beta1 : 0.9f;
  const float b2 = o->beta2 > 0.0f ? o->beta2 : 0.999f;
  const float eps = o->eps > 0.0f ? o->eps : 1e-8f;
  const unsigned long long un = static_cast<unsigned long long>(n);
  for (int ep = 0; ep < o->epochs; ++ep) {
    unsigned long long a = 1, c = 0;
    if (n > 1) {
      do {
        a = 1 + gen() % (un - 1);
      } while (gcd_u(a, un) != 1);
      c = gen() % un;
    }
    HIP_TRY(ctx, hipMemsetAsync(m->loss_acc, 0, sizeof(double), s));
    for (long first = 0; first < n; first += o->batch) {
      const int count = static_cast<int>(std::min<long>(o->batch, n - first));
      int rc = batch_step(ctx, m, x, t, n, a, c, first, count, true, o->lr, b1, b2, eps, s);
      if (rc) return rc;
    }
    double loss = 0.0;
    HIP_TRY(ctx, hipMemcpyAsync(&loss, m->loss_acc, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (epoch_loss) epoch_loss[ep] = loss / static_cast<double>(n);
  }
  int rc = publish(ctx, m, s);
  if (rc) return rc;
  HIP_TRY(ctx, hipStreamSynchronize(s));
  end_work(ctx);
  return ODEHIP_OK;
}

int odehip_mlp_infer(odehip_context* ctx, const odehip_mlp* m, const float* x,
                     long n, float* y) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!m || !x || !y) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (m->device != ctx->device)
    return fail(ctx, ODEHIP_ERR_INVALID, "the network lives on another device");
  if (n <= 0 || n > (1L << 34)) return fail(ctx, ODEHIP_ERR_INVALID, "bad sample count");
  hipStream_t s = begin_work(ctx);
  const long blocks = (n + 15) / 16;
  hipLaunchKernelGGL(k_mlp_infer, dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s,
                     m->infer, x, n, y);
  HIP_TRY(ctx, hipGetLastError());
  end_work(ctx);
  return ODEHIP_OK;
}

}  // extern "C"
