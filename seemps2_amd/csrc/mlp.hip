// Rank-predictor network of the learned truncation: a dense FP32 MLP
//   x (F) -> relu(W1 x + b1) (H) -> relu(W2 h1 + b2) (H) -> w3 . h2 + b3 (1)
// trained with a mean-squared-error loss and Adam, entirely on the device.
// Parameters live in one FP32 buffer, in the order W1 (H x F, row-major), b1,
// W2 (H x H), b2, w3 (H), b3.  Activations are kept from the forward pass for
// the backward pass.  Batches are small (thousands of spectra), so every
// kernel is a plain grid over (sample, unit) or (unit, input) pairs; FP32
// accumulation throughout.
#include <cmath>

#include "common.h"

namespace seemps {
namespace {

unsigned grid_for(int64_t n) {
  int64_t b = ceil_div(n, 256);
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// out[b][j] = act(sum_i in[b][i] W[j][i] + bias[j]); relu when relu != 0.
__global__ void dense_forward(int B, int nin, int nout, const float* __restrict__ in,
                              const float* __restrict__ W, const float* __restrict__ bias,
                              float* __restrict__ out, int relu) {
  const int64_t total = (int64_t)B * nout;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nout), j = (int)(e % nout);
    float acc = bias[j];
    for (int i = 0; i < nin; ++i) acc = fmaf(in[(int64_t)b * nin + i], W[(int64_t)j * nin + i], acc);
    out[e] = (relu && acc < 0.0f) ? 0.0f : acc;
  }
}

// dy[b] = 2 (y[b] - t[b]) / B, loss partial sums.
__global__ void mse_grad(int B, const float* __restrict__ y, const float* __restrict__ t,
                         float* __restrict__ dy, float* __restrict__ loss) {
  __shared__ float red[256];
  float acc = 0.0f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float d = y[b] - t[b];
    dy[b] = 2.0f * d / (float)B;
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / (float)B;
}

// gW[j][i] = sum_b dout[b][j] in[b][i]; gb[j] = sum_b dout[b][j].
__global__ void dense_weight_grad(int B, int nin, int nout, const float* __restrict__ in,
                                  const float* __restrict__ dout, float* __restrict__ gW,
                                  float* __restrict__ gb) {
  const int64_t total = (int64_t)nout * (nin + 1);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / (nin + 1)), i = (int)(e % (nin + 1));
    float acc = 0.0f;
    if (i < nin) {
      for (int b = 0; b < B; ++b)
        acc = fmaf(dout[(int64_t)b * nout + j], in[(int64_t)b * nin + i], acc);
      gW[(int64_t)j * nin + i] = acc;
    } else {
      for (int b = 0; b < B; ++b) acc += dout[(int64_t)b * nout + j];
      gb[j] = acc;
    }
  }
}

// din[b][i] = (sum_j dout[b][j] W[j][i]) * (act[b][i] > 0)  (ReLU of the input layer).
__global__ void dense_input_grad(int B, int nin, int nout, const float* __restrict__ dout,
                                 const float* __restrict__ W, const float* __restrict__ act,
                                 float* __restrict__ din) {
  const int64_t total = (int64_t)B * nin;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nin), i = (int)(e % nin);
    float acc = 0.0f;
    for (int j = 0; j < nout; ++j) acc = fmaf(dout[(int64_t)b * nout + j], W[(int64_t)j * nin + i], acc);
    din[e] = act[e] > 0.0f ? acc : 0.0f;
  }
}

// Adam step with bias correction (step counts from 1).
__global__ void adam_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.0f - b1) * gi;
    const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

struct MlpParams {
  const float *W1, *b1, *W2, *b2, *w3, *b3;
};

MlpParams carve_params(const float* p, int F, int H) {
  MlpParams q;
  q.W1 = p;
  q.b1 = q.W1 + (int64_t)H * F;
  q.W2 = q.b1 + H;
  q.b2 = q.W2 + (int64_t)H * H;
  q.w3 = q.b2 + H;
  q.b3 = q.w3 + H;
  return q;
}

}  // namespace
}  // namespace seemps

extern "C" int64_t seemps_mlp_param_count(int F, int H) {
  return (int64_t)H * F + H + (int64_t)H * H + H + H + 1;
}

extern "C" int seemps_mlp_forward(int B, int F, int H, const float* params, const float* x,
                                  float* h1, float* h2, float* y, void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(B >= 1 && F >= 1 && H >= 1, "mlp_forward: invalid sizes");
  hipStream_t s = (hipStream_t)stream;
  const MlpParams p = carve_params(params, F, H);
  int st;
  hipLaunchKernelGGL(dense_forward, dim3(grid_for((int64_t)B * H)), dim3(256), 0, s, B, F, H, x,
                     p.W1, p.b1, h1, 1);
  if ((st = check_launch("mlp layer 1"))) return st;
  hipLaunchKernelGGL(dense_forward, dim3(grid_for((int64_t)B * H)), dim3(256), 0, s, B, H, H, h1,
                     p.W2, p.b2, h2, 1);
  if ((st = check_launch("mlp layer 2"))) return st;
  hipLaunchKernelGGL(dense_forward, dim3(grid_for(B)), dim3(256), 0, s, B, H, 1, h2, p.w3, p.b3, y,
                     0);
  return check_launch("mlp output");
}

extern "C" int seemps_mlp_backward(int B, int F, int H, const float* params, const float* x,
                                   const float* h1, const float* h2, const float* y,
                                   const float* target, float* grads, float* loss, float* scratch,
                                   void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(B >= 1 && F >= 1 && H >= 1, "mlp_backward: invalid sizes");
  hipStream_t s = (hipStream_t)stream;
  const MlpParams p = carve_params(params, F, H);
  float* gW1 = grads;
  float* gb1 = gW1 + (int64_t)H * F;
  float* gW2 = gb1 + H;
  float* gb2 = gW2 + (int64_t)H * H;
  float* gw3 = gb2 + H;
  float* gb3 = gw3 + H;
  float* dy = scratch;                   // B
  float* dh2 = dy + B;                   // B x H
  float* dh1 = dh2 + (int64_t)B * H;     // B x H
  int st;
  hipLaunchKernelGGL(mse_grad, dim3(1), dim3(256), 0, s, B, y, target, dy, loss);
  if ((st = check_launch("mlp loss"))) return st;
  hipLaunchKernelGGL(dense_weight_grad, dim3(grid_for(H + 1)), dim3(256), 0, s, B, H, 1, h2, dy,
                     gw3, gb3);
  if ((st = check_launch("mlp grad 3"))) return st;
  hipLaunchKernelGGL(dense_input_grad, dim3(grid_for((int64_t)B * H)), dim3(256), 0, s, B, H, 1,
                     dy, p.w3, h2, dh2);
  if ((st = check_launch("mlp back 3"))) return st;
  hipLaunchKernelGGL(dense_weight_grad, dim3(grid_for((int64_t)H * (H + 1))), dim3(256), 0, s, B,
                     H, H, h1, dh2, gW2, gb2);
  if ((st = check_launch("mlp grad 2"))) return st;
  hipLaunchKernelGGL(dense_input_grad, dim3(grid_for((int64_t)B * H)), dim3(256), 0, s, B, H, H,
                     dh2, p.W2, h1, dh1);
  if ((st = check_launch("mlp back 2"))) return st;
  hipLaunchKernelGGL(dense_weight_grad, dim3(grid_for((int64_t)H * (F + 1))), dim3(256), 0, s, B,
                     F, H, x, dh1, gW1, gb1);
  return check_launch("mlp grad 1");
}

extern "C" int64_t seemps_mlp_scratch_count(int B, int H) { return (int64_t)B * (1 + 2 * H); }

extern "C" int seemps_adam_step(int64_t n, float* params, const float* grads, float* m, float* v,
                                float lr, float beta1, float beta2, float eps, int64_t step,
                                void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(n >= 0 && step >= 1, "adam_step: invalid arguments");
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, params,
                     grads, m, v, lr, beta1, beta2, eps, c1, c2);
  return check_launch("adam");
}
