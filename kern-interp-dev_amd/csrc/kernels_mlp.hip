// Learned ID-rank predictor: a 12-64-64-1 perceptron (tanh, tanh, sigmoid)
// mapping a box's features to its skeleton-rank fraction k / n.
//
//   mlp_grad_kernel   forward + backward of the mean squared error over 32
//                     samples per workgroup (FP32), per-workgroup partial
//                     gradients; mlp_reduce_kernel sums them
//   mlp_adam_kernel   Adam update of the flat FP32 parameter vector
//   mlp_infer_kernel  inference: FP32, or both hidden-layer products on the
//                     matrix cores in BF16 (v_mfma_f32_16x16x16_bf16, FP32
//                     accumulation), 16 samples per wave
//
// Flat parameter layout (FP32): W1 [64][12] (row = hidden unit), b1 [64],
// W2 [64][64], b2 [64], W3 [64], b3.
#include <cmath>

#include "ki_internal.h"
#include "mfma.h"

namespace ki {

namespace {

constexpr int NF = MLP_NF, NH = MLP_NH, SB = 32;
constexpr int OW1 = 0, OB1 = OW1 + NH * NF, OW2 = OB1 + NH,
              OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH,
              NP = OB3 + 1;
static_assert(NP == MLP_NPARAM, "parameter count");

__device__ __forceinline__ float sigmoid(float z) {
  return 1.0f / (1.0f + expf(-z));
}

// Samples idx[s0 + s] (or s0 + s when idx is null), s < SB.
__global__ void __launch_bounds__(256)
mlp_grad_kernel(const float* __restrict__ params, const float* __restrict__ X,
                const float* __restrict__ y, const int* __restrict__ idx,
                int n, float inv_n, float* __restrict__ part) {
  __shared__ float W[NP];
  __shared__ float xs[SB][NF];
  __shared__ float h1[SB][NH + 1];
  __shared__ float h2[SB][NH + 1];
  __shared__ float dd[SB][NH + 1];
  __shared__ float dz[SB], ls[SB];
  const int tid = threadIdx.x, s0 = blockIdx.x * SB;
  const int ns = min(SB, n - s0);
  for (int e = tid; e < NP; e += 256) W[e] = params[e];
  for (int e = tid; e < SB * NF; e += 256) {
    const int s = e / NF, f = e % NF;
    const int g = s < ns ? (idx ? idx[s0 + s] : s0 + s) : 0;
    xs[s][f] = s < ns ? X[(int64_t)g * NF + f] : 0.0f;
  }
  __syncthreads();
  for (int e = tid; e < SB * NH; e += 256) {
    const int s = e / NH, j = e % NH;
    float a = W[OB1 + j];
    for (int f = 0; f < NF; f++) a += W[OW1 + j * NF + f] * xs[s][f];
    h1[s][j] = tanhf(a);
  }
  __syncthreads();
  for (int e = tid; e < SB * NH; e += 256) {
    const int s = e / NH, j = e % NH;
    float a = W[OB2 + j];
    for (int i = 0; i < NH; i++) a += W[OW2 + j * NH + i] * h1[s][i];
    h2[s][j] = tanhf(a);
  }
  __syncthreads();
  if (tid < SB) {
    const int s = tid;
    float l = 0.0f, d = 0.0f;
    if (s < ns) {
      float z = W[OB3];
      for (int j = 0; j < NH; j++) z += W[OW3 + j] * h2[s][j];
      const float yh = sigmoid(z);
      const int g = idx ? idx[s0 + s] : s0 + s;
      const float diff = yh - y[g];
      l = diff * diff;
      d = 2.0f * diff * yh * (1.0f - yh) * inv_n;
    }
    ls[s] = l;
    dz[s] = d;
  }
  __syncthreads();
  for (int e = tid; e < SB * NH; e += 256) {
    const int s = e / NH, j = e % NH;
    dd[s][j] = dz[s] * W[OW3 + j] * (1.0f - h2[s][j] * h2[s][j]);
  }
  __syncthreads();
  float* P = part + (int64_t)blockIdx.x * (NP + 1);
  if (tid < NH) {
    float g3 = 0.0f, g2 = 0.0f;
    for (int s = 0; s < SB; s++) {
      g3 += dz[s] * h2[s][tid];
      g2 += dd[s][tid];
    }
    P[OW3 + tid] = g3;
    P[OB2 + tid] = g2;
  }
  if (tid == 0) {
    float gb = 0.0f, lt = 0.0f;
    for (int s = 0; s < SB; s++) {
      gb += dz[s];
      lt += ls[s];
    }
    P[OB3] = gb;
    P[NP] = lt;
  }
  for (int e = tid; e < NH * NH; e += 256) {
    const int j = e / NH, i = e % NH;
    float g = 0.0f;
    for (int s = 0; s < SB; s++) g += dd[s][j] * h1[s][i];
    P[OW2 + e] = g;
  }
  __syncthreads();
  // da1 = (W2^T da2) * (1 - h1^2), kept in h2 (no longer needed)
  for (int e = tid; e < SB * NH; e += 256) {
    const int s = e / NH, i = e % NH;
    float a = 0.0f;
    for (int j = 0; j < NH; j++) a += W[OW2 + j * NH + i] * dd[s][j];
    h2[s][i] = a * (1.0f - h1[s][i] * h1[s][i]);
  }
  __syncthreads();
  for (int e = tid; e < NH * NF; e += 256) {
    const int j = e / NF, f = e % NF;
    float g = 0.0f;
    for (int s = 0; s < SB; s++) g += h2[s][j] * xs[s][f];
    P[OW1 + e] = g;
  }
  if (tid < NH) {
    float g = 0.0f;
    for (int s = 0; s < SB; s++) g += h2[s][tid];
    P[OB1 + tid] = g;
  }
}

// out[e] = sum over workgroups of part[w][e], e <= NP (NP: loss sum).
__global__ void mlp_reduce_kernel(const float* __restrict__ part, int nwg,
                                  float* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e > NP) return;
  float s = 0.0f;
  for (int w = 0; w < nwg; w++) s += part[(int64_t)w * (NP + 1) + e];
  out[e] = s;
}

__global__ void mlp_adam_kernel(float* __restrict__ params,
                                const float* __restrict__ grad,
                                float* __restrict__ m, float* __restrict__ v,
                                float lr, float b1, float b2, float eps,
                                float c1, float c2) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= NP) return;
  const float g = grad[e];
  const float mm = b1 * m[e] + (1.0f - b1) * g;
  const float vv = b2 * v[e] + (1.0f - b2) * g * g;
  m[e] = mm;
  v[e] = vv;
  params[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// FP32 -> BF16 with round-to-nearest-even (finite inputs).
__device__ __forceinline__ short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));

// 64 samples per workgroup, 16 per wave.  BF16 mode: A = activations
// (16 samples x 16 inputs per MFMA), B = W^T; the accumulator's element v of
// lane l is (sample 4 (l / 16) + v, unit l % 16) of the 16 x 16 tile.
__global__ void __launch_bounds__(256)
mlp_infer_kernel(const float* __restrict__ params, const float* __restrict__ X,
                 int n, float* __restrict__ out, int bf16) {
  __shared__ float W[NP];
  __shared__ float xs[64][NF];
  __shared__ float h1[64][NH + 1];
  __shared__ short w1b[NH][16];      // W1 (inputs padded to 16), BF16
  __shared__ short w2b[NH][NH + 4];  // W2, BF16
  __shared__ short h1b[64][NH + 4];  // h1, BF16
  const int tid = threadIdx.x, s0 = blockIdx.x * 64;
  for (int e = tid; e < NP; e += 256) W[e] = params[e];
  for (int e = tid; e < 64 * NF; e += 256) {
    const int s = e / NF, f = e % NF;
    xs[s][f] = s0 + s < n ? X[(int64_t)(s0 + s) * NF + f] : 0.0f;
  }
  __syncthreads();
  if (!bf16) {
    for (int e = tid; e < 64 * NH; e += 256) {
      const int s = e / NH, j = e % NH;
      float a = W[OB1 + j];
      for (int f = 0; f < NF; f++) a += W[OW1 + j * NF + f] * xs[s][f];
      h1[s][j] = tanhf(a);
    }
    __syncthreads();
    if (tid < 64 && s0 + tid < n) {
      const int s = tid;
      float z = W[OB3];
      for (int j = 0; j < NH; j++) {
        float a = W[OB2 + j];
        for (int i = 0; i < NH; i++) a += W[OW2 + j * NH + i] * h1[s][i];
        z += W[OW3 + j] * tanhf(a);
      }
      out[s0 + s] = sigmoid(z);
    }
    return;
  }
  for (int e = tid; e < NH * 16; e += 256) {
    const int j = e / 16, f = e % 16;
    w1b[j][f] = f < NF ? to_bf16(W[OW1 + j * NF + f]) : (short)0;
  }
  for (int e = tid; e < NH * NH; e += 256)
    w2b[e / NH][e % NH] = to_bf16(W[OW2 + e]);
  __syncthreads();
  const int lane = tid & 63, wid = tid >> 6;
  const int sb = wid * 16;           // this wave's first sample
  const int ar = lane & 15, kq = 4 * (lane >> 4);
  // layer 1: [16 x 16] x [16 x 64]
  bf16x4 a;
  for (int e = 0; e < 4; e++)
    a[e] = kq + e < NF ? to_bf16(xs[sb + ar][kq + e]) : (short)0;
  for (int t = 0; t < 4; t++) {
    bf16x4 b;
    for (int e = 0; e < 4; e++) b[e] = w1b[16 * t + ar][kq + e];
    f32x4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    for (int v = 0; v < 4; v++) {
      const int s = kq + v, j = 16 * t + ar;
      h1b[sb + s][j] = to_bf16(tanhf(acc[v] + W[OB1 + j]));
    }
  }
  __syncthreads();
  // layer 2: [16 x 64] x [64 x 64], then the output unit in FP32
  float zp[4] = {0, 0, 0, 0};
  for (int t = 0; t < 4; t++) {
    f32x4 acc = {0, 0, 0, 0};
    for (int kk = 0; kk < NH; kk += 16) {
      bf16x4 a2, b2;
      for (int e = 0; e < 4; e++) {
        a2[e] = h1b[sb + ar][kk + kq + e];
        b2[e] = w2b[16 * t + ar][kk + kq + e];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b2, acc, 0, 0, 0);
    }
    const int j = 16 * t + ar;
    for (int v = 0; v < 4; v++) zp[v] += W[OW3 + j] * tanhf(acc[v] + W[OB2 + j]);
  }
  // sum over the 16 units held by lanes with the same lane / 16
  for (int v = 0; v < 4; v++)
    for (int o = 1; o < 16; o <<= 1) zp[v] += __shfl_xor(zp[v], o, 64);
  if (ar == 0)
    for (int v = 0; v < 4; v++) {
      const int s = s0 + sb + kq + v;
      if (s < n) out[s] = sigmoid(zp[v] + W[OB3]);
    }
}

}  // namespace

void mlp_grad(const float* params, const float* X, const float* y,
              const int* idx, int n, float* part, float* grad,
              hipStream_t s) {
  const int nwg = (n + SB - 1) / SB;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(nwg), dim3(256), 0, s, params, X,
                     y, idx, n, 1.0f / (float)n, part);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((NP + 1 + 255) / 256), dim3(256),
                     0, s, (const float*)part, nwg, grad);
  KI_HIP(hipGetLastError());
}

void mlp_adam(float* params, const float* grad, float* m, float* v, float lr,
              int t, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)t);
  const float c2 = 1.0f - std::pow(b2, (float)t);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((NP + 255) / 256), dim3(256), 0, s,
                     params, grad, m, v, lr, b1, b2, 1e-8f, c1, c2);
  KI_HIP(hipGetLastError());
}

void mlp_infer(const float* params, const float* X, int n, float* out,
               int bf16, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((n + 63) / 64), dim3(256), 0, s,
                     params, X, n, out, bf16);
  KI_HIP(hipGetLastError());
}

int mlp_partials(int n) { return (n + SB - 1) / SB; }

}  // namespace ki
