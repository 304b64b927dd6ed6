// Learned proposal for the STZ slip-rate root: a 4 -> 64 -> 64 -> 1 tanh MLP
// (FP32) that maps (locked shear traction phi, normal stress sn, effective
// temperature chi, radiation-damping impedance eta) to log(V), V the slip
// rate of the root of stz::calc_mu (stz.cpp:263-356).  Inference runs inside
// the fault kernel (one point per thread); training runs on the GPU with the
// kernels below (per-sample forward and deltas, weight gradients as
// reductions over the batch, Adam).  Nothing here exists in the reference.
//
// Parameter vector (floats), in this order:
//   W1[64][4] b1[64] W2[64][64] b2[64] W3[64] b3
//   in_mean[4] in_std[4] out_mean out_std          (normalisation)
#pragma once

#include <hip/hip_runtime.h>

namespace fdf {

constexpr int MLP_IN = 4, MLP_H = 64;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN, MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H,
              MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H, MLP_NW = MLP_B3 + 1;  // 4545 trained parameters
constexpr int MLP_NORM = MLP_NW, MLP_NP = MLP_NORM + 2 * MLP_IN + 2;      // + normalisation

// tanh(x) = 1 - 2 / (exp(2x) + 1) with the hardware exponential and
// reciprocal (relative error ~1e-6; exact limits +-1 for large |x|); training
// and inference use the same function
__device__ __forceinline__ float mlp_tanh(float x) { return 1.f - __fdividef(2.f, __expf(2.f * x) + 1.f); }

// normalised network output for one normalised input (FP32); the second
// layer runs four hidden units at a time (independent accumulation chains)
__device__ __forceinline__ float mlp_eval(const float* __restrict__ P, const float* x) {
  float h1[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    float s = P[MLP_B1 + j];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) s += P[MLP_W1 + j * MLP_IN + i] * x[i];
    h1[j] = mlp_tanh(s);
  }
  float y = P[MLP_B3];
  for (int j = 0; j < MLP_H; j += 4) {
    float s0 = P[MLP_B2 + j], s1 = P[MLP_B2 + j + 1], s2 = P[MLP_B2 + j + 2], s3 = P[MLP_B2 + j + 3];
#pragma unroll
    for (int k = 0; k < MLP_H; k++) {
      s0 += P[MLP_W2 + j * MLP_H + k] * h1[k];
      s1 += P[MLP_W2 + (j + 1) * MLP_H + k] * h1[k];
      s2 += P[MLP_W2 + (j + 2) * MLP_H + k] * h1[k];
      s3 += P[MLP_W2 + (j + 3) * MLP_H + k] * h1[k];
    }
    y += P[MLP_W3 + j] * mlp_tanh(s0) + P[MLP_W3 + j + 1] * mlp_tanh(s1) + P[MLP_W3 + j + 2] * mlp_tanh(s2) +
         P[MLP_W3 + j + 3] * mlp_tanh(s3);
  }
  return y;
}

// The network for the 64 points of a wave (one per lane) with the 64 x 64
// hidden layer on the matrix cores: H2 = W2 H1 as a (64 neurons) x (64
// points) product of four 32 x 32 tiles, eight v_mfma_f32_32x32x8f16 steps
// each (FP16 operands, FP32 accumulation).  Operand layouts of that
// instruction (wave64): A[m][k] sits in lane m + 32 (k / 4), element k % 4;
// B[k][n] in lane n + 32 (k / 4), element k % 4; D[m][n] in lane
// n + 32 ((m / 4) % 2), register (m % 4) + 4 (m / 8).  So lane l (column
// c = l % 32, half h = l / 32) computes the first-layer activations of
// neurons 8 kb + 4 h + i for the points c and c + 32, holds rows 32 mt + c of
// W2, and finishes the output layer over its 32 rows of the two point
// columns; the two lanes sharing a column add their halves.  Every lane of
// the wave must be active (the inputs and halves move between lanes).
typedef _Float16 mlp_h4 __attribute__((ext_vector_type(4)));
typedef float mlp_f16x __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float mlp_eval_wave(const float* __restrict__ P, const float* x) {
  const int lane = threadIdx.x & 63, col = lane & 31, kh = lane >> 5;
  float xin[2][MLP_IN];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) xin[t][i] = __shfl(x[i], col + 32 * t, 64);
  mlp_h4 B[2][8];
#pragma unroll
  for (int kb = 0; kb < 8; kb++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = 8 * kb + 4 * kh + i;
#pragma unroll
      for (int t = 0; t < 2; t++) {
        float s = P[MLP_B1 + k];
#pragma unroll
        for (int c = 0; c < MLP_IN; c++) s += P[MLP_W1 + k * MLP_IN + c] * xin[t][c];
        B[t][kb][i] = (_Float16)mlp_tanh(s);
      }
    }
  mlp_f16x acc[2][2];
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int j = 0; j < 16; j++) acc[mt][t][j] = 0.f;
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int kb = 0; kb < 8; kb++) {
      mlp_h4 a;
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = (_Float16)P[MLP_W2 + (32 * mt + col) * MLP_H + 8 * kb + 4 * kh + i];
#pragma unroll
      for (int t = 0; t < 2; t++) acc[mt][t] = __builtin_amdgcn_mfma_f32_32x32x8f16(a, B[t][kb], acc[mt][t], 0, 0, 0);
    }
  float s[2] = {0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int m = 32 * mt + 8 * (j / 4) + 4 * kh + (j % 4);
        s[t] += P[MLP_W3 + m] * mlp_tanh(acc[mt][t][j] + P[MLP_B2 + m]);
      }
  const float other = __shfl_xor(kh ? s[0] : s[1], 32, 64);
  return (kh ? s[1] : s[0]) + other + P[MLP_B3];
}

// normalised inputs of a raw (phi, sn, chi, eta)
__device__ __forceinline__ void mlp_inputs(const float* __restrict__ P, double phi, double sn, double chi, double eta,
                                           float* x) {
  const double raw[MLP_IN] = {phi, sn, chi, eta};
#pragma unroll
  for (int i = 0; i < MLP_IN; i++)
    x[i] = (float)((raw[i] - (double)P[MLP_NORM + i]) / (double)P[MLP_NORM + MLP_IN + i]);
}

// log(V) from a normalised network output
__device__ __forceinline__ double mlp_out(const float* __restrict__ P, float y) {
  return (double)y * (double)P[MLP_NORM + 2 * MLP_IN + 1] + (double)P[MLP_NORM + 2 * MLP_IN];
}

// log(V) proposed for raw inputs (phi, sn, chi, eta)
__device__ __forceinline__ double mlp_logv(const float* __restrict__ P, double phi, double sn, double chi, double eta) {
  const double raw[MLP_IN] = {phi, sn, chi, eta};
  float x[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; i++)
    x[i] = (float)((raw[i] - (double)P[MLP_NORM + i]) / (double)P[MLP_NORM + MLP_IN + i]);
  const float y = mlp_eval(P, x);
  return (double)y * (double)P[MLP_NORM + 2 * MLP_IN + 1] + (double)P[MLP_NORM + 2 * MLP_IN];
}

// ---- training ---------------------------------------------------------------
// Buffers (B = batch size): X [n][4] normalised inputs and T [n] normalised
// targets (a batch picks rows through idx), H1/H2 [B][64] activations and
// D1/D2 [B][64] deltas of the pre-activations (sample-major: at a fixed
// sample the 64 lanes of a wave read 64 consecutive floats), DY [B] output
// delta.  Loss = mean over the batch of (y - t)^2 / 2.

// forward and backward deltas, one sample per thread
__global__ void __launch_bounds__(256) mlp_fwd_bwd_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                          const float* __restrict__ T, const int* __restrict__ idx,
                                                          int B, float* __restrict__ H1, float* __restrict__ H2,
                                                          float* __restrict__ D1, float* __restrict__ D2,
                                                          float* __restrict__ DY, float* __restrict__ loss) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (b < B) {
    const int s = idx ? idx[b] : b;
    float x[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) x[i] = X[(long long)s * MLP_IN + i];
    float h1[MLP_H], h2[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
      float a = P[MLP_B1 + j];
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) a += P[MLP_W1 + j * MLP_IN + i] * x[i];
      h1[j] = mlp_tanh(a);
      H1[(long long)b * MLP_H + j] = h1[j];
    }
    float y = P[MLP_B3];
    for (int j = 0; j < MLP_H; j++) {
      float a = P[MLP_B2 + j];
#pragma unroll
      for (int k = 0; k < MLP_H; k++) a += P[MLP_W2 + j * MLP_H + k] * h1[k];
      h2[j] = mlp_tanh(a);
      H2[(long long)b * MLP_H + j] = h2[j];
      y += P[MLP_W3 + j] * h2[j];
    }
    const float r = y - T[s];
    l = 0.5f * r * r;
    const float dy = r / (float)B;
    DY[b] = dy;
    float d2[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
      d2[j] = dy * P[MLP_W3 + j] * (1.f - h2[j] * h2[j]);
      D2[(long long)b * MLP_H + j] = d2[j];
    }
    for (int k = 0; k < MLP_H; k++) {
      float g = 0.f;
#pragma unroll
      for (int j = 0; j < MLP_H; j++) g += d2[j] * P[MLP_W2 + j * MLP_H + k];
      D1[(long long)b * MLP_H + k] = g * (1.f - h1[k] * h1[k]);
    }
  }
  // batch loss (sum; the host divides by B)
  __shared__ float red[256];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

// weight gradients as reductions over the batch, split into MLP_GS slices
// (blockIdx.y) written to GP[slice][4545]; mlp_gsum_kernel adds the slices in
// a fixed order.  Blocks 0..63: row g of W2 and b2 (4 waves split the slice,
// lane k owns column k); blocks 64..127: row j of W1, b1 j and W3 j (256
// threads split the slice, tree reduction in LDS); block 128: b3.
constexpr int MLP_GS = 8;

__global__ void __launch_bounds__(256) mlp_grad_kernel(const float* __restrict__ X, const int* __restrict__ idx, int B,
                                                       const float* __restrict__ H1, const float* __restrict__ H2,
                                                       const float* __restrict__ D1, const float* __restrict__ D2,
                                                       const float* __restrict__ DY, float* __restrict__ GP) {
  const int g = blockIdx.x, sl = blockIdx.y, t = threadIdx.x;
  const int b0 = (int)((long long)B * sl / MLP_GS), b1 = (int)((long long)B * (sl + 1) / MLP_GS);
  float* __restrict__ G = GP + (long long)sl * MLP_NW;
  __shared__ float part[4][MLP_H + 1];
  __shared__ float red[6][256];
  if (g < MLP_H) {
    const int k = t & 63, wv = t >> 6;
    float s = 0.f, sb = 0.f;
    for (int b = b0 + wv; b < b1; b += 4) {
      const float d = D2[(long long)b * MLP_H + g];
      s += d * H1[(long long)b * MLP_H + k];
      sb += d;
    }
    part[wv][k] = s;
    if (k == 0) part[wv][MLP_H] = sb;
    __syncthreads();
    if (t < MLP_H) G[MLP_W2 + g * MLP_H + t] = part[0][t] + part[1][t] + part[2][t] + part[3][t];
    if (t == 0) G[MLP_B2 + g] = part[0][MLP_H] + part[1][MLP_H] + part[2][MLP_H] + part[3][MLP_H];
    return;
  }
  const int j = g - MLP_H;  // (block 128: j = 64, the b3 block)
  float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int b = b0 + t; b < b1; b += 256) {
    if (j < MLP_H) {
      const float d = D1[(long long)b * MLP_H + j];
      const int s = idx ? idx[b] : b;
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) v[i] += d * X[(long long)s * MLP_IN + i];
      v[4] += d;
      v[5] += DY[b] * H2[(long long)b * MLP_H + j];
    } else {
      v[0] += DY[b];
    }
  }
#pragma unroll
  for (int q = 0; q < 6; q++) red[q][t] = v[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
#pragma unroll
      for (int q = 0; q < 6; q++) red[q][t] += red[q][t + o];
    __syncthreads();
  }
  if (t == 0) {
    if (j < MLP_H) {
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) G[MLP_W1 + j * MLP_IN + i] = red[i][0];
      G[MLP_B1 + j] = red[4][0];
      G[MLP_W3 + j] = red[5][0];
    } else {
      G[MLP_B3] = red[0][0];
    }
  }
}

// G = sum of the MLP_GS slices of GP, in slice order
__global__ void mlp_gsum_kernel(const float* __restrict__ GP, float* __restrict__ G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NW) return;
  float s = 0.f;
#pragma unroll
  for (int sl = 0; sl < MLP_GS; sl++) s += GP[(long long)sl * MLP_NW + i];
  G[i] = s;
}

// Adam (Kingma & Ba 2015) on the trained parameters; step counts from 1
__global__ void mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                                float* __restrict__ V, float lr, float b1, float b2, float eps, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NW) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  const float mh = m / (1.f - powf(b1, (float)step));
  const float vh = v / (1.f - powf(b2, (float)step));
  P[i] -= lr * mh / (sqrtf(vh) + eps);
}

// normalised predictions for a set of normalised inputs
__global__ void mlp_predict_kernel(const float* __restrict__ P, const float* __restrict__ X, int n,
                                   float* __restrict__ Y) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float x[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) x[i] = X[(long long)b * MLP_IN + i];
  Y[b] = mlp_eval(P, x);
}

// the matrix-core evaluation for n normalised inputs (a whole wave per 64
// points; lanes past n evaluate a zero input and store nothing)
__global__ void __launch_bounds__(256) mlp_predict_wave_kernel(const float* __restrict__ Pg,
                                                               const float* __restrict__ X, int n,
                                                               float* __restrict__ Y) {
  __shared__ float P[MLP_NP];
  for (int i = threadIdx.x; i < MLP_NP; i += blockDim.x) P[i] = Pg[i];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float x[MLP_IN] = {0.f, 0.f, 0.f, 0.f};
  if (b < n)
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) x[i] = X[(long long)b * MLP_IN + i];
  const float y = mlp_eval_wave(P, x);
  if (b < n) Y[b] = y;
}

}  // namespace fdf
