// imexsw_mlp.h -- the learned Newton warm start: a per-cell MLP
// 9 -> 64 -> 64 -> 64 -> 2 (GELU) that predicts the converged momenta of the
// stage-2 implicit solve from the state that enters it, with BF16 operands
// and FP32 accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16).
//
// Activations are kept transposed (features x samples), so the output tile
// of one layer is, register for register, the B operand of the next:
//   mfma 16x16x16, lane l, i = l % 16, g = l / 16, element j = 0..3:
//     A(i, 4g+j) = a[j],  B(4g+j, i) = b[j],  D(4g+j, i) = d[j]
//   layer:  Z^T(out x samples) = W(out x in) . A^T(in x samples)
// A wavefront evaluates 64 samples (four 16-sample column blocks).
//
// Features (all from the stage-2 predictor that the plain method hands to
// Newton): log h, u, v (the semi-implicit predictor's velocity, m/s), T,
// three solid volume fractions, log dt, |grad B|; each standardised with
// the training set's mean and deviation.  Targets: the converged momenta
// relative to the predictor's, q*_x / q_si_x - 1 and q*_y / q_si_y - 1,
// standardised the same way.
#pragma once

#include "imexsw_newton.h"

namespace imexsw {

constexpr int kMlpIn = 9, kMlpInPad = 16, kMlpH = 64, kMlpOut = 2, kMlpOutPad = 16;
// parameter block offsets (floats): W1[64][16] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[16][64] b4[16]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpH * kMlpInPad, kOffW2 = kOffB1 + kMlpH,
              kOffB2 = kOffW2 + kMlpH * kMlpH, kOffW3 = kOffB2 + kMlpH,
              kOffB3 = kOffW3 + kMlpH * kMlpH, kOffW4 = kOffB3 + kMlpH,
              kOffB4 = kOffW4 + kMlpOutPad * kMlpH, kMlpParams = kOffB4 + kMlpOutPad;
constexpr int kMlpNorm = 2 * kMlpIn + 2 * kMlpOut;  // in mean, in std, out mean, out std
constexpr int kHarvestWidth = kMlpIn + kMlpOut;    // floats per harvested pair

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) {
  return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// GELU, tanh form, and its derivative
__device__ __forceinline__ float gelu(float x) {
  const float t = tanhf(0.7978845608f * (x + 0.044715f * x * x * x));
  return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_d(float x) {
  const float t = tanhf(0.7978845608f * (x + 0.044715f * x * x * x));
  return 0.5f * (1.0f + t) +
         0.5f * x * (1.0f - t * t) * 0.7978845608f * (1.0f + 3.0f * 0.044715f * x * x);
}

struct MlpDev {
  const float* w;       // fp32 master parameters (kMlpParams)
  const short* wb;      // bf16 copies of the weights (same offsets)
  const float* norm;    // kMlpNorm
};

// A operand of W (rows out, cols in, row-major, ld = in): block (mb, kb)
__device__ __forceinline__ bf16x4 load_a(const short* W, int ld, int mb, int kb) {
  const int l = threadIdx.x & 63, i = l & 15, g = l >> 4;
  const short* p = W + (size_t)(mb * 16 + i) * ld + kb * 16 + 4 * g;
  return bf16x4{p[0], p[1], p[2], p[3]};
}
// A operand of W^T (rows in, cols out): A(i, 4g+j) = W[kb*16+4g+j][mb*16+i]
__device__ __forceinline__ bf16x4 load_at(const short* W, int ld, int mb, int kb) {
  const int l = threadIdx.x & 63, i = l & 15, g = l >> 4;
  const short* p = W + (size_t)(kb * 16 + 4 * g) * ld + mb * 16 + i;
  return bf16x4{p[0], p[ld], p[2 * ld], p[3 * ld]};
}

// Forward of one wavefront's 64 samples.  xb[nb] = B operand of the input
// block nb (features 4g..4g+3 of sample nb*16+i).  Returns the outputs in
// D layout (y[nb][j]: output 4g+j of sample nb*16+i).  When `keep` is
// non-null the pre-activations and activations are written for the backward
// pass: z_l / a_l as [sample][feature] for the wavefront's samples.
struct MlpKeep {
  float* z[3];   // hidden pre-activations, [64][kMlpH] per wavefront
  short* a[4];   // a[0]: inputs [64][16]; a[1..3]: hidden activations [64][kMlpH]
};

__device__ __forceinline__ void mlp_forward_wave(const MlpDev& M, const bf16x4* xb,
                                                 f32x4* y, const MlpKeep* keep) {
  const int l = threadIdx.x & 63, i = l & 15, g = l >> 4;
  bf16x4 cur[4][4];  // [kb][nb] B operands of the current layer's input
  bf16x4 nxt[4][4];
  // layer 1: one K block
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    const bf16x4 a = load_a(M.wb + kOffW1, kMlpInPad, mb, 0);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma16(a, xb[nb], acc);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int o = mb * 16 + 4 * g + j;
        const float z = acc[j] + M.w[kOffB1 + o];
        const float v = gelu(z);
        const short vb = f2bf(v);
        nxt[mb][nb][j] = vb;
        if (keep) {
          keep->z[0][(nb * 16 + i) * kMlpH + o] = z;
          keep->a[1][(nb * 16 + i) * kMlpH + o] = vb;
        }
      }
    }
  }
  // layers 2 and 3
#pragma unroll
  for (int layer = 0; layer < 2; layer++) {
    const int offW = layer == 0 ? kOffW2 : kOffW3, offB = layer == 0 ? kOffB2 : kOffB3;
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
#pragma unroll
      for (int nb = 0; nb < 4; nb++) cur[kb][nb] = nxt[kb][nb];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      f32x4 acc[4];
#pragma unroll
      for (int nb = 0; nb < 4; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        const bf16x4 a = load_a(M.wb + offW, kMlpH, mb, kb);
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[nb] = mfma16(a, cur[kb][nb], acc[nb]);
      }
#pragma unroll
      for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int o = mb * 16 + 4 * g + j;
          const float z = acc[nb][j] + M.w[offB + o];
          const short vb = f2bf(gelu(z));
          nxt[mb][nb][j] = vb;
          if (keep) {
            keep->z[layer + 1][(nb * 16 + i) * kMlpH + o] = z;
            keep->a[layer + 2][(nb * 16 + i) * kMlpH + o] = vb;
          }
        }
    }
  }
  // output layer: one M block
#pragma unroll
  for (int nb = 0; nb < 4; nb++) y[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    const bf16x4 a = load_a(M.wb + kOffW4, kMlpH, 0, kb);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) y[nb] = mfma16(a, nxt[kb][nb], y[nb]);
  }
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int j = 0; j < 4; j++) y[nb][j] += M.w[kOffB4 + 4 * g + j];
}

// Gather B operands of the input block from a [64][16] bf16 LDS tile.
__device__ __forceinline__ void input_operands(const short (*xs)[kMlpInPad], bf16x4* xb) {
  const int l = threadIdx.x & 63, i = l & 15, g = l >> 4;
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
    xb[nb] = bf16x4{xs[nb * 16 + i][4 * g], xs[nb * 16 + i][4 * g + 1],
                    xs[nb * 16 + i][4 * g + 2], xs[nb * 16 + i][4 * g + 3]};
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// layout check of the matrix-core operands: D = A(16x16) . B(16x16), both
// row-major bf16, one wavefront
template <int = 0>
__global__ void __launch_bounds__(64) k_mfma_probe(const short* A, const short* B, float* D) {
  const int l = threadIdx.x, i = l & 15, g = l >> 4;
  const bf16x4 a{A[i * 16 + 4 * g], A[i * 16 + 4 * g + 1], A[i * 16 + 4 * g + 2],
                 A[i * 16 + 4 * g + 3]};
  const bf16x4 b{B[(4 * g) * 16 + i], B[(4 * g + 1) * 16 + i], B[(4 * g + 2) * 16 + i],
                 B[(4 * g + 3) * 16 + i]};
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = mfma16(a, b, d);
#pragma unroll
  for (int j = 0; j < 4; j++) D[(4 * g + j) * 16 + i] = d[j];
}

// inference on standardised inputs x[n][16] -> y[n][2] (standardised)
template <int = 0>
__global__ void __launch_bounds__(64) k_mlp_forward(MlpDev M, const float* x, int n, float* y) {
  __shared__ short xs[64][kMlpInPad];
  const int l = threadIdx.x;
  const int s0 = blockIdx.x * 64;
  for (int f = 0; f < kMlpInPad; f++)
    xs[l][f] = (s0 + l < n) ? f2bf(x[(size_t)(s0 + l) * kMlpInPad + f]) : (short)0;
  __syncthreads();
  bf16x4 xb[4];
  input_operands(xs, xb);
  f32x4 out[4];
  mlp_forward_wave(M, xb, out, nullptr);
  const int i = l & 15, g = l >> 4;
  if (g == 0)
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int s = s0 + nb * 16 + i;
      if (s < n)
        for (int o = 0; o < kMlpOut; o++) y[(size_t)s * kMlpOut + o] = out[nb][o];
    }
}

// Training buffers for one minibatch of up to `cap` samples.
struct MlpTrainBuf {
  float* z[3];     // [cap][64]
  short* a[4];     // a[0]: [cap][16], a[1..3]: [cap][64]
  short* dz[4];    // dz[0..2]: layers 1..3 [cap][64]; dz[3]: output [cap][16]
  float* grad;     // kMlpParams
  float* loss;     // 1
};

// forward + backward of the activations for a minibatch x[n][16], t[n][2]
// (standardised); loss = mean over samples of the squared error summed over
// the two outputs.  One wavefront per 64 samples.
template <int = 0>
__global__ void __launch_bounds__(64) k_mlp_fwd_bwd(MlpDev M, MlpTrainBuf T, const float* x,
                                                    const float* t, int n) {
  __shared__ short xs[64][kMlpInPad];
  const int l = threadIdx.x, i = l & 15, g = l >> 4;
  const int s0 = blockIdx.x * 64;
  for (int f = 0; f < kMlpInPad; f++) {
    const short v = (s0 + l < n) ? f2bf(x[(size_t)(s0 + l) * kMlpInPad + f]) : (short)0;
    xs[l][f] = v;
    T.a[0][(size_t)(s0 + l) * kMlpInPad + f] = v;
  }
  __syncthreads();
  bf16x4 xb[4];
  input_operands(xs, xb);
  MlpKeep keep;
  for (int k = 0; k < 3; k++) keep.z[k] = T.z[k] + (size_t)s0 * kMlpH;
  keep.a[0] = T.a[0] + (size_t)s0 * kMlpInPad;
  for (int k = 1; k < 4; k++) keep.a[k] = T.a[k] + (size_t)s0 * kMlpH;
  f32x4 y[4];
  mlp_forward_wave(M, xb, y, &keep);
  // output gradient and loss
  const float inv_n = 1.0f / (float)n;
  float lsum = 0.f;
  bf16x4 d[4][4];  // [kb][nb] B operands of the gradient flowing down
  bf16x4 dout[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const int s = s0 + nb * 16 + i;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int o = 4 * g + j;
      float gr = 0.f;
      if (o < kMlpOut && s < n) {
        const float e = y[nb][j] - t[(size_t)s * kMlpOut + o];
        lsum += e * e;
        gr = 2.0f * e * inv_n;
      }
      const short gb = f2bf(gr);
      dout[nb][j] = gb;
      T.dz[3][(size_t)(nb * 16 + i + s0) * kMlpOutPad + o] = gb;
    }
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (l == 0) atomicAdd(T.loss, lsum * inv_n);
  // dA3^T = W4^T . dZ4 (one K block), dZ3 = dA3 * gelu'(Z3)
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    const bf16x4 a = load_at(M.wb + kOffW4, kMlpH, mb, 0);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma16(a, dout[nb], acc);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const size_t e = (size_t)(s0 + nb * 16 + i) * kMlpH + mb * 16 + 4 * g + j;
        const short v = f2bf(acc[j] * gelu_d(T.z[2][e]));
        d[mb][nb][j] = v;
        T.dz[2][e] = v;
      }
    }
  }
  // layers 3 -> 2 -> 1
#pragma unroll
  for (int layer = 1; layer >= 0; layer--) {
    const short* W = M.wb + (layer == 1 ? kOffW3 : kOffW2);
    bf16x4 cur[4][4];
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
#pragma unroll
      for (int nb = 0; nb < 4; nb++) cur[kb][nb] = d[kb][nb];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      f32x4 acc[4];
#pragma unroll
      for (int nb = 0; nb < 4; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        const bf16x4 a = load_at(W, kMlpH, mb, kb);
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[nb] = mfma16(a, cur[kb][nb], acc[nb]);
      }
#pragma unroll
      for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const size_t e = (size_t)(s0 + nb * 16 + i) * kMlpH + mb * 16 + 4 * g + j;
          const short v = f2bf(acc[nb][j] * gelu_d(T.z[layer][e]));
          d[mb][nb][j] = v;
          T.dz[layer][e] = v;
        }
    }
  }
}

// Weight and bias gradients over the minibatch: one wavefront per 16x16
// tile of a weight matrix, K = samples.  dW_l = dZ_l^T(out x samples) .
// A_{l-1}(samples x in).  Tiles: W1 4, W2 16, W3 16, W4 4.
template <int = 0>
__global__ void __launch_bounds__(64) k_mlp_wgrad(MlpTrainBuf T, int n) {
  const int l = threadIdx.x, i = l & 15, g = l >> 4;
  int tile = blockIdx.x, layer, mb, nb;
  if (tile < 4) { layer = 0; mb = tile; nb = 0; }
  else if (tile < 20) { layer = 1; mb = (tile - 4) / 4; nb = (tile - 4) % 4; }
  else if (tile < 36) { layer = 2; mb = (tile - 20) / 4; nb = (tile - 20) % 4; }
  else { layer = 3; mb = 0; nb = tile - 36; }
  const short* dz = T.dz[layer];
  const int ldz = (layer == 3) ? kMlpOutPad : kMlpH;
  const short* a = T.a[layer];
  const int lda = (layer == 0) ? kMlpInPad : kMlpH;
  const int ldw = lda;
  const int offW = layer == 0 ? kOffW1 : layer == 1 ? kOffW2 : layer == 2 ? kOffW3 : kOffW4;
  const int offB = layer == 0 ? kOffB1 : layer == 1 ? kOffB2 : layer == 2 ? kOffB3 : kOffB4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const int nk = (n + 15) / 16;
  for (int kb = 0; kb < nk; kb++) {
    bf16x4 A, B;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int s = kb * 16 + 4 * g + j;
      const bool in = s < n;
      A[j] = in ? dz[(size_t)s * ldz + mb * 16 + i] : (short)0;
      B[j] = in ? a[(size_t)s * lda + nb * 16 + i] : (short)0;
      bsum += in ? bf2f(A[j]) : 0.f;
    }
    acc = mfma16(A, B, acc);
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
    T.grad[offW + (size_t)(mb * 16 + 4 * g + j) * ldw + nb * 16 + i] = acc[j];
  if (nb == 0) {
    bsum += __shfl_xor(bsum, 16, 64);
    bsum += __shfl_xor(bsum, 32, 64);
    if (g == 0) T.grad[offB + mb * 16 + i] = bsum;
  }
}

// Adam on the fp32 master weights, bf16 copies refreshed.  Padded rows and
// columns (inputs 9..15, outputs 2..15) have zero gradients and stay zero.
template <int = 0>
__global__ void k_mlp_adam(float* w, short* wb, const float* grad, float* m1, float* m2,
                           float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const float gr = grad[p];
  const float a = b1 * m1[p] + (1.f - b1) * gr;
  const float b = b2 * m2[p] + (1.f - b2) * gr * gr;
  m1[p] = a;
  m2[p] = b;
  const float v = w[p] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[p] = v;
  wb[p] = f2bf(v);
}

template <int = 0>
__global__ void k_mlp_cast(const float* w, short* wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) wb[p] = f2bf(w[p]);
}

}  // namespace imexsw
