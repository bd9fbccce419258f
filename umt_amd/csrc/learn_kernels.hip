// Learned scalar-flux correction: a per-(zone, group) 3-layer MLP
// (16 feature slots -> 64 -> 64 -> 1, tanh) whose output r multiplies the
// change a sweep made, phi <- phi_half + r (phi_half - phi_prev).
//
// Inference and training run on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulate). Every layer is computed transposed (hidden x samples), so a
// layer's 16x16 output tiles, which leave lane l holding samples l%16 and
// hidden units 16T + 4(l/16) + j, are consumed by the next product as its B
// operand without a round trip through LDS: the 32 k-slots of a lane are
// assigned to exactly the hidden units it already holds, and the weight
// operands are packed with the same k-slot assignment (mlp_pack). Weight
// gradients reduce over samples, so there the samples are the k dimension
// and the activations pass through LDS once per 32 samples.
//
// Parameters are FP32 masters (Adam); MFMA operands are BF16 rounded to
// nearest even. The rounding points are: features, W1, W2, H1 (as the
// layer-2 operand), dZ2 and dZ1 (as operands of the backward products).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace umt {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kLd = 40;  // LDS row stride (bf16) of the sample-major staging tiles

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void learn_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ inline unsigned bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                 // round to nearest even
  return u >> 16;
}

__device__ inline unsigned pack2(float lo, float hi) { return bf16_bits(lo) | (bf16_bits(hi) << 16); }

__device__ inline bf16x8 pack8(const float* v) {
  u32x4 u;
  u[0] = pack2(v[0], v[1]);
  u[1] = pack2(v[2], v[3]);
  u[2] = pack2(v[4], v[5]);
  u[3] = pack2(v[6], v[7]);
  return __builtin_bit_cast(bf16x8, u);
}

__device__ inline bf16x8 as_bf16x8(u32x4 u) { return __builtin_bit_cast(bf16x8, u); }

__device__ inline floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// k-slot (block kb, lane group q, item i) of a lane <-> the hidden unit that
// lane holds in item (i % 4) of output tile 2 kb + i / 4.
__host__ __device__ inline int slot_hidden(int kb, int q, int i) {
  return 16 * (2 * kb + i / 4) + 4 * q + (i & 3);
}

// Packed operands (u32x4 per lane): W1^T tiles [T][lane], W2^T [kb][To][lane],
// W2 [kb][Ti][lane].
constexpr int kOpW1T = 0;
constexpr int kOpW2T = 4 * 64;
constexpr int kOpW2N = kOpW2T + 8 * 64;
static_assert(kMlpPacked == (kOpW2N + 8 * 64) * 8, "packed operand size");

struct Net {
  const u32x4* op;  // LDS copy of the packed operands
  const float* b1;
  const float* b2;
  const float* w3;
  float b3;
};

// Forward pass of 16 samples (lane l: sample l%16). x: the lane's B operand
// of layer 1 (features 8(l/16) .. +7; zero for l/16 >= 2). Leaves h1, h2
// (tile T, item j: hidden 16T + 4(l/16) + j) and r in every lane.
__device__ inline float forward16(const Net& n, bf16x8 x, int lane, float h1[4][4], float h2[4][4]) {
  const int q = lane >> 4;
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(as_bf16x8(n.op[kOpW1T + T * 64 + lane]), x, acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) h1[T][j] = tanhf(acc[j] + n.b1[16 * T + 4 * q + j]);
  }
  bf16x8 hb[2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = h1[2 * kb + i / 4][i & 3];
    hb[kb] = pack8(v);
  }
  float part = 0.f;
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) acc = mfma(as_bf16x8(n.op[kOpW2T + (kb * 4 + T) * 64 + lane]), hb[kb], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      h2[T][j] = tanhf(acc[j] + n.b2[16 * T + 4 * q + j]);
      part += n.w3[16 * T + 4 * q + j] * h2[T][j];
    }
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  return part + n.b3;
}

__device__ inline bf16x8 load_features(const float* __restrict__ X, long s, bool valid, int q, float* xf) {
#pragma unroll
  for (int i = 0; i < 8; ++i) xf[i] = 0.f;
  if (valid && q < 2) {
    const float4* p = reinterpret_cast<const float4*>(X + s * kMlpFeatures + 8 * q);
    const float4 u = p[0], w = p[1];
    xf[0] = u.x; xf[1] = u.y; xf[2] = u.z; xf[3] = u.w;
    xf[4] = w.x; xf[5] = w.y; xf[6] = w.z; xf[7] = w.w;
  }
  return pack8(xf);
}

__device__ inline void load_net(const unsigned short* __restrict__ packed,
                                const float* __restrict__ params, u32x4* sOp, float* sB1,
                                float* sB2, float* sW3, Net& n) {
  const u32x4* src = reinterpret_cast<const u32x4*>(packed);
  for (int i = threadIdx.x; i < kMlpPacked / 8; i += blockDim.x) sOp[i] = src[i];
  for (int i = threadIdx.x; i < kMlpHidden; i += blockDim.x) {
    sB1[i] = params[kMlpB1 + i];
    sB2[i] = params[kMlpB2 + i];
    sW3[i] = params[kMlpW3 + i];
  }
  __syncthreads();
  n.op = sOp;
  n.b1 = sB1;
  n.b2 = sB2;
  n.w3 = sW3;
  n.b3 = params[kMlpB3];
}

}  // namespace

// Gathers each packed operand slot from the FP32 parameters.
__global__ __launch_bounds__(256) void mlp_pack(const float* __restrict__ params,
                                                unsigned short* __restrict__ packed) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kMlpPacked) return;
  const int i = idx & 7;
  const int slot = idx >> 3;  // u32x4 index
  const int lane = slot & 63, q = lane >> 4, n16 = lane & 15;
  float v;
  if (slot < kOpW2T) {
    const int T = slot >> 6;
    const int f = 8 * q + i;
    v = f < kMlpFeatures ? params[kMlpW1 + f * kMlpHidden + 16 * T + n16] : 0.f;
  } else if (slot < kOpW2N) {
    const int t = (slot - kOpW2T) >> 6, kb = t >> 2, To = t & 3;
    v = params[kMlpW2 + slot_hidden(kb, q, i) * kMlpHidden + 16 * To + n16];
  } else {
    const int t = (slot - kOpW2N) >> 6, kb = t >> 2, Ti = t & 3;
    v = params[kMlpW2 + (16 * Ti + n16) * kMlpHidden + slot_hidden(kb, q, i)];
  }
  packed[idx] = (unsigned short)bf16_bits(v);
}

// r(s) for n samples; 4 waves per block, 16 samples per wave per step.
__global__ __launch_bounds__(256) void mlp_forward(const float* __restrict__ X, long n,
                                                   const unsigned short* __restrict__ packed,
                                                   const float* __restrict__ params,
                                                   float* __restrict__ r) {
  __shared__ u32x4 sOp[kMlpPacked / 8];
  __shared__ float sB1[kMlpHidden], sB2[kMlpHidden], sW3[kMlpHidden];
  Net net;
  load_net(packed, params, sOp, sB1, sB2, sW3, net);
  const int lane = threadIdx.x & 63, q = lane >> 4;
  const long nTiles = (n + 15) / 16;
  // The tile index is uniform across a wave, so every MFMA runs with all lanes.
  for (long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6); t < nTiles; t += (long)gridDim.x * 4) {
    const long s = 16 * t + (lane & 15);
    const bool valid = s < n;
    float xf[8], h1[4][4], h2[4][4];
    const bf16x8 x = load_features(X, s, valid, q, xf);
    const float out = forward16(net, x, lane, h1, h2);
    if (valid && q == 0) r[s] = out;
  }
}

// Loss (1/B) sum (r a - b)^2 over a batch and its gradient. Wave w handles
// stepsPerWave steps of 32 samples, positions start + (w * stepsPerWave + j)
// * 32 + [0, 32) below end, sample ids idx[position] (identity when idx is
// null). Writes partial[w][0 .. kMlpParams) = the wave's gradient and
// partial[w][kMlpParams] = its sum of squared errors.
__global__ __launch_bounds__(256) void mlp_train(const float* __restrict__ X,
                                                 const float* __restrict__ a,
                                                 const float* __restrict__ b,
                                                 const int* __restrict__ idx, long start, long end,
                                                 int stepsPerWave, float invB,
                                                 const unsigned short* __restrict__ packed,
                                                 const float* __restrict__ params,
                                                 float* __restrict__ partial) {
  __shared__ u32x4 sOp[kMlpPacked / 8];
  __shared__ float sB1[kMlpHidden], sB2[kMlpHidden], sW3[kMlpHidden];
  __shared__ __align__(16) unsigned short sStage[4][(3 * kMlpHidden + kMlpFeatures) * kLd];
  Net net;
  load_net(packed, params, sOp, sB1, sB2, sW3, net);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, q = lane >> 4, n16 = lane & 15;
  unsigned short* H1s = sStage[wv];
  unsigned short* dZ2s = H1s + kMlpHidden * kLd;
  unsigned short* dZ1s = dZ2s + kMlpHidden * kLd;
  unsigned short* Xs = dZ1s + kMlpHidden * kLd;
  const long gw = (long)blockIdx.x * 4 + wv;

  floatx4 gW2[4][4], gW1[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    gW1[i] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 4; ++j) gW2[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  float gB1[16], gB2[16], gW3[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) gB1[i] = gB2[i] = gW3[i] = 0.f;
  float gB3 = 0.f, sse = 0.f;

  for (int j = 0; j < stepsPerWave; ++j) {
    const long base = start + (gw * stepsPerWave + j) * 32;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const long p = base + 16 * u + n16;
      const bool valid = p < end;
      const long s = valid ? (idx ? (long)idx[p] : p) : 0;
      float xf[8], h1[4][4], h2[4][4];
      const bf16x8 x = load_features(X, s, valid, q, xf);
      const int col = 16 * u + n16;
      if (q < 2) {
#pragma unroll
        for (int i = 0; i < 8; ++i) Xs[(8 * q + i) * kLd + col] = (unsigned short)bf16_bits(xf[i]);
      }
      const float r = forward16(net, x, lane, h1, h2);
      const float av = valid ? a[s] : 0.f, bv = valid ? b[s] : 0.f;
      const float e = r * av - bv;
      const float g = 2.f * av * e * invB;
      if (q == 0) {
        sse += e * e;
        gB3 += g;
      }
      float dZ2[4][4];
#pragma unroll
      for (int T = 0; T < 4; ++T)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int h = 16 * T + 4 * q + jj;
          dZ2[T][jj] = g * net.w3[h] * (1.f - h2[T][jj] * h2[T][jj]);
          gW3[4 * T + jj] += g * h2[T][jj];
          gB2[4 * T + jj] += dZ2[T][jj];
          H1s[h * kLd + col] = (unsigned short)bf16_bits(h1[T][jj]);
          dZ2s[h * kLd + col] = (unsigned short)bf16_bits(dZ2[T][jj]);
        }
      bf16x8 db[2];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        float v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = dZ2[2 * kb + i / 4][i & 3];
        db[kb] = pack8(v);
      }
#pragma unroll
      for (int Ti = 0; Ti < 4; ++Ti) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) acc = mfma(as_bf16x8(net.op[kOpW2N + (kb * 4 + Ti) * 64 + lane]), db[kb], acc);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int h = 16 * Ti + 4 * q + jj;
          const float dz1 = acc[jj] * (1.f - h1[Ti][jj] * h1[Ti][jj]);
          gB1[4 * Ti + jj] += dz1;
          dZ1s[h * kLd + col] = (unsigned short)bf16_bits(dz1);
        }
      }
    }
    __syncthreads();
    // Weight gradients over the step's 32 samples (k-slot (q, i) = sample 8q + i).
#pragma unroll
    for (int Ti = 0; Ti < 4; ++Ti) {
      const bf16x8 A = as_bf16x8(*reinterpret_cast<const u32x4*>(H1s + (16 * Ti + n16) * kLd + 8 * q));
#pragma unroll
      for (int To = 0; To < 4; ++To) {
        const bf16x8 B = as_bf16x8(*reinterpret_cast<const u32x4*>(dZ2s + (16 * To + n16) * kLd + 8 * q));
        gW2[Ti][To] = mfma(A, B, gW2[Ti][To]);
      }
    }
    {
      const bf16x8 A = as_bf16x8(*reinterpret_cast<const u32x4*>(Xs + n16 * kLd + 8 * q));
#pragma unroll
      for (int To = 0; To < 4; ++To) {
        const bf16x8 B = as_bf16x8(*reinterpret_cast<const u32x4*>(dZ1s + (16 * To + n16) * kLd + 8 * q));
        gW1[To] = mfma(A, B, gW1[To]);
      }
    }
    __syncthreads();
  }

  float* out = partial + gw * (kMlpParams + 1);
#pragma unroll
  for (int Ti = 0; Ti < 4; ++Ti)
#pragma unroll
    for (int To = 0; To < 4; ++To)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
        out[kMlpW2 + (16 * Ti + 4 * q + jj) * kMlpHidden + 16 * To + n16] = gW2[Ti][To][jj];
#pragma unroll
  for (int To = 0; To < 4; ++To)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) out[kMlpW1 + (4 * q + jj) * kMlpHidden + 16 * To + n16] = gW1[To][jj];
  // Per-hidden sums: reduce over the 16 lanes (samples) of each lane group.
#pragma unroll
  for (int k = 0; k < 16; ++k) {
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
      gB1[k] += __shfl_xor(gB1[k], m, 64);
      gB2[k] += __shfl_xor(gB2[k], m, 64);
      gW3[k] += __shfl_xor(gW3[k], m, 64);
    }
  }
#pragma unroll
  for (int m = 1; m < 16; m <<= 1) {
    gB3 += __shfl_xor(gB3, m, 64);
    sse += __shfl_xor(sse, m, 64);
  }
  if (n16 == 0) {
#pragma unroll
    for (int T = 0; T < 4; ++T)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int h = 16 * T + 4 * q + jj;
        out[kMlpB1 + h] = gB1[4 * T + jj];
        out[kMlpB2 + h] = gB2[4 * T + jj];
        out[kMlpW3 + h] = gW3[4 * T + jj];
      }
    if (q == 0) {
      out[kMlpB3] = gB3;
      out[kMlpParams] = sse;
    }
  }
}

// Sums the wave partials in wave order (gradient and squared-error sum), then
// optionally takes one Adam step (lr > 0) on the FP32 parameters.
__global__ __launch_bounds__(256) void mlp_reduce_adam(const float* __restrict__ partial, int nWaves,
                                                       float* __restrict__ grad,
                                                       float* __restrict__ params,
                                                       float* __restrict__ m, float* __restrict__ v,
                                                       float lr, float beta1, float beta2, float eps,
                                                       float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpParams) return;
  float s = 0.f;
  for (int w = 0; w < nWaves; ++w) s += partial[(long)w * (kMlpParams + 1) + p];
  grad[p] = s;
  if (p == kMlpParams || lr <= 0.f) return;
  const float mm = beta1 * m[p] + (1.f - beta1) * s;
  const float vv = beta2 * v[p] + (1.f - beta2) * s * s;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// ---------------------------------------------------------------------------
// Features of the correction network, per (zone, group), from zone averages
// of the state a sweep started from (pp), of its result (ph) and of the
// previous sweep's change (dp, may be null).

// rel = d / |ph|, ratio = clamp(d / dp, -2, 2), per-group max |rel|.
__global__ __launch_bounds__(256) void learn_prep(const double* __restrict__ ph,
                                                  const double* __restrict__ pp,
                                                  const double* __restrict__ dp, int G, long n,
                                                  double* __restrict__ rel,
                                                  double* __restrict__ ratio,
                                                  unsigned long long* __restrict__ maxRel) {
  extern __shared__ unsigned long long sMax[];
  for (int g = threadIdx.x; g < G; g += blockDim.x) sMax[g] = 0ull;
  __syncthreads();
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double d = ph[i] - pp[i];
    const double rl = d / fmax(fabs(ph[i]), kLearnTiny);
    double rt = 0.0;
    if (dp) {
      const double den = fabs(dp[i]) > kLearnTiny ? dp[i] : kLearnTiny;
      rt = fmin(fmax(d / den, -2.0), 2.0);
    }
    rel[i] = rl;
    ratio[i] = rt;
    atomicMax(&sMax[i % G], (unsigned long long)__double_as_longlong(fabs(rl)));
  }
  __syncthreads();
  for (int g = threadIdx.x; g < G; g += blockDim.x)
    if (sMax[g]) atomicMax(&maxRel[g], sMax[g]);
}

__global__ __launch_bounds__(256) void learn_features(
    const double* __restrict__ ph, const double* __restrict__ rel, const double* __restrict__ ratio,
    const unsigned long long* __restrict__ maxRel, const double* __restrict__ sigt,
    const double* __restrict__ sigs, const double* __restrict__ q,
    const int* __restrict__ zoneNbr, int G, long n, float* __restrict__ X) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long z = i / G;
  const int g = (int)(i - z * G);
  const double st = sigt[i], ss = sigs[i], c = ss / st, sa = st - ss;
  const double M = __longlong_as_double((long long)maxRel[g]) + kLearnTiny;
  double sRel = 0.0, sRatio = 0.0, sPh = 0.0, sSt = 0.0;
  int cnt = 0;
  for (int f = 0; f < 6; ++f) {
    const int j = zoneNbr[6 * z + f];
    if (j < 0) continue;
    const long k = (long)j * G + g;
    sRel += rel[k];
    sRatio += ratio[k];
    sPh += ph[k];
    sSt += sigt[k];
    ++cnt;
  }
  const double nRel = cnt ? sRel / cnt : rel[i];
  const double nRatio = cnt ? sRatio / cnt : ratio[i];
  const double nPh = cnt ? sPh / cnt : ph[i];
  const double nSt = cnt ? sSt / cnt : st;
  float* x = X + i * kMlpFeatures;
  x[0] = (float)log10(st);
  x[1] = (float)c;
  x[2] = (float)log10(fmax(1.0 - c, 1e-12));
  x[3] = (float)fmin(fmax(log(fmax(ph[i] * sa, kLearnTiny) / fmax(q[i], kLearnTiny)), -30.0), 30.0);
  x[4] = (float)(rel[i] / M);
  x[5] = (float)(nRel / M);
  x[6] = (float)ratio[i];
  x[7] = (float)nRatio;
  x[8] = (float)fmin(fmax(log(fmax(nPh, kLearnTiny) / fmax(ph[i], kLearnTiny)), -30.0), 30.0);
  x[9] = (float)log10(fmax(nSt, kLearnTiny));
#pragma unroll
  for (int k = 10; k < kMlpFeatures; ++k) x[k] = 0.f;
}

// phi <- phi + r (phi - phiPrev), r per (zone, group); phi holds the sweep's
// result. A non-finite r (or corrected value) leaves that entry uncorrected
// and raises *bad, so the iteration can stop correcting.
__global__ __launch_bounds__(256) void learn_apply(double* __restrict__ phi,
                                                   const double* __restrict__ phiPrev,
                                                   const float* __restrict__ r, int G, long n,
                                                   unsigned long long* __restrict__ bad) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c = i / G;
  const int g = (int)(i - c * G);
  const double h = phi[i];
  const double v = h + (double)r[(c >> 3) * G + g] * (h - phiPrev[i]);
  if (isfinite(v)) phi[i] = v;
  else *bad = 1ull;
}

__global__ __launch_bounds__(256) void learn_delta(const double* __restrict__ ph,
                                                   const double* __restrict__ pp, long n,
                                                   double* __restrict__ d) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = ph[i] - pp[i];
}

// Training targets of one collected state: a = d / ph, b = (pstar - ph) / ph
// (zero where ph <= 0), and per-block sums of a^2.
__global__ __launch_bounds__(256) void learn_targets(const double* __restrict__ ph,
                                                     const double* __restrict__ d,
                                                     const double* __restrict__ pstar, long n,
                                                     double* __restrict__ a,
                                                     double* __restrict__ b,
                                                     double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double av = 0.0, bv = 0.0;
    if (ph[i] > 0.0) {
      av = d[i] / ph[i];
      bv = (pstar[i] - ph[i]) / ph[i];
    }
    a[i] = av;
    b[i] = bv;
    acc += av * av;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Scales a and b by 1 / rms(a) (the block partials summed in order) into the
// FP32 training arrays.
__global__ __launch_bounds__(256) void learn_scale(const double* __restrict__ a,
                                                   const double* __restrict__ b,
                                                   const double* __restrict__ partial, int nPartial,
                                                   long n, float* __restrict__ aOut,
                                                   float* __restrict__ bOut) {
  double s = 0.0;
  for (int k = 0; k < nPartial; ++k) s += partial[k];
  const double scale = s > 0.0 ? 1.0 / sqrt(s / (double)n) : 0.0;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  aOut[i] = (float)(a[i] * scale);
  bOut[i] = (float)(b[i] * scale);
}

// ---------------------------------------------------------------------------

hipError_t launch_mlp_pack(const float* params, unsigned short* packed, hipStream_t s) {
  hipLaunchKernelGGL(mlp_pack, dim3(blocks_for(kMlpPacked, 256)), dim3(256), 0, s, params, packed);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* X, long n, const unsigned short* packed,
                              const float* params, float* r, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long tiles = (n + 15) / 16;
  const long blocks = std::min<long>((tiles + 3) / 4, 2048);
  hipLaunchKernelGGL(mlp_forward, dim3((unsigned)blocks), dim3(256), 0, s, X, n, packed, params, r);
  return hipGetLastError();
}

int mlp_train_waves(long count, int stepsPerWave) {
  const long per = 32L * stepsPerWave;
  const long waves = (count + per - 1) / per;
  return (int)((waves + 3) / 4 * 4);
}

hipError_t launch_mlp_train(const float* X, const float* a, const float* b, const int* idx,
                            long start, long count, int stepsPerWave,
                            const unsigned short* packed, const float* params, float* partial,
                            hipStream_t s) {
  const int waves = mlp_train_waves(count, stepsPerWave);
  const float invB = count > 0 ? 1.0f / (float)count : 0.f;
  hipLaunchKernelGGL(mlp_train, dim3(waves / 4), dim3(256), 0, s, X, a, b, idx, start,
                     start + count, stepsPerWave, invB, packed, params, partial);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce_adam(const float* partial, int nWaves, float* grad, float* params,
                                  float* m, float* v, float lr, float beta1, float beta2, float eps,
                                  float bc1, float bc2, hipStream_t s) {
  hipLaunchKernelGGL(mlp_reduce_adam, dim3(blocks_for(kMlpParams + 1, 256)), dim3(256), 0, s,
                     partial, nWaves, grad, params, m, v, lr, beta1, beta2, eps, bc1, bc2);
  return hipGetLastError();
}

hipError_t launch_learn_features(const double* ph, const double* pp, const double* dp,
                                 const double* sigt, const double* sigs, const double* q,
                                 const int* zoneNbr, int G, long n, double* rel, double* ratio,
                                 unsigned long long* maxRel, float* X, hipStream_t s) {
  hipLaunchKernelGGL(learn_fill_u64, dim3(blocks_for(G, 64)), dim3(64), 0, s, maxRel, G, 0ull);
  const long blocks = std::min<long>(blocks_for(n, 256), 1024);
  hipLaunchKernelGGL(learn_prep, dim3((unsigned)blocks), dim3(256), G * sizeof(unsigned long long),
                     s, ph, pp, dp, G, n, rel, ratio, maxRel);
  hipLaunchKernelGGL(learn_features, dim3(blocks_for(n, 256)), dim3(256), 0, s, ph, rel, ratio,
                     maxRel, sigt, sigs, q, zoneNbr, G, n, X);
  return hipGetLastError();
}

hipError_t launch_learn_apply(double* phi, const double* phiPrev, const float* r, int G, long n,
                              unsigned long long* bad, hipStream_t s) {
  hipLaunchKernelGGL(learn_apply, dim3(blocks_for(n, 256)), dim3(256), 0, s, phi, phiPrev, r, G, n,
                     bad);
  return hipGetLastError();
}

hipError_t launch_learn_delta(const double* ph, const double* pp, long n, double* d, hipStream_t s) {
  hipLaunchKernelGGL(learn_delta, dim3(blocks_for(n, 256)), dim3(256), 0, s, ph, pp, n, d);
  return hipGetLastError();
}

hipError_t launch_learn_targets(const double* ph, const double* d, const double* pstar, long n,
                                double* a, double* b, double* partial, float* aOut, float* bOut,
                                hipStream_t s) {
  hipLaunchKernelGGL(learn_targets, dim3(kLearnTargetBlocks), dim3(256), 0, s, ph, d, pstar, n, a,
                     b, partial);
  hipLaunchKernelGGL(learn_scale, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, b, partial,
                     kLearnTargetBlocks, n, aOut, bOut);
  return hipGetLastError();
}

}  // namespace umt
