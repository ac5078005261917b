// nn_kernels.h -- the learned shading surrogate for Cook-Torrance + IBL.
//
// A 12-64-64-64-4 ReLU network replaces the analytic evaluation of a PBR
// material (tina/matr/material.py:322-362, 701-706) under SkyboxLighting
// (tina/core/lighting.py:4-21) plus one directional light: outputs 0-2 are the
// ACES-mapped colour, output 3 estimates that colour's error, and pixels whose
// estimate exceeds a threshold are shaded analytically instead.
//
//   k_nn_samples   training pairs from the analytic shader (the same code path
//                  as k_shade<PROG_PBR, true>) at hashed random inputs
//   k_mlp_grad     forward + backward of one minibatch slice per workgroup, in
//                  float32, with the slice's gradient written to its own row
//   k_mlp_adam     sums the rows in a fixed order and takes one Adam step
//   k_mlp_pack     float32 master weights -> bf16 MFMA fragments + f32 biases
//   k_shade_nn     deferred shading through the network on MFMA
//                  (v_mfma_f32_16x16x32_bf16), fallback pixels queued
//   k_shade_list   the analytic shader over the queued pixels
//
// Parameter layout (float32, row-major): W1[64][12] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[4][64] b4[4].
#pragma once
#include "shade_kernels.h"

namespace tina {

constexpr int kNnIn = 12, kNnHid = 64, kNnOut = 4;
constexpr int kOffW1 = 0, kOffB1 = 768, kOffW2 = 832, kOffB2 = 4928, kOffW3 = 4992, kOffB3 = 9088;
constexpr int kOffW4 = 9152, kOffB4 = 9408, kNnParams = 9412;
constexpr int kNnBatchPerBlock = 64;      // samples per k_mlp_grad workgroup (eight waves)
constexpr float kHuberBeta = 0.01f;
constexpr float kErrWeight = 1.0f;

// Packed inference weights: 22 A-operand fragments (64 lanes x 8 bf16) --
// layer 1: 4 row blocks x 1 K block, layers 2 and 3: 4 x 2, layer 4: 1 x 2 --
// followed by the f32 biases b1, b2, b3 (64 each) and b4 (4).
constexpr int kNnFrags = 22;
constexpr int kNnPackBias = kNnFrags * 64 * 8;  // offset of the biases, in bf16 units
constexpr int kNnPackBytes = kNnFrags * 64 * 16 + (3 * 64 + 4) * 4;
// Fallback queue of k_shade_nn: pixel chunk c appends to segment c % kFbSeg,
// whose counter sits alone on a 128-byte line (one global atomic per chunk).
// kFbSeg counters kFbLine words apart, then the last pass's total, then the
// tickets k_shade_list's workgroups take when they finish (kFbTick of them, one
// per workgroup residue, and a top one): the last workgroup empties the counters
// and the tickets, so every call starts from zeros (a captured call replays as
// it ran).  Spreading the tickets keeps ~500 workgroups that find the queue
// empty and finish together from serialising on one word.
constexpr int kFbSeg = 32, kFbLine = 32, kFbTick = 16;
constexpr int kFbWords = (kFbSeg + 2 + kFbTick) * kFbLine;

// ---- random inputs -----------------------------------------------------------

// 32-bit integer hash (xorshift-multiply, "lowbias32")
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// uniform in [0, 1): the k-th draw of sample i under seed
__device__ __forceinline__ float urand(uint32_t seed, uint64_t i, uint32_t k) {
  const uint32_t h = hash32(seed ^ hash32((uint32_t)i ^ hash32((uint32_t)(i >> 32) + 0x9e3779b9u * (k + 1))));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ V3 unit_sphere(float u1, float u2) {
  const float z = 1.0f - 2.0f * u1;
  const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  const float phi = 6.2831855f * u2;
  return v3(r * cosf(phi), r * sinf(phi), z);
}

// The 12 network inputs: n, R = 2 (n.V) n - V (the view direction mirrored about
// n), R.L, log(roughness), metallic, albedo.
__device__ __forceinline__ void nn_features(V3 n, V3 V, V3 L, float r, float m, V3 alb, float* x) {
  const float t = 2.0f * vdot(n, V);
  const V3 R = v3(t * n.x - V.x, t * n.y - V.y, t * n.z - V.z);
  x[0] = n.x; x[1] = n.y; x[2] = n.z;
  x[3] = R.x; x[4] = R.y; x[5] = R.z;
  x[6] = vdot(R, L);
  x[7] = logf(r);
  x[8] = m;
  x[9] = alb.x; x[10] = alb.y; x[11] = alb.z;
}

// FresnelFactor of PBR (material.py:68-83, 701-706): metallic * albedo +
// (1 - metallic) * 0.16 * specular^2, in the op order of the material program.
__device__ __forceinline__ V3 pbr_f0(float m, V3 alb, float spec) {
  const V3 mt = vs(m), sp = vs(spec);
  return vadd(vmul(mt, alb), vmul(vsmul(0.16f, vsub(vs(1.0f), mt)), vmul(sp, sp)));
}

// Training pairs.  X[i] = 12 features, Y[i] = (ACES colour, 0).  n uniform on the
// sphere, V uniform on the hemisphere around n, roughness U[r_lo, r_hi],
// metallic Bernoulli(p_metal) (uniform when p_metal < 0), albedo U[0,1]^3.
__global__ void __launch_bounds__(256)
k_nn_samples(const ShadeArgs a_by_value, uint32_t seed, long nsamples, float r_lo, float r_hi, float p_metal,
             float spec, float* __restrict__ X, float* __restrict__ Y) {
  (void)a_by_value;
#if defined(__HIP_DEVICE_COMPILE__)
  const ShadeArgs& a = *(const ShadeArgs*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const ShadeArgs& a = a_by_value;
#endif
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsamples) return;
  const V3 n = unit_sphere(urand(seed, i, 0), urand(seed, i, 1));
  V3 V = unit_sphere(urand(seed, i, 2), urand(seed, i, 3));
  if (vdot(V, n) < 0.0f) V = vneg(V);
  const float r = r_lo + (r_hi - r_lo) * urand(seed, i, 4);
  const float u = urand(seed, i, 5);
  const float m = p_metal < 0.0f ? u : (u < p_metal ? 1.0f : 0.0f);
  const V3 alb = v3(urand(seed, i, 6), urand(seed, i, 7), urand(seed, i, 8));
  const V3 f0 = pbr_f0(m, alb, spec);
  V3 c = shade_core<PROG_PBR, true>(a, n, vs(0.0f), V, alb, f0, f0, r);
  c = aces(c);
  const V3 L = light_at(a.lights, 0).xyz;
  float x[kNnIn];
  nn_features(n, V, L, r, m, alb, x);
#pragma unroll
  for (int k = 0; k < kNnIn; k++) X[i * kNnIn + k] = x[k];
  Y[i * 4 + 0] = c.x;
  Y[i * 4 + 1] = c.y;
  Y[i * 4 + 2] = c.z;
  Y[i * 4 + 3] = 0.0f;
}

// The analytic shader on feature rows: the Cook-Torrance + IBL evaluation the
// network replaces, on the same inputs (n, V = 2 (n.R) n - R, roughness =
// exp(x7), metallic, albedo), for comparing the two evaluations alone.
__global__ void __launch_bounds__(256)
k_nn_analytic_rows(const ShadeArgs a_by_value, const float* __restrict__ X, long n, float spec,
                   float* __restrict__ O) {
  (void)a_by_value;
#if defined(__HIP_DEVICE_COMPILE__)
  const ShadeArgs& a = *(const ShadeArgs*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const ShadeArgs& a = a_by_value;
#endif
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* x = X + i * kNnIn;
  const V3 nrm = v3(x[0], x[1], x[2]), R = v3(x[3], x[4], x[5]);
  const float t = 2.0f * vdot(nrm, R);
  const V3 V = v3(t * nrm.x - R.x, t * nrm.y - R.y, t * nrm.z - R.z);
  const float r = expf(x[7]), m = x[8];
  const V3 alb = v3(x[9], x[10], x[11]);
  const V3 f0 = pbr_f0(m, alb, spec);
  const V3 c = aces(shade_core<PROG_PBR, true>(a, nrm, vs(0.0f), V, alb, f0, f0, r));
  O[i * 4 + 0] = c.x;
  O[i * 4 + 1] = c.y;
  O[i * 4 + 2] = c.z;
  O[i * 4 + 3] = 0.0f;
}

// ---- training ------------------------------------------------------------------
//
// One workgroup (eight waves) per 64 samples; every product of the forward and
// backward pass is a 16 x 16 tile of v_mfma_f32_16x16x4_f32 (float32 operands
// and accumulation), lane l holding A[l % 16][l / 16], B[l / 16][l % 16] and
// D[4 (l / 16) + q][l % 16], q = 0..3.  Activations live in LDS as one row per
// sample (stride kGradRow floats: a row-block read of 16-byte vectors is
// bank-conflict free), and each delta overwrites the activation rows it no
// longer needs.  The K order inside a product is free as long as both operands
// use it: layers 2-4 take input neuron c = 16 (l / 16) + kk at step kk, so a
// lane reads its B values as four 16-byte vectors, and the weight gradients sum
// over samples in the order of grad_sample.  Each wave keeps the weights and
// biases it uses in registers, loaded once before the first product.  The workgroup's gradient
// goes to row blockIdx.x of `partial`.
//
// Loss per sample: mean_c smoothL1(o_c - y_c) + kErrWeight/2 (o_3 - max_c|o_c - y_c|)^2,
// averaged over the `batch` samples of the step; the error term trains W4[3], b4[3] only.

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGradRow = 68;  // LDS floats per sample row of an activation / delta matrix

__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Hout[s][16m + n] = relu(b0 + sum_c W[16m + n][c] Hin[s][c]) for the 32 samples
// of tiles sb0, sb0 + 1, K = 64 in the order c = 16 k + kk.
__device__ __forceinline__ void grad_dense(const float (&w)[16], const f32x4 b0, const float* Hin, float* Hout, int m,
                                           int sb0, int i, int k) {
  f32x4 acc[2] = {b0, b0};
#pragma unroll
  for (int v = 0; v < 4; v++) {
    f32x4 r[2];
#pragma unroll
    for (int t = 0; t < 2; t++) r[t] = *(const f32x4*)(Hin + (16 * (sb0 + t) + i) * kGradRow + 16 * k + 4 * v);
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int t = 0; t < 2; t++) acc[t] = mfma_f32(w[4 * v + e], r[t][e], acc[t]);
  }
#pragma unroll
  for (int t = 0; t < 2; t++) {
    f32x4 h;
#pragma unroll
    for (int q = 0; q < 4; q++) h[q] = fmaxf(acc[t][q], 0.0f);
    *(f32x4*)(Hout + (16 * (sb0 + t) + i) * kGradRow + 16 * m + 4 * k) = h;
  }
}

// D[s][16m + n] = (H[s][16m + n] > 0) * sum_c W[c][16m + n] Dn[s][c] (A = W^T from
// registers) for the samples of tiles sb0, sb0 + 1, written over H.
__device__ __forceinline__ void grad_delta(const float (&wt)[16], const float* Dn, float* H, int m, int sb0, int i,
                                           int k) {
  f32x4 acc[2];
#pragma unroll
  for (int t = 0; t < 2; t++) acc[t] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int v = 0; v < 4; v++) {
    f32x4 r[2];
#pragma unroll
    for (int t = 0; t < 2; t++) r[t] = *(const f32x4*)(Dn + (16 * (sb0 + t) + i) * kGradRow + 16 * k + 4 * v);
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int t = 0; t < 2; t++) acc[t] = mfma_f32(wt[4 * v + e], r[t][e], acc[t]);
  }
#pragma unroll
  for (int t = 0; t < 2; t++) {
    f32x4* hp = (f32x4*)(H + (16 * (sb0 + t) + i) * kGradRow + 16 * m + 4 * k);
    const f32x4 h = *hp;
    f32x4 d;
#pragma unroll
    for (int q = 0; q < 4; q++) d[q] = h[q] > 0.0f ? acc[t][q] : 0.0f;
    *hp = d;
  }
}

// Sample of step kk for lanes l / 16 = k in the weight-gradient sums: the
// rotation by 4k puts the four k-groups' rows on disjoint LDS banks.
__device__ __forceinline__ int grad_sample(int kk, int k) { return 16 * k + ((kk + 4 * k) & 15); }

// dW[16n + j][16m + r] = sum_s D[s][16n + j] H[s][16m + r] for input block m and
// output blocks n0, n0 + 1, written transposed from the D layout (float4 per
// lane) to the weight's rows of `outW` (row length kin).
__device__ __forceinline__ void grad_weights(const float* H, int hstride, const float* D, float* __restrict__ outW,
                                             int kin, int m, int n0, int i, int k) {
  f32x4 acc[2];
#pragma unroll
  for (int t = 0; t < 2; t++) acc[t] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kk = 0; kk < 16; kk++) {
    const int s = grad_sample(kk, k);
    const float a = H[s * hstride + 16 * m + i];
#pragma unroll
    for (int t = 0; t < 2; t++) acc[t] = mfma_f32(a, D[s * kGradRow + 16 * (n0 + t) + i], acc[t]);
  }
  const int r = 16 * m + 4 * k;  // input neurons r..r+3 of output neuron 16n + i
#pragma unroll
  for (int t = 0; t < 2; t++)
    if (r < kin) *(f32x4*)(outW + (16 * (n0 + t) + i) * kin + r) = acc[t];
}

// sum over the 16 sample steps of one tile as four interleaved chains (steps
// kk = 4t + c into chain c), added in chain order
__device__ __forceinline__ f32x4 chains4(const f32x4 (&a)[4]) { return ((a[0] + a[1]) + a[2]) + a[3]; }

// the bias gradient sum_s D[s][o] of outputs 8w..8w+7: lane l sums samples
// (l >> 3) + 8t, t = 0..7, of output 8w + (l & 7) in order, then the eight lane
// groups are added by a fixed butterfly (deterministic)
__device__ __forceinline__ void grad_bias(const float* D, float* __restrict__ outB, int w, int lane) {
  const int o = 8 * w + (lane & 7), s0 = lane >> 3;
  float v[8];
#pragma unroll
  for (int t = 0; t < 8; t++) v[t] = D[(s0 + 8 * t) * kGradRow + o];
  float sb = v[0];
#pragma unroll
  for (int t = 1; t < 8; t++) sb += v[t];
  sb += __shfl_xor(sb, 8);
  sb += __shfl_xor(sb, 16);
  sb += __shfl_xor(sb, 32);
  if (lane < 8) outB[o] = sb;
}

constexpr int kGradBlock = 512;  // eight waves: two per SIMD, each with half of a layer's tiles
static_assert(64 * 16 == 2 * kGradBlock, "the feature and gradient staging takes two passes");

__global__ void __launch_bounds__(kGradBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_mlp_grad(const float* __restrict__ params, const float* __restrict__ X, const float* __restrict__ Y,
           long first, long nsamples, int batch, float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) float sX[64 * 16];   // [s][16]: features, 12..15 zero
  __shared__ __attribute__((aligned(16))) float sH1[64 * kGradRow], sH2[64 * kGradRow], sH3[64 * kGradRow];
  __shared__ __attribute__((aligned(16))) float sGo[64 * 16];  // [s][16]: output gradients, 4..15 zero
  __shared__ __attribute__((aligned(16))) float sP[kNnParams];   // the parameters, staged once
  const int tid = threadIdx.x, lane = tid & 63, i = lane & 15, k = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m = w & 3, h2 = 2 * (w >> 2);  // this wave's neuron block and first tile of the other dimension
  const long s0 = first + (long)blockIdx.x * 64;
  float* __restrict__ P = partial + (size_t)blockIdx.x * kNnParams;

  // parameters, features and targets in with coalesced loads, all issued before the first store
  constexpr int kPv = (kNnParams / 4 + kGradBlock - 1) / kGradBlock;  // float4 loads per thread
  f32x4 pv[kPv];
#pragma unroll
  for (int t = 0; t < kPv; t++) {
    const int e = tid + t * kGradBlock;
    if (e < kNnParams / 4) pv[t] = ((const f32x4*)params)[e];
  }
  float xv[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int e = tid + t * kGradBlock, s = e >> 4, f = e & 15;
    xv[t] = (f < kNnIn && s0 + s < nsamples) ? X[(s0 + s) * kNnIn + f] : 0.0f;
  }
#pragma unroll
  for (int t = 0; t < kPv; t++) {
    const int e = tid + t * kGradBlock;
    if (e < kNnParams / 4) ((f32x4*)sP)[e] = pv[t];
  }
#pragma unroll
  for (int t = 0; t < 2; t++) {
    sX[tid + t * kGradBlock] = xv[t];
    sGo[tid + t * kGradBlock] = 0.0f;
  }
  // the targets of the 16 samples of waves 0-3 (their lanes k == 0 compute the output gradient)
  const long sg = s0 + 16 * m + i;
  const bool valid = sg < nsamples;
  float y[3] = {0.0f, 0.0f, 0.0f};
  if (w < 4 && k == 0 && valid)
#pragma unroll
    for (int c = 0; c < 3; c++) y[c] = Y[sg * 4 + c];
  __syncthreads();

  // the weights and biases this wave uses, in the K order of each product
  const float* params_l = sP;
  float w1[3], w2[16], w3[16], w4[16], w4t[1], w3t[16], w2t[16];
#pragma unroll
  for (int kk = 0; kk < 3; kk++) w1[kk] = params_l[kOffW1 + (16 * m + i) * kNnIn + 4 * kk + k];
#pragma unroll
  for (int kk = 0; kk < 16; kk++) {
    w2[kk] = params_l[kOffW2 + (16 * m + i) * kNnHid + 16 * k + kk];
    w3[kk] = params_l[kOffW3 + (16 * m + i) * kNnHid + 16 * k + kk];
    w4[kk] = i < kNnOut ? params_l[kOffW4 + i * kNnHid + 16 * k + kk] : 0.0f;
    w3t[kk] = params_l[kOffW3 + (16 * k + kk) * kNnHid + 16 * m + i];
    w2t[kk] = params_l[kOffW2 + (16 * k + kk) * kNnHid + 16 * m + i];
  }
  w4t[0] = k < 3 ? params_l[kOffW4 + k * kNnHid + 16 * m + i] : 0.0f;  // the colour rows of W4 only
  const f32x4 bias1 = *(const f32x4*)(params_l + kOffB1 + 16 * m + 4 * k);
  const f32x4 bias2 = *(const f32x4*)(params_l + kOffB2 + 16 * m + 4 * k);
  const f32x4 bias3 = *(const f32x4*)(params_l + kOffB3 + 16 * m + 4 * k);
  const f32x4 bias4 = *(const f32x4*)(params_l + kOffB4);
  // forward: layer 1 (K = 12 in the order f = 4 kk + k), layers 2 and 3
  {
    f32x4 a1[2] = {bias1, bias1};
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
#pragma unroll
      for (int t = 0; t < 2; t++) a1[t] = mfma_f32(w1[kk], sX[(16 * (h2 + t) + i) * 16 + 4 * kk + k], a1[t]);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      f32x4 hv;
#pragma unroll
      for (int q = 0; q < 4; q++) hv[q] = fmaxf(a1[t][q], 0.0f);
      *(f32x4*)(sH1 + (16 * (h2 + t) + i) * kGradRow + 16 * m + 4 * k) = hv;
    }
  }
  __syncthreads();
  grad_dense(w2, bias2, sH1, sH2, m, h2, i, k);
  __syncthreads();
  grad_dense(w3, bias3, sH2, sH3, m, h2, i, k);
  __syncthreads();
  // layer 4 on the 16 samples of tile m (waves 0-3), then the output gradient
  // (lanes k == 0 hold outputs 0..3)
  if (w < 4) {
    f32x4 ch[4];
#pragma unroll
    for (int c = 0; c < 4; c++) ch[c] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    if (k == 0) ch[0] = bias4;
    const f32x4* row = (const f32x4*)(sH3 + (16 * m + i) * kGradRow + 16 * k);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const f32x4 r = row[v];
#pragma unroll
      for (int e = 0; e < 4; e++) ch[e] = mfma_f32(w4[4 * v + e], r[e], ch[e]);
    }
    const f32x4 acc = chains4(ch);
    if (k == 0) {
      const float inv_b = 1.0f / (float)batch;
      float e = 0.0f;
      f32x4 go;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const float d = acc[c] - y[c];
        e = fmaxf(e, fabsf(d));
        const float hg = fabsf(d) < kHuberBeta ? d / kHuberBeta : (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f));
        go[c] = valid ? hg * (inv_b / 3.0f) : 0.0f;
      }
      go[3] = valid ? kErrWeight * (acc[3] - e) * inv_b : 0.0f;
      *(f32x4*)(sGo + (16 * m + i) * 16) = go;
    }
  }
  __syncthreads();

  // backward: W4, b4 from the output gradient and h3 (waves 0-3: input block m)
  if (w < 4) {
    f32x4 ch[4];
#pragma unroll
    for (int c = 0; c < 4; c++) ch[c] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      const int s = grad_sample(kk, k);
      ch[kk & 3] = mfma_f32(sH3[s * kGradRow + 16 * m + i], sGo[s * 16 + i], ch[kk & 3]);
    }
    const f32x4 acc = chains4(ch);
    // D[4k + q][i] = dW4[i][16m + 4k + q]
    if (i < kNnOut) *(f32x4*)(P + kOffW4 + i * kNnHid + 16 * m + 4 * k) = acc;
  } else if (tid - 256 < kNnOut) {
    float sb = 0.0f;
    for (int s = 0; s < 64; s++) sb += sGo[s * 16 + (tid - 256)];
    P[kOffB4 + tid - 256] = sb;
  }
  __syncthreads();
  // delta3 = (W4[0:3]^T go) * (h3 > 0), over h3 (K = 4: output c = k)
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int sb = h2 + t;
    const f32x4 acc = mfma_f32(w4t[0], sGo[(16 * sb + i) * 16 + k], (f32x4){0.0f, 0.0f, 0.0f, 0.0f});
    f32x4* hp = (f32x4*)(sH3 + (16 * sb + i) * kGradRow + 16 * m + 4 * k);
    const f32x4 hv = *hp;
    f32x4 d;
#pragma unroll
    for (int q = 0; q < 4; q++) d[q] = hv[q] > 0.0f ? acc[q] : 0.0f;
    *hp = d;
  }
  __syncthreads();
  grad_weights(sH2, kGradRow, sH3, P + kOffW3, kNnHid, m, h2, i, k);
  grad_bias(sH3, P + kOffB3, w, lane);
  __syncthreads();
  grad_delta(w3t, sH3, sH2, m, h2, i, k);
  __syncthreads();
  grad_weights(sH1, kGradRow, sH2, P + kOffW2, kNnHid, m, h2, i, k);
  grad_bias(sH2, P + kOffB2, w, lane);
  __syncthreads();
  grad_delta(w2t, sH2, sH1, m, h2, i, k);
  __syncthreads();
  // W1 (input block 0: features 0..15, the first 12 kept; waves 0-3: output block m), b1
  grad_bias(sH1, P + kOffB1, w, lane);
  if (w < 4) {
    f32x4 ch[4];
#pragma unroll
    for (int c = 0; c < 4; c++) ch[c] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      const int s = grad_sample(kk, k);
      ch[kk & 3] = mfma_f32(sX[s * 16 + i], sH1[s * kGradRow + 16 * m + i], ch[kk & 3]);
    }
    const f32x4 acc = chains4(ch);
    // D[4k + q][i] = dW1[16m + i][4k + q]
    if (4 * k < kNnIn) *(f32x4*)(P + kOffW1 + (16 * m + i) * kNnIn + 4 * k) = acc;
  }
}

// g = the sum of the `nrows` rows of partial; Adam with bias corrections
// c1 = 1/(1 - b1^t), c2 = 1/(1 - b2^t) supplied by the host.  A workgroup takes
// 64 parameters: its kAdamGroups waves sum consecutive runs of the rows, each in
// row order, and wave 0 adds the runs' sums in order (deterministic).
constexpr int kAdamGroups = 16;

__global__ void __launch_bounds__(64 * kAdamGroups)
k_mlp_adam(float* __restrict__ params, float* __restrict__ m1, float* __restrict__ m2,
           const float* __restrict__ partial, int nrows, float lr, float b1, float b2, float eps, float c1,
           float c2, float* __restrict__ grad_out) {
  __shared__ float s_sum[kAdamGroups][64];
  const int lane = threadIdx.x & 63, wq = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + lane;
  const int per = (nrows + kAdamGroups - 1) / kAdamGroups, r0 = wq * per, r1 = min(nrows, r0 + per);
  float g = 0.0f;
  if (p < kNnParams)
    for (int r = r0; r < r1; r++) g += partial[(size_t)r * kNnParams + p];
  s_sum[wq][lane] = g;
  __syncthreads();
  if (wq != 0 || p >= kNnParams) return;
  g = s_sum[0][lane];
#pragma unroll
  for (int q = 1; q < kAdamGroups; q++) g += s_sum[q][lane];
  if (grad_out) grad_out[p] = g;
  if (lr == 0.0f) return;  // gradient only
  const float m = b1 * m1[p] + (1.0f - b1) * g;
  const float v = b2 * m2[p] + (1.0f - b2) * (g * g);
  m1[p] = m;
  m2[p] = v;
  const float mh = m * c1, vh = v * c2;
  params[p] = params[p] - lr * (mh / (sqrtf(vh) + eps));
}

// ---- inference -------------------------------------------------------------------
//
// v_mfma_f32_16x16x32_bf16 operand layout (lane l, group g = l / 16, j = l % 16):
//   A (16 x 32):  A[j][8g + e], e = 0..7
//   B (32 x 16):  B[8g + e][j]
//   D (16 x 16):  D[4g + q][j], q = 0..3
// Neurons are the M dimension and pixels the N dimension, so a layer's D
// fragments hold, per lane, neurons 16mb + 4g + q of pixel j.  The next layer
// takes K block kb from the D fragments of row blocks 2kb and 2kb + 1 as they
// lie: K slot 8g + e of block kb is neuron 32kb + 4g + e (e < 4) or
// 32kb + 16 + 4g + (e - 4) (e >= 4), and the weights are packed in that K order.

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// neuron feeding K slot (g, e) of K block kb in layers 2-4
__device__ __forceinline__ int nn_kperm(int kb, int g, int e) {
  return 32 * kb + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

// One thread per packed element: fragment f, lane l, element e.
__global__ void __launch_bounds__(256) k_mlp_pack(const float* __restrict__ params, __bf16* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < kNnFrags * 64 * 8) {
    const int f = idx / 512, l = (idx / 8) % 64, e = idx % 8;
    const int g = l / 16, j = l % 16;
    float w = 0.0f;
    if (f < 4) {  // layer 1: rows 16f + j, K = 8g + e < 12
      const int k = 8 * g + e;
      if (k < kNnIn) w = params[kOffW1 + (16 * f + j) * kNnIn + k];
    } else if (f < 20) {  // layers 2, 3: frag = 4 + 8 (layer - 2) + 2 mb + kb
      const int q = f - 4, layer = q / 8, mb = (q % 8) / 2, kb = q % 2;
      const int off = layer == 0 ? kOffW2 : kOffW3;
      w = params[off + (16 * mb + j) * kNnHid + nn_kperm(kb, g, e)];
    } else {  // layer 4: rows j < 4
      const int kb = f - 20;
      if (j < kNnOut) w = params[kOffW4 + j * kNnHid + nn_kperm(kb, g, e)];
    }
    out[idx] = (__bf16)w;
  } else if (idx < kNnFrags * 64 * 8 + 3 * 64 + 4) {
    const int b = idx - kNnFrags * 64 * 8;
    float v;
    if (b < 64) v = params[kOffB1 + b];
    else if (b < 128) v = params[kOffB2 + b - 64];
    else if (b < 192) v = params[kOffB3 + b - 128];
    else v = params[kOffB4 + b - 192];
    float* bias = (float*)(out + kNnPackBias);
    bias[b] = v;
  }
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// bf16(a), bf16(b) (round to nearest even) packed in one word: v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t bf16_pair(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// relu(bf16(a)), relu(bf16(b)) = bf16(relu(a)), bf16(relu(b)): rounding keeps the
// sign, so the clamp can follow it, on the bit patterns -- a bf16 with its sign
// bit set is a negative int16, and max with 0 makes it +0 (one v_pk_max_i16 for
// two values, with none of the NaN quieting that fmaxf costs on floats).
__device__ __forceinline__ uint32_t relu_bf16_pair(float a, float b) {
  const i16x2 v = __builtin_bit_cast(i16x2, bf16_pair(a, b));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, (i16x2){0, 0}));
}

// relu + bf16 of the D fragments of row blocks 2kb, 2kb+1: the B operand of K block kb
__device__ __forceinline__ bf16x8 nn_next_b(const f32x4& lo, const f32x4& hi) {
  const u32x4 r = {relu_bf16_pair(lo[0], lo[1]), relu_bf16_pair(lo[2], lo[3]), relu_bf16_pair(hi[0], hi[1]),
                   relu_bf16_pair(hi[2], hi[3])};
  return __builtin_bit_cast(bf16x8, r);
}

// The analytic shader over the pixels k_shade_nn queued: segment k holds
// fbn[k * kFbLine] pixels at fbq + k * segcap.  Stores the total after the
// counters (fbn[kFbSeg * kFbLine]); the last workgroup to finish empties the
// counters for the next k_shade_nn launch (every workgroup has read them
// before it takes its ticket).
template <int PROG, bool IBL>
__global__ void __launch_bounds__(kBlock)
k_shade_list(const ShadeArgs a_by_value, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ fbq,
             long segcap, uint32_t* __restrict__ fbn, float* __restrict__ img) {
  (void)a_by_value;
#if defined(__HIP_DEVICE_COMPILE__)
  const ShadeArgs& a = *(const ShadeArgs*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const ShadeArgs& a = a_by_value;
#endif
  __shared__ uint32_t s_pre[kFbSeg + 1];  // s_pre[k]: queued pixels in segments before k
  const int tid = threadIdx.x;
  if (tid < 64) {
    const uint32_t cnt = tid < kFbSeg ? fbn[tid * kFbLine] : 0u;
    const int incl = wave_inclusive_scan((int)cnt, tid);
    if (tid < kFbSeg) s_pre[tid + 1] = (uint32_t)incl;
    if (tid == 0) s_pre[0] = 0;
  }
  __syncthreads();
  const uint32_t total = s_pre[kFbSeg];
  if (blockIdx.x == 0 && tid == 0) fbn[kFbSeg * kFbLine] = total;
  for (uint32_t i = blockIdx.x * kBlock + tid; i < total; i += gridDim.x * kBlock) {
    int k = 0;  // s_pre[k] <= i < s_pre[k + 1]
#pragma unroll
    for (int step = kFbSeg / 2; step > 0; step >>= 1)
      if (s_pre[k + step] <= i) k += step;
    const long P = fbq[k * segcap + (i - s_pre[k])];
    const uint32_t f = (uint32_t)keys[P] - a.mesh.id_base;
    const int px = (int)(P % a.cam.W), py = (int)(P / a.cam.W);
    store_rgb(img, P, shade_pixel<PROG, IBL>(a, f, px, py));
  }
  if (tid == 0) {
    const unsigned sub = blockIdx.x % kFbTick, nsub = min(gridDim.x, (unsigned)kFbTick);
    uint32_t* const tick = fbn + (kFbSeg + 2 + sub) * kFbLine;
    if (atomicAdd(tick, 1u) == (gridDim.x - sub + kFbTick - 1) / kFbTick - 1) {
      *tick = 0u;  // every workgroup of this residue has taken its ticket
      if (atomicAdd(fbn + (kFbSeg + 1) * kFbLine, 1u) == nsub - 1) {
        for (int k = 0; k < kFbSeg; k++) fbn[k * kFbLine] = 0u;
        fbn[(kFbSeg + 1) * kFbLine] = 0u;
      }
    }
  }
}

}  // namespace tina
