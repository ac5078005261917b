// nn_shade.h -- deferred shading through the learned surrogate (k_shade_nn).
//
// Persistent workgroups: each stages the packed network in LDS once (A-operand
// fragments and biases) and then walks kNnShadePix-pixel chunks.
// Per chunk it paints the empty pixels with the background, compacts the
// object's covered pixels into an LDS list with a wave ballot, and each wave
// takes the list 64 entries at a time: network inputs per lane, the network on
// four 16-pixel MFMA blocks, the ACES colour stored where the error estimate is
// at most `threshold` and the pixel queued for the analytic kernel otherwise
// (per chunk in LDS, then one global atomic reserves its queue segment);
// pixels with a NaN or infinite feature (degenerate faces) are queued too.
//
// The inputs are computed with v_rcp / v_rsq in place of correctly rounded
// division and square root: the surrogate's colour is approximate by design,
// and every pixel it does not shade within the threshold is shaded exactly.
#pragma once
#include "nn_kernels.h"

namespace tina {

// Fast-math fragment inputs of face f at pixel (px, py): unit normal, unit
// direction toward the camera, albedo.  Same construction as fragment() and
// view_ray() (tina/core/triangle.py:31-49, 140-151; shader.py:81-93).
__device__ __forceinline__ void nn_inputs(const ShadeArgs& a, uint32_t f, int px, int py, V3& n, V3& V,
                                          V3& alb) {
  const float* M = a.cam.W2V;
  gf32* vp = gptr(a.mesh.verts) + 9 * (size_t)f;
  float sx[3], sy[3], rw[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float x = vp[3 * k], y = vp[3 * k + 1], z = vp[3 * k + 2];
    const float cx = M[3] + M[0] * x + M[1] * y + M[2] * z;
    const float cy = M[7] + M[4] * x + M[5] * y + M[6] * z;
    const float cw = M[15] + M[12] * x + M[13] * y + M[14] * z;
    rw[k] = __builtin_amdgcn_rcpf(cw);
    sx[k] = (cx * rw[k] * 0.5f + 0.5f) * (float)a.cam.W;
    sy[k] = (cy * rw[k] * 0.5f + 0.5f) * (float)a.cam.H;
  }
  const float rn = __builtin_amdgcn_rcpf((sx[1] - sx[0]) * (sy[2] - sy[0]) - (sy[1] - sy[0]) * (sx[2] - sx[0]));
  const float p_x = (float)px + a.cam.bias[0], p_y = (float)py + a.cam.bias[1];
  const float w_bc = ((p_x - sx[1]) * (sy[1] - sy[2]) - (p_y - sy[1]) * (sx[1] - sx[2])) * rn;
  const float w_ca = ((p_x - sx[2]) * (sy[2] - sy[0]) - (p_y - sy[2]) * (sx[2] - sx[0])) * rn;
  float w0 = w_bc * rw[0], w1 = w_ca * rw[1], w2 = (1.0f - w_bc - w_ca) * rw[2];
  const float rs = __builtin_amdgcn_rcpf(w0 + w1 + w2);
  w0 *= rs;
  w1 *= rs;
  w2 *= rs;
  V3 nr;
  if ((a.flags & TINA_SMOOTH) && a.mesh.norms) {
    gf32* np = gptr(a.mesh.norms) + 9 * (size_t)f;
    nr = vadd(vadd(vsmul(w0, ld3(np)), vsmul(w1, ld3(np + 3))), vsmul(w2, ld3(np + 6)));
  } else {
    const V3 A = ld3(vp), B = ld3(vp + 3), C = ld3(vp + 6);
    nr = vcross(vsub(B, A), vsub(C, A));
  }
  n = vsmul(__builtin_amdgcn_rsqf(vdot(nr, nr)), nr);
  alb = vs(1.0f);
  if ((a.flags & TINA_VCOLOR) && a.mesh.colors) {
    gf32* cp = gptr(a.mesh.colors) + 9 * (size_t)f;
    alb = vadd(vadd(vsmul(w0, ld3(cp)), vsmul(w1, ld3(cp + 3))), vsmul(w2, ld3(cp + 6)));
  }
  // view ray: V2W (u, v, +-1), difference, normalised, negated
  const float* Q = a.cam.V2W;
  const float ux = (p_x * __builtin_amdgcn_rcpf((float)a.cam.W)) * 2.0f - 1.0f;
  const float uy = (p_y * __builtin_amdgcn_rcpf((float)a.cam.H)) * 2.0f - 1.0f;
  V3 e[2];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const float z = s ? 1.0f : -1.0f;
    const float r = __builtin_amdgcn_rcpf(Q[15] + Q[12] * ux + Q[13] * uy + Q[14] * z);
    e[s] = v3((Q[3] + Q[0] * ux + Q[1] * uy + Q[2] * z) * r, (Q[7] + Q[4] * ux + Q[5] * uy + Q[6] * z) * r,
              (Q[11] + Q[8] * ux + Q[9] * uy + Q[10] * z) * r);
  }
  const V3 d = vsub(e[1], e[0]);
  V = vsmul(-__builtin_amdgcn_rsqf(vdot(d, d)), d);
}

// The network on four 16-pixel blocks at once, layer by layer, so that every
// layer issues 16 independent MFMAs (4 row blocks x 4 pixel blocks) back to
// back instead of one dependent chain per block.  The 22 weight fragments are
// read from LDS (sw: fragment q of this lane at sw[64 q + lane]) where they are
// used: kept in VGPRs they cost 88 registers a lane and the occupancy that
// hides the per-pixel gathers.  `lane` is laundered by the caller, so the
// reads stay inside its pixel loop.
__device__ __forceinline__ void nn_eval64(const bf16x8* __restrict__ sw, uint32_t lane,
                                          const float* __restrict__ sb, int g, const bf16x8 (&x)[4],
                                          f32x4 (&o)[4]) {
  auto wf = [&](int q) { return sw[q * 64 + lane]; };
  auto bias = [&](int base, int mb) { return *(const f32x4*)(sb + base + 16 * mb + 4 * g); };
  f32x4 h[4][4];
  bf16x8 k0[4], k1[4];
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    const f32x4 b = bias(0, mb);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) h[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf(mb), x[nb], b, 0, 0, 0);
  }
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    k0[nb] = nn_next_b(h[nb][0], h[nb][1]);
    k1[nb] = nn_next_b(h[nb][2], h[nb][3]);
  }
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    const f32x4 b = bias(64, mb);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) h[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf(4 + 2 * mb), k0[nb], b, 0, 0, 0);
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
      h[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf(5 + 2 * mb), k1[nb], h[nb][mb], 0, 0, 0);
  }
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    k0[nb] = nn_next_b(h[nb][0], h[nb][1]);
    k1[nb] = nn_next_b(h[nb][2], h[nb][3]);
  }
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    const f32x4 b = bias(128, mb);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) h[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf(12 + 2 * mb), k0[nb], b, 0, 0, 0);
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
      h[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf(13 + 2 * mb), k1[nb], h[nb][mb], 0, 0, 0);
  }
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    k0[nb] = nn_next_b(h[nb][0], h[nb][1]);
    k1[nb] = nn_next_b(h[nb][2], h[nb][3]);
  }
  const f32x4 b4 = bias(192, 0);
#pragma unroll
  for (int nb = 0; nb < 4; nb++) o[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf(20), k0[nb], b4, 0, 0, 0);
#pragma unroll
  for (int nb = 0; nb < 4; nb++) o[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf(21), k1[nb], o[nb], 0, 0, 0);
}

// Copies the packed network into LDS: the 22 weight fragments (sw) and the
// biases b1, b2, b3, then b4 padded to 16 rows (sb).
__device__ __forceinline__ void nn_stage(const __bf16* __restrict__ pk, bf16x8* sw, float* sb) {
  const float* bias = (const float*)(pk + kNnPackBias);
  for (int i = threadIdx.x; i < 3 * 64 + 16; i += blockDim.x) sb[i] = i < 3 * 64 + kNnOut ? bias[i] : 0.0f;
  const bf16x8* fr = (const bf16x8*)pk;
  for (int i = threadIdx.x; i < kNnFrags * 64; i += blockDim.x) sw[i] = fr[i];
}

// This lane's index, opaque to the compiler: keeps LDS weight reads inside a loop.
__device__ __forceinline__ uint32_t laundered_lane(int lane) {
  uint32_t l = (uint32_t)lane;
  asm volatile("" : "+v"(l));
  return l;
}

// Raw buffer resource over `bytes` bytes at p (gfx9 layout, no stride): loads
// at offsets outside the range return 0 and stores there are dropped, so a lane
// with nothing to move needs no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t nn_rsrc(const void* p, long bytes) {
  const int nr = bytes <= 0 ? 0 : bytes >= 0x7fffffffL ? 0x7fffffff : (int)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, nr, 0x00020000);
}

// Layer 1's B operands of rows [64c, 64c + 64) of X, raw: lane (g, j) of block nb
// takes features 8g..8g+7 of row 16 nb + j -- two 16-byte loads for g = 0, one
// (features 8-11, then zeros) for g = 1, zeros for g >= 2 and rows past n.
__device__ __forceinline__ void nn_rows_load(const float* __restrict__ X, long n, long c, int g, int j, u32x4 (&p)[4],
                                             u32x4 (&q)[4]) {
  const long r0 = c * 64;
  const __amdgpu_buffer_rsrc_t R = nn_rsrc(X + r0 * kNnIn, (n - r0) * (kNnIn * 4));
  constexpr uint32_t kOut = 0x80000000u;  // past any range: loads 0
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const uint32_t row = (uint32_t)(16 * nb + j) * (kNnIn * 4);
    p[nb] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(R, g < 2 ? row + 32u * g : kOut, 0, 0));
    q[nb] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(R, g == 0 ? row + 16u : kOut, 0, 0));
  }
}

// The network on n feature rows X[n][12] -> O[n][4] (max(0, colour), error
// estimate), with k_shade_nn's arithmetic: persistent workgroups, each wave
// taking 64 rows at a time.  Each wave loads its next rows before it evaluates
// the current ones; the loads and stores are raw buffer operations whose
// out-of-range lanes read zeros and write nothing, so the loop has no branch
// around them and the wait before the next rows leaves the stores in flight.
// Three waves per SIMD: the register budget that keeps three 64-row passes in
// flight on each SIMD.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_mlp_rows(const __bf16* __restrict__ pk, const float* __restrict__ X, long n, float* __restrict__ O) {
  __shared__ __attribute__((aligned(16))) float s_b[3 * 64 + 16];
  __shared__ __attribute__((aligned(16))) bf16x8 s_w[kNnFrags * 64];
  const int lane = __lane_id(), g = lane / 16, j = lane % 16;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);  // uniform: the buffer ranges live in SGPRs
  nn_stage(pk, s_w, s_b);
  __syncthreads();
  const long nchunks = (n + 63) / 64;
  const long stride = (long)gridDim.x * (kBlock / 64);
  long c = (long)blockIdx.x * (kBlock / 64) + wv;
  u32x4 p[4], q[4];
  nn_rows_load(X, n, c, g, j, p, q);
  {
    // as many stores (to an empty range: dropped) as every pass issues after its
    // loads, so the loop entry leaves the same operations in flight as the back
    // edge and the wait at the top of the loop covers the loads only
    const __amdgpu_buffer_rsrc_t W0 = nn_rsrc(O, 0);
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
      __builtin_amdgcn_raw_buffer_store_b128((u32x4){0u, 0u, 0u, 0u}, W0, (uint32_t)(16 * nb + j) * 16u, 0, 0);
  }
  do {  // the grid gives every wave at least one chunk; a wave past the end loads and stores nothing
    bf16x8 xb[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const f32x4 a = __builtin_bit_cast(f32x4, p[nb]), b = __builtin_bit_cast(f32x4, q[nb]);
      const u32x4 v = {bf16_pair(a[0], a[1]), bf16_pair(a[2], a[3]), bf16_pair(b[0], b[1]), bf16_pair(b[2], b[3])};
      xb[nb] = __builtin_bit_cast(bf16x8, v);
    }
    nn_rows_load(X, n, c + stride, g, j, p, q);
    f32x4 ob[4];
    nn_eval64(s_w, laundered_lane(lane), s_b, g, xb, ob);
    const long r0 = c * 64;
    const __amdgpu_buffer_rsrc_t W = nn_rsrc(O + r0 * kNnOut, (n - r0) * (kNnOut * 4));
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const f32x4 v = {fmaxf(ob[nb][0], 0.0f), fmaxf(ob[nb][1], 0.0f), fmaxf(ob[nb][2], 0.0f), ob[nb][3]};
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), W,
                                             g == 0 ? (uint32_t)(16 * nb + j) * 16u : 0x80000000u, 0, 0);
    }
    c += stride;
  } while (c < nchunks);
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_shade_nn(const ShadeArgs a_by_value, const uint64_t* __restrict__ keys, const __bf16* __restrict__ pk,
           float threshold, uint32_t* __restrict__ fbq, long segcap, uint32_t* __restrict__ fbn,
           float* __restrict__ img) {
  (void)a_by_value;
#if defined(__HIP_DEVICE_COMPILE__)
  const ShadeArgs& a = *(const ShadeArgs*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const ShadeArgs& a = a_by_value;
#endif
  __shared__ uint32_t s_pix[kNnShadePix];
  __shared__ uint32_t s_face[kNnShadePix];
  __shared__ uint32_t s_fb[kNnShadePix];  // the chunk's pixels for the analytic kernel
  __shared__ uint32_t s_n, s_fbn, s_fbbase;
  __shared__ __attribute__((aligned(16))) __bf16 s_x[kBlock / 64][64][16];
  __shared__ __attribute__((aligned(16))) float s_o[kBlock / 64][64][4];
  __shared__ __attribute__((aligned(16))) float s_b[3 * 64 + 16];
  __shared__ __attribute__((aligned(16))) bf16x8 s_w[kNnFrags * 64];
  const int tid = threadIdx.x;
  const int lane = __lane_id(), wv = tid / 64, g = lane / 16, j = lane % 16;
  const long npix = (long)a.cam.W * a.cam.H;
  const long nchunks = (npix + kNnShadePix - 1) / kNnShadePix;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  nn_stage(pk, s_w, s_b);
  const tina_material& m = a.mat;
  const int fres = m.ops[2].b;  // FresnelFactor slot: metallic = its first operand
  const V3 L = light_at(a.lights, 0).xyz;
  for (long chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const long base = chunk * kNnShadePix;
    __syncthreads();  // the previous chunk's lists are consumed
    if (tid == 0) {
      s_n = 0;
      s_fbn = 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kNnShadePix / kBlock; r++) {
      const long P = base + r * kBlock + tid;
      bool take = false;
      uint32_t f = 0;
      if (P < npix) {
        const uint64_t k = keys[P];
        if (key_depth(k) >= TINA_MAXDEPTH) {
          if (a.flags & TINA_BG) store_rgb(img, P, background(a, (int)(P % a.cam.W), (int)(P / a.cam.W)));
        } else {
          f = (uint32_t)k - a.mesh.id_base;
          take = f < (uint32_t)a.mesh.nfaces;
        }
      }
      const unsigned long long mask = __ballot(take);
      uint32_t off = 0;
      if (lane == 0 && mask) off = atomicAdd(&s_n, (uint32_t)__popcll(mask));
      off = __shfl(off, 0);
      if (take) {
        const uint32_t q = off + (uint32_t)__popcll(mask & lt_mask);
        s_pix[q] = (uint32_t)(P - base);
        s_face[q] = f;
      }
    }
    __syncthreads();
    const uint32_t n = s_n;
    for (uint32_t c0 = wv * 64; c0 < n; c0 += kBlock) {
      const uint32_t q = c0 + lane;
      const bool live = q < n;
      long P = 0;
      float x[kNnIn];
#pragma unroll
      for (int k = 0; k < kNnIn; k++) x[k] = 0.0f;
      bool finite = true;  // a NaN or infinite feature (a degenerate face) goes to the analytic kernel
      if (live) {
        P = base + s_pix[q];
        const uint32_t f = s_face[q];
        const int px = (int)(P % a.cam.W), py = (int)(P / a.cam.W);
        V3 nrm, V, alb;
        nn_inputs(a, f, px, py, nrm, V, alb);
        Frag fr;  // slots of the PBR program read only the face index (per-face arrays) or constants
        fr.pos = vs(0.0f);
        fr.nrm = nrm;
        fr.tc = vs(0.0f);
        fr.color = alb;
        const V3 albedo = eval_slot(m, m.ops[1].a, f, fr);
        const float rough = eval_slot(m, m.ops[2].a, f, fr).x;
        const float metal = eval_leaf(m, param_at(m, fres).a, f, fr).x;
        const float t = 2.0f * vdot(nrm, V);
        const V3 R = v3(t * nrm.x - V.x, t * nrm.y - V.y, t * nrm.z - V.z);
        x[0] = nrm.x; x[1] = nrm.y; x[2] = nrm.z;
        x[3] = R.x; x[4] = R.y; x[5] = R.z;
        x[6] = vdot(R, L);
        x[7] = __logf(rough);
        x[8] = metal;
        x[9] = albedo.x; x[10] = albedo.y; x[11] = albedo.z;
        float sum = 0.0f;
#pragma unroll
        for (int k = 0; k < kNnIn; k++) sum += x[k] * 0.0f;  // NaN iff some feature is NaN or infinite
        finite = sum == 0.0f;
      }
      u32x4* row = (u32x4*)s_x[wv][lane];
      row[0] = (u32x4){bf16_pair(x[0], x[1]), bf16_pair(x[2], x[3]), bf16_pair(x[4], x[5]), bf16_pair(x[6], x[7])};
      row[1] = (u32x4){bf16_pair(x[8], x[9]), bf16_pair(x[10], x[11]), 0u, 0u};
      wave_sync_lds();
      bf16x8 xb[4];
      f32x4 ob[4];
#pragma unroll
      for (int nb = 0; nb < 4; nb++) {
        if (g < 2) {
          xb[nb] = *(const bf16x8*)&s_x[wv][16 * nb + j][8 * g];
        } else {
#pragma unroll
          for (int e = 0; e < 8; e++) xb[nb][e] = (__bf16)0.0f;
        }
      }
      nn_eval64(s_w, laundered_lane(lane), s_b, g, xb, ob);
      if (g == 0) {
#pragma unroll
        for (int nb = 0; nb < 4; nb++) *(f32x4*)s_o[wv][16 * nb + j] = ob[nb];
      }
      wave_sync_lds();
      if (live) {
        const f32x4 o = *(const f32x4*)s_o[wv][lane];
        const bool fall = !(o[3] <= threshold) || !finite;
        const unsigned long long fm = __ballot(fall);
        uint32_t at = 0;
        if (lane == 0 && fm) at = atomicAdd(&s_fbn, (uint32_t)__popcll(fm));
        at = __shfl(at, 0);
        if (fall) s_fb[at + (uint32_t)__popcll(fm & lt_mask)] = (uint32_t)P;
        else store_rgb(img, P, v3(fmaxf(o[0], 0.0f), fmaxf(o[1], 0.0f), fmaxf(o[2], 0.0f)));
      }
      wave_sync_lds();
    }
    // hand the chunk's fallback pixels to its queue segment with one global atomic
    __syncthreads();
    const uint32_t nfb = s_fbn;
    if (nfb == 0) continue;
    const long seg = chunk % kFbSeg;
    if (tid == 0) s_fbbase = atomicAdd(fbn + seg * kFbLine, nfb);
    __syncthreads();
    uint32_t* __restrict__ q = fbq + seg * segcap + s_fbbase;
    for (uint32_t i = tid; i < nfb; i += kBlock) q[i] = s_fb[i];
  }
}

}  // namespace tina
