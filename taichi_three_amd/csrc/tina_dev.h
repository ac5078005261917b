// tina_dev.h -- device-side building blocks shared by the raster and shade kernels.
//
// Every float expression keeps the grouping of the Tina source it follows and is
// compiled with -ffp-contract=off, so each multiply and add rounds separately:
// the coverage test and the packed depth are then bit-identical to the binary32
// evaluation of tina/core/triangle.py:88-131.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/tina_hip.h"

#define TINA_MAXDEPTH 1073741824  // Engine.maxdepth = 2**30 (tina/core/engine.py:12)

namespace tina {

// float -> int32 with saturation and NaN -> 0 (the v_cvt_i32_f32 rule, spelled out
// so the compiler never treats an out-of-range value as undefined).
__device__ __forceinline__ int32_t f2i_sat(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}

__device__ __forceinline__ uint64_t pack_key(int32_t depth, uint32_t id) {
  return ((uint64_t)((uint32_t)depth ^ 0x80000000u) << 32) | (uint64_t)id;
}
__host__ __device__ __forceinline__ uint64_t empty_key() {
  return ((uint64_t)((uint32_t)TINA_MAXDEPTH ^ 0x80000000u) << 32);
}
__device__ __forceinline__ int32_t key_depth(uint64_t k) {
  return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
}

// Buffers that reach a kernel as integers inside an argument struct (mesh
// attributes, textures, IBL maps) are global memory: addressing them through a
// global-address-space pointer gives global_load instructions instead of flat
// loads, which would also occupy the LDS counter.
typedef const __attribute__((address_space(1))) float gf32;
__device__ __forceinline__ gf32* gptr(uint64_t p) { return (gf32*)p; }

// ---- correctly rounded binary32 division with a shared divisor ----------------
//
// The arithmetic core of the compiler's IEEE f32 division: v_rcp_f32, one
// Newton step on the reciprocal, the quotient, and two fused corrections (the
// sequence between v_div_scale and v_div_fixup).  The scaling and the special-
// case fix-up around it change nothing while |a| and |d| lie in [2^-60, 2^60]:
// the quotient stays in [2^-120, 2^120] and every remainder above 2^-84, so no
// intermediate is subnormal or overflows.  Outside that range (zero, infinities
// and NaN included) the compiler's division runs.  The reciprocal step depends
// on d alone, so the quotients by one divisor share it: a quotient then costs
// one multiply and four FMAs.  A device self-test (tina_selftest_div) compares
// it bit for bit with the compiler's division.
struct RcpRN {
  float d, r;
  bool ok;
};

__device__ __forceinline__ RcpRN rcp_rn(float d) {
  RcpRN R;
  R.d = d;
  const float ad = fabsf(d);
  R.ok = (ad >= 0x1p-60f) & (ad <= 0x1p60f);
  const float r0 = __builtin_amdgcn_rcpf(d);
  const float e = __builtin_fmaf(-d, r0, 1.0f);
  R.r = __builtin_fmaf(e, r0, r0);
  return R;
}

// RN(a / R.d), bit-identical to the compiler's a / d.
__device__ __forceinline__ float div_rn(float a, const RcpRN& R) {
  const float aa = fabsf(a);
  if (R.ok & (aa >= 0x1p-60f) & (aa <= 0x1p60f)) {
    const float q = a * R.r;
    const float f2 = __builtin_fmaf(-R.d, q, a);
    const float q3 = __builtin_fmaf(f2, R.r, q);
    const float f4 = __builtin_fmaf(-R.d, q3, a);
    return __builtin_fmaf(f4, R.r, q3);
  }
  return a / R.d;
}

__device__ __forceinline__ float fdiv_rn(float a, float d) { return div_rn(a, rcp_rn(d)); }

// Camera passed by value as a kernel argument (lives in the kernarg segment).
struct Cam {
  float W2V[16];
  float V2W[16];
  float bias[2];
  int W, H;
};

// mapply(mat, pos, 1) (tina/common.py:169-177)
__device__ __forceinline__ void mapply1(const float* __restrict__ M, float p0, float p1, float p2,
                                        float& r0, float& r1, float& r2, float& w) {
  float a = M[3] * 1.0f;  a = a + M[0] * p0;  a = a + M[1] * p1;  a = a + M[2] * p2;  r0 = a;
  float b = M[7] * 1.0f;  b = b + M[4] * p0;  b = b + M[5] * p1;  b = b + M[6] * p2;  r1 = b;
  float c = M[11] * 1.0f; c = c + M[8] * p0;  c = c + M[9] * p1;  c = c + M[10] * p2; r2 = c;
  float d = M[15] * 1.0f; d = d + M[12] * p0; d = d + M[13] * p1; d = d + M[14] * p2; w = d;
}

// The per-face state render_occup keeps (triangle.py:127-131) plus what the
// coverage loop needs: viewport b, c; bcn, can; wscale; view-space z of A, B, C.
struct FaceSetup {
  float bx, by, cx, cy;
  float bcnx, bcny, canx, cany;
  float ws0, ws1, ws2;
  float Az, Bz, Cz;
  int32_t botx, boty, topx, topy;
};

// Per-face part of render_occup (triangle.py:93-113).  Returns false when the
// face is culled or clipped.  v points at 9 floats (A, B, C).
__device__ __forceinline__ bool setup_face(const Cam& cam, const float* v, uint32_t flags,
                                           FaceSetup& s) {
  float Ar0, Ar1, Ar2, Aw, Br0, Br1, Br2, Bw, Cr0, Cr1, Cr2, Cw;
  mapply1(cam.W2V, v[0], v[1], v[2], Ar0, Ar1, Ar2, Aw);
  mapply1(cam.W2V, v[3], v[4], v[5], Br0, Br1, Br2, Bw);
  mapply1(cam.W2V, v[6], v[7], v[8], Cr0, Cr1, Cr2, Cw);
  const RcpRN RA = rcp_rn(Aw), RB = rcp_rn(Bw), RC = rcp_rn(Cw);
  const float Avx = div_rn(Ar0, RA), Avy = div_rn(Ar1, RA), Avz = div_rn(Ar2, RA);
  const float Bvx = div_rn(Br0, RB), Bvy = div_rn(Br1, RB), Bvz = div_rn(Br2, RB);
  const float Cvx = div_rn(Cr0, RC), Cvy = div_rn(Cr1, RC), Cvz = div_rn(Cr2, RC);

  const float facing = (Bvx - Avx) * (Cvy - Avy) - (Bvy - Avy) * (Cvx - Avx);  // :95
  if ((flags & TINA_CULL) && facing <= 0.0f) return false;                     // :96-98
  if (flags & TINA_CLIP) {                                                      // :100-104
    const bool ai = (-1.0f <= Avx) & (Avx <= 1.0f) & (-1.0f <= Avy) & (Avy <= 1.0f) &
                    (-1.0f <= Avz) & (Avz <= 1.0f);
    const bool bi = (-1.0f <= Bvx) & (Bvx <= 1.0f) & (-1.0f <= Bvy) & (Bvy <= 1.0f) &
                    (-1.0f <= Bvz) & (Bvz <= 1.0f);
    const bool ci = (-1.0f <= Cvx) & (Cvx <= 1.0f) & (-1.0f <= Cvy) & (Cvy <= 1.0f) &
                    (-1.0f <= Cvz) & (Cvz <= 1.0f);
    if (!ai && !bi && !ci) return false;
  }
  const float resx = (float)cam.W, resy = (float)cam.H;
  const float ax = (Avx * 0.5f + 0.5f) * resx, ay = (Avy * 0.5f + 0.5f) * resy;  // to_viewport
  const float bx = (Bvx * 0.5f + 0.5f) * resx, by = (Bvy * 0.5f + 0.5f) * resy;
  const float cx = (Cvx * 0.5f + 0.5f) * resx, cy = (Cvy * 0.5f + 0.5f) * resy;

  const int32_t bo_x = f2i_sat(floorf(fminf(ax, fminf(bx, cx))));               // :108-109
  const int32_t bo_y = f2i_sat(floorf(fminf(ay, fminf(by, cy))));
  const int32_t to_x = f2i_sat(ceilf(fmaxf(ax, fmaxf(bx, cx))));
  const int32_t to_y = f2i_sat(ceilf(fmaxf(ay, fmaxf(by, cy))));
  s.botx = bo_x > 0 ? bo_x : 0;
  s.boty = bo_y > 0 ? bo_y : 0;
  s.topx = to_x < cam.W - 1 ? to_x : cam.W - 1;
  s.topy = to_y < cam.H - 1 ? to_y : cam.H - 1;

  const float n = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);               // :110-112
  const RcpRN Rn = rcp_rn(n);
  s.bcnx = div_rn(bx - cx, Rn);
  s.bcny = div_rn(by - cy, Rn);
  s.canx = div_rn(cx - ax, Rn);
  s.cany = div_rn(cy - ay, Rn);
  s.bx = bx; s.by = by; s.cx = cx; s.cy = cy;
  s.ws0 = div_rn(1.0f, RA); s.ws1 = div_rn(1.0f, RB); s.ws2 = div_rn(1.0f, RC);  // :113
  s.Az = Avz; s.Bz = Bvz; s.Cz = Cvz;
  return true;
}

// Weights (:115-119) for pixel (px, py); returns the coverage test (:120).
__device__ __forceinline__ bool pixel_weights(const FaceSetup& s, float posx, float posy,
                                              float& w0, float& w1, float& w2) {
  const float w_bc = (posx - s.bx) * s.bcny - (posy - s.by) * s.bcnx;
  const float w_ca = (posx - s.cx) * s.cany - (posy - s.cy) * s.canx;
  const float t = (1.0f - w_bc) - w_ca;
  const float x = w_bc * s.ws0, y = w_ca * s.ws1, z = t * s.ws2;
  const float sum = (x + y) + z;
  const RcpRN Rs = rcp_rn(sum);
  w0 = div_rn(x, Rs); w1 = div_rn(y, Rs); w2 = div_rn(z, Rs);
  return (w0 >= 0.0f) & (w1 >= 0.0f) & (w2 >= 0.0f);
}

// The same test and weights, dividing only when the pixel is covered.
//
// wei_i = x_i / sum >= 0 is decided from signs when that is exact: for finite x_i
// and a finite sum with 2**-126 <= |sum| <= 2**100, x_i / sum can only round to
// zero if |x_i| < |sum| * 2**-126, so whenever every nonzero |x_i| is at least
// that, "x_i / sum >= 0" equals "x_i == 0 or x_i has the sign of sum".  Anything
// else (zero or tiny sums, NaN, infinities, tiny x_i) takes the division path,
// which is the reference's expression itself.  Returns the coverage; when covered
// w0..w2 hold the divided weights, bit-identical to pixel_weights.
__device__ __forceinline__ bool pixel_cover_fast(const FaceSetup& s, float posx, float posy,
                                                 float& w0, float& w1, float& w2) {
  const float w_bc = (posx - s.bx) * s.bcny - (posy - s.by) * s.bcnx;
  const float w_ca = (posx - s.cx) * s.cany - (posy - s.cy) * s.canx;
  const float t = (1.0f - w_bc) - w_ca;
  const float x = w_bc * s.ws0, y = w_ca * s.ws1, z = t * s.ws2;
  const float sum = (x + y) + z;
  const float as = fabsf(sum);
  const float lo = as * 1.1754944e-38f;  // |sum| * 2**-126, exact (power of two, no underflow below)
  const bool sane = (as >= 1.1754944e-38f) & (as <= 1.2676506e30f) &
                    ((x == 0.0f) | (fabsf(x) >= lo)) & ((y == 0.0f) | (fabsf(y) >= lo)) &
                    ((z == 0.0f) | (fabsf(z) >= lo)) & (fabsf(x) <= 1.2676506e30f) &
                    (fabsf(y) <= 1.2676506e30f) & (fabsf(z) <= 1.2676506e30f);
  if (sane) {
    const bool pos = sum > 0.0f;
    const bool cx_ = (x == 0.0f) | ((x > 0.0f) == pos);
    const bool cy_ = (y == 0.0f) | ((y > 0.0f) == pos);
    const bool cz_ = (z == 0.0f) | ((z > 0.0f) == pos);
    if (!(cx_ & cy_ & cz_)) return false;
  }
  const RcpRN Rs = rcp_rn(sum);
  w0 = div_rn(x, Rs); w1 = div_rn(y, Rs); w2 = div_rn(z, Rs);
  return (w0 >= 0.0f) & (w1 >= 0.0f) & (w2 >= 0.0f);
}

// A face is well-conditioned when its 1/w are in [2**-20, 2**20], its viewport
// b, c are within 2**20 and its edge coefficients bcn, can within 2**40 (NaN
// fails every test).  Then, for pixel centres within 2**20, |w_bc|, |w_ca| <= 2**62,
// every x_i = w_i * ws_i and the sum stay below 2**86 (no overflow), and a pixel
// with an unscaled edge value w_bc, w_ca or t below -2**-40 is not covered:
//   * its x_i is below -2**-60, so x_i / sum (|.| >= 2**-146) keeps its sign and
//     passes ">= 0" only if sum < 0;
//   * w_bc + w_ca + t = 1 up to rounding, with t = (1 - w_bc) - w_ca: when one of
//     them is negative another is at least 1/2, its x_j >= 2**-21 and
//     x_j / sum (|.| >= 2**-107) is negative for sum < 0 -- so the pixel fails.
__device__ __forceinline__ bool face_well_conditioned(const FaceSetup& s) {
  const float lo = 9.5367432e-7f, hi = 1048576.0f;  // 2**-20, 2**20
  const bool ws = (s.ws0 >= lo) & (s.ws0 <= hi) & (s.ws1 >= lo) & (s.ws1 <= hi) & (s.ws2 >= lo) & (s.ws2 <= hi);
  const float mp = fmaxf(fmaxf(fabsf(s.bx), fabsf(s.by)), fmaxf(fabsf(s.cx), fabsf(s.cy)));
  const float me = fmaxf(fmaxf(fabsf(s.bcnx), fabsf(s.bcny)), fmaxf(fabsf(s.canx), fabsf(s.cany)));
  return ws & (mp <= hi) & (me <= 1.0995116e12f);  // 2**40
}

// The cheap rejection for a well-conditioned face (see above).  True: not covered.
__device__ __forceinline__ bool pixel_rejects(const FaceSetup& s, float posx, float posy) {
  const float w_bc = (posx - s.bx) * s.bcny - (posy - s.by) * s.bcnx;
  const float w_ca = (posx - s.cx) * s.cany - (posy - s.cy) * s.canx;
  const float t = (1.0f - w_bc) - w_ca;
  return fminf(fminf(w_bc, w_ca), t) < -9.0949470e-13f;  // -2**-40
}

// depth = int((wei.x*Av.z + wei.y*Bv.z + wei.z*Cv.z) * 2**30)   (:121-122)
__device__ __forceinline__ int32_t pixel_depth(const FaceSetup& s, float w0, float w1, float w2) {
  const float d = ((w0 * s.Az) + (w1 * s.Bz)) + (w2 * s.Cz);
  return f2i_sat(d * 1073741824.0f);
}

}  // namespace tina
