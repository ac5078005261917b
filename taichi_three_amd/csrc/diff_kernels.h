// diff_kernels.h -- differentiable deferred shading and its adjoint.
//
// The material is Tina's PBR (tina/matr/material.py:322-362, 387-396, 701-706)
// under point lights (tina/core/lighting.py:84-98), with albedo, smooth normal,
// roughness and metallic interpolated per vertex of an indexed mesh and a scalar
// intensity per light.  Visibility comes from the packed depth|id buffer of
// the raster pass and is held fixed (no gradient through coverage).
//
//   k_dshade_fwd  colour of every covered pixel
//   k_dshade_bwd  given dLoss/dcolour per pixel: recomputes the face's
//                 perspective-correct weights from the saved id buffer (nothing
//                 per pixel is stored by the forward pass), differentiates the
//                 shading with 8-way forward-mode duals (albedo, normal,
//                 roughness, metallic), contracts with dLoss/dcolour, applies the
//                 normalisation and interpolation Jacobians by hand and adds the
//                 result to the three vertices (float atomics); light-intensity
//                 gradients are reduced per wave first.
#pragma once
#include "shade_kernels.h"

namespace tina {

// ---- forward-mode duals ------------------------------------------------------

template <int N>
struct Dual {
  float v;
  float d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(float v) {
  Dual<N> r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = 0.0f;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dvar(float v, int k) {
  Dual<N> r = dconst<N>(v);
  r.d[k] = 1.0f;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v / b.v;
  const float ib = 1.0f / b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(float s, const Dual<N>& a) {
  Dual<N> r;
  r.v = s * a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = s * a.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(float s, const Dual<N>& a) {
  Dual<N> r = a;
  r.v = s + a.v;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(float s, const Dual<N>& a) {
  Dual<N> r;
  r.v = s - a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, float s) { return s + a; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, float s) {
  Dual<N> r = a;
  r.v = a.v - s;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, float s) { return s * a; }
template <int N>
__device__ __forceinline__ Dual<N> operator/(float s, const Dual<N>& b) {
  Dual<N> r;
  r.v = s / b.v;
  const float q = -r.v / b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = q * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = sqrtf(a.v);
  const float h = 0.5f / r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * h;
  return r;
}
// max(c, a) / min(c, a) with a constant: the derivative of the branch taken
template <int N>
__device__ __forceinline__ Dual<N> dmax(float c, const Dual<N>& a) { return a.v > c ? a : dconst<N>(c); }
template <int N>
__device__ __forceinline__ Dual<N> dmin(float c, const Dual<N>& a) { return a.v < c ? a : dconst<N>(c); }

__device__ __forceinline__ float tval(float x) { return x; }
template <int N>
__device__ __forceinline__ float tval(const Dual<N>& x) { return x.v; }
__device__ __forceinline__ float tsqrt(float x) { return sqrtf(x); }
template <int N>
__device__ __forceinline__ Dual<N> tsqrt(const Dual<N>& x) { return dsqrt(x); }
__device__ __forceinline__ float tmax(float c, float x) { return fmaxf(c, x); }
template <int N>
__device__ __forceinline__ Dual<N> tmax(float c, const Dual<N>& x) { return dmax(c, x); }
__device__ __forceinline__ float tmin(float c, float x) { return fminf(c, x); }
template <int N>
__device__ __forceinline__ Dual<N> tmin(float c, const Dual<N>& x) { return dmin(c, x); }
template <class T> __device__ __forceinline__ T tconst(float v);
template <> __device__ __forceinline__ float tconst<float>(float v) { return v; }
template <> __device__ __forceinline__ Dual<8> tconst<Dual<8>>(float v) { return dconst<8>(v); }

template <class T>
struct T3 { T x, y, z; };
template <class T>
__device__ __forceinline__ T tdot(const T3<T>& a, const V3& b) { return ((b.x * a.x) + (b.y * a.y)) + (b.z * a.z); }

// ---- the shaded colour, generic over float / Dual<8> ----------------------------

struct DiffArgs {
  Cam cam;
  uint64_t soup;     // [F][3][3] f32 positions (the raster's input)
  uint64_t faces;    // [F][3] i32 vertex indices
  uint64_t nrm, col; // [V][3] f32
  uint64_t rough, metal;  // [V] f32
  uint64_t inten;    // [L] f32
  int nfaces;
  uint32_t id_base;
  int nlights;
  float lpos[TINA_MAX_LIGHTS][3];
  float lcol[TINA_MAX_LIGHTS][3];
  float amb[3];
  float f0_dielectric;  // 0.16 specular^2
};

// PBR under the point lights for one pixel: albedo a, unit normal n, roughness r,
// metallic m (each T), at position pos seen from V.  xl[l] receives light l's
// contribution per unit intensity (values only), for the intensity gradients.
template <class T>
__device__ __forceinline__ T3<T> diff_color(const DiffArgs& A, gf32* inten, T3<T> a, T3<T> n, T r, T m, V3 pos,
                                            V3 V, V3* xl) {
  const float inv_pi = (float)(1.0 / 3.141592653589793);
  const float EPS = 1e-10f, ONE_M_EPS = (float)(1.0 - 1e-10);
  // f0 = m a + (1 - m) 0.16 spec^2 ; PBR = Mix(Lambert a, CT(r, f0), f0)
  const T om = 1.0f - m;
  const T3<T> f0 = {m * a.x + om * tconst<T>(A.f0_dielectric), m * a.y + om * tconst<T>(A.f0_dielectric),
                    m * a.z + om * tconst<T>(A.f0_dielectric)};
  // ambient: amb * ((1 - f0) a + f0)
  T3<T> c = {A.amb[0] * ((1.0f - f0.x) * a.x + f0.x), A.amb[1] * ((1.0f - f0.y) * a.y + f0.y),
             A.amb[2] * ((1.0f - f0.z) * a.z + f0.z)};
  const T r2 = tmax(1e-6f, r * r);
  const T k = 0.5f * r2;
  for (int l = 0; l < A.nlights; l++) {
    V3 L = v3(A.lpos[l][0] - pos.x, A.lpos[l][1] - pos.y, A.lpos[l][2] - pos.z);
    const float dist = sqrtf(vdot(L, L));
    L = vdivs(L, dist);
    const T cosv = tdot(n, L);
    if (!(tval(cosv) > 0.0f)) {
      if (xl) xl[l] = vs(0.0f);
      continue;
    }
    const V3 h = vnormalized(vadd(L, V));
    const T NoH = tmax(EPS, tdot(n, h));
    const T NoL = tmax(EPS, cosv);
    const T NoV = tmax(EPS, tdot(n, V));
    const float VoH = fminf(ONE_M_EPS, fmaxf(EPS, vdot(h, V)));
    const T den = 1.0f - (NoH * NoH) * (1.0f - r2);
    const T ndf = r2 / (den * den);
    const T vdf = (1.0f / ((NoV * k + 1.0f) - k)) * (1.0f / ((NoL * k + 1.0f) - k)) /
                  ((1.0f - r2) + (float)(4.0 * 3.141592653589793) * r2);
    const float x1 = 1.0f - VoH, x2 = x1 * x1, p5 = (x2 * x2) * x1;
    const T nv = ndf * vdf;
    const float sc = 1.0f / (dist * dist);
    const T3<T> br = {(1.0f - f0.x) * a.x * tconst<T>(inv_pi) + f0.x * nv * (f0.x + p5 * (1.0f - f0.x)),
                      (1.0f - f0.y) * a.y * tconst<T>(inv_pi) + f0.y * nv * (f0.y + p5 * (1.0f - f0.y)),
                      (1.0f - f0.z) * a.z * tconst<T>(inv_pi) + f0.z * nv * (f0.z + p5 * (1.0f - f0.z))};
    const T3<T> u = {cosv * ((A.lcol[l][0] * sc) * br.x), cosv * ((A.lcol[l][1] * sc) * br.y),
                     cosv * ((A.lcol[l][2] * sc) * br.z)};
    const float I = inten[l];
    if (xl) xl[l] = v3(tval(u.x), tval(u.y), tval(u.z));
    c.x = c.x + I * u.x;
    c.y = c.y + I * u.y;
    c.z = c.z + I * u.z;
  }
  return c;
}

// Weights, position and interpolated inputs of pixel (px, py) on face f.
struct DiffFrag {
  float w[3];
  int vi[3];
  V3 pos, nraw, a;
  float r, m;
};

__device__ __forceinline__ void diff_fragment(const DiffArgs& A, uint32_t f, int px, int py, DiffFrag& q) {
  ShadeArgs s;  // only the camera and positions are read by fragment()
  s.cam = A.cam;
  s.mesh.verts = A.soup;
  s.mesh.norms = 0;
  s.mesh.coors = 0;
  s.mesh.colors = 0;
  s.flags = 0;
  Frag fr;
  fragment(s, f, px, py, fr);
  q.w[0] = fr.w0;
  q.w[1] = fr.w1;
  q.w[2] = fr.w2;
  q.pos = fr.pos;
  const int* fc = (const int*)gptr(A.faces) + 3 * (size_t)f;
  gf32* nrm = gptr(A.nrm);
  gf32* col = gptr(A.col);
  gf32* rgh = gptr(A.rough);
  gf32* mtl = gptr(A.metal);
  q.nraw = vs(0.0f);
  q.a = vs(0.0f);
  q.r = 0.0f;
  q.m = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int v = fc[k];
    q.vi[k] = v;
    q.nraw = vadd(q.nraw, vsmul(q.w[k], ld3(nrm + 3 * (size_t)v)));
    q.a = vadd(q.a, vsmul(q.w[k], ld3(col + 3 * (size_t)v)));
    q.r = q.r + q.w[k] * rgh[v];
    q.m = q.m + q.w[k] * mtl[v];
  }
}

__global__ void __launch_bounds__(kBlock)
k_dshade_fwd(const DiffArgs A, const uint64_t* __restrict__ keys, float* __restrict__ img) {
  const long npix = (long)A.cam.W * A.cam.H;
  const long P = (long)blockIdx.x * kBlock + threadIdx.x;
  if (P >= npix) return;
  const uint64_t key = keys[P];
  V3 c = vs(0.0f);
  const uint32_t f = (uint32_t)key - A.id_base;
  if (key_depth(key) < TINA_MAXDEPTH && f < (uint32_t)A.nfaces) {
    const int px = (int)(P % A.cam.W), py = (int)(P / A.cam.W);
    DiffFrag q;
    diff_fragment(A, f, px, py, q);
    const V3 n = vnormalized(q.nraw);
    const V3 V = vneg(view_ray(A.cam, px, py));
    const T3<float> a = {q.a.x, q.a.y, q.a.z}, nn = {n.x, n.y, n.z};
    const T3<float> r = diff_color<float>(A, gptr(A.inten), a, nn, q.r, q.m, q.pos, V, nullptr);
    c = v3(r.x, r.y, r.z);
  } else if (key_depth(key) < TINA_MAXDEPTH) {
    return;  // another object's pixel
  }
  store_rgb(img, P, c);
}

__global__ void __launch_bounds__(kBlock)
k_dshade_bwd(const DiffArgs A, const uint64_t* __restrict__ keys, const float* __restrict__ gimg,
             float* __restrict__ g_col, float* __restrict__ g_nrm, float* __restrict__ g_rough,
             float* __restrict__ g_metal, float* __restrict__ g_inten) {
  const long npix = (long)A.cam.W * A.cam.H;
  const long P = (long)blockIdx.x * kBlock + threadIdx.x;
  V3 xl[TINA_MAX_LIGHTS];
  V3 g = vs(0.0f);
  bool live = false;
  if (P < npix) {
    const uint64_t key = keys[P];
    const uint32_t f = (uint32_t)key - A.id_base;
    if (key_depth(key) < TINA_MAXDEPTH && f < (uint32_t)A.nfaces) {
      g = ld3(gimg + 3 * P);
      live = true;
      const int px = (int)(P % A.cam.W), py = (int)(P / A.cam.W);
      DiffFrag q;
      diff_fragment(A, f, px, py, q);
      const float inv = 1.0f / sqrtf(vdot(q.nraw, q.nraw));
      const V3 n = vsmul(inv, q.nraw);
      const V3 V = vneg(view_ray(A.cam, px, py));
      typedef Dual<8> D;
      const T3<D> a = {dvar<8>(q.a.x, 0), dvar<8>(q.a.y, 1), dvar<8>(q.a.z, 2)};
      const T3<D> nn = {dvar<8>(n.x, 3), dvar<8>(n.y, 4), dvar<8>(n.z, 5)};
      const T3<D> c = diff_color<D>(A, gptr(A.inten), a, nn, dvar<8>(q.r, 6), dvar<8>(q.m, 7), q.pos, V, xl);
      float gi[8];
#pragma unroll
      for (int i = 0; i < 8; i++) gi[i] = (g.x * c.x.d[i] + g.y * c.y.d[i]) + g.z * c.z.d[i];
      // d n / d nraw = (I - n n^T) / |nraw|
      const V3 gn = v3(gi[3], gi[4], gi[5]);
      const float ndg = vdot(n, gn);
      const V3 graw = vsmul(inv, vsub(gn, vsmul(ndg, n)));
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int v = q.vi[k];
        const float wk = q.w[k];
        atomicAdd(g_col + 3 * (size_t)v + 0, wk * gi[0]);
        atomicAdd(g_col + 3 * (size_t)v + 1, wk * gi[1]);
        atomicAdd(g_col + 3 * (size_t)v + 2, wk * gi[2]);
        atomicAdd(g_nrm + 3 * (size_t)v + 0, wk * graw.x);
        atomicAdd(g_nrm + 3 * (size_t)v + 1, wk * graw.y);
        atomicAdd(g_nrm + 3 * (size_t)v + 2, wk * graw.z);
        atomicAdd(g_rough + v, wk * gi[6]);
        atomicAdd(g_metal + v, wk * gi[7]);
      }
    }
  }
  // dLoss/dI_l = sum over pixels of g . x_l, reduced across the wave first
  for (int l = 0; l < A.nlights; l++) {
    float s = live ? vdot(g, xl[l]) : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (__lane_id() == 0 && s != 0.0f) atomicAdd(g_inten + l, s);
  }
}

}  // namespace tina
