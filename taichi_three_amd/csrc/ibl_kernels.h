// ibl_kernels.h -- prefiltered environment maps for image-based lighting.
//
//   k_bake_ibl  Lambert.cook_for_ibl (tina/matr/material.py:407-438) and
//               CookTorrance.cook_for_ibl (material.py:279-320): one output
//               texel per lane, its direction from Skybox.unmapcoor
//               (tina/skybox.py:131-136), nsamples directions around it in
//               tangentspace (tina/advans.py:96-108), each weighted by its
//               cosine u; out = sum(env(dir) * u) / sum(u).
//
// The reference draws (u, v) from ti.random and denoises the Lambert map with a
// KNN filter.  Here the points are the Hammersley set ((s + 0.5)/n, radical
// inverse of s), so a bake is a deterministic function of the environment and
// needs no denoiser, and the arithmetic is binary64 with the sample sum taken
// in sample order: a one-time bake, where precision is worth more than rate.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tina_dev.h"

namespace tina {

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 dadd(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 dsmul(double s, D3 a) { return d3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
  return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Matrix.normalized(): v * (1 / sqrt(v.v))
__device__ __forceinline__ D3 dnormalized(D3 a) { return dsmul(1.0 / sqrt(ddot(a, a)), a); }

// bilerp (tina/common.py:140-149) of a [w][h][3] f32 map, indices clamped to the map
__device__ __forceinline__ D3 bilerp_f64(gf32* __restrict__ img, int w, int h, double px, double py) {
  const double fx = floor(px), fy = floor(py);
  const int ix = f2i_sat((float)fx), iy = f2i_sat((float)fy);
  const double x0 = px - (double)ix, x1 = py - (double)iy;
  const double y0 = 1.0 - x0, y1 = 1.0 - x1;
  const int xa = min(max(ix, 0), w - 1), xb = min(max(ix + 1, 0), w - 1);
  const int ya = min(max(iy, 0), h - 1), yb = min(max(iy + 1, 0), h - 1);
  auto at = [&](int x, int y) {
    gf32* t = img + ((long)x * h + y) * 3;
    return d3((double)t[0], (double)t[1], (double)t[2]);
  };
  D3 r = dsmul(x1, dsmul(x0, at(xb, yb)));
  r = dadd(r, dsmul(y1, dsmul(x0, at(xb, ya))));
  r = dadd(r, dsmul(y1, dsmul(y0, at(xa, ya))));
  r = dadd(r, dsmul(x1, dsmul(y0, at(xa, yb))));
  return r;
}

// Skybox.sample for an equirect map (tina/skybox.py:23-28, 138-145)
__device__ __forceinline__ D3 skybox_sample_f64(gf32* __restrict__ img, int w, int h, D3 dir) {
  const double p = atan2(-dir.z, dir.x) / 6.283185307179586;
  const double cx = p - floor(p);
  const double cy = dir.y * 0.5 + 0.5;
  return bilerp_f64(img, w, h, (double)(w - 1) * cx, (double)(h - 1) * cy);
}

// Van der Corput radical inverse in base 2 (bit reversal of s, scaled by 2^-32)
__device__ __forceinline__ double radical_inverse2(uint32_t s) {
  return (double)__builtin_bitreverse32(s) * 2.3283064365386963e-10;
}

// Skybox.unmapcoor + unspheremap + spherical (tina/skybox.py:31-35, 131-136;
// tina/advans.py:104-108) for texel (ix, iy) of a w x h map
__device__ __forceinline__ D3 unmapcoor_f64(int ix, int iy, int w, int h) {
  const double tau = 6.283185307179586;
  const double v = (double)ix / (double)(w - 1), u = (double)iy / (double)(h - 1);
  const double hh = u * 2.0 - 1.0;
  const double r = sqrt(fmax(0.0, 1.0 - hh * hh));
  const double dx = r * cos(v * tau), dy = r * sin(v * tau), dz = hh;
  return d3(dx, dz, -dy);
}

// tangentspace(nrm) @ spherical(u, v) (tina/advans.py:96-108)
__device__ __forceinline__ D3 tangent_dir_f64(D3 nrm, double u, double v) {
  const double tau = 6.283185307179586;
  const D3 bitan = dnormalized(dcross(nrm, d3(233.0, 666.0, 512.0)));
  const D3 tan = dcross(bitan, nrm);
  const double r = sqrt(fmax(0.0, 1.0 - u * u));
  const double sx = r * cos(v * tau), sy = r * sin(v * tau), sz = u;
  return dadd(dadd(dsmul(sx, tan), dsmul(sy, bitan)), dsmul(sz, nrm));
}

// alpha2 < 0 selects the cosine (Lambert) bake; otherwise the GGX-distributed
// cosines u = sqrt((1 - u0) / (1 - u0 (1 - alpha2))) of CookTorrance.cook_for_ibl.
__global__ void __launch_bounds__(256) k_bake_ibl(const float* __restrict__ env, int ew, int eh,
                                                  float* __restrict__ out, int ow, int oh, double alpha2,
                                                  int nsamples) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ow * oh) return;
  const int ix = (int)(t / oh), iy = (int)(t % oh);
  gf32* e = gptr((uint64_t)env);
  const D3 dir = unmapcoor_f64(ix, iy, ow, oh);
  D3 res = d3(0.0, 0.0, 0.0);
  double dem = 0.0;
  const double n = (double)nsamples;
  for (int s = 0; s < nsamples; s++) {
    const double u0 = ((double)s + 0.5) / n, v = radical_inverse2((uint32_t)s);
    const double u = alpha2 < 0.0 ? u0 : sqrt((1.0 - u0) / (1.0 - u0 * (1.0 - alpha2)));
    const D3 wei = skybox_sample_f64(e, ew, eh, tangent_dir_f64(dir, u, v));
    res = dadd(res, dsmul(u, wei));
    dem += u;
  }
  float* o = out + t * 3;
  o[0] = (float)(res.x / dem);
  o[1] = (float)(res.y / dem);
  o[2] = (float)(res.z / dem);
}

}  // namespace tina
