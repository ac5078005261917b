// postp_kernels.h -- image post-processing on the [H][W][3] f32 image.
//
//   k_bloom_down / k_bloom_blur / k_bloom_add   Blooming.apply
//                       (tina/postp/blooming.py:45-68): bright pass averaged
//                       over 2x2 blocks into a half-resolution image, a
//                       separable Gaussian blur (clamped at the borders), and
//                       the bilinearly upsampled result added to the image.
//   k_fxaa_lumi / k_fxaa   FXAA.apply (tina/postp/fxaa.py:26-68): luminance and
//                       a copy of the image, then per pixel the contrast test,
//                       the edge direction and a bilinear re-sample along it.
//   k_accumulate        Accumator.update (tina/util/accumator.py:17-23): the
//                       running mean of the frames of a jittered (TAA) sequence.
//
// Every expression keeps the reference's grouping (the library is compiled
// with -ffp-contract=off).  Neighbour reads past the image clamp to the edge;
// the reference reads outside its fields there (undefined), so the one-pixel
// border of FXAA and the last row/column of the bloom upsample are the places
// this differs from it.  The image is [H][W][3] row-major with y = 0 at the
// bottom, so the reference's [x, y] is img[(y * W + x) * 3 + c].
#pragma once
#include "shade_dev.h"

namespace tina {

__device__ __forceinline__ V3 ld_px(const float* __restrict__ img, int W, int x, int y) {
  const float* p = img + ((long)y * W + x) * 3;
  return v3(p[0], p[1], p[2]);
}

// bilerp (tina/common.py:140-149) over a [h][w][3] image with clamped texels.
__device__ __forceinline__ V3 bilerp_img(const float* __restrict__ img, int w, int h, float px, float py) {
  const int ix = f2i_sat(floorf(px)), iy = f2i_sat(floorf(py));
  const float x0 = px - (float)ix, x1 = py - (float)iy;
  const float y0 = 1.0f - x0, y1 = 1.0f - x1;
  const int xa = min(max(ix, 0), w - 1), xb = min(max(ix + 1, 0), w - 1);
  const int ya = min(max(iy, 0), h - 1), yb = min(max(iy + 1, 0), h - 1);
  const V3 f00 = ld_px(img, w, xa, ya), f01 = ld_px(img, w, xa, yb);
  const V3 f10 = ld_px(img, w, xb, ya), f11 = ld_px(img, w, xb, yb);
  V3 r = vsmul(x1, vsmul(x0, f11));
  r = vadd(r, vsmul(y1, vsmul(x0, f10)));
  r = vadd(r, vsmul(y1, vsmul(y0, f00)));
  r = vadd(r, vsmul(x1, vsmul(y0, f01)));
  return r;
}

// Blooming.filter: factor * (1 - 1 / (1 + scale * max(0, x - thresh))), per channel
__device__ __forceinline__ float bloom_filter(float x, float thresh, float scale, float factor) {
  float t = fmaxf(0.0f, x - thresh);
  t = 1.0f - 1.0f / (1.0f + scale * t);
  return factor * t;
}

__global__ void __launch_bounds__(kBlock)
k_bloom_down(const float* __restrict__ img, int W, float* __restrict__ half, int hw, int hh, float thresh,
             float scale, float factor) {
  const long n = (long)hw * hh;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
    const int x = (int)(i % hw), y = (int)(i / hw);
    float r[3] = {0.0f, 0.0f, 0.0f};
    // ti.ndrange(2, 2): (0, 0), (0, 1), (1, 0), (1, 1)
#pragma unroll
    for (int jx = 0; jx < 2; jx++)
#pragma unroll
      for (int jy = 0; jy < 2; jy++) {
        const float* p = img + ((long)(2 * y + jy) * W + 2 * x + jx) * 3;
#pragma unroll
        for (int c = 0; c < 3; c++) r[c] = r[c] + bloom_filter(p[c], thresh, scale, factor);
      }
    float* o = half + i * 3;
#pragma unroll
    for (int c = 0; c < 3; c++) o[c] = r[c] / 4.0f;
  }
}

// One pass of the separable blur: along x (vert = 0) or y (vert = 1), taps
// clamped to the image: out = in * g0 + sum_i (in[-i] + in[+i]) * g_i.
__global__ void __launch_bounds__(kBlock)
k_bloom_blur(const float* __restrict__ src, float* __restrict__ dst, int hw, int hh, const float* __restrict__ gwei,
             int radius, int vert) {
  const long n = (long)hw * hh;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
    const int x = (int)(i % hw), y = (int)(i / hw);
    const float g0 = gwei[0];
    const V3 c0 = ld_px(src, hw, x, y);
    V3 res = v3(c0.x * g0, c0.y * g0, c0.z * g0);
    for (int k = 1; k <= radius; k++) {
      V3 val;
      if (vert) {
        val = ld_px(src, hw, x, max(0, y - k));
        val = vadd(val, ld_px(src, hw, x, min(hh - 1, y + k)));
      } else {
        val = ld_px(src, hw, max(0, x - k), y);
        val = vadd(val, ld_px(src, hw, min(hw - 1, x + k), y));
      }
      const float g = gwei[k];
      res = vadd(res, v3(val.x * g, val.y * g, val.z * g));
    }
    float* o = dst + i * 3;
    o[0] = res.x;
    o[1] = res.y;
    o[2] = res.z;
  }
}

// The same blur with the taps read from LDS.  Along x a workgroup stages 256
// pixels of one row plus `radius` on each side; along y it stages a tile of TX
// columns by 32 rows plus `radius` rows above and below (TX = 32 or 16, chosen
// so the tile fits 64 KB).  Staged texels are the clamped ones the direct
// kernel reads, kept as three planes (no bank conflicts), and every output sums
// its taps in the same order: identical results, one HBM/L2 read per staged
// texel instead of 2 radius + 1 per output.
constexpr int kBlurRow = 256;
constexpr int kBlurRows = 32;

__global__ void __launch_bounds__(kBlock)
k_bloom_blur_x(const float* __restrict__ src, float* __restrict__ dst, int hw, int hh, const float* __restrict__ gwei,
               int radius) {
  extern __shared__ uint32_t lds_words[];
  float* lds = (float*)lds_words;
  const int n = kBlurRow + 2 * radius;
  const int y = blockIdx.y, x0 = blockIdx.x * kBlurRow, t = threadIdx.x;
  for (int i = t; i < n; i += kBlock) {
    const int xs = min(max(x0 - radius + i, 0), hw - 1);
    const float* p = src + ((long)y * hw + xs) * 3;
    lds[i] = p[0];
    lds[n + i] = p[1];
    lds[2 * n + i] = p[2];
  }
  __syncthreads();
  const int x = x0 + t;
  if (x >= hw) return;
  const int i = t + radius;
  const float g0 = gwei[0];
  V3 res = v3(lds[i] * g0, lds[n + i] * g0, lds[2 * n + i] * g0);
  for (int k = 1; k <= radius; k++) {
    const V3 val = v3(lds[i - k] + lds[i + k], lds[n + i - k] + lds[n + i + k], lds[2 * n + i - k] + lds[2 * n + i + k]);
    const float g = gwei[k];
    res = vadd(res, v3(val.x * g, val.y * g, val.z * g));
  }
  float* o = dst + ((long)y * hw + x) * 3;
  o[0] = res.x;
  o[1] = res.y;
  o[2] = res.z;
}

template <int TX>
__global__ void __launch_bounds__(kBlock)
k_bloom_blur_y(const float* __restrict__ src, float* __restrict__ dst, int hw, int hh, const float* __restrict__ gwei,
               int radius) {
  extern __shared__ uint32_t lds_words[];
  float* lds = (float*)lds_words;
  const int rows = kBlurRows + 2 * radius, n = rows * TX;
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * kBlurRows, t = threadIdx.x;
  for (int idx = t; idx < n; idx += kBlock) {
    const int r = idx / TX, c = idx - r * TX;
    const int xs = min(x0 + c, hw - 1), ys = min(max(y0 - radius + r, 0), hh - 1);
    const float* p = src + ((long)ys * hw + xs) * 3;
    lds[idx] = p[0];
    lds[n + idx] = p[1];
    lds[2 * n + idx] = p[2];
  }
  __syncthreads();
  const int c = t % TX, x = x0 + c;
  if (x >= hw) return;
  const float g0 = gwei[0];
  for (int ry = t / TX; ry < kBlurRows; ry += kBlock / TX) {
    const int y = y0 + ry;
    if (y >= hh) break;
    const int i = (ry + radius) * TX + c;
    V3 res = v3(lds[i] * g0, lds[n + i] * g0, lds[2 * n + i] * g0);
    for (int k = 1; k <= radius; k++) {
      const int a = i - k * TX, b = i + k * TX;
      const V3 val = v3(lds[a] + lds[b], lds[n + a] + lds[n + b], lds[2 * n + a] + lds[2 * n + b]);
      const float g = gwei[k];
      res = vadd(res, v3(val.x * g, val.y * g, val.z * g));
    }
    float* o = dst + ((long)y * hw + x) * 3;
    o[0] = res.x;
    o[1] = res.y;
    o[2] = res.z;
  }
}

// image[I] += bilerp(half, I / 2)
__global__ void __launch_bounds__(kBlock)
k_bloom_add(float* __restrict__ img, int W, int H, const float* __restrict__ half, int hw, int hh) {
  const long n = (long)W * H;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
    const int x = (int)(i % W), y = (int)(i / W);
    const V3 b = bilerp_img(half, hw, hh, (float)x / 2.0f, (float)y / 2.0f);
    float* p = img + i * 3;
    p[0] = p[0] + b.x;
    p[1] = p[1] + b.y;
    p[2] = p[2] + b.z;
  }
}

__global__ void __launch_bounds__(kBlock)
k_fxaa_lumi(const float* __restrict__ img, float* __restrict__ lumi, float* __restrict__ copy, long n) {
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
    const float r = img[3 * i], g = img[3 * i + 1], b = img[3 * i + 2];
    const float l = ((0.2989f * r) + (0.587f * g)) + (0.114f * b);
    lumi[i] = fminf(1.0f, fmaxf(0.0f, l));
    copy[3 * i] = r;
    copy[3 * i + 1] = g;
    copy[3 * i + 2] = b;
  }
}

__global__ void __launch_bounds__(kBlock)
k_fxaa(float* __restrict__ img, const float* __restrict__ lumi, const float* __restrict__ copy, int W, int H,
       float abs_thresh, float rel_thresh, float factor) {
  const long npix = (long)W * H;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < npix; i += (long)gridDim.x * kBlock) {
    const int x = (int)(i % W), y = (int)(i / W);
    auto L = [&](int dx, int dy) {
      return lumi[(long)min(max(y + dy, 0), H - 1) * W + min(max(x + dx, 0), W - 1)];
    };
    const float m = L(0, 0), n = L(0, 1), e = L(1, 0), s = L(0, -1), w = L(-1, 0);
    const float ne = L(1, 1), nw = L(-1, 1), se = L(1, -1), sw = L(-1, -1);
    const float hi = fmaxf(fmaxf(fmaxf(fmaxf(m, n), e), s), w);
    const float lo = fminf(fminf(fminf(fminf(m, n), e), s), w);
    const float c = hi - lo;
    if (c < abs_thresh || c < rel_thresh * hi) continue;
    float filt = 2.0f * (((n + e) + s) + w);
    filt = filt + (((ne + nw) + se) + sw);
    filt = fabsf(filt / 12.0f - m);
    filt = fminf(1.0f, fmaxf(0.0f, filt / c));
    // smoothstep(filt, 0, 1) ** 2 * factor
    const float t = fminf(1.0f, fmaxf(0.0f, (filt - 0.0f) / (1.0f - 0.0f)));
    const float sm = (t * t) * (3.0f - 2.0f * t);
    float blend = (sm * sm) * factor;
    float hori = fabsf((n + s) - 2.0f * m) * 2.0f;
    hori = hori + fabsf((ne + se) - 2.0f * e);
    hori = hori + fabsf((nw + sw) - 2.0f * w);
    float vert = fabsf((e + w) - 2.0f * m) * 2.0f;
    vert = vert + fabsf((ne + nw) - 2.0f * n);
    vert = vert + fabsf((se + sw) - 2.0f * s);
    const bool is_hori = hori >= vert;
    const float plumi = is_hori ? n : e, nlumi = is_hori ? s : w;
    if (fabsf(plumi - m) < fabsf(nlumi - m)) blend = -blend;
    const float px = (float)x + blend * (is_hori ? 0.0f : 1.0f);
    const float py = (float)y + blend * (is_hori ? 1.0f : 0.0f);
    const V3 r = bilerp_img(copy, W, H, px, py);
    img[3 * i] = r.x;
    img[3 * i + 1] = r.y;
    img[3 * i + 2] = r.z;
  }
}

// Accumator.update (tina/util/accumator.py:17-23): acc = acc * (1 - inv) + src * inv
__global__ void __launch_bounds__(kBlock)
k_accumulate(float* __restrict__ acc, const float* __restrict__ src, long n, float inv) {
  const float keep = 1.0f - inv;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
    float a = acc[i] * keep;
    a = a + src[i] * inv;
    acc[i] = a;
  }
}

// ---- SSAO (tina/postp/ssao.py) and the normal buffer it reads ---------------

// The per-pixel buffers of Tina's auxiliary shaders (tina/core/shader.py:21-85,
// 96-109), written for every pixel won by an id of the mesh (particles: from
// their sphere fragment), `channels` floats per pixel:
//   0 NormalShader      normal                   5 DepthShader     engine.depth[P]
//   1 ViewNormalShader  normalize(mapply_dir(W2V, normal))
//   2 TexcoordShader    texcoord                 6 ConstShader     value
//   3 ColorShader       color                    7 ViewdirShader   viewdir * 0.5 + 0.5
//   4 PositionShader    pos                      8 SimpleShader    |normal . viewdir|
//   9 ChessboardShader  lerp((p // size).sum() % 2, 0.4, 0.9), p = P + bias
// MODE >= 0 fixes the mode at compile time (the normal buffer SSAO reads takes
// MODE 0); MODE = -1 reads it from `mode`.
template <bool PART, int MODE>
__global__ void __launch_bounds__(kBlock)
k_gbuffer(const ShadeArgs a, const uint64_t* __restrict__ keys, float* __restrict__ out, int mode_arg, int channels,
          float c0, float c1, float c2, float size) {
  const int mode = MODE >= 0 ? MODE : mode_arg;
  const long npix = (long)a.cam.W * a.cam.H;
  for (long P = (long)blockIdx.x * kBlock + threadIdx.x; P < npix; P += (long)gridDim.x * kBlock) {
    const uint64_t k = keys[P];
    if (key_depth(k) >= TINA_MAXDEPTH) continue;
    const uint32_t f = (uint32_t)k - a.mesh.id_base;
    if (f >= (uint32_t)a.mesh.nfaces) continue;
    const int px = (int)(P % a.cam.W), py = (int)(P / a.cam.W);
    V3 v = v3(c0, c1, c2);
    if (mode == 5) {
      v = vs((float)key_depth(k));
    } else if (mode == 9) {
      const float ux = floorf(((float)px + a.cam.bias[0]) / size), uy = floorf(((float)py + a.cam.bias[1]) / size);
      const float fac = fmodf(ux + uy, 2.0f);
      v = vs(0.4f * (1.0f - fac) + 0.9f * fac);
    } else if (mode != 6) {
      Frag fr;
      if (PART) particle_fragment(a, f, px, py, fr);
      else fragment(a, f, px, py, fr);
      switch (mode) {
        case 0: v = fr.nrm; break;
        case 1: v = vnormalized(mapply_dir(a.cam.W2V, fr.nrm)); break;
        case 2: v = fr.tc; break;
        case 3: v = fr.color; break;
        case 4: v = fr.pos; break;
        default: {
          const V3 viewdir = vneg(view_ray(a.cam, px, py));
          if (mode == 7) v = v3(viewdir.x * 0.5f + 0.5f, viewdir.y * 0.5f + 0.5f, viewdir.z * 0.5f + 0.5f);
          else v = vs(fabsf(vdot(fr.nrm, viewdir)));
        }
      }
    }
    float* o = out + channels * P;
    o[0] = v.x;
    if (channels > 1) o[1] = v.y;
    if (channels > 2) o[2] = v.z;
  }
}

// SSAO.render_at (ssao.py:67-96) with the fixed sample set and rotation tile of
// the non-TAA mode (:28-35, 75-82): ao = clamp(factor * (occ / n - thresh)).
__global__ void __launch_bounds__(kBlock)
k_ssao(const Cam cam, const uint64_t* __restrict__ keys, const float* __restrict__ norm,
       const float* __restrict__ samples, int nsamples, const float* __restrict__ rot, int nrot, float radius,
       float thresh, float factor, float* __restrict__ ao_out) {
  const long npix = (long)cam.W * cam.H;
  const float resx = (float)cam.W, resy = (float)cam.H;
  for (long P = (long)blockIdx.x * kBlock + threadIdx.x; P < npix; P += (long)gridDim.x * kBlock) {
    const int x = (int)(P % cam.W), y = (int)(P / cam.W);
    const V3 normal = ld_px(norm, cam.W, x, y);
    const float p_x = (float)x + cam.bias[0], p_y = (float)y + cam.bias[1];
    const V3 vpos = v3(((p_x / resx) * 2.0f) - 1.0f, ((p_y / resy) * 2.0f) - 1.0f,
                       (float)key_depth(keys[P]) / (float)TINA_MAXDEPTH);
    const V3 pos = mapply_pos(cam.V2W, vpos);
    const V3 viewdir = vneg(view_ray(cam, x, y));
    const float vradius = mapply_pos(cam.W2V, vsub(pos, vsmul(radius, viewdir))).z - vpos.z;
    // tangentspace(normal) (tina/advans.py:97-100): columns tan, bitan, normal
    const V3 bitan = vnormalized(vcross(normal, v3(233.0f, 666.0f, 512.0f)));
    const V3 tan = vcross(bitan, normal);
    const float* r = rot + 2 * ((long)(x % nrot) * nrot + (y % nrot));
    const float rx = r[0], ry = r[1];
    float occ = 0.0f;
    // a zero normal (no object, or a wire) makes the tangent frame NaN and every
    // sample fails its screen test, so the loop would add nothing: skip it
    const bool frame = !(normal.x == 0.0f && normal.y == 0.0f && normal.z == 0.0f);
    for (int i = 0; frame && i < nsamples; i++) {
      const float* sp = samples + 3 * i;
      const float sx = (rx * sp[0]) + (ry * sp[1]), sy = (-rx * sp[0]) + (ry * sp[1]), sz = sp[2];
      V3 smp = v3(((tan.x * sx) + (bitan.x * sy)) + (normal.x * sz), ((tan.y * sx) + (bitan.y * sy)) + (normal.y * sz),
                  ((tan.z * sx) + (bitan.z * sy)) + (normal.z * sz));
      smp = vadd(pos, v3(smp.x * radius, smp.y * radius, smp.z * radius));
      smp = mapply_pos(cam.W2V, smp);
      const float Dx = (smp.x * 0.5f + 0.5f) * resx, Dy = (smp.y * 0.5f + 0.5f) * resy;
      if (!(0.0f <= Dx && Dx < resx && 0.0f <= Dy && Dy < resy)) continue;
      const long Q = (long)(int)Dy * cam.W + (int)Dx;
      const float depth = (float)key_depth(keys[Q]) / (float)TINA_MAXDEPTH;
      if (depth < smp.z) {
        float rc = vradius / (vpos.z - depth);
        const float t = fminf(1.0f, fmaxf(0.0f, fabsf(rc)));  // smoothstep(|rc|, 0, 1)
        rc = (t * t) * (3.0f - 2.0f * t);
        occ = occ + rc;
      }
    }
    float ao = occ / (float)nsamples;
    ao = factor * (ao - thresh);
    ao_out[P] = fminf(1.0f, fmaxf(0.0f, ao));
  }
}

// SSAO.apply (ssao.py:37-48): the image times 1 - (the ao averaged over the
// nrot x nrot box at offset -nrot / 2, reads clamped to the screen), or times
// 1 - ao without the box (box = 0, the TAA mode).
__global__ void __launch_bounds__(kBlock)
k_ssao_apply(float* __restrict__ img, const float* __restrict__ ao, int W, int H, int nrot, int box) {
  const long npix = (long)W * H;
  for (long P = (long)blockIdx.x * kBlock + threadIdx.x; P < npix; P += (long)gridDim.x * kBlock) {
    const int x = (int)(P % W), y = (int)(P / W);
    float m;
    if (box) {
      const int offs = nrot / 2;
      float r = 0.0f;
      for (int k = 0; k < nrot; k++)
        for (int l = 0; l < nrot; l++) {
          const int xx = min(max(x + k - offs, 0), W - 1), yy = min(max(y + l - offs, 0), H - 1);
          r = r + ao[(long)yy * W + xx];
        }
      m = 1.0f - r / (float)(nrot * nrot);
    } else {
      m = 1.0f - ao[P];
    }
    float* p = img + 3 * P;
    p[0] = p[0] * m;
    p[1] = p[1] * m;
    p[2] = p[2] * m;
  }
}

// k_ssao_apply's box form with the occlusion staged in LDS: a 16 x 16 block of
// pixels reads its (16 + nrot - 1)^2 window once (clamped reads, as above) and
// sums each box in the same order.  nrot <= 16.
constexpr int kAoTile = 16;

__global__ void __launch_bounds__(kBlock)
k_ssao_apply_box(float* __restrict__ img, const float* __restrict__ ao, int W, int H, int nrot) {
  __shared__ float s_ao[(kAoTile + 15) * (kAoTile + 15)];
  const int offs = nrot / 2, side = kAoTile + nrot - 1;
  const int x0 = blockIdx.x * kAoTile, y0 = blockIdx.y * kAoTile, t = threadIdx.x;
  for (int i = t; i < side * side; i += kBlock) {
    const int r = i / side, c = i - r * side;
    const int xx = min(max(x0 + c - offs, 0), W - 1), yy = min(max(y0 + r - offs, 0), H - 1);
    s_ao[i] = ao[(long)yy * W + xx];
  }
  __syncthreads();
  const int tx = t % kAoTile, ty = t / kAoTile;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= W || y >= H) return;
  float r = 0.0f;
  for (int k = 0; k < nrot; k++)
    for (int l = 0; l < nrot; l++) r = r + s_ao[(ty + l) * side + tx + k];
  const float m = 1.0f - r / (float)(nrot * nrot);
  float* p = img + 3 * ((long)y * W + x);
  p[0] = p[0] * m;
  p[1] = p[1] * m;
  p[2] = p[2] * m;
}

}  // namespace tina
