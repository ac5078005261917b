// shade_dev.h -- the deferred per-pixel shading of Tina as device code.
//
// Follows tina/core/triangle.py:31-49,133-153 (render_color, interpolate),
// tina/core/shader.py:81-93,112-131 (calc_viewdir, Shader.shade_color),
// tina/core/lighting.py:4-21,84-98 (SkyboxLighting, Lighting), the material
// classes of tina/matr/material.py and the parameter nodes of
// tina/matr/nodes.py.  Grouping of every float expression follows the Python
// source; the library is compiled with -ffp-contract=off.
#pragma once
#include "tina_dev.h"

namespace tina {

struct V3 { float x, y, z; };
__device__ __forceinline__ V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ V3 vs(float s) { return v3(s, s, s); }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 vmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 vsmul(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 vdivs(V3 a, float s) {
  const RcpRN R = rcp_rn(s);
  return v3(div_rn(a.x, R), div_rn(a.y, R), div_rn(a.z, R));
}
__device__ __forceinline__ V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
// a / d as a times the hardware reciprocal of d (v_rcp_f32, within 1 ulp), for
// the quotients that only scale a light's contribution -- its 1/d^2 falloff and
// the Cook-Torrance D and G terms: the colour moves by a few ulps of its own
// value.  The normalisations, the view ray, the interpolation weights and ACES
// keep the correctly rounded division (post passes such as FXAA branch on the
// tone-mapped values, so those stay the binary32 oracle's to the bit).
__device__ __forceinline__ float div_scale(float a, float d) { return a * __builtin_amdgcn_rcpf(d); }
__device__ __forceinline__ float vdot(V3 a, V3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
// Matrix.normalized(): v * (1 / sqrt(sum v^2))
__device__ __forceinline__ V3 vnormalized(V3 a) { const float inv = fdiv_rn(1.0f, sqrtf(vdot(a, a))); return vsmul(inv, a); }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// reflect(I, N) = I - 2 * N.dot(I) * N   (tina/common.py:197-199)
__device__ __forceinline__ V3 vreflect(V3 I, V3 N) { const float d = 2.0f * vdot(N, I); return vsub(I, vsmul(d, N)); }
__device__ __forceinline__ V3 ld3(const float* __restrict__ p) { return v3(p[0], p[1], p[2]); }
__device__ __forceinline__ V3 ld3(gf32* __restrict__ p) { return v3(p[0], p[1], p[2]); }

// mapply_pos (tina/common.py:180-183)
__device__ __forceinline__ V3 mapply_pos(const float* __restrict__ M, V3 p, float* wout = nullptr) {
  float r0, r1, r2, w;
  mapply1(M, p.x, p.y, p.z, r0, r1, r2, w);
  if (wout) *wout = w;
  const RcpRN R = rcp_rn(w);
  return v3(div_rn(r0, R), div_rn(r1, R), div_rn(r2, R));
}

// bilerp (tina/common.py:140-149) over a [w][h][c] array; the indices are
// clamped to the array (the reference reads one texel past the edge with
// weight 0, or out of bounds for coordinates outside [0, shape-1]).
template <class PTR>
__device__ __forceinline__ V3 bilerp(PTR __restrict__ img, int w, int h, int c, int ch,
                                     float px, float py) {
  const int ix = f2i_sat(floorf(px)), iy = f2i_sat(floorf(py));
  const float x0 = px - (float)ix, x1 = py - (float)iy;
  const float y0 = 1.0f - x0, y1 = 1.0f - x1;
  const int xa = min(max(ix, 0), w - 1), xb = min(max(ix + 1, 0), w - 1);
  const int ya = min(max(iy, 0), h - 1), yb = min(max(iy + 1, 0), h - 1);
  const auto t00 = img + ((long)xa * h + ya) * c + ch;
  const auto t01 = img + ((long)xa * h + yb) * c + ch;
  const auto t10 = img + ((long)xb * h + ya) * c + ch;
  const auto t11 = img + ((long)xb * h + yb) * c + ch;
  V3 f00, f01, f10, f11;
  if (c >= 3) { f00 = ld3(t00); f01 = ld3(t01); f10 = ld3(t10); f11 = ld3(t11); }
  else { f00 = vs(t00[0]); f01 = vs(t01[0]); f10 = vs(t10[0]); f11 = vs(t11[0]); }
  V3 r = vsmul(x1, vsmul(x0, f11));
  r = vadd(r, vsmul(y1, vsmul(x0, f10)));
  r = vadd(r, vsmul(y1, vsmul(y0, f00)));
  r = vadd(r, vsmul(x1, vsmul(y0, f01)));
  return r;
}

// Skybox.sample for an equirect map (tina/skybox.py:23-28, 138-145)
__device__ __forceinline__ V3 skybox_sample(gf32* __restrict__ img, int w, int h, V3 dir) {
  const float p = fdiv_rn(atan2f(-dir.z, dir.x), 6.2831855f);
  const float cx = p - floorf(p);
  const float cy = dir.y * 0.5f + 0.5f;
  return bilerp(img, w, h, 3, 0, (float)(w - 1) * cx, (float)(h - 1) * cy);
}

// ---- lobes -----------------------------------------------------------------

// Phong.brdf (material.py:449-454)
__device__ __forceinline__ V3 phong(float m, V3 n, V3 L, V3 V) {
  const V3 rdir = vreflect(vneg(L), n);
  const float VoR = fmaxf(0.0f, vdot(V, rdir));
  return vs((powf(VoR, m) * (m + 2.0f)) * 0.5f);  // x * 0.5 is x / 2 to the bit
}

// BlinnPhong (extension): max(0, N.normalize(L + V))**m * (m + 8) / 8
__device__ __forceinline__ V3 blinnphong(float m, V3 n, V3 L, V3 V) {
  const V3 h = vnormalized(vadd(L, V));
  const float NoH = fmaxf(0.0f, vdot(n, h));
  return vs((powf(NoH, m) * (m + 8.0f)) * 0.125f);  // x * 0.125 is x / 8 to the bit
}

// CookTorrance.sub_brdf / brdf (material.py:322-362)
__device__ __forceinline__ V3 cooktorrance(float roughness, V3 f0, V3 n, V3 L, V3 V) {
  const float EPS = 1e-10f;
  const float ONE_M_EPS = (float)(1.0 - 1e-10);
  const V3 half = vnormalized(vadd(L, V));
  const float NoH = fmaxf(EPS, vdot(half, n));
  const float NoL = fmaxf(EPS, vdot(L, n));
  const float NoV = fmaxf(EPS, vdot(V, n));
  const float VoH = fminf(ONE_M_EPS, fmaxf(EPS, vdot(half, V)));
  const float alpha2 = fmaxf(1e-6f, roughness * roughness);
  const float denom = 1.0f - (NoH * NoH) * (1.0f - alpha2);
  const float ndf = div_scale(alpha2, denom * denom);
  const float k = alpha2 * 0.5f;  // x * 0.5 is x / 2 to the bit
  float vdf = div_scale(1.0f, ((NoV * k) + 1.0f) - k);
  vdf = vdf * div_scale(1.0f, ((NoL * k) + 1.0f) - k);
  vdf = div_scale(vdf, (1.0f * (1.0f - alpha2)) + ((float)(4.0 * 3.141592653589793) * alpha2));
  // (1 - VoH)**5 as ((x * x) * (x * x)) * x: the integer power restated by
  // squaring, as the binary32 oracle does
  const float x1 = 1.0f - VoH, x2 = x1 * x1;
  const float p5 = (x2 * x2) * x1;
  const V3 fdf = vadd(f0, vsmul(p5, vsub(vs(1.0f), f0)));
  return vsmul(ndf, vsmul(vdf, fdf));
}

// CookTorrance.sample_ibl (material.py:255-277)
__device__ __forceinline__ V3 cooktorrance_ibl(const tina_ibl& ibl, float roughness, V3 f0, V3 idir, V3 nrm) {
  const float EPS = 1e-10f;
  const float ONE_M_EPS = (float)(1.0 - 1e-10);
  const V3 rdir = vreflect(vneg(idir), nrm);
  V3 wei = vs(1.0f);
#pragma unroll
  for (int id = 0; id < 6; id++) {
    const float lo = ibl.rough_lo[id], hi = ibl.rough_hi[id], st = ibl.rough_step[id];
    if (lo <= roughness && roughness <= hi) {
      const V3 w1 = skybox_sample(gptr(ibl.spec[id]), ibl.spec_w[id], ibl.spec_h[id], rdir);
      const V3 w2 = skybox_sample(gptr(ibl.spec[id + 1]), ibl.spec_w[id + 1], ibl.spec_h[id + 1], rdir);
      const float fac2 = fdiv_rn(hi - roughness, st);
      wei = vadd(vsmul(1.0f - fac2, w2), vsmul(fac2, w1));  // lerp(fac2, wei2, wei1)
    }
  }
  const V3 half = vnormalized(vadd(idir, rdir));
  const float VoH = fminf(ONE_M_EPS, fmaxf(EPS, vdot(half, rdir)));
  const float A = bilerp(&ibl.lut[0][0][0], 3, 3, 2, 0, VoH, roughness).x;
  const float B = bilerp(&ibl.lut[0][0][0], 3, 3, 2, 1, VoH, roughness).x;
  return vmul(wei, vadd(vsmul(A, f0), vs(B)));  // wei *= f0 * AB.x + AB.y
}

enum { M_BRDF = 0, M_AMB = 1, M_EMI = 2, M_IBL = 3 };

// The material, light and IBL tables arrive by value in the kernel-argument
// segment.  Indexing them with a run-time index would make the compiler copy the
// whole 2 KB argument block to per-lane scratch, so every access goes through a
// switch: each case reads the kernarg segment at a fixed offset (the index is
// uniform across the launch, so the switch is a scalar branch).
#define TINA_CASE16(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)

__device__ __forceinline__ tina_param param_at(const tina_material& m, int i) {
  switch (i) {
#define TINA_P(k) case k: if (k < TINA_MAX_PARAMS) return m.params[k < TINA_MAX_PARAMS ? k : 0]; break;
    TINA_CASE16(TINA_P)
#undef TINA_P
  }
  return m.params[0];
}

__device__ __forceinline__ tina_op op_at(const tina_material& m, int i) {
  switch (i) {
#define TINA_O(k) case k: return m.ops[k];
    TINA_CASE16(TINA_O)
#undef TINA_O
  }
  return m.ops[0];
}

struct Light { V3 xyz; float w; V3 color; };
__device__ __forceinline__ Light light_at(const tina_lights& lt, int l) {
  Light r;
  switch (l) {
#define TINA_L(k) case k: r.xyz = v3(lt.dirs[k][0], lt.dirs[k][1], lt.dirs[k][2]); r.w = lt.dirs[k][3]; \
      r.color = v3(lt.colors[k][0], lt.colors[k][1], lt.colors[k][2]); break;
    TINA_CASE16(TINA_L)
#undef TINA_L
    default: r.xyz = vs(0.0f); r.w = 0.0f; r.color = vs(0.0f); break;
  }
  return r;
}

// Parameter slot values of one pixel, in registers.
struct Slots {
  V3 v[TINA_MAX_PARAMS];
  __device__ __forceinline__ V3 get(int i) const {
    V3 r = v[0];
#pragma unroll
    for (int k = 1; k < TINA_MAX_PARAMS; k++)
      if (i == k) r = v[k];
    return r;
  }
};

// The material program (post-order walk of the material tree): leaves push,
// combinators pop.  The program is uniform across the launch, so every branch
// here is a scalar branch.  The 4-deep stack lives in registers and shifts.
__device__ __forceinline__ V3 run_material(const tina_material& m, const Slots& pv, int mode,
                                           V3 n, V3 L, V3 V, const tina_ibl& ibl) {
  V3 s0 = vs(0.0f), s1 = vs(0.0f), s2 = vs(0.0f), s3 = vs(0.0f);
  const float inv_pi = (float)(1.0 / 3.141592653589793);
  for (int i = 0; i < m.nops; i++) {
    const tina_op o = op_at(m, i);
    const int op = o.op, a = o.a, b = o.b;
    if (op <= TINA_OP_EMISSION) {
      V3 v = vs(0.0f);
      if (op == TINA_OP_LAMBERT) {
        if (mode == M_BRDF) v = vs(inv_pi);
        else if (mode == M_AMB) v = vs(1.0f);
        else if (mode == M_IBL) v = skybox_sample(gptr(ibl.diff), ibl.diff_w, ibl.diff_h, n);
      } else if (op == TINA_OP_PHONG || op == TINA_OP_BLINNPHONG) {
        if (mode == M_BRDF)
          v = op == TINA_OP_PHONG ? phong(pv.get(a).x, n, L, V) : blinnphong(pv.get(a).x, n, L, V);
        else if (mode == M_AMB) v = vs(1.0f);
      } else if (op == TINA_OP_COOKTORRANCE) {
        if (mode == M_BRDF) v = cooktorrance(pv.get(a).x, pv.get(b), n, L, V);
        else if (mode == M_AMB) v = vs(1.0f);
        else if (mode == M_IBL) v = cooktorrance_ibl(ibl, pv.get(a).x, pv.get(b), V, n);
      } else {  // EMISSION
        if (mode == M_EMI) v = vs(1.0f);
      }
      s3 = s2; s2 = s1; s1 = s0; s0 = v;
    } else if (op == TINA_OP_SCALE) {  // fac * wei
      s0 = vmul(pv.get(a), s0);
    } else if (op == TINA_OP_MIX) {    // (1 - fac) * wei1 + fac * wei2
      const V3 fac = pv.get(a);
      s0 = vadd(vmul(vsub(vs(1.0f), fac), s1), vmul(fac, s0));
      s1 = s2; s2 = s3;
    } else {                           // ADD
      s0 = vadd(s1, s0);
      s1 = s2; s2 = s3;
    }
  }
  return s0;
}

// Lighting.shade_color (lighting.py:84-98); with TINA_IBL the SkyboxLighting
// form (lighting.py:20-21) replaces emission + ambient.
__device__ __forceinline__ V3 lighting(const tina_material& m, const Slots& pv, const tina_lights& lt,
                                       const tina_ibl& ibl, uint32_t flags, V3 pos, V3 n, V3 V) {
  V3 res = vs(0.0f);
  if (flags & TINA_IBL) {
    res = run_material(m, pv, M_IBL, n, n, V, ibl);
    res = vadd(res, run_material(m, pv, M_EMI, n, n, V, ibl));
  } else {
    res = vadd(res, run_material(m, pv, M_EMI, n, n, V, ibl));
    const V3 amb = v3(lt.ambient[0], lt.ambient[1], lt.ambient[2]);
    res = vadd(res, vmul(amb, run_material(m, pv, M_AMB, n, n, V, ibl)));
  }
  if (flags & TINA_NOLIGHT) return res;
  for (int l = 0; l < lt.n; l++) {
    const Light li = light_at(lt, l);
    const V3 lxyz = li.xyz;
    const float lw = li.w;
    V3 ldir = vsub(lxyz, v3(pos.x * lw, pos.y * lw, pos.z * lw));
    const float dist = sqrtf(vdot(ldir, ldir));
    ldir = vdivs(ldir, dist);
    const float cos_i = vdot(n, ldir);
    if (cos_i > 0.0f) {
      V3 lcolor = li.color;
      lcolor = vsmul(div_scale(1.0f, dist * dist), lcolor);
      const V3 mcolor = run_material(m, pv, M_BRDF, n, ldir, V, ibl);
      res = vadd(res, vmul(vsmul(cos_i, lcolor), mcolor));
    }
  }
  return res;
}

// aces_tonemap (tina/advans.py:31-34)
__device__ __forceinline__ float aces1(float c) {
  return fdiv_rn(c * ((2.51f * c) + 0.03f), (c * ((2.43f * c) + 0.59f)) + 0.14f);
}
__device__ __forceinline__ V3 aces(V3 c) { return v3(aces1(c.x), aces1(c.y), aces1(c.z)); }

// Parameter nodes (nodes.py:42-111, FresnelFactor material.py:68-83) for one pixel.
__device__ __forceinline__ void eval_params(const tina_material& m, uint32_t f, V3 pos, V3 nrm, V3 tc,
                                            V3 color, Slots& pv) {
#pragma unroll
  for (int i = 0; i < TINA_MAX_PARAMS; i++) {
    if (i < m.nparams) {
      const tina_param p = param_at(m, i);
      V3 v = vs(0.0f);
      switch (p.kind) {
        case TINA_P_CONST: v = p.dim == 1 ? vs(p.value[0]) : v3(p.value[0], p.value[1], p.value[2]); break;
        case TINA_P_PARAM: {
          gf32* q = gptr(p.ptr);
          v = p.dim == 1 ? vs(q[0]) : v3(q[0], q[1], q[2]);
          break;
        }
        case TINA_P_COLOR: v = color; break;
        case TINA_P_POS: v = pos; break;
        case TINA_P_NORMAL: v = nrm; break;
        case TINA_P_TEXCOORD: v = tc; break;
        case TINA_P_TEXTURE: {
          const V3 t = pv.get(p.a);
          v = bilerp(gptr(p.ptr), p.tex_w, p.tex_h, p.tex_c, 0, t.x * (float)(p.tex_w - 1),
                     t.y * (float)(p.tex_h - 1));
          break;
        }
        case TINA_P_FACE: {
          gf32* q = gptr(p.ptr);
          v = p.dim == 1 ? vs(q[f]) : v3(q[3 * (size_t)f], q[3 * (size_t)f + 1], q[3 * (size_t)f + 2]);
          break;
        }
        case TINA_P_FRESNEL: {  // metallic * albedo + (1 - metallic) * 0.16 * specular**2
          const V3 mt = pv.get(p.a), al = pv.get(p.b), sp = pv.get(p.c);
          v = vadd(vmul(mt, al), vmul(vsmul(0.16f, vsub(vs(1.0f), mt)), vmul(sp, sp)));
          break;
        }
        default: break;
      }
      pv.v[i] = v;
    }
  }
}

// Everything the deferred pass reads, passed by value.
struct ShadeArgs {
  tina_mesh mesh;
  tina_material mat;
  tina_lights lights;
  tina_ibl ibl;
  Cam cam;
  float bg[3];
  uint32_t flags;
};

// Geometry stage of render_color for face f at pixel (px, py): the setup of
// triangle.py:93-113 recomputed from the vertices (bit-identical to the raster's),
// the weights of :140-151, and interpolate (:31-49).
struct Frag {
  V3 pos, nrm, tc, color;
  float w0, w1, w2;
};

__device__ __forceinline__ void fragment(const ShadeArgs& a, uint32_t f, int px, int py, Frag& fr) {
  gf32* vp = gptr(a.mesh.verts) + 9 * (size_t)f;
  const V3 A = ld3(vp), B = ld3(vp + 3), C = ld3(vp + 6);
  float wA, wB, wC;
  const V3 Av = mapply_pos(a.cam.W2V, A, &wA);
  const V3 Bv = mapply_pos(a.cam.W2V, B, &wB);
  const V3 Cv = mapply_pos(a.cam.W2V, C, &wC);
  const float resx = (float)a.cam.W, resy = (float)a.cam.H;
  const float ax = (Av.x * 0.5f + 0.5f) * resx, ay = (Av.y * 0.5f + 0.5f) * resy;
  const float bx = (Bv.x * 0.5f + 0.5f) * resx, by = (Bv.y * 0.5f + 0.5f) * resy;
  const float cx = (Cv.x * 0.5f + 0.5f) * resx, cy = (Cv.y * 0.5f + 0.5f) * resy;
  const float nn = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
  const RcpRN Rn = rcp_rn(nn);
  const float bcnx = div_rn(bx - cx, Rn), bcny = div_rn(by - cy, Rn);
  const float canx = div_rn(cx - ax, Rn), cany = div_rn(cy - ay, Rn);
  const float ws0 = fdiv_rn(1.0f, wA), ws1 = fdiv_rn(1.0f, wB), ws2 = fdiv_rn(1.0f, wC);
  const float p_x = (float)px + a.cam.bias[0], p_y = (float)py + a.cam.bias[1];
  const float w_bc = (p_x - bx) * bcny - (p_y - by) * bcnx;
  const float w_ca = (p_x - cx) * cany - (p_y - cy) * canx;
  const float t = (1.0f - w_bc) - w_ca;
  float wx = w_bc * ws0, wy = w_ca * ws1, wz = t * ws2;
  const float s = (wx + wy) + wz;
  const RcpRN Rs = rcp_rn(s);
  wx = div_rn(wx, Rs); wy = div_rn(wy, Rs); wz = div_rn(wz, Rs);
  fr.w0 = wx; fr.w1 = wy; fr.w2 = wz;

  fr.pos = vadd(vadd(vsmul(wx, A), vsmul(wy, B)), vsmul(wz, C));
  V3 nrm;
  if ((a.flags & TINA_SMOOTH) && a.mesh.norms) {
    gf32* np = gptr(a.mesh.norms) + 9 * (size_t)f;
    nrm = vadd(vadd(vsmul(wx, ld3(np)), vsmul(wy, ld3(np + 3))), vsmul(wz, ld3(np + 6)));
  } else {
    nrm = vcross(vsub(B, A), vsub(C, A));
  }
  fr.nrm = vnormalized(nrm);
  fr.tc = vs(0.0f);
  if ((a.flags & TINA_TEX) && a.mesh.coors) {
    gf32* q = gptr(a.mesh.coors) + 6 * (size_t)f;
    fr.tc = v3(((wx * q[0]) + (wy * q[2])) + (wz * q[4]), ((wx * q[1]) + (wy * q[3])) + (wz * q[5]), 0.0f);
  }
  fr.color = vs(1.0f);
  if ((a.flags & TINA_VCOLOR) && a.mesh.colors) {
    gf32* cp = gptr(a.mesh.colors) + 9 * (size_t)f;
    fr.color = vadd(vadd(vsmul(wx, ld3(cp)), vsmul(wy, ld3(cp + 3))), vsmul(wz, ld3(cp + 6)));
  }
}

// ---- particles (tina/core/particle.py) ----------------------------------------

// mapply_dir (tina/common.py:186-189): mapply with weight 0, no division.
__device__ __forceinline__ V3 mapply_dir(const float* __restrict__ M, V3 d) {
  float r[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float acc = M[4 * i + 3] * 0.0f;
    acc = acc + M[4 * i + 0] * d.x;
    acc = acc + M[4 * i + 1] * d.y;
    acc = acc + M[4 * i + 2] * d.z;
    r[i] = acc;
  }
  return v3(r[0], r[1], r[2]);
}

// What render_occup keeps of particle i (particle.py:82-104): centre Al, Rl**2,
// view-space z, the depth int(z * 2**30) and the pixel box clamped to the screen.
struct PSetup {
  V3 Al;
  float r2, vz;
  int32_t depth, botx, boty, topx, topy;
};

// false when the particle is clipped (z outside [-1, 1] with clip) or its box is empty
__device__ __forceinline__ bool particle_setup(const Cam& cam, gf32* pos, gf32* rad, long i, bool clip, PSetup& s) {
  const V3 Al = ld3(pos + 3 * i);
  const float Rl = rad[i];
  const V3 Av = mapply_pos(cam.W2V, Al);
  if (clip && !(-1.0f <= Av.z && Av.z <= 1.0f)) return false;
  const V3 DXl = vnormalized(mapply_dir(cam.V2W, v3(1.0f, 0.0f, 0.0f)));
  const V3 DYl = vnormalized(mapply_dir(cam.V2W, v3(0.0f, 1.0f, 0.0f)));
  const float Rvx = mapply_pos(cam.W2V, vadd(Al, vsmul(Rl, DXl))).x - Av.x;
  const float Rvy = mapply_pos(cam.W2V, vadd(Al, vsmul(Rl, DYl))).y - Av.y;
  const float W = (float)cam.W, H = (float)cam.H;
  // to_viewport of Av -/+ (Rv.x, 0, 0) and Av -/+ (0, Rv.y, 0)
  const float b0x = ((Av.x - Rvx) * 0.5f + 0.5f) * W, b0y = ((Av.y - 0.0f) * 0.5f + 0.5f) * H;
  const float b1x = ((Av.x + Rvx) * 0.5f + 0.5f) * W, b1y = ((Av.y + 0.0f) * 0.5f + 0.5f) * H;
  const float b2x = ((Av.x - 0.0f) * 0.5f + 0.5f) * W, b2y = ((Av.y - Rvy) * 0.5f + 0.5f) * H;
  const float b3x = ((Av.x + 0.0f) * 0.5f + 0.5f) * W, b3y = ((Av.y + Rvy) * 0.5f + 0.5f) * H;
  s.botx = max(f2i_sat(floorf(fminf(b0x, b2x))), 0);
  s.boty = max(f2i_sat(floorf(fminf(b0y, b2y))), 0);
  s.topx = min(f2i_sat(ceilf(fmaxf(b1x, b3x))), cam.W - 1);
  s.topy = min(f2i_sat(ceilf(fmaxf(b1y, b3y))), cam.H - 1);
  s.Al = Al;
  s.r2 = Rl * Rl;
  s.vz = Av.z;
  s.depth = f2i_sat(Av.z * (float)TINA_MAXDEPTH);
  return s.topx >= s.botx && s.topy >= s.boty;
}

// Pl = from_viewspace(from_viewport(P + bias), z) (particle.py:106-108, 134-136)
__device__ __forceinline__ V3 particle_plane_point(const Cam& cam, int px, int py, float vz) {
  const float p_x = (float)px + cam.bias[0], p_y = (float)py + cam.bias[1];
  const V3 Pv = v3((fdiv_rn(p_x, (float)cam.W) * 2.0f) - 1.0f, (fdiv_rn(p_y, (float)cam.H) * 2.0f) - 1.0f, vz);
  return mapply_pos(cam.V2W, Pv);
}

// the disc test of particle.py:109: covered unless |Pl - Al|^2 > Rl**2
__device__ __forceinline__ bool particle_covers(const Cam& cam, const PSetup& s, int px, int py) {
  const V3 d = vsub(particle_plane_point(cam, px, py, s.vz), s.Al);
  return !(vdot(d, d) > s.r2);
}

// render_color of particle f at pixel (px, py) (particle.py:125-146): the point
// of the front hemisphere behind the pixel, its normal, the particle's colour.
__device__ __forceinline__ void particle_fragment(const ShadeArgs& a, uint32_t f, int px, int py, Frag& fr) {
  const V3 Al = ld3(gptr(a.mesh.verts) + 3 * (size_t)f);
  const float Rl = gptr(a.mesh.norms)[f];
  const V3 Av = mapply_pos(a.cam.W2V, Al);
  const V3 Pl = particle_plane_point(a.cam, px, py, Av.z);
  V3 Dl = vdivs(vsub(Pl, Al), Rl);
  const V3 Zl = vnormalized(mapply_dir(a.cam.V2W, v3(0.0f, 0.0f, 1.0f)));
  Dl = vsub(Dl, vsmul(sqrtf(1.0f - vdot(Dl, Dl)), Zl));
  Dl = vnormalized(Dl);
  fr.nrm = Dl;
  fr.pos = vadd(Al, vsmul(Rl, Dl));
  fr.tc = vs(0.0f);
  fr.color = vs(1.0f);
  if ((a.flags & TINA_VCOLOR) && a.mesh.colors) fr.color = ld3(gptr(a.mesh.colors) + 3 * (size_t)f);
  fr.w0 = fr.w1 = fr.w2 = 0.0f;
}

// calc_viewdir (shader.py:81-93): -normalize(V2W(uv, +1) - V2W(uv, -1))
__device__ __forceinline__ V3 view_ray(const Cam& cam, int px, int py) {
  const float p_x = (float)px + cam.bias[0], p_y = (float)py + cam.bias[1];
  const float ux = (fdiv_rn(p_x, (float)cam.W) * 2.0f) - 1.0f, uy = (fdiv_rn(p_y, (float)cam.H) * 2.0f) - 1.0f;
  const V3 ro = mapply_pos(cam.V2W, v3(ux, uy, -1.0f));
  const V3 ro1 = mapply_pos(cam.V2W, v3(ux, uy, 1.0f));
  return vnormalized(vsub(ro1, ro));
}

// ---- specialised material programs ------------------------------------------
//
// The presets of tina/matr/material.py:684-706 lower to four fixed programs:
//   Diffuse  [LAMBERT, SCALE a]
//   Classic  [LAMBERT, SCALE a, PHONG m, MIX s]
//   Blinn    [LAMBERT, SCALE a, BLINNPHONG m, MIX s]
//   PBR      [LAMBERT, SCALE a, COOKTORRANCE r f, MIX c]
// For these the kernel is compiled with the program fixed, so no parameter
// array and no interpreter stack are needed; the arithmetic (and its order) is
// the one run_material performs for the same program.
// PROG_PARTICLE is the generic program over particle fragments.
enum Prog { PROG_GENERIC = 0, PROG_DIFFUSE = 1, PROG_CLASSIC = 2, PROG_BLINN = 3, PROG_PBR = 4, PROG_PARTICLE = 5 };

// Value of a non-derived parameter slot (everything but FRESNEL and TEXTURE).
__device__ __forceinline__ V3 eval_basic(const tina_param& p, uint32_t f, const Frag& fr) {
  switch (p.kind) {
    case TINA_P_CONST: return p.dim == 1 ? vs(p.value[0]) : v3(p.value[0], p.value[1], p.value[2]);
    case TINA_P_PARAM: {
      gf32* q = gptr(p.ptr);
      return p.dim == 1 ? vs(q[0]) : v3(q[0], q[1], q[2]);
    }
    case TINA_P_COLOR: return fr.color;
    case TINA_P_POS: return fr.pos;
    case TINA_P_NORMAL: return fr.nrm;
    case TINA_P_TEXCOORD: return fr.tc;
    case TINA_P_FACE: {
      gf32* q = gptr(p.ptr);
      return p.dim == 1 ? vs(q[f]) : v3(q[3 * (size_t)f], q[3 * (size_t)f + 1], q[3 * (size_t)f + 2]);
    }
    default: return vs(0.0f);
  }
}

__device__ __forceinline__ V3 eval_leaf(const tina_material& m, int i, uint32_t f, const Frag& fr) {
  const tina_param p = param_at(m, i);
  if (p.kind == TINA_P_TEXTURE) {
    const V3 t = eval_basic(param_at(m, p.a), f, fr);
    return bilerp(gptr(p.ptr), p.tex_w, p.tex_h, p.tex_c, 0, t.x * (float)(p.tex_w - 1),
                  t.y * (float)(p.tex_h - 1));
  }
  return eval_basic(p, f, fr);
}

__device__ __forceinline__ V3 eval_slot(const tina_material& m, int i, uint32_t f, const Frag& fr) {
  const tina_param p = param_at(m, i);
  if (p.kind == TINA_P_FRESNEL) {  // metallic * albedo + (1 - metallic) * 0.16 * specular**2
    const V3 mt = eval_leaf(m, p.a, f, fr), al = eval_leaf(m, p.b, f, fr), sp = eval_leaf(m, p.c, f, fr);
    return vadd(vmul(mt, al), vmul(vsmul(0.16f, vsub(vs(1.0f), mt)), vmul(sp, sp)));
  }
  return eval_leaf(m, i, f, fr);
}

// The lighting of a fixed program once its slots are known: albedo `alb`, mix
// factor `fac`, the Cook-Torrance f0 and the lobe parameter (roughness or
// shininess), at position `pos` with unit normal n, seen from direction V.
template <int PROG, bool IBL>
__device__ __forceinline__ V3 shade_core(const ShadeArgs& a, V3 n, V3 pos, V3 V, V3 alb, V3 fac, V3 f0,
                                         float lobe_p);

template <int PROG, bool IBL>
__device__ __forceinline__ V3 shade_fixed(const ShadeArgs& a, uint32_t f, const Frag& fr, V3 V) {
  const tina_material& m = a.mat;
  const V3 alb = eval_slot(m, m.ops[1].a, f, fr);
  V3 fac = vs(0.0f), f0 = vs(0.0f);
  float lobe_p = 0.0f;
  if constexpr (PROG != PROG_DIFFUSE) {
    fac = eval_slot(m, m.ops[3].a, f, fr);
    if constexpr (PROG == PROG_PBR) {
      lobe_p = eval_slot(m, m.ops[2].a, f, fr).x;  // roughness
      f0 = m.ops[2].b == m.ops[3].a ? fac : eval_slot(m, m.ops[2].b, f, fr);
    } else {
      lobe_p = eval_slot(m, m.ops[2].a, f, fr).x;  // shineness
    }
  }
  return shade_core<PROG, IBL>(a, fr.nrm, fr.pos, V, alb, fac, f0, lobe_p);
}

template <int PROG, bool IBL>
__device__ __forceinline__ V3 shade_core(const ShadeArgs& a, V3 n, V3 pos, V3 V, V3 alb, V3 fac, V3 f0,
                                         float lobe_p) {
  const float inv_pi = (float)(1.0 / 3.141592653589793);
  const V3 omf = vsub(vs(1.0f), fac);
  V3 res;
  if constexpr (IBL) {
    // sample_ibl: Scale -> wei * fac, Lambert -> diff.sample(nrm), Mix -> (1-f)*w1 + f*w2
    const V3 d = vmul(alb, skybox_sample(gptr(a.ibl.diff), a.ibl.diff_w, a.ibl.diff_h, n));
    if constexpr (PROG == PROG_DIFFUSE) res = d;
    else res = vadd(vmul(omf, d), vmul(fac, cooktorrance_ibl(a.ibl, lobe_p, f0, V, n)));
  } else {
    // emission (0) + ambient_color * ambient(): Scale -> fac * 1, Mix -> (1-f)*w1 + f*1
    const V3 amb = v3(a.lights.ambient[0], a.lights.ambient[1], a.lights.ambient[2]);
    V3 ma = vmul(alb, vs(1.0f));
    if constexpr (PROG != PROG_DIFFUSE) ma = vadd(vmul(omf, ma), vmul(fac, vs(1.0f)));
    res = vadd(vs(0.0f), vmul(amb, ma));
  }
  if (a.flags & TINA_NOLIGHT) return res;
  const V3 diff_brdf = vmul(alb, vs(inv_pi));
  for (int l = 0; l < a.lights.n; l++) {
    const Light li = light_at(a.lights, l);
    const V3 lxyz = li.xyz;
    const float lw = li.w;
    V3 ldir = vsub(lxyz, v3(pos.x * lw, pos.y * lw, pos.z * lw));
    const float dist = sqrtf(vdot(ldir, ldir));
    ldir = vdivs(ldir, dist);
    const float cos_i = vdot(n, ldir);
    if (cos_i > 0.0f) {
      V3 lcolor = li.color;
      lcolor = vsmul(div_scale(1.0f, dist * dist), lcolor);
      V3 mcolor;
      if constexpr (PROG == PROG_DIFFUSE) {
        mcolor = diff_brdf;
      } else {
        V3 spec;
        if constexpr (PROG == PROG_CLASSIC) spec = phong(lobe_p, n, ldir, V);
        else if constexpr (PROG == PROG_BLINN) spec = blinnphong(lobe_p, n, ldir, V);
        else spec = cooktorrance(lobe_p, f0, n, ldir, V);
        mcolor = vadd(vmul(omf, diff_brdf), vmul(fac, spec));
      }
      res = vadd(res, vmul(vsmul(cos_i, lcolor), mcolor));
    }
  }
  return res;
}

template <int PROG, bool IBL>
__device__ __forceinline__ V3 shade_pixel(const ShadeArgs& a, uint32_t f, int px, int py) {
  Frag fr;
  if constexpr (PROG == PROG_PARTICLE) particle_fragment(a, f, px, py, fr);
  else fragment(a, f, px, py, fr);
  const V3 viewdir = vneg(view_ray(a.cam, px, py));
  V3 res;
  if constexpr (PROG == PROG_GENERIC || PROG == PROG_PARTICLE) {
    Slots pv;
    eval_params(a.mat, f, fr.pos, fr.nrm, fr.tc, fr.color, pv);
    res = lighting(a.mat, pv, a.lights, a.ibl, IBL ? (a.flags | TINA_IBL) : (a.flags & ~TINA_IBL), fr.pos,
                   fr.nrm, viewdir);
  } else {
    res = shade_fixed<PROG, IBL>(a, f, fr, viewdir);
  }
  if (a.flags & TINA_TONEMAP) res = aces(res);
  return res;
}

// Host-side pattern match of a lowered program onto one of the fixed kernels.
inline int classify_program(const tina_material& m, bool ibl) {
  const tina_op* o = m.ops;
  if (m.nops == 2 && o[0].op == TINA_OP_LAMBERT && o[1].op == TINA_OP_SCALE) return PROG_DIFFUSE;
  if (m.nops == 4 && o[0].op == TINA_OP_LAMBERT && o[1].op == TINA_OP_SCALE && o[3].op == TINA_OP_MIX) {
    if (o[2].op == TINA_OP_COOKTORRANCE) return PROG_PBR;
    if (!ibl && o[2].op == TINA_OP_PHONG) return PROG_CLASSIC;
    if (!ibl && o[2].op == TINA_OP_BLINNPHONG) return PROG_BLINN;
  }
  return PROG_GENERIC;
}

__device__ __forceinline__ V3 background(const ShadeArgs& a, int px, int py) {
  V3 c = v3(a.bg[0], a.bg[1], a.bg[2]);
  if (a.flags & TINA_IBL)  // Engine.render_background (engine.py:41-49)
    c = skybox_sample(gptr(a.ibl.spec[0]), a.ibl.spec_w[0], a.ibl.spec_h[0], view_ray(a.cam, px, py));
  if (a.flags & TINA_TONEMAP) c = aces(c);
  return c;
}

}  // namespace tina
