// ssr_kernels.h -- screen-space reflections (tina/postp/ssr.py:6-103).
//
//   k_ssr        SSR.render_at for the pixels one object won: the view position
//                from the depth buffer, nsamples directions drawn from the
//                object's material (IMaterial.sample, material.py:25-32,
//                124-138, 181-184, 228-238, 365-384, 398-405, 467-477), each
//                marched nsteps steps through the depth buffer; a step landing
//                within the tolerance behind the depth there takes the image
//                colour at that point times the sample weight.
//   k_ssr_apply  SSR.apply: the blur x blur box mean of the reflection buffer
//                (at offset -blur / 2), image * (1 - a) + rgb.
//
// The random stream is the reference's WangHashRNG (tina/random.py:17-64)
// seeded with the pixel's position modulo the box size, so the non-TAA mode
// draws the reference's own numbers.  cos and sin are evaluated in binary64 and
// rounded (the reference's ti.cos / ti.sin codegen is not pinned); every other
// expression keeps the reference's grouping.  Reads past the screen clamp to
// it (the reference reads outside its fields there).
#pragma once
#include "postp_kernels.h"

namespace tina {

// WangHashRNG._noise (random.py:28-35), in uint32 arithmetic
__device__ __forceinline__ uint32_t wang_noise(uint32_t v) {
  v = (v ^ 61u) ^ (v >> 16);
  v *= 9u;
  v ^= v << 4;
  v *= 0x27d4eb2du;
  v ^= v >> 15;
  return v;
}

struct WangRNG {
  uint32_t seed;
  // random(): noise_int(seed) >> 1, times 2 / 2^32 (random.py:38-41, 54-57)
  __device__ __forceinline__ float random() {
    const uint32_t u = wang_noise(seed) >> 1;
    seed += 1u;
    return (float)u * 4.656612873077393e-10f;
  }
  __device__ __forceinline__ uint32_t random_int() {
    const uint32_t u = wang_noise(seed);
    seed += 1u;
    return u;
  }
};

// spherical(h, p) (tina/advans.py:105-108): (sqrt(max(0, 1 - h^2)) (cos, sin)(p tau), h)
__device__ __forceinline__ V3 spherical(float h, float p) {
  const float a = p * 6.2831855f;
  const float c = (float)cos((double)a), s = (float)sin((double)a);
  const float r = sqrtf(fmaxf(0.0f, 1.0f - h * h));
  return v3(r * c, r * s, h);
}

// tangentspace(n) @ d (advans.py:97-100): columns tan, bitan, n
__device__ __forceinline__ V3 tangent_apply(V3 n, V3 d) {
  const V3 bitan = vnormalized(vcross(n, v3(233.0f, 666.0f, 512.0f)));
  const V3 tan = vcross(bitan, n);
  return v3(((tan.x * d.x) + (bitan.x * d.y)) + (n.x * d.z), ((tan.y * d.x) + (bitan.y * d.y)) + (n.y * d.z),
            ((tan.z * d.x) + (bitan.z * d.y)) + (n.z * d.z));
}

// CookTorrance.sub_brdf (material.py:323-357): the Fresnel and visibility terms
__device__ __forceinline__ void ct_sub(float roughness, V3 f0, V3 n, V3 L, V3 V, V3& fdf, float& vdf) {
  const float EPS = 1e-10f;
  const float ONE_M_EPS = (float)(1.0 - 1e-10);
  const V3 half = vnormalized(vadd(L, V));
  const float NoL = fmaxf(EPS, vdot(L, n));
  const float NoV = fmaxf(EPS, vdot(V, n));
  const float VoH = fminf(ONE_M_EPS, fmaxf(EPS, vdot(half, V)));
  const float alpha2 = fmaxf(1e-6f, roughness * roughness);
  const float k = alpha2 * 0.5f;  // x * 0.5 is x / 2 to the bit
  vdf = fdiv_rn(1.0f, ((NoV * k) + 1.0f) - k);
  vdf = vdf * fdiv_rn(1.0f, ((NoL * k) + 1.0f) - k);
  vdf = fdiv_rn(vdf, (1.0f * (1.0f - alpha2)) + ((float)(4.0 * 3.141592653589793) * alpha2));
  const float x1 = 1.0f - VoH, x2 = x1 * x1;
  const float p5 = (x2 * x2) * x1;
  fdf = vadd(f0, vsmul(p5, vsub(vs(1.0f), f0)));
}

// Subtree extents of a post-order material program: size[i] = number of ops in
// the subtree whose root is op i.
__device__ __forceinline__ int program_size(const tina_material& m, int root, const int* size) {
  (void)m;
  return size[root];
}

// One draw of material.sample(idir, nrm, 1, rng) for the program m: the
// combinators pick a branch (MixMaterial with probability smoothlerp(avg(fac),
// .12, .88), AddMaterial by a random integer's parity), the leaf samples its
// lobe, and the branch factors multiply the leaf's weight innermost first.
__device__ __forceinline__ void material_sample(const tina_material& m, const Slots& pv, V3 idir, V3 nrm,
                                                WangRNG& rng, V3& odir, V3& wei) {
  int size[TINA_MAX_OPS];
  for (int i = 0; i < m.nops; i++) {
    const tina_op o = op_at(m, i);
    if (o.op <= TINA_OP_EMISSION) size[i] = 1;
    else if (o.op == TINA_OP_SCALE) size[i] = 1 + size[i - 1];
    else size[i] = 1 + size[i - 1] + size[i - 1 - size[i - 1]];
  }
  V3 fac_stack[TINA_MAX_OPS];
  int nfac = 0;
  int node = m.nops - 1;
  for (;;) {
    const tina_op o = op_at(m, node);
    if (o.op <= TINA_OP_EMISSION) break;
    if (o.op == TINA_OP_SCALE) {             // ScaleMaterial.sample: wei * fac
      fac_stack[nfac++] = pv.get(o.a);
      node = node - 1;
    } else if (o.op == TINA_OP_MIX) {        // MixMaterial.sample (material.py:124-138)
      const V3 fac = pv.get(o.a);
      const bool scalar = param_at(m, o.a).dim == 1;
      float f = scalar ? fac.x : fdiv_rn((fac.x + fac.y) + fac.z, 3.0f);   // Vavg
      if (f != 0.0f && f != 1.0f) {          // smoothlerp(f, 0.12, 0.88)
        const float t = fminf(1.0f, fmaxf(0.0f, fdiv_rn(f - 0.0f, 1.0f - 0.0f)));
        const float s = (t * t) * (3.0f - 2.0f * t);
        f = 0.12f * (1.0f - s) + 0.88f * s;
      }
      if (rng.random() < f) {
        fac_stack[nfac++] = scalar ? vs(fdiv_rn(fac.x, f)) : vdivs(fac, f);
        node = node - 1;
      } else {
        fac_stack[nfac++] = scalar ? vs(fdiv_rn(1.0f - fac.x, 1.0f - f)) : vdivs(vsub(vs(1.0f), fac), 1.0f - f);
        node = node - 1 - size[node - 1];
      }
    } else {                                 // AddMaterial.sample (material.py:228-238)
      const bool first = (rng.random_int() % 2u) == 0u;
      fac_stack[nfac++] = vs(2.0f);
      node = first ? node - 1 - size[node - 1] : node - 1;
    }
  }
  const tina_op leaf = op_at(m, node);
  wei = vs(0.0f);
  odir = idir;
  if (leaf.op == TINA_OP_LAMBERT) {          // Lambert.sample (material.py:398-405)
    float u = rng.random();
    const float v = rng.random();
    u = sqrtf(u);
    odir = vnormalized(tangent_apply(nrm, spherical(u, v)));
    wei = vs(1.0f);
  } else if (leaf.op == TINA_OP_COOKTORRANCE) {  // CookTorrance.sample (material.py:365-384)
    const float roughness = pv.get(leaf.a).x;
    const V3 f0 = pv.get(leaf.b);
    const float alpha2 = fmaxf(0.0f, roughness * roughness);
    float u = rng.random();
    const float v = rng.random();
    u = sqrtf(fdiv_rn(1.0f - u, 1.0f - u * (1.0f - alpha2)));
    const V3 rdir = vreflect(vneg(idir), nrm);
    odir = tangent_apply(rdir, spherical(u, v));
    V3 fdf;
    float vdf;
    ct_sub(roughness, f0, nrm, idir, odir, fdf, vdf);
    if (vdot(odir, nrm) < 0.0f) {
      odir = vneg(odir);
      fdf = vs(0.0f);
    }
    wei = vsmul(vdf, fdf);
  } else if (leaf.op == TINA_OP_PHONG || leaf.op == TINA_OP_BLINNPHONG) {
    // Phong.sample (material.py:467-477); the Blinn-Phong lobe (an extension with
    // no sample of its own) draws the same directions
    const float mm = pv.get(leaf.a).x;
    float u = rng.random();
    const float v = rng.random();
    u = (float)pow((double)u, (double)fdiv_rn(1.0f, mm + 1.0f));
    const V3 rdir = vreflect(vneg(idir), nrm);
    odir = tangent_apply(rdir, spherical(u, v));
    float w = 1.0f;
    if (vdot(odir, nrm) < 0.0f) {
      odir = vneg(odir);
      w = 0.0f;
    }
    wei = vs(w);
  }  // Emission.sample: (idir, 0, 0)
  for (int i = nfac - 1; i >= 0; i--) wei = vmul(wei, fac_stack[i]);
}

// The SSR arguments: the object's mesh and material and the camera (ShadeArgs),
// the screen buffers and the march parameters.
struct SsrArgs {
  const float* norm;   // [H][W][3] NormalShader buffer
  const float* coor;   // [H][W][2] TexcoordShader buffer, or null (texcoord 0)
  const float* image;  // [H][W][3] the linear image the reflections sample
  float* out;          // [H][W][4] rgb + hit fraction
  int nsamples, nsteps, blur, taa;
  float stepsize, tolerance;
  uint32_t seed;       // the TAA mode's frame seed
};

__global__ void __launch_bounds__(kBlock)
k_ssr(const ShadeArgs a_by_value, const uint64_t* __restrict__ keys, const SsrArgs s) {
  (void)a_by_value;
#if defined(__HIP_DEVICE_COMPILE__)
  const ShadeArgs& a = *(const ShadeArgs*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const ShadeArgs& a = a_by_value;
#endif
  const Cam& cam = a.cam;
  const long npix = (long)cam.W * cam.H;
  const float resx = (float)cam.W, resy = (float)cam.H;
  for (long P = (long)blockIdx.x * kBlock + threadIdx.x; P < npix; P += (long)gridDim.x * kBlock) {
    const uint64_t k = keys[P];
    if (key_depth(k) >= TINA_MAXDEPTH) continue;
    const uint32_t f = (uint32_t)k - a.mesh.id_base;
    if (f >= (uint32_t)a.mesh.nfaces) continue;
    const int x = (int)(P % cam.W), y = (int)(P / cam.W);
    const V3 normal = ld_px(s.norm, cam.W, x, y);
    float* o = s.out + 4 * P;
    if (vdot(normal, normal) < 1e-6f) {      // norm_sqr() < eps: nothing to reflect
      o[0] = o[1] = o[2] = o[3] = 0.0f;
      continue;
    }
    const V3 tc = s.coor ? v3(s.coor[2 * P], s.coor[2 * P + 1], 0.0f) : vs(0.0f);
    const float p_x = (float)x + cam.bias[0], p_y = (float)y + cam.bias[1];
    const V3 vpos = v3((fdiv_rn(p_x, resx) * 2.0f) - 1.0f, (fdiv_rn(p_y, resy) * 2.0f) - 1.0f,
                       fdiv_rn((float)key_depth(k), (float)TINA_MAXDEPTH));
    const V3 pos = mapply_pos(cam.V2W, vpos);
    const V3 viewdir = vneg(view_ray(cam, x, y));
    Slots pv;
    eval_params(a.mat, f, pos, normal, tc, vs(1.0f), pv);   // Input('color') is (1, 1, 1) here
    WangRNG rng;
    if (!s.taa) {
      // WangHashRNG(P % blurring): noise_int of the 2-vector (random.py:43-52)
      rng.seed = wang_noise((uint32_t)(y % s.blur) ^ wang_noise((uint32_t)(x % s.blur)));
    } else {
      rng.seed = wang_noise((uint32_t)y ^ wang_noise((uint32_t)x ^ wang_noise(s.seed)));
    }
    const float nst = (float)s.nsteps;
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
    const V3 vpz = mapply_pos(cam.W2V, pos);
    const float vtol =
        s.tolerance * (mapply_pos(cam.W2V, vsub(pos, v3(fdiv_rn(viewdir.x, nst), fdiv_rn(viewdir.y, nst),
                                                        fdiv_rn(viewdir.z, nst)))).z - vpz.z);
    for (int i = 0; i < s.nsamples; i++) {
      V3 odir, wei;
      material_sample(a.mat, pv, viewdir, normal, rng, odir, wei);
      const float c = vdot(odir, viewdir);
      const float step = fdiv_rn(s.stepsize, sqrtf(1.0f - c * c) * nst);
      const float jit = rng.random();
      V3 ro = vadd(pos, vsmul(step, vsmul(jit, odir)));
      for (int j = 0; j < s.nsteps; j++) {
        ro = vadd(ro, vsmul(step, odir));
        const V3 vro = mapply_pos(cam.W2V, ro);
        if (!((-1.0f <= vro.x) & (vro.x <= 1.0f) & (-1.0f <= vro.y) & (vro.y <= 1.0f) & (-1.0f <= vro.z) &
              (vro.z <= 1.0f)))
          break;
        const float Dx = (vro.x * 0.5f + 0.5f) * resx, Dy = (vro.y * 0.5f + 0.5f) * resy;
        const int ix = min((int)Dx, cam.W - 1), iy = min((int)Dy, cam.H - 1);
        const float depth = fdiv_rn((float)key_depth(keys[(long)iy * cam.W + ix]), (float)TINA_MAXDEPTH);
        if ((vro.z - vtol < depth) & (depth < vro.z)) {
          const V3 clr = vmul(bilerp_img(s.image, cam.W, cam.H, Dx, Dy), wei);
          r0 = r0 + clr.x;
          r1 = r1 + clr.y;
          r2 = r2 + clr.z;
          r3 = r3 + 1.0f;
          break;
        }
      }
    }
    const float ns = (float)s.nsamples;
    o[0] = fdiv_rn(r0, ns);
    o[1] = fdiv_rn(r1, ns);
    o[2] = fdiv_rn(r2, ns);
    o[3] = fdiv_rn(r3, ns);
  }
}

// SSR.apply (ssr.py:31-45): res = the blur x blur mean of the buffer at offset
// -blur / 2 (reads clamped; the pixel itself in the TAA mode); image *= 1 - res.w;
// image += res.xyz.
__global__ void __launch_bounds__(kBlock)
k_ssr_apply(float* __restrict__ img, const float* __restrict__ buf, int W, int H, int blur, int taa) {
  const long npix = (long)W * H;
  for (long P = (long)blockIdx.x * kBlock + threadIdx.x; P < npix; P += (long)gridDim.x * kBlock) {
    const int x = (int)(P % W), y = (int)(P / W);
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
    if (taa) {
      const float* b = buf + 4 * P;
      r0 = b[0]; r1 = b[1]; r2 = b[2]; r3 = b[3];
    } else {
      const int offs = blur / 2;
      for (int k = 0; k < blur; k++)
        for (int l = 0; l < blur; l++) {
          const int xx = min(max(x + k - offs, 0), W - 1), yy = min(max(y + l - offs, 0), H - 1);
          const float* b = buf + 4 * ((long)yy * W + xx);
          r0 = r0 + b[0]; r1 = r1 + b[1]; r2 = r2 + b[2]; r3 = r3 + b[3];
        }
      const float n = (float)(blur * blur);
      r0 = fdiv_rn(r0, n); r1 = fdiv_rn(r1, n); r2 = fdiv_rn(r2, n); r3 = fdiv_rn(r3, n);
    }
    float* p = img + 3 * P;
    const float m = 1.0f - r3;
    p[0] = p[0] * m + r0;
    p[1] = p[1] * m + r1;
    p[2] = p[2] * m + r2;
  }
}

}  // namespace tina
