// wire_kernels.h -- WireframeRaster (tina/core/wireframe.py:49-95).
//
//   k_wires   the samples draw_line steps along each wire's screen segment
//             (one pixel per step on the major axis, :50-64) enter the depth|id
//             min with the depth interpolated by the 1/w weights of :83-91.
//             Wires go in groups of 64 like particles: a lane sets one up, the
//             wave scans the sample counts, and four waves walk the samples.
//             Only the samples whose major-axis position can land on the screen
//             are walked (the range is found in binary64 with a margin far above
//             the binary32 rounding of a + k i + bias); each walked sample is
//             evaluated and tested exactly as the reference evaluates it.
//   k_paint   blend_color with factor 1 (tina/core/shader.py:134-135): the pixels
//             a range of ids won take one colour (the wire colour).
#pragma once
#include "shade_dev.h"

namespace tina {

// What one wire keeps: viewport start a, step k, 1 / w of the two ends, their
// view-space z, the step count siz of draw_line and the first walked step i0.
struct WSetup {
  float ax, ay, kx, ky, ws0, ws1, Az, Bz;
  int32_t siz, i0;
};

// Steps i in [lo, hi] of a line whose major-axis position a + k i + bias can fall
// in [0, res): bounds in binary64 widened by 2 + 1e-6 of the magnitudes involved.
__device__ __forceinline__ void wire_major_range(float a, float k, float bias, int res, int32_t siz, long& lo,
                                                 long& hi) {
  lo = 0;
  hi = siz;
  if (!(a == a)) return;  // a NaN start comes with siz = 0: the one step is tested as is
  const double ad = a, bd = bias;
  const double err = 2.0 + (fabs(ad) + (double)siz + fabs(bd) + (double)res) * 1e-6;
  double l, h;
  if (k > 0.0f) {
    l = -ad - bd - err;
    h = (double)res - ad - bd + err;
  } else {
    l = ad + bd - (double)res - err;
    h = ad + bd + err;
  }
  l = fmax(l, 0.0);
  h = fmin(h, (double)siz);
  if (!(l <= h)) {
    hi = -1;
    return;
  }
  lo = (long)ceil(l);
  hi = (long)floor(h);
}

// false when the wire is clipped (wireframe.py:75-78) or no step can reach the screen
__device__ __forceinline__ bool wire_setup(const Cam& cam, gf32* verts, long f, bool clip, WSetup& s,
                                           uint64_t& count) {
  const V3 Al = ld3(verts + 6 * f), Bl = ld3(verts + 6 * f + 3);
  float wA, wB;
  const V3 Av = mapply_pos(cam.W2V, Al, &wA);
  const V3 Bv = mapply_pos(cam.W2V, Bl, &wB);
  if (clip) {
    const bool ina = (-1.0f <= Av.x && Av.x <= 1.0f) && (-1.0f <= Av.y && Av.y <= 1.0f) && (-1.0f <= Av.z && Av.z <= 1.0f);
    const bool inb = (-1.0f <= Bv.x && Bv.x <= 1.0f) && (-1.0f <= Bv.y && Bv.y <= 1.0f) && (-1.0f <= Bv.z && Bv.z <= 1.0f);
    if (!ina && !inb) return false;
  }
  const float W = (float)cam.W, H = (float)cam.H;
  const float ax = (Av.x * 0.5f + 0.5f) * W, ay = (Av.y * 0.5f + 0.5f) * H;
  const float bx = (Bv.x * 0.5f + 0.5f) * W, by = (Bv.y * 0.5f + 0.5f) * H;
  s.ws0 = 1.0f / wA;  // wscale = 1 / (w_A, w_B)   (:83)
  s.ws1 = 1.0f / wB;
  // draw_line (:50-61): unit step on the major axis, the slope on the other
  const float dx = bx - ax, dy = by - ay;
  const float adx = fabsf(dx), ady = fabsf(dy);
  long lo, hi;
  if (adx >= ady) {
    s.kx = dx >= 0.0f ? 1.0f : -1.0f;
    s.ky = (s.kx * dy) / dx;
    s.siz = f2i_sat(adx);
    wire_major_range(ax, s.kx, cam.bias[0], cam.W, s.siz, lo, hi);
  } else {
    s.ky = dy >= 0.0f ? 1.0f : -1.0f;
    s.kx = (s.ky * dx) / dy;
    s.siz = f2i_sat(ady);
    wire_major_range(ay, s.ky, cam.bias[1], cam.H, s.siz, lo, hi);
  }
  s.ax = ax;
  s.ay = ay;
  s.Az = Av.z;
  s.Bz = Bv.z;
  s.i0 = (int32_t)lo;
  count = hi >= lo ? (uint64_t)(hi - lo + 1) : 0;
  return count > 0;
}

// Step i of a wire (:62-64, 84-91): false when its pixel is off the screen.
__device__ __forceinline__ bool wire_step(const Cam& cam, const WSetup& s, long i, int& X, int& Y, int32_t& depth) {
  const float fi = (float)i;
  const float px = s.ax + s.kx * fi, py = s.ay + s.ky * fi;
  const float cor = fi / (float)s.siz;
  X = f2i_sat(floorf(px + cam.bias[0]));
  Y = f2i_sat(floorf(py + cam.bias[1]));
  if (X < 0 || X >= cam.W || Y < 0 || Y >= cam.H) return false;
  float w0 = (1.0f - cor) * s.ws0, w1 = cor * s.ws1;
  const float sum = w0 + w1;
  w0 = w0 / sum;
  w1 = w1 / sum;
  depth = f2i_sat(((w0 * s.Az) + (w1 * s.Bz)) * (float)TINA_MAXDEPTH);
  return true;
}

constexpr int kWireWalkWaves = 4;

__global__ void __launch_bounds__(kBlock)
k_wires(const float* __restrict__ verts, int n, uint32_t id_base, uint32_t flags, const Cam cam,
        unsigned long long* __restrict__ keys) {
  constexpr int GW = kWireWalkWaves, NG = (kBlock / 64) / GW;
  __shared__ uint64_t s_pre[64 * NG];
  __shared__ uint64_t s_tot[NG];
  __shared__ float s_f[8][64 * NG];  // ax ay kx ky ws0 ws1 Az Bz
  __shared__ int32_t s_siz[64 * NG], s_i0[64 * NG];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const bool clip = (flags & TINA_CLIP) != 0;
  gf32* gv = (gf32*)verts;
  for (long bb = (long)blockIdx.x * 64 * NG; bb < n; bb += (long)gridDim.x * 64 * NG) {
    WSetup s;
    uint64_t cnt = 0;
    if (t < 64 * NG && bb + t < n && !wire_setup(cam, gv, bb + t, clip, s, cnt)) cnt = 0;
    uint64_t inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t u = __shfl_up(inc, d, 64);
      if (lane >= d) inc += u;
    }
    __syncthreads();  // the previous pass's readers are done with the tables
    if (t < 64 * NG) {
      s_pre[t] = inc - cnt;
      if (lane == 63) s_tot[wave] = inc;
      if (cnt) {
        s_f[0][t] = s.ax;
        s_f[1][t] = s.ay;
        s_f[2][t] = s.kx;
        s_f[3][t] = s.ky;
        s_f[4][t] = s.ws0;
        s_f[5][t] = s.ws1;
        s_f[6][t] = s.Az;
        s_f[7][t] = s.Bz;
        s_siz[t] = s.siz;
        s_i0[t] = s.i0;
      }
    }
    __syncthreads();
    const int gb = (wave / GW) * 64;
    const uint64_t total = s_tot[wave / GW];
    for (uint64_t k = (uint64_t)((wave % GW) * 64 + lane); k < total; k += 64 * GW) {
      int lo = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (s_pre[gb + lo + step] <= k) lo += step;
      const int j = gb + lo;
      WSetup q;
      q.ax = s_f[0][j];
      q.ay = s_f[1][j];
      q.kx = s_f[2][j];
      q.ky = s_f[3][j];
      q.ws0 = s_f[4][j];
      q.ws1 = s_f[5][j];
      q.Az = s_f[6][j];
      q.Bz = s_f[7][j];
      q.siz = s_siz[j];
      int X, Y;
      int32_t depth;
      if (!wire_step(cam, q, (long)s_i0[j] + (long)(k - s_pre[j]), X, Y, depth)) continue;
      const unsigned long long key = pack_key(depth, id_base + (uint32_t)(bb + j));
      unsigned long long* kp = keys + (long)Y * cam.W + X;
      if (key < __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(kp, key);
    }
  }
}

// Pixels won by ids [id_base, id_base + n) take color; with bg_on, empty pixels take bg.
__global__ void __launch_bounds__(kBlock)
k_paint(const uint64_t* __restrict__ keys, long npix, uint32_t id_base, uint32_t n, float cr, float cg, float cb,
        float br, float bgg, float bb, int bg_on, int tonemap, float* __restrict__ img) {
  V3 c = v3(cr, cg, cb), b = v3(br, bgg, bb);
  if (tonemap) {
    c = aces(c);
    b = aces(b);
  }
  for (long P = (long)blockIdx.x * kBlock + threadIdx.x; P < npix; P += (long)gridDim.x * kBlock) {
    const uint64_t k = keys[P];
    if (key_depth(k) >= TINA_MAXDEPTH) {
      if (bg_on) {
        img[3 * P] = b.x;
        img[3 * P + 1] = b.y;
        img[3 * P + 2] = b.z;
      }
    } else if ((uint32_t)k - id_base < n) {
      img[3 * P] = c.x;
      img[3 * P + 1] = c.y;
      img[3 * P + 2] = c.z;
    }
  }
}

}  // namespace tina
