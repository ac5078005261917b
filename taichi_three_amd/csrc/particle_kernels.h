// particle_kernels.h -- ParticleRaster.render_occup (tina/core/particle.py:77-116).
//
// Particles go in groups of 64: a lane computes the setup of one (particle_setup,
// the arithmetic of particle.py:82-104), the wave scans the group's pixel-box
// sizes, and then GW waves walk the concatenated boxes 64 * GW pixels at a time,
// each lane finding its particle by a binary search of the prefix in LDS.  A covered pixel (the disc test of :106-110) takes the
// depth|id key of the particle through a 64-bit atomic min, the packed form of
// the atomic_min + occup update of :112-116, with ties to the lower id (the
// earlier particle, as in a serial walk).
#pragma once
#include "shade_dev.h"

namespace tina {

// GW waves walk each group of 64 particles (1, 2 or 4); PRE reads the key before
// the atomic and skips it when the particle cannot win.
constexpr int kParWalkWaves = 4;
constexpr bool kParPrecheck = true;

template <int GW, bool PRE>
__global__ void __launch_bounds__(kBlock)
k_particles(const float* __restrict__ pos, const float* __restrict__ rad, int n, uint32_t id_base, uint32_t flags,
            const Cam cam, unsigned long long* __restrict__ keys) {
  static_assert(GW == 1 || GW == 2 || GW == 4, "waves per group");
  constexpr int NG = (kBlock / 64) / GW;  // groups of 64 particles per block pass
  __shared__ uint64_t s_pre[64 * NG];     // exclusive prefix of the box sizes within the group
  __shared__ uint64_t s_tot[NG];
  __shared__ float s_ax[64 * NG], s_ay[64 * NG], s_az[64 * NG], s_r2[64 * NG], s_vz[64 * NG];
  __shared__ int32_t s_depth[64 * NG], s_botx[64 * NG], s_boty[64 * NG], s_bw[64 * NG];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const bool clip = (flags & TINA_CLIP) != 0;
  gf32* gpos = (gf32*)pos;
  gf32* grad = (gf32*)rad;
  // the block walks particles 64 * NG at a time, so every barrier below is uniform
  for (long bb = (long)blockIdx.x * 64 * NG; bb < n; bb += (long)gridDim.x * 64 * NG) {
    PSetup s;
    uint64_t cnt = 0;
    if (t < 64 * NG) {
      const long i = bb + t;
      if (i < n && particle_setup(cam, gpos, grad, i, clip, s))
        cnt = (uint64_t)(s.topx - s.botx + 1) * (uint64_t)(s.topy - s.boty + 1);
    }
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
        s_ax[t] = s.Al.x;
        s_ay[t] = s.Al.y;
        s_az[t] = s.Al.z;
        s_r2[t] = s.r2;
        s_vz[t] = s.vz;
        s_depth[t] = s.depth;
        s_botx[t] = s.botx;
        s_boty[t] = s.boty;
        s_bw[t] = s.topx - s.botx + 1;
      }
    }
    __syncthreads();
    const int gb = (wave / GW) * 64;
    const uint64_t total = s_tot[wave / GW];
    for (uint64_t k = (uint64_t)((wave % GW) * 64 + lane); k < total; k += 64 * GW) {
      // the last slot whose prefix is <= k owns pixel k (it has a non-empty box)
      int lo = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (s_pre[gb + lo + step] <= k) lo += step;
      const int j = gb + lo;
      const uint32_t local = (uint32_t)(k - s_pre[j]);
      const uint32_t bw = (uint32_t)s_bw[j];
      const uint32_t dy = local / bw, dx = local - dy * bw;
      const int px = s_botx[j] + (int)dx, py = s_boty[j] + (int)dy;
      PSetup q;
      q.Al = v3(s_ax[j], s_ay[j], s_az[j]);
      q.r2 = s_r2[j];
      q.vz = s_vz[j];
      if (!particle_covers(cam, q, px, py)) continue;
      const unsigned long long key = pack_key(s_depth[j], id_base + (uint32_t)(bb + j));
      unsigned long long* kp = keys + (long)py * cam.W + px;
      if (PRE) {
        // keys only decrease while the kernel runs: a key not below one read here loses anyway
        if (key < __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(kp, key);
      } else {
        atomicMin(kp, key);
      }
    }
  }
}

}  // namespace tina
