// shade_kernels.h -- deferred shading and the image passes.
//
//   k_shade<PROG, IBL>  render_color + Shader + Lighting + material (+ ACES, + bg)
//                       (tina/core/triangle.py:133-153, shader.py:112-131,
//                       lighting.py:84-98): each workgroup owns a chunk of pixels;
//                       it reads their keys, paints the empty ones with the
//                       background, compacts the covered ones into an LDS list
//                       with a wave ballot, then shades that list densely.
//   k_fill              Scene.image.fill(bgcolor) (tina/scene/raster.py:176)
//   k_aces              ToneMapping.apply (tina/postp/tonemap.py:9-12)
//   k_resolve           unpacks keys into Engine.depth / face ids
#pragma once
#include "bin_kernels.h"
#include "shade_dev.h"

namespace tina {

// Pixels per chunk of a shade workgroup.  The covered pixels of a chunk are
// shaded kBlock at a time, so the last pass of each chunk runs part-full; one
// workgroup per chunk (k_shade) also ends with a partial round of workgroups.
// Measured at 1080p/2M: chunks of 256 / 512 / 1024 pixels shade in 57 / 62 /
// 70 us (the smaller last round wins); the persistent batch kernel has no last
// round and prefers 2048 (1.29 ms per 64-view batch against 1.37 ms for 1024).
constexpr int kShadePix = 256;        // k_shade
constexpr int kBatchShadePix = 2048;  // k_bshade
constexpr int kNnShadePix = 1024;     // k_shade_nn and its fallback queue

__device__ __forceinline__ void store_rgb(float* __restrict__ img, long P, V3 c) {
  img[3 * P] = c.x;
  img[3 * P + 1] = c.y;
  img[3 * P + 2] = c.z;
}

// Pixels [chunk * PIX, (chunk + 1) * PIX) of one shade call (the body of k_shade
// and k_bshade).
template <int PROG, bool IBL, int PIX>
__device__ __forceinline__ void shade_chunk(const ShadeArgs& a, const uint64_t* __restrict__ keys,
                                            float* __restrict__ img, long chunk) {
  static_assert(PIX % kBlock == 0, "whole rows of the workgroup");
  __shared__ uint32_t s_pix[PIX];
  __shared__ uint32_t s_face[PIX];
  __shared__ uint32_t s_n;
  const int tid = threadIdx.x;
  const long npix = (long)a.cam.W * a.cam.H;
  const long base = chunk * PIX;
  if (tid == 0) s_n = 0;
  __syncthreads();
  const int lane = __lane_id();
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < PIX / kBlock; r++) {
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
    // wave-ballot compaction of the covered pixels into the LDS list
    const unsigned long long mask = __ballot(take);
    uint32_t off = 0;
    if (lane == 0 && mask) off = atomicAdd(&s_n, (uint32_t)__popcll(mask));
    off = __shfl(off, 0);
    if (take) {
      const uint32_t j = off + (uint32_t)__popcll(mask & lt_mask);
      s_pix[j] = (uint32_t)(P - base);
      s_face[j] = f;
    }
  }
  __syncthreads();
  const uint32_t n = s_n;
  for (uint32_t j = tid; j < n; j += kBlock) {
    const long P = base + s_pix[j];
    const int px = (int)(P % a.cam.W), py = (int)(P / a.cam.W);
    store_rgb(img, P, shade_pixel<PROG, IBL>(a, s_face[j], px, py));
  }
}

template <int PROG, bool IBL>
__global__ void __launch_bounds__(kBlock)
k_shade(const ShadeArgs a_by_value, const uint64_t* __restrict__ keys, float* __restrict__ img) {
  // Read the 2 KB argument block in place (it is the first kernel argument, at
  // offset 0 of the kernarg segment): naming the by-value parameter would copy it
  // to per-lane scratch.
  (void)a_by_value;
#if defined(__HIP_DEVICE_COMPILE__)
  const ShadeArgs& a = *(const ShadeArgs*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const ShadeArgs& a = a_by_value;  // the host pass only needs the signature
#endif
  shade_chunk<PROG, IBL, kShadePix>(a, keys, img, blockIdx.x);
}

// The same pass held to eight waves per SIMD, for the fixed programs without
// IBL: they fit 64 VGPRs with no spills once the compiler is asked to (78
// SGPRs instead of 106, which held seven waves); the 1080p/2M PBR shade takes
// 55.5 us instead of 57.8.  The IBL, generic and particle programs would spill
// at that budget and keep k_shade.
template <int PROG>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_shade8(const ShadeArgs a_by_value, const uint64_t* __restrict__ keys, float* __restrict__ img) {
  (void)a_by_value;
#if defined(__HIP_DEVICE_COMPILE__)
  const ShadeArgs& a = *(const ShadeArgs*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const ShadeArgs& a = a_by_value;
#endif
  shade_chunk<PROG, false, kShadePix>(a, keys, img, blockIdx.x);
}

// One scene of a batched shade: its arguments, key buffer and image.
struct BatchShade {
  ShadeArgs a;
  const uint64_t* keys;
  float* img;
};

// Persistent workgroups over the pixel chunks of every scene of a batch; chunk
// item belongs to scene s with prefix[s] <= item < prefix[s + 1].
template <int PROG, bool IBL>
__global__ void __launch_bounds__(kBlock)
k_bshade(const BatchShade* __restrict__ scenes, const int* __restrict__ prefix, int nscenes) {
  const long total = prefix[nscenes];
  for (long item = blockIdx.x; item < total; item += gridDim.x) {
    int lo = 0, hi = nscenes;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (prefix[mid] <= item) lo = mid; else hi = mid;
    }
    const BatchShade& sc = scenes[lo];
    shade_chunk<PROG, IBL, kBatchShadePix>(sc.a, sc.keys, sc.img, item - prefix[lo]);
    __syncthreads();
  }
}

template <int PROG, bool IBL>
void launch_shade(unsigned grid, hipStream_t s, const ShadeArgs& a, const uint64_t* keys, float* img) {
  if constexpr (!IBL && PROG != PROG_GENERIC && PROG != PROG_PARTICLE)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade8<PROG>), dim3(grid), dim3(kBlock), 0, s, a, keys, img);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<PROG, IBL>), dim3(grid), dim3(kBlock), 0, s, a, keys, img);
}

inline void dispatch_shade(int prog, bool ibl, unsigned grid, hipStream_t s, const ShadeArgs& a,
                           const uint64_t* keys, float* img) {
  if (ibl) {
    switch (prog) {
      case PROG_DIFFUSE: launch_shade<PROG_DIFFUSE, true>(grid, s, a, keys, img); break;
      case PROG_PBR: launch_shade<PROG_PBR, true>(grid, s, a, keys, img); break;
      case PROG_PARTICLE: launch_shade<PROG_PARTICLE, true>(grid, s, a, keys, img); break;
      default: launch_shade<PROG_GENERIC, true>(grid, s, a, keys, img); break;
    }
  } else {
    switch (prog) {
      case PROG_DIFFUSE: launch_shade<PROG_DIFFUSE, false>(grid, s, a, keys, img); break;
      case PROG_CLASSIC: launch_shade<PROG_CLASSIC, false>(grid, s, a, keys, img); break;
      case PROG_BLINN: launch_shade<PROG_BLINN, false>(grid, s, a, keys, img); break;
      case PROG_PBR: launch_shade<PROG_PBR, false>(grid, s, a, keys, img); break;
      case PROG_PARTICLE: launch_shade<PROG_PARTICLE, false>(grid, s, a, keys, img); break;
      default: launch_shade<PROG_GENERIC, false>(grid, s, a, keys, img); break;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_fill(const ShadeArgs a, float* __restrict__ img) {
  const long npix = (long)a.cam.W * a.cam.H;
  for (long P = (long)blockIdx.x * kBlock + threadIdx.x; P < npix; P += (long)gridDim.x * kBlock)
    store_rgb(img, P, background(a, (int)(P % a.cam.W), (int)(P / a.cam.W)));
}

__global__ void __launch_bounds__(kBlock) k_aces(float* __restrict__ img, long n) {
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
    img[i] = aces1(img[i]);
}

__global__ void __launch_bounds__(kBlock)
k_resolve(const uint64_t* __restrict__ keys, int32_t* __restrict__ id, int32_t* __restrict__ depth, long n) {
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
    const uint64_t k = keys[i];
    const int32_t d = key_depth(k);
    if (depth) depth[i] = d;
    if (id) id[i] = d >= TINA_MAXDEPTH ? -1 : (int32_t)(uint32_t)k;
  }
}

}  // namespace tina
