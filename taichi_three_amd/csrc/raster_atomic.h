// raster_kernels.h -- coverage kernels of render_occup (tina/core/triangle.py:88-131).
//
//   k_clear         Engine.clear_depth + occup = -1 (engine.py:67-70, triangle.py:90-91):
//                   every packed key becomes (depth 2**30, no face).
//   k_raster_small  one lane per face: setup, cull, clip, bbox.  Faces whose bbox has
//                   at most kSmallArea pixels are rasterised cooperatively by their
//                   wave: the 64 bboxes are laid end to end (prefix sum of the areas)
//                   and each lane takes one pixel of that list per step, so a wave
//                   runs ceil(sum of areas / 64) steps instead of max(area) steps
//                   with most lanes idle.  Larger faces are appended to the big or
//                   giant queue (wave-ballot compaction, one atomic per wave).
//   k_raster_queue  big faces: one wave per face, lanes stride over its bbox;
//                   giant faces: every lane of the grid takes a slice of each one.
//
// Every covered pixel is reduced with a 64-bit atomicMin of
// ((depth ^ 0x80000000) << 32) | face_id, which is the serial rule of the
// reference ("depth < depth_buf -> take it", faces in ascending order) without
// its race on `occup`.
#pragma once
#include "tina_dev.h"

namespace tina {

constexpr int kBlock = 256;
constexpr int kSmallArea = 256;     // bbox pixels rasterised by the wave-cooperative pass
constexpr int kGiantArea = 16384;   // bbox pixels above which all lanes share a face
constexpr int kQueueGrid = 256;     // persistent grid of the queue pass (1024 waves)

struct BigRec {
  FaceSetup s;  // 72 B
  uint32_t id;
  uint32_t pad;
};
static_assert(sizeof(BigRec) == 80, "BigRec layout");

struct Counters {
  uint32_t big, giant, done, pad;
};

__global__ void __launch_bounds__(kBlock) k_clear(uint64_t* __restrict__ keys, long n) {
  const uint64_t e = empty_key();
  const long stride = (long)gridDim.x * kBlock;
  const long n2 = n >> 1;
  ulonglong2* k2 = reinterpret_cast<ulonglong2*>(keys);
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n2; i += stride) k2[i] = make_ulonglong2(e, e);
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) keys[n - 1] = e;
}

__device__ __forceinline__ void cover_pixel(const FaceSetup& s, const Cam& cam, int px, int py, uint32_t id,
                                            uint64_t* __restrict__ keys) {
  float w0, w1, w2;
  if (pixel_cover_fast(s, (float)px + cam.bias[0], (float)py + cam.bias[1], w0, w1, w2)) {
    const uint64_t k = pack_key(pixel_depth(s, w0, w1, w2), id);
    atomicMin((unsigned long long*)(keys + (size_t)py * cam.W + px), (unsigned long long)k);
  }
}

// Appends the taking lanes' faces to a queue: one atomic per wave.
__device__ __forceinline__ void push_rec(bool take, const FaceSetup& s, uint32_t id, BigRec* __restrict__ q,
                                         uint32_t* __restrict__ count) {
  const unsigned long long mask = __ballot(take);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (take) {
    const uint32_t idx = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    BigRec r;
    r.s = s;
    r.id = id;
    r.pad = 0;
    q[idx] = r;
  }
}

// Wave-wide inclusive prefix sum (64 lanes).
__device__ __forceinline__ int wave_inclusive_scan(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

__global__ void __launch_bounds__(kBlock)
k_raster_small(const float* __restrict__ verts, int T, uint32_t id_base, uint32_t flags, Cam cam,
               uint64_t* __restrict__ keys, BigRec* __restrict__ bigq, BigRec* __restrict__ giantq,
               Counters* __restrict__ ctr) {
  __shared__ float sv[kBlock * 9];
  // small-face setup, 16-byte records so one lane fetches a face with 4 LDS reads
  __shared__ float4 R0[kBlock], R1[kBlock], R2[kBlock], R3[kBlock];
  __shared__ float2 R4[kBlock];
  __shared__ int32_t Lend[kBlock];  // inclusive prefix sum of the bbox areas within each wave
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wbase = tid & ~63;
  const long f0 = (long)blockIdx.x * kBlock;
  const int nf = (int)min((long)kBlock, (long)T - f0);
  // stage this block's faces through LDS with coalesced loads (36 B per face)
  const float* src = verts + f0 * 9;
  for (int i = tid; i < nf * 9; i += kBlock) sv[i] = src[i];
  __syncthreads();

  bool big = false, giant = false;
  int area = 0;
  FaceSetup s;
  if (tid < nf && setup_face(cam, sv + tid * 9, flags, s)) {
    const int w = s.topx - s.botx + 1, h = s.topy - s.boty + 1;
    if (w > 0 && h > 0) {
      const long a = (long)w * h;
      if (a <= kSmallArea) {
        area = (int)a;
        R0[tid] = make_float4(s.bx, s.by, s.cx, s.cy);
        R1[tid] = make_float4(s.bcnx, s.bcny, s.canx, s.cany);
        R2[tid] = make_float4(s.ws0, s.ws1, s.ws2, s.Az);
        R3[tid] = make_float4(s.Bz, s.Cz, __int_as_float(s.botx), __int_as_float(s.boty));
        R4[tid] = make_float2(__int_as_float(w), 1.0f / (float)w);
      } else if (a <= kGiantArea) {
        big = true;
      } else {
        giant = true;
      }
    }
  }
  const uint32_t id = id_base + (uint32_t)(f0 + tid);
  push_rec(big, s, id, bigq, &ctr->big);
  push_rec(giant, s, id, giantq, &ctr->giant);

  const int end = wave_inclusive_scan(area, lane);
  Lend[tid] = end;
  const int total = __shfl(end, 63, 64);
  __syncthreads();
  const uint32_t id0 = id_base + (uint32_t)(f0 + wbase);
  for (int g = lane; g < total; g += 64) {
    // face k of fragment g: the number of lanes whose prefix end is <= g
    int k = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
      if (Lend[wbase + k + step - 1] <= g) k += step;
    const int j = wbase + k;
    const int start = k > 0 ? Lend[j - 1] : 0;
    const float2 r4 = R4[j];
    const int wj = __float_as_int(r4.x);
    const int li = g - start;
    // li / wj exactly: (li + 1/2) / wj is at least 1/(2 wj) away from an integer
    const int dy = (int)(((float)li + 0.5f) * r4.y);
    const int dx = li - dy * wj;
    const float4 r0 = R0[j], r1 = R1[j], r2 = R2[j], r3 = R3[j];
    FaceSetup q;
    q.bx = r0.x; q.by = r0.y; q.cx = r0.z; q.cy = r0.w;
    q.bcnx = r1.x; q.bcny = r1.y; q.canx = r1.z; q.cany = r1.w;
    q.ws0 = r2.x; q.ws1 = r2.y; q.ws2 = r2.z; q.Az = r2.w;
    q.Bz = r3.x; q.Cz = r3.y;
    cover_pixel(q, cam, __float_as_int(r3.z) + dx, __float_as_int(r3.w) + dy, id0 + (uint32_t)k, keys);
  }
}

__global__ void __launch_bounds__(kBlock)
k_raster_queue(const BigRec* __restrict__ bigq, const BigRec* __restrict__ giantq, Counters* __restrict__ ctr,
               Cam cam, uint64_t* __restrict__ keys) {
  __shared__ uint32_t s_nb, s_ng;
  const int tid = threadIdx.x, lane = __lane_id();
  if (tid == 0) {
    s_nb = __atomic_load_n(&ctr->big, __ATOMIC_RELAXED);
    s_ng = __atomic_load_n(&ctr->giant, __ATOMIC_RELAXED);
  }
  __syncthreads();
  const uint32_t nb = s_nb, ng = s_ng;
  // big faces: one wave per face
  const uint32_t nwaves = gridDim.x * (kBlock / 64);
  for (uint32_t r = blockIdx.x * (kBlock / 64) + (tid >> 6); r < nb; r += nwaves) {
    const BigRec rec = bigq[r];
    const uint32_t w = (uint32_t)(rec.s.topx - rec.s.botx + 1);
    const uint32_t area = w * (uint32_t)(rec.s.topy - rec.s.boty + 1);
    // i / w as a multiply: with magic = ceil(2**32 / w), floor(i * magic / 2**32)
    // equals floor(i / w) whenever i * w < 2**32 (here i, w <= kGiantArea)
    const uint64_t magic = ((1ull << 32) + w - 1) / w;
    for (uint32_t i = lane; i < area; i += 64) {
      const uint32_t dy = (uint32_t)(((uint64_t)i * magic) >> 32);
      cover_pixel(rec.s, cam, rec.s.botx + (int)(i - dy * w), rec.s.boty + (int)dy, rec.id, keys);
    }
  }
  // giant faces: every lane of the grid takes a slice of each one
  const long stride = (long)gridDim.x * kBlock;
  for (uint32_t r = 0; r < ng; r++) {
    const BigRec rec = giantq[r];
    const long w = rec.s.topx - rec.s.botx + 1;
    const long area = w * (long)(rec.s.topy - rec.s.boty + 1);
    for (long i = (long)blockIdx.x * kBlock + tid; i < area; i += stride) {
      const long dy = i / w;
      cover_pixel(rec.s, cam, rec.s.botx + (int)(i - dy * w), rec.s.boty + (int)dy, rec.id, keys);
    }
  }
  // The last workgroup to finish empties the queues for the next raster call;
  // every workgroup has read the counts before it takes its ticket.
  __syncthreads();
  if (tid == 0) {
    const uint32_t t = atomicAdd(&ctr->done, 1u);
    if (t == gridDim.x - 1) {
      __atomic_store_n(&ctr->big, 0u, __ATOMIC_RELAXED);
      __atomic_store_n(&ctr->giant, 0u, __ATOMIC_RELAXED);
      __atomic_store_n(&ctr->done, 0u, __ATOMIC_RELAXED);
    }
  }
}

}  // namespace tina
