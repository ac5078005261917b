// tina_kernels.hip -- the MI355X kernels of the rasterise-and-shade path and the
// C ABI declared in include/tina_hip.h.
//
//   k_clear         Engine.clear_depth + occup = -1 (engine.py:67-70, triangle.py:90-91)
//   k_raster_small  render_occup (triangle.py:92-131), one lane per face: setup,
//                   cull, clip, bbox; faces with a bbox of <= SMALL_AREA pixels are
//                   scanned by their own lane with a 64-bit depth|id atomicMin;
//                   larger ones are appended (wave-ballot compaction, one atomic
//                   per wave) to the big or giant queue.
//   k_raster_big    one workgroup per queued face, pixel-parallel over its bbox.
//   k_raster_giant  faces whose bbox exceeds GIANT_AREA: every workgroup takes a
//                   slice of every giant face, so one huge face spreads over the chip.
//   k_shade         render_color + Shader + Lighting + material (+ ACES, + bg):
//                   per workgroup 1024 pixels, covered ones compacted with a wave
//                   ballot into an LDS list, then shaded densely.
//   k_fill, k_aces, k_resolve   image fill, ToneMapping.apply, key unpacking.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <new>

#include "shade_dev.h"

using namespace tina;

namespace {

constexpr int kBlock = 256;
constexpr int kSmallArea = 64;      // bbox pixels scanned by the face's own lane
constexpr int kGiantArea = 16384;   // bbox pixels above which all workgroups share a face
constexpr int kBigGrid = 2048;      // persistent grid of the big / giant passes
constexpr int kShadePix = 1024;     // pixels per shade workgroup

struct BigRec {
  FaceSetup s;     // 72 B
  uint32_t id;
  uint32_t pad;
};
static_assert(sizeof(BigRec) == 80, "BigRec layout");

struct Counters {
  uint32_t big, giant, done_big, done_giant;
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
  if (pixel_weights(s, (float)px + cam.bias[0], (float)py + cam.bias[1], w0, w1, w2)) {
    const uint64_t k = pack_key(pixel_depth(s, w0, w1, w2), id);
    atomicMin((unsigned long long*)(keys + (size_t)py * cam.W + px), (unsigned long long)k);
  }
}

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

__global__ void __launch_bounds__(kBlock)
k_raster_small(const float* __restrict__ verts, int T, uint32_t id_base, uint32_t flags, Cam cam,
               uint64_t* __restrict__ keys, BigRec* __restrict__ bigq, BigRec* __restrict__ giantq,
               Counters* __restrict__ ctr) {
  __shared__ float sv[kBlock * 9];
  const int tid = threadIdx.x;
  const long f0 = (long)blockIdx.x * kBlock;
  const int nf = (int)min((long)kBlock, (long)T - f0);
  // stage this block's faces through LDS with coalesced loads (36 B per face)
  const float* src = verts + f0 * 9;
  for (int i = tid; i < nf * 9; i += kBlock) sv[i] = src[i];
  __syncthreads();

  bool big = false, giant = false;
  FaceSetup s;
  const uint32_t id = id_base + (uint32_t)(f0 + tid);
  if (tid < nf && setup_face(cam, sv + tid * 9, flags, s)) {
    const int w = s.topx - s.botx + 1, h = s.topy - s.boty + 1;
    if (w > 0 && h > 0) {
      const long area = (long)w * h;
      if (area <= kSmallArea) {
        for (int py = s.boty; py <= s.topy; py++)
          for (int px = s.botx; px <= s.topx; px++) cover_pixel(s, cam, px, py, id, keys);
      } else if (area <= kGiantArea) {
        big = true;
      } else {
        giant = true;
      }
    }
  }
  push_rec(big, s, id, bigq, &ctr->big);
  push_rec(giant, s, id, giantq, &ctr->giant);
}

// One workgroup per big face (persistent grid); lanes stride over the bbox.
__global__ void __launch_bounds__(kBlock)
k_raster_big(const BigRec* __restrict__ q, Counters* __restrict__ ctr, Cam cam, uint64_t* __restrict__ keys) {
  __shared__ uint32_t s_count;
  if (threadIdx.x == 0) s_count = __atomic_load_n(&ctr->big, __ATOMIC_RELAXED);
  __syncthreads();
  const uint32_t count = s_count;
  for (uint32_t r = blockIdx.x; r < count; r += gridDim.x) {
    const BigRec rec = q[r];
    const int w = rec.s.topx - rec.s.botx + 1;
    const int area = w * (rec.s.topy - rec.s.boty + 1);
    for (int i = threadIdx.x; i < area; i += kBlock) {
      const int dy = i / w;
      cover_pixel(rec.s, cam, rec.s.botx + (i - dy * w), rec.s.boty + dy, rec.id, keys);
    }
  }
  // the last workgroup to finish resets the queue for the next raster call;
  // every workgroup has read the count before it takes a ticket
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = atomicAdd(&ctr->done_big, 1u);
    if (t == gridDim.x - 1) {
      __atomic_store_n(&ctr->big, 0u, __ATOMIC_RELAXED);
      __atomic_store_n(&ctr->done_big, 0u, __ATOMIC_RELAXED);
    }
  }
}

// Giant faces: every workgroup takes a slice of every giant face.
__global__ void __launch_bounds__(kBlock)
k_raster_giant(const BigRec* __restrict__ q, Counters* __restrict__ ctr, Cam cam, uint64_t* __restrict__ keys) {
  __shared__ uint32_t s_count;
  if (threadIdx.x == 0) s_count = __atomic_load_n(&ctr->giant, __ATOMIC_RELAXED);
  __syncthreads();
  const uint32_t count = s_count;
  const long stride = (long)gridDim.x * kBlock;
  for (uint32_t r = 0; r < count; r++) {
    const BigRec rec = q[r];
    const long w = rec.s.topx - rec.s.botx + 1;
    const long area = w * (long)(rec.s.topy - rec.s.boty + 1);
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < area; i += stride) {
      const long dy = i / w;
      cover_pixel(rec.s, cam, rec.s.botx + (int)(i - dy * w), rec.s.boty + (int)dy, rec.id, keys);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = atomicAdd(&ctr->done_giant, 1u);
    if (t == gridDim.x - 1) {
      __atomic_store_n(&ctr->giant, 0u, __ATOMIC_RELAXED);
      __atomic_store_n(&ctr->done_giant, 0u, __ATOMIC_RELAXED);
    }
  }
}

__device__ __forceinline__ void store_rgb(float* __restrict__ img, long P, V3 c) {
  img[3 * P] = c.x;
  img[3 * P + 1] = c.y;
  img[3 * P + 2] = c.z;
}

__global__ void __launch_bounds__(kBlock)
k_shade(const ShadeArgs a, const uint64_t* __restrict__ keys, float* __restrict__ img) {
  __shared__ uint32_t s_pix[kShadePix];
  __shared__ uint32_t s_face[kShadePix];
  __shared__ uint32_t s_n;
  const int tid = threadIdx.x;
  const long npix = (long)a.cam.W * a.cam.H;
  const long base = (long)blockIdx.x * kShadePix;
  if (tid == 0) s_n = 0;
  __syncthreads();
  const int lane = __lane_id();
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < kShadePix / kBlock; r++) {
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
    store_rgb(img, P, shade_pixel(a, s_face[j], px, py));
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

}  // namespace

// ============================================================== C ABI

struct tina_ctx {
  int W = 0, H = 0, cap = 0;
  uint64_t* keys = nullptr;
  BigRec* bigq = nullptr;
  BigRec* giantq = nullptr;
  Counters* ctr = nullptr;
  tina_camera cam{};
  std::string err;
};

namespace {

thread_local std::string g_create_err;

int fail(tina_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_err = msg;
  return code;
}

int hip_check(tina_ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return TINA_OK;
  return fail(c, TINA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define TINA_HIP(ctx, call)                                   \
  do {                                                        \
    int _rc = hip_check((ctx), (call), #call);                \
    if (_rc != TINA_OK) return _rc;                           \
  } while (0)

#define TINA_LAUNCH(ctx, name)                                \
  do {                                                        \
    int _rc = hip_check((ctx), hipGetLastError(), name);      \
    if (_rc != TINA_OK) return _rc;                           \
  } while (0)

Cam make_cam(const tina_ctx* c) {
  Cam k;
  memcpy(k.W2V, c->cam.W2V, sizeof(k.W2V));
  memcpy(k.V2W, c->cam.V2W, sizeof(k.V2W));
  k.bias[0] = c->cam.bias[0];
  k.bias[1] = c->cam.bias[1];
  k.W = c->W;
  k.H = c->H;
  return k;
}

unsigned grid_for(long n, long per_block, unsigned cap = 65535u * 8u) {
  long g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > (long)cap) g = cap;
  return (unsigned)g;
}

void free_scratch(tina_ctx* c) {
  if (c->bigq) (void)hipFree(c->bigq);
  if (c->giantq) (void)hipFree(c->giantq);
  c->bigq = c->giantq = nullptr;
}

int alloc_scratch(tina_ctx* c, int cap) {
  free_scratch(c);
  size_t bytes = sizeof(BigRec) * (size_t)(cap > 0 ? cap : 1);
  if (hipMalloc(&c->bigq, bytes) != hipSuccess || hipMalloc(&c->giantq, bytes) != hipSuccess) {
    free_scratch(c);
    (void)hipGetLastError();
    return fail(c, TINA_ERR_ALLOC, "hipMalloc of raster queues failed");
  }
  c->cap = cap;
  return TINA_OK;
}

int check_material(tina_ctx* c, const tina_material* m) {
  if (!m) return fail(c, TINA_ERR_INVALID, "material is NULL");
  if (m->nparams < 0 || m->nparams > TINA_MAX_PARAMS) return fail(c, TINA_ERR_CAPACITY, "too many material parameters");
  if (m->nops < 1 || m->nops > TINA_MAX_OPS) return fail(c, TINA_ERR_CAPACITY, "material program size out of range");
  int depth = 0;
  for (int i = 0; i < m->nops; i++) {
    const tina_op& o = m->ops[i];
    if (o.op < TINA_OP_LAMBERT || o.op > TINA_OP_ADD) return fail(c, TINA_ERR_INVALID, "unknown material op");
    auto slot_ok = [&](int s) { return s >= 0 && s < m->nparams; };
    if ((o.op == TINA_OP_PHONG || o.op == TINA_OP_BLINNPHONG || o.op == TINA_OP_SCALE || o.op == TINA_OP_MIX ||
         o.op == TINA_OP_COOKTORRANCE) && !slot_ok(o.a))
      return fail(c, TINA_ERR_INVALID, "material op refers to a missing parameter slot");
    if (o.op == TINA_OP_COOKTORRANCE && !slot_ok(o.b))
      return fail(c, TINA_ERR_INVALID, "CookTorrance fresnel slot missing");
    if (o.op <= TINA_OP_EMISSION) depth++;
    else if (o.op == TINA_OP_MIX || o.op == TINA_OP_ADD) depth--;
    if (depth < 1 || depth > TINA_MAX_STACK) return fail(c, TINA_ERR_INVALID, "material program stack out of range");
  }
  if (depth != 1) return fail(c, TINA_ERR_INVALID, "material program must leave one value");
  for (int i = 0; i < m->nparams; i++) {
    const tina_param& p = m->params[i];
    if (p.kind < TINA_P_CONST || p.kind > TINA_P_FRESNEL) return fail(c, TINA_ERR_INVALID, "unknown parameter kind");
    if ((p.kind == TINA_P_PARAM || p.kind == TINA_P_FACE || p.kind == TINA_P_TEXTURE) && !p.ptr)
      return fail(c, TINA_ERR_INVALID, "parameter pointer is NULL");
    if (p.kind == TINA_P_FRESNEL && !(p.a >= 0 && p.a < i && p.b >= 0 && p.b < i && p.c >= 0 && p.c < i))
      return fail(c, TINA_ERR_INVALID, "FresnelFactor operands must precede it");
    if (p.kind == TINA_P_TEXTURE && !(p.a >= 0 && p.a < i && p.tex_w > 0 && p.tex_h > 0 && p.tex_c > 0))
      return fail(c, TINA_ERR_INVALID, "texture parameter malformed");
  }
  return TINA_OK;
}

}  // namespace

extern "C" {

int tina_version(void) { return 1; }

const char* tina_last_error(const tina_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int tina_ctx_create(int width, int height, int max_faces, tina_ctx** out) {
  if (!out) return fail(nullptr, TINA_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (width <= 0 || height <= 0 || (long)width * height > (1L << 31) - 1)
    return fail(nullptr, TINA_ERR_INVALID, "resolution must be positive and below 2**31 pixels");
  if (max_faces < 0) return fail(nullptr, TINA_ERR_INVALID, "max_faces must be >= 0");
  tina_ctx* c = new (std::nothrow) tina_ctx();
  if (!c) return fail(nullptr, TINA_ERR_ALLOC, "host allocation failed");
  c->W = width;
  c->H = height;
  // default camera: W2V = V2W = diag(1, 1, -1, 1), bias 0.5 (engine.py:20-26)
  for (int i = 0; i < 16; i++) c->cam.W2V[i] = c->cam.V2W[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  c->cam.W2V[10] = c->cam.V2W[10] = -1.0f;
  c->cam.bias[0] = c->cam.bias[1] = 0.5f;
  if (hipMalloc(&c->keys, sizeof(uint64_t) * (size_t)width * height) != hipSuccess ||
      hipMalloc(&c->ctr, sizeof(Counters)) != hipSuccess) {
    (void)hipGetLastError();
    if (c->keys) (void)hipFree(c->keys);
    delete c;
    return fail(nullptr, TINA_ERR_ALLOC, "hipMalloc of the depth|id buffer failed");
  }
  if (hipMemset(c->ctr, 0, sizeof(Counters)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(c->keys);
    (void)hipFree(c->ctr);
    delete c;
    return fail(nullptr, TINA_ERR_HIP, "counter initialisation failed");
  }
  int rc = alloc_scratch(c, max_faces);
  if (rc != TINA_OK) {
    g_create_err = c->err;
    (void)hipFree(c->keys);
    (void)hipFree(c->ctr);
    delete c;
    return rc;
  }
  *out = c;
  return TINA_OK;
}

int tina_ctx_destroy(tina_ctx* c) {
  if (!c) return TINA_OK;
  free_scratch(c);
  if (c->keys) (void)hipFree(c->keys);
  if (c->ctr) (void)hipFree(c->ctr);
  delete c;
  return TINA_OK;
}

int tina_ctx_reserve(tina_ctx* c, int max_faces) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (max_faces < 0) return fail(c, TINA_ERR_INVALID, "max_faces must be >= 0");
  if (max_faces <= c->cap) return TINA_OK;
  TINA_HIP(c, hipDeviceSynchronize());
  return alloc_scratch(c, max_faces);
}

int tina_set_camera(tina_ctx* c, const tina_camera* cam) {
  if (!c || !cam) return fail(c, TINA_ERR_INVALID, "ctx or camera is NULL");
  c->cam = *cam;
  return TINA_OK;
}

int tina_keys(tina_ctx* c, uint64_t** d_keys) {
  if (!c || !d_keys) return fail(c, TINA_ERR_INVALID, "ctx or out is NULL");
  *d_keys = c->keys;
  return TINA_OK;
}

int tina_clear(tina_ctx* c, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  const long n = (long)c->W * c->H;
  hipLaunchKernelGGL(k_clear, dim3(grid_for(n / 2 + 1, kBlock, 4096)), dim3(kBlock), 0, (hipStream_t)stream, c->keys, n);
  TINA_LAUNCH(c, "k_clear");
  return TINA_OK;
}

int tina_raster(tina_ctx* c, const float* d_verts, int nfaces, uint32_t id_base, uint32_t flags, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (nfaces < 0) return fail(c, TINA_ERR_INVALID, "nfaces must be >= 0");
  if (nfaces == 0) return TINA_OK;
  if (!d_verts) return fail(c, TINA_ERR_INVALID, "verts is NULL");
  if (nfaces > c->cap)
    return fail(c, TINA_ERR_CAPACITY, "nfaces " + std::to_string(nfaces) + " exceeds the context capacity " +
                                          std::to_string(c->cap));
  if ((uint64_t)id_base + (uint64_t)nfaces > 0xFFFFFFFFull)
    return fail(c, TINA_ERR_INVALID, "face ids overflow 32 bits");
  hipStream_t s = (hipStream_t)stream;
  const Cam cam = make_cam(c);
  hipLaunchKernelGGL(k_raster_small, dim3(grid_for(nfaces, kBlock, 0x7fffffff)), dim3(kBlock), 0, s, d_verts,
                     nfaces, id_base, flags, cam, c->keys, c->bigq, c->giantq, c->ctr);
  TINA_LAUNCH(c, "k_raster_small");
  hipLaunchKernelGGL(k_raster_big, dim3(kBigGrid), dim3(kBlock), 0, s, c->bigq, c->ctr, cam, c->keys);
  TINA_LAUNCH(c, "k_raster_big");
  hipLaunchKernelGGL(k_raster_giant, dim3(kBigGrid), dim3(kBlock), 0, s, c->giantq, c->ctr, cam, c->keys);
  TINA_LAUNCH(c, "k_raster_giant");
  return TINA_OK;
}

static int make_shade_args(tina_ctx* c, const tina_mesh* mesh, const tina_material* mat, const tina_lights* lights,
                           const tina_ibl* ibl, const float bg[3], uint32_t flags, ShadeArgs& a) {
  memset(&a, 0, sizeof(a));
  if (mesh) a.mesh = *mesh;
  if (mat) a.mat = *mat;
  if (lights) {
    if (lights->n < 0 || lights->n > TINA_MAX_LIGHTS) return fail(c, TINA_ERR_CAPACITY, "light count out of range");
    a.lights = *lights;
  }
  if (ibl) a.ibl = *ibl;
  if ((flags & TINA_IBL)) {
    if (!ibl || !ibl->spec[0]) return fail(c, TINA_ERR_INVALID, "TINA_IBL needs the IBL tables");
  }
  a.cam = make_cam(c);
  if (bg) { a.bg[0] = bg[0]; a.bg[1] = bg[1]; a.bg[2] = bg[2]; }
  a.flags = flags;
  return TINA_OK;
}

int tina_shade(tina_ctx* c, const tina_mesh* mesh, const tina_material* mat, const tina_lights* lights,
               const tina_ibl* ibl, float* d_img, const float bg[3], uint32_t flags, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!mesh || !d_img || !mesh->verts) return fail(c, TINA_ERR_INVALID, "mesh, verts or image is NULL");
  if (mesh->nfaces < 0) return fail(c, TINA_ERR_INVALID, "nfaces must be >= 0");
  if ((flags & TINA_SMOOTH) && !mesh->norms) return fail(c, TINA_ERR_INVALID, "TINA_SMOOTH without normals");
  if ((flags & TINA_TEX) && !mesh->coors) return fail(c, TINA_ERR_INVALID, "TINA_TEX without texcoords");
  if ((flags & TINA_VCOLOR) && !mesh->colors) return fail(c, TINA_ERR_INVALID, "TINA_VCOLOR without colours");
  int rc = check_material(c, mat);
  if (rc) return rc;
  if (!lights && !(flags & TINA_NOLIGHT)) return fail(c, TINA_ERR_INVALID, "lights is NULL");
  ShadeArgs a;
  rc = make_shade_args(c, mesh, mat, lights, ibl, bg, flags, a);
  if (rc) return rc;
  const long npix = (long)c->W * c->H;
  hipLaunchKernelGGL(k_shade, dim3(grid_for(npix, kShadePix, 0x7fffffff)), dim3(kBlock), 0, (hipStream_t)stream, a,
                     c->keys, d_img);
  TINA_LAUNCH(c, "k_shade");
  return TINA_OK;
}

int tina_fill(tina_ctx* c, float* d_img, const float bg[3], uint32_t flags, void* stream) {
  if (!c || !d_img) return fail(c, TINA_ERR_INVALID, "ctx or image is NULL");
  if (flags & TINA_IBL) return fail(c, TINA_ERR_INVALID, "tina_fill: use tina_shade for the IBL background");
  ShadeArgs a;
  int rc = make_shade_args(c, nullptr, nullptr, nullptr, nullptr, bg, flags & TINA_TONEMAP, a);
  if (rc) return rc;
  const long npix = (long)c->W * c->H;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(npix, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, a, d_img);
  TINA_LAUNCH(c, "k_fill");
  return TINA_OK;
}

int tina_tonemap_aces(float* d_img, int npixels, void* stream) {
  if (!d_img || npixels < 0) return fail(nullptr, TINA_ERR_INVALID, "image is NULL or size negative");
  if (npixels == 0) return TINA_OK;
  const long n = 3L * npixels;
  hipLaunchKernelGGL(k_aces, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, d_img, n);
  return hip_check(nullptr, hipGetLastError(), "k_aces");
}

int tina_resolve(tina_ctx* c, int32_t* d_id, int32_t* d_depth, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!d_id && !d_depth) return TINA_OK;
  const long n = (long)c->W * c->H;
  hipLaunchKernelGGL(k_resolve, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, c->keys, d_id,
                     d_depth, n);
  TINA_LAUNCH(c, "k_resolve");
  return TINA_OK;
}

}  // extern "C"
