// tina_kernels.hip -- the C ABI of include/tina_hip.h over the MI355X kernels of
// raster_kernels.h (coverage) and shade_kernels.h (deferred shading, image passes).
//
// Every entry point validates its arguments on the host before it enqueues
// anything, returns a status code, and records a message for tina_last_error().
// The per-frame calls only launch kernels on the caller's stream: no allocation,
// no synchronisation, no memset (so a caller may capture them in a HIP graph).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "bin_kernels.h"
#include "diff_kernels.h"
#include "ibl_kernels.h"
#include "nn_kernels.h"
#include "nn_shade.h"
#include "shade_kernels.h"
#include "postp_kernels.h"
#include "ssr_kernels.h"
#include "particle_kernels.h"
#include "wire_kernels.h"

using namespace tina;

struct tina_ctx {
  int W = 0, H = 0, cap = 0;
  int tiles_x = 0, tiles_y = 0, groups_cap = 0;
  uint64_t* keys = nullptr;
  BigRec* bigq = nullptr;
  BigRec* giantq = nullptr;
  Counters* ctr = nullptr;
  // binned path: per (tile, binning group) chunk records, the bin entries, the
  // giant faces by group and each group's giant count
  uint32_t* chunk_off = nullptr;
  uint32_t* chunk_cnt = nullptr;
  uint32_t* entries = nullptr;
  uint32_t* gids = nullptr;
  uint32_t* gcnt = nullptr;
  uint32_t* cursor = nullptr;     // the binning groups' entry cursor (zero between calls)
  FaceRec* recs = nullptr;
  int ncu = 0;                    // compute units of the device
  // learned-shading fallback queue (kFbSeg segments of fb_segcap pixel indices),
  // its segment counters, and k_shade_nn's resident grid
  uint32_t* fbq = nullptr;
  uint32_t* fbn = nullptr;
  long fb_segcap = 0;
  int nn_grid = 0;
  // tina_clear only records the clear; the next call that touches the keys applies it
  // (the binned raster by starting its tiles from empty keys, anything else with
  // k_clear first)
  bool clear_pending = false;
  hipStream_t clear_stream = nullptr;
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

// Workgroups of `kernel` (block threads, no dynamic LDS) that the current device
// holds at once: the grid of a persistent kernel, so that no workgroup waits for
// another to finish.  `fallback` if the runtime cannot say.
int resident_grid(const void* kernel, int block, int fallback) {
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || ncu <= 0 ||
      per_cu <= 0) {
    (void)hipGetLastError();
    return fallback;
  }
  return ncu * per_cu;
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

unsigned grid_for(long n, long per_block, unsigned cap) {
  long g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > (long)cap) g = cap;
  return (unsigned)g;
}

// Lets the binning kernels take more than 64 KB of dynamic LDS (once per process).
void allow_big_lds() {
  static const bool done = [] {
    (void)hipFuncSetAttribute((const void*)k_bin, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBudget);
    (void)hipFuncSetAttribute((const void*)k_bbin, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBudget);
    (void)hipGetLastError();
    return true;
  }();
  (void)done;
}

// Applies a clear recorded by tina_clear, on `stream`.
int flush_clear(tina_ctx* c, hipStream_t stream) {
  if (!c->clear_pending) return TINA_OK;
  c->clear_pending = false;
  const long n = (long)c->W * c->H;
  hipLaunchKernelGGL(k_clear, dim3(grid_for(n / 2 + 1, kBlock, 4096)), dim3(kBlock), 0, stream, c->keys, n);
  return hip_check(c, hipGetLastError(), "k_clear");
}


void free_scratch(tina_ctx* c) {
  for (void* p : {(void*)c->bigq, (void*)c->giantq, (void*)c->chunk_off, (void*)c->chunk_cnt, (void*)c->entries,
                  (void*)c->gids, (void*)c->gcnt, (void*)c->recs})
    if (p) (void)hipFree(p);
  c->bigq = c->giantq = nullptr;
  c->chunk_off = c->chunk_cnt = c->entries = c->gids = c->gcnt = nullptr;
  c->recs = nullptr;
  c->groups_cap = 0;
}

// The binned path needs its per-tile counters and the group's face spans in LDS.
bool binned_fits(const tina_ctx* c, int nfaces) {
  const long ntiles = (long)c->tiles_x * c->tiles_y;
  const int bf = bin_faces_call(ntiles, nfaces, c->ncu);
  if (bf == 0) return false;
  const long groups = (nfaces + bf - 1) / bf;
  return groups <= c->groups_cap && groups <= kMaxBinGroups &&
         (long)nfaces * kBinMaxTiles <= 0xFFFFFFFFL &&  // 32-bit entry indices
         c->W < 32768 && c->H < 32768;                  // FaceRec packs bbox corners in 16 bits
}

int alloc_scratch(tina_ctx* c, int cap) {
  free_scratch(c);
  const size_t n = (size_t)(cap > 0 ? cap : 1);
  const size_t ntiles = (size_t)c->tiles_x * c->tiles_y;
  const int bf = bin_faces_for((long)ntiles);
  const size_t groups = (size_t)bin_groups_cap((long)ntiles, (long)n, c->ncu);
  bool ok = hipMalloc(&c->bigq, sizeof(BigRec) * n) == hipSuccess &&
            hipMalloc(&c->giantq, sizeof(BigRec) * n) == hipSuccess;
  if (ok && bf && groups <= (size_t)kMaxBinGroups) {
    ok = hipMalloc(&c->chunk_off, 4 * ntiles * groups) == hipSuccess &&
         hipMalloc(&c->chunk_cnt, 4 * ntiles * groups) == hipSuccess &&
         hipMalloc(&c->entries, 4 * n * kBinMaxTiles) == hipSuccess &&
         hipMalloc(&c->gids, sizeof(uint32_t) * n) == hipSuccess &&
         hipMalloc(&c->gcnt, sizeof(uint32_t) * groups) == hipSuccess &&
         hipMalloc(&c->recs, sizeof(FaceRec) * n) == hipSuccess;
    if (ok) c->groups_cap = (int)groups;
  }
  if (!ok) {
    free_scratch(c);
    (void)hipGetLastError();
    return fail(c, TINA_ERR_ALLOC, "hipMalloc of the raster bins and queues failed");
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
    if (p.kind == TINA_P_FRESNEL) {
      if (!(p.a >= 0 && p.a < i && p.b >= 0 && p.b < i && p.c >= 0 && p.c < i))
        return fail(c, TINA_ERR_INVALID, "FresnelFactor operands must precede it");
      for (int q : {p.a, p.b, p.c})
        if (m->params[q].kind == TINA_P_FRESNEL)
          return fail(c, TINA_ERR_INVALID, "FresnelFactor operands must be plain parameters");
    }
    if (p.kind == TINA_P_TEXTURE) {
      if (!(p.a >= 0 && p.a < i && p.tex_w > 0 && p.tex_h > 0 && p.tex_c > 0))
        return fail(c, TINA_ERR_INVALID, "texture parameter malformed");
      const int k = m->params[p.a].kind;
      if (k == TINA_P_TEXTURE || k == TINA_P_FRESNEL)
        return fail(c, TINA_ERR_INVALID, "texture coordinates must be a plain parameter");
    }
  }
  return TINA_OK;
}

// The IBL tables a shade reads: the irradiance map, every GGX level with its
// size, and roughness brackets with a positive step (cooktorrance_ibl divides by it).
int check_ibl(tina_ctx* c, const tina_ibl* ibl) {
  if (!ibl || !ibl->diff || !ibl->spec[0]) return fail(c, TINA_ERR_INVALID, "the IBL tables are missing");
  if (ibl->diff_w < 1 || ibl->diff_h < 1) return fail(c, TINA_ERR_INVALID, "IBL irradiance map has no size");
  for (int i = 0; i < TINA_IBL_LEVELS; i++)
    if (!ibl->spec[i] || ibl->spec_w[i] < 1 || ibl->spec_h[i] < 1)
      return fail(c, TINA_ERR_INVALID, "IBL level missing");
  for (int i = 0; i < TINA_IBL_LEVELS - 1; i++)
    if (!(ibl->rough_step[i] > 0.0f)) return fail(c, TINA_ERR_INVALID, "IBL roughness step must be > 0");
  return TINA_OK;
}

int make_shade_args(tina_ctx* c, const tina_mesh* mesh, const tina_material* mat, const tina_lights* lights,
                    const tina_ibl* ibl, const float bg[3], uint32_t flags, ShadeArgs& a) {
  memset(&a, 0, sizeof(a));
  if (mesh) a.mesh = *mesh;
  if (mat) a.mat = *mat;
  if (lights) {
    if (lights->n < 0 || lights->n > TINA_MAX_LIGHTS) return fail(c, TINA_ERR_CAPACITY, "light count out of range");
    a.lights = *lights;
  }
  if (ibl) a.ibl = *ibl;
  if (flags & TINA_IBL) {
    const int rc = check_ibl(c, ibl);
    if (rc) return rc;
  }
  a.cam = make_cam(c);
  if (bg) {
    a.bg[0] = bg[0];
    a.bg[1] = bg[1];
    a.bg[2] = bg[2];
  }
  a.flags = flags;
  return TINA_OK;
}

}  // namespace

// The message tina_last_error(NULL) returns, for the entry points of the other
// translation units (mesh_kernels.hip), which take no context.
namespace tina {
void set_global_error(const std::string& msg) { g_create_err = msg; }
}  // namespace tina

// ---- device self-test of the shared-divisor division --------------------------
//
// Pair i of mode 0: numerator mantissa i mod 2^23 in [1, 2) against the hashed
// divisor number i >> 23 in [1, 2) (every numerator against each divisor, so the
// quotients fall at every distance from the rounding midpoints); mode 1: both
// operands random bit patterns over the whole binary32 range, specials
// included; mode 2: random mantissas with exponents in [-70, 70].  counts[0]
// += pairs that took the shared-divisor path, counts[1] += pairs whose result
// differs in any bit from the compiler's a / d (NaN matches NaN).
__global__ void __launch_bounds__(kBlock)
k_selftest_div(int mode, uint64_t npairs, uint32_t seed, unsigned long long* __restrict__ counts) {
  auto h = [](uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
  };
  unsigned long long fast = 0, bad = 0;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < npairs; i += stride) {
    uint32_t ab, db;
    if (mode == 0) {
      ab = 0x3f800000u | (uint32_t)(i & 0x7fffffu);
      db = 0x3f800000u | (h(seed ^ h((uint32_t)(i >> 23) + 0x9e3779b9u)) & 0x7fffffu);
    } else {
      const uint32_t k = h(seed + (uint32_t)i * 2654435761u) ^ (uint32_t)(i >> 32);
      ab = h(k ^ 0x85ebca6bu);
      db = h(k ^ 0xc2b2ae35u);
      if (mode == 2) {
        ab = (ab & 0x807fffffu) | ((uint32_t)(127 - 70 + (int)(h(ab) % 141u)) << 23);
        db = (db & 0x807fffffu) | ((uint32_t)(127 - 70 + (int)(h(db) % 141u)) << 23);
      }
    }
    const float a = __uint_as_float(ab), d = __uint_as_float(db);
    const RcpRN R = rcp_rn(d);
    const float aa = fabsf(a);
    fast += (R.ok & (aa >= 0x1p-60f) & (aa <= 0x1p60f)) ? 1u : 0u;
    const float got = div_rn(a, R);
    const float want = a / d;
    const bool same = (__float_as_uint(got) == __float_as_uint(want)) || (got != got && want != want);
    bad += same ? 0u : 1u;
  }
  atomicAdd(&counts[0], fast);
  atomicAdd(&counts[1], bad);
}

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
  c->tiles_x = (width + kTile - 1) / kTile;
  c->tiles_y = (height + kTile - 1) / kTile;
  // default camera: W2V = V2W = diag(1, 1, -1, 1), bias 0.5 (engine.py:20-26)
  for (int i = 0; i < 16; i++) c->cam.W2V[i] = c->cam.V2W[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  c->cam.W2V[10] = c->cam.V2W[10] = -1.0f;
  c->cam.bias[0] = c->cam.bias[1] = 0.5f;
  auto release = [c]() {
    for (void* p : {(void*)c->keys, (void*)c->ctr, (void*)c->cursor, (void*)c->fbq, (void*)c->fbn})
      if (p) (void)hipFree(p);
    delete c;
  };
  // segment k takes the fallback pixels of chunks k, k + kFbSeg, ...
  const long nchunks = ((long)width * height + kNnShadePix - 1) / kNnShadePix;
  c->fb_segcap = (nchunks + kFbSeg - 1) / kFbSeg * kNnShadePix;
  if (hipMalloc(&c->keys, sizeof(uint64_t) * (size_t)width * height) != hipSuccess ||
      hipMalloc(&c->ctr, sizeof(Counters)) != hipSuccess || hipMalloc(&c->cursor, sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&c->fbq, sizeof(uint32_t) * (size_t)c->fb_segcap * kFbSeg) != hipSuccess ||
      hipMalloc(&c->fbn, sizeof(uint32_t) * kFbWords) != hipSuccess) {
    (void)hipGetLastError();
    release();
    return fail(nullptr, TINA_ERR_ALLOC, "hipMalloc of the depth|id buffer failed");
  }
  if (hipMemset(c->ctr, 0, sizeof(Counters)) != hipSuccess || hipMemset(c->cursor, 0, sizeof(uint32_t)) != hipSuccess ||
      hipMemset(c->fbn, 0, sizeof(uint32_t) * kFbWords) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    release();
    return fail(nullptr, TINA_ERR_HIP, "counter initialisation failed");
  }
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      c->ncu = 256;
    (void)hipGetLastError();
  }
  int rc = alloc_scratch(c, max_faces);
  if (rc != TINA_OK) {
    g_create_err = c->err;
    release();
    return rc;
  }
  c->nn_grid = resident_grid((const void*)k_shade_nn, kBlock, 768);
  *out = c;
  return TINA_OK;
}

int tina_ctx_destroy(tina_ctx* c) {
  if (!c) return TINA_OK;
  free_scratch(c);
  for (void* p : {(void*)c->keys, (void*)c->ctr, (void*)c->cursor, (void*)c->fbq, (void*)c->fbn})
    if (p) (void)hipFree(p);
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
  int rc = flush_clear(c, c->clear_stream);
  if (rc) return rc;
  *d_keys = c->keys;
  return TINA_OK;
}

int tina_clear(tina_ctx* c, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  c->clear_pending = true;
  c->clear_stream = (hipStream_t)stream;
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
  if (binned_fits(c, nfaces)) {
    BinJob j;
    j.verts = d_verts;
    j.keys = c->keys;
    j.chunk_off = c->chunk_off;
    j.chunk_cnt = c->chunk_cnt;
    j.entries = c->entries;
    j.recs = c->recs;
    j.cursor = c->cursor;
    j.giantq = c->giantq;
    j.gids = c->gids;
    j.gcnt = c->gcnt;
    j.cam = cam;
    j.T = nfaces;
    j.id_base = id_base;
    j.flags = flags;
    j.tiles_x = c->tiles_x;
    j.ntiles = c->tiles_x * c->tiles_y;
    j.bin_faces = bin_faces_call(j.ntiles, nfaces, c->ncu);
    j.ngroups = (nfaces + j.bin_faces - 1) / j.bin_faces;
    j.max_tiles = kBinMaxTiles;
    // a clear recorded by tina_clear is applied by k_tile itself: it starts every
    // tile from empty keys and writes all of them
    j.load_keys = c->clear_pending ? 0 : 1;
    c->clear_pending = false;
    allow_big_lds();
    hipLaunchKernelGGL(k_bin, dim3(j.ngroups), dim3(kBinBlock), (size_t)(j.ntiles + j.bin_faces) * 4, s, j);
    TINA_LAUNCH(c, "k_bin");
    hipLaunchKernelGGL(k_tile, dim3(j.ntiles), dim3(kTileBlock), tile_lds_bytes(j.ngroups), s, j);
    TINA_LAUNCH(c, "k_tile");
    return TINA_OK;
  }
  // resolutions or face counts beyond the binned path's LDS tables: global atomics
  int rc = flush_clear(c, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_raster_small, dim3(grid_for(nfaces, kBlock, 0x7fffffff)), dim3(kBlock), 0, s, d_verts,
                     nfaces, id_base, flags, cam, c->keys, c->bigq, c->giantq, c->ctr);
  TINA_LAUNCH(c, "k_raster_small");
  hipLaunchKernelGGL(k_raster_queue, dim3(kQueueGrid), dim3(kBlock), 0, s, c->bigq, c->giantq, c->ctr, cam,
                     c->keys);
  TINA_LAUNCH(c, "k_raster_queue");
  return TINA_OK;
}

int tina_raster_particles(tina_ctx* c, const float* d_pos, const float* d_radius, int n, uint32_t id_base,
                          uint32_t flags, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (n < 0) return fail(c, TINA_ERR_INVALID, "n must be >= 0");
  if (n == 0) return TINA_OK;
  if (!d_pos || !d_radius) return fail(c, TINA_ERR_INVALID, "particle positions or radii are NULL");
  if ((uint64_t)id_base + (uint64_t)n > 0xFFFFFFFFull) return fail(c, TINA_ERR_INVALID, "particle ids overflow 32 bits");
  if (flags & ~TINA_CLIP) return fail(c, TINA_ERR_INVALID, "particle raster flags: only TINA_CLIP applies");
  hipStream_t s = (hipStream_t)stream;
  int rc = flush_clear(c, s);
  if (rc) return rc;
  constexpr int per_block = kBlock / kParWalkWaves;  // particles per block pass
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_particles<kParWalkWaves, kParPrecheck>), dim3(grid_for(n, per_block, 16384)),
                     dim3(kBlock), 0, s, d_pos, d_radius, n, id_base, flags, make_cam(c), (unsigned long long*)c->keys);
  TINA_LAUNCH(c, "k_particles");
  return TINA_OK;
}

int tina_raster_wires(tina_ctx* c, const float* d_verts, int n, uint32_t id_base, uint32_t flags, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (n < 0) return fail(c, TINA_ERR_INVALID, "n must be >= 0");
  if (n == 0) return TINA_OK;
  if (!d_verts) return fail(c, TINA_ERR_INVALID, "wire vertices are NULL");
  if ((uint64_t)id_base + (uint64_t)n > 0xFFFFFFFFull) return fail(c, TINA_ERR_INVALID, "wire ids overflow 32 bits");
  if (flags & ~TINA_CLIP) return fail(c, TINA_ERR_INVALID, "wire raster flags: only TINA_CLIP applies");
  hipStream_t s = (hipStream_t)stream;
  int rc = flush_clear(c, s);
  if (rc) return rc;
  constexpr int per_block = kBlock / kWireWalkWaves;
  hipLaunchKernelGGL(k_wires, dim3(grid_for(n, per_block, 16384)), dim3(kBlock), 0, s, d_verts, n, id_base, flags,
                     make_cam(c), (unsigned long long*)c->keys);
  TINA_LAUNCH(c, "k_wires");
  return TINA_OK;
}

int tina_paint_ids(tina_ctx* c, uint32_t id_base, int n, const float color[3], const float bg[3], uint32_t flags,
                   float* d_img, void* stream) {
  if (!c || !d_img || !color) return fail(c, TINA_ERR_INVALID, "ctx, image or colour is NULL");
  if (n < 0) return fail(c, TINA_ERR_INVALID, "n must be >= 0");
  if (flags & ~(TINA_BG | TINA_TONEMAP)) return fail(c, TINA_ERR_INVALID, "paint flags: TINA_BG and TINA_TONEMAP only");
  if ((flags & TINA_BG) && !bg) return fail(c, TINA_ERR_INVALID, "TINA_BG without a background colour");
  hipStream_t s = (hipStream_t)stream;
  int rc = flush_clear(c, s);
  if (rc) return rc;
  const long npix = (long)c->W * c->H;
  const float* b = (flags & TINA_BG) ? bg : color;
  hipLaunchKernelGGL(k_paint, dim3(grid_for(npix, kBlock, 8192)), dim3(kBlock), 0, s, (const uint64_t*)c->keys, npix,
                     id_base, (uint32_t)n, color[0], color[1], color[2], b[0], b[1], b[2], (flags & TINA_BG) ? 1 : 0,
                     (flags & TINA_TONEMAP) ? 1 : 0, d_img);
  TINA_LAUNCH(c, "k_paint");
  return TINA_OK;
}

int tina_gbuffer(tina_ctx* c, const tina_mesh* mesh, uint32_t flags, int mode, int channels, const float value[3],
                 float size, float* d_out, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!mesh || !d_out) return fail(c, TINA_ERR_INVALID, "mesh or output buffer is NULL");
  if (flags & ~(TINA_SMOOTH | TINA_TEX | TINA_VCOLOR | TINA_PARTICLES))
    return fail(c, TINA_ERR_INVALID, "gbuffer flags: TINA_SMOOTH, TINA_TEX, TINA_VCOLOR, TINA_PARTICLES");
  if (mode < 0 || mode > 9 || channels < 1 || channels > 3)
    return fail(c, TINA_ERR_INVALID, "gbuffer: mode 0..9 and 1..3 channels");
  if (mode == 9 && !(size > 0.0f)) return fail(c, TINA_ERR_INVALID, "gbuffer: the chessboard size must be > 0");
  if (mesh->nfaces < 0) return fail(c, TINA_ERR_INVALID, "nfaces must be >= 0");
  if (mesh->nfaces == 0) return TINA_OK;
  if (!mesh->verts) return fail(c, TINA_ERR_INVALID, "verts is NULL");
  if ((flags & (TINA_SMOOTH | TINA_PARTICLES)) && !mesh->norms)
    return fail(c, TINA_ERR_INVALID, "TINA_SMOOTH / TINA_PARTICLES need mesh->norms (normals / radii)");
  if ((flags & TINA_TEX) && !mesh->coors) return fail(c, TINA_ERR_INVALID, "TINA_TEX without texcoords");
  if ((flags & TINA_VCOLOR) && !mesh->colors) return fail(c, TINA_ERR_INVALID, "TINA_VCOLOR without colours");
  ShadeArgs a;
  memset(&a, 0, sizeof(a));
  a.mesh = *mesh;
  a.cam = make_cam(c);
  a.flags = flags;
  hipStream_t s = (hipStream_t)stream;
  int rc = flush_clear(c, s);
  if (rc) return rc;
  const long npix = (long)c->W * c->H;
  const float v0 = value ? value[0] : 0.0f, v1 = value ? value[1] : 0.0f, v2 = value ? value[2] : 0.0f;
  const dim3 grid(grid_for(npix, kBlock, 8192));
  const bool part = (flags & TINA_PARTICLES) != 0;
  if (mode == 0 && channels == 3) {
    if (part)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gbuffer<true, 0>), grid, dim3(kBlock), 0, s, a, (const uint64_t*)c->keys,
                         d_out, mode, channels, v0, v1, v2, size);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gbuffer<false, 0>), grid, dim3(kBlock), 0, s, a, (const uint64_t*)c->keys,
                         d_out, mode, channels, v0, v1, v2, size);
  } else if (part) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gbuffer<true, -1>), grid, dim3(kBlock), 0, s, a, (const uint64_t*)c->keys,
                       d_out, mode, channels, v0, v1, v2, size);
  } else {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gbuffer<false, -1>), grid, dim3(kBlock), 0, s, a, (const uint64_t*)c->keys,
                       d_out, mode, channels, v0, v1, v2, size);
  }
  TINA_LAUNCH(c, "k_gbuffer");
  return TINA_OK;
}

int tina_normals(tina_ctx* c, const tina_mesh* mesh, uint32_t flags, float* d_norm, void* stream) {
  if (flags & ~(TINA_SMOOTH | TINA_PARTICLES)) return fail(c, TINA_ERR_INVALID, "normal flags: TINA_SMOOTH, TINA_PARTICLES");
  return tina_gbuffer(c, mesh, flags, 0, 3, nullptr, 1.0f, d_norm, stream);
}

int tina_ssao(tina_ctx* c, const float* d_norm, const float* d_samples, int nsamples, const float* d_rot, int nrot,
              float radius, float thresh, float factor, float* d_ao, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!d_norm || !d_samples || !d_rot || !d_ao) return fail(c, TINA_ERR_INVALID, "tina_ssao: buffer is NULL");
  if (nsamples < 1 || nrot < 1) return fail(c, TINA_ERR_INVALID, "tina_ssao: nsamples and nrot must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  int rc = flush_clear(c, s);
  if (rc) return rc;
  const long npix = (long)c->W * c->H;
  hipLaunchKernelGGL(k_ssao, dim3(grid_for(npix, kBlock, 8192)), dim3(kBlock), 0, s, make_cam(c),
                     (const uint64_t*)c->keys, d_norm, d_samples, nsamples, d_rot, nrot, radius, thresh, factor, d_ao);
  TINA_LAUNCH(c, "k_ssao");
  return TINA_OK;
}

int tina_ssao_apply(float* d_img, int width, int height, const float* d_ao, int nrot, int box, void* stream) {
  if (!d_img || !d_ao) return fail(nullptr, TINA_ERR_INVALID, "tina_ssao_apply: buffer is NULL");
  if (width < 1 || height < 1 || nrot < 1) return fail(nullptr, TINA_ERR_INVALID, "tina_ssao_apply: bad size");
  const long n = (long)width * height;
  if (box && nrot <= 16)
    hipLaunchKernelGGL(k_ssao_apply_box, dim3((width + kAoTile - 1) / kAoTile, (height + kAoTile - 1) / kAoTile),
                       dim3(kBlock), 0, (hipStream_t)stream, d_img, d_ao, width, height, nrot);
  else
    hipLaunchKernelGGL(k_ssao_apply, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, d_img,
                       d_ao, width, height, nrot, box ? 1 : 0);
  return hip_check(nullptr, hipGetLastError(), "k_ssao_apply");
}

int tina_ssr(tina_ctx* c, const tina_mesh* mesh, const tina_material* mat, const float* d_norm, const float* d_coor,
             const float* d_image, int nsamples, int nsteps, float stepsize, float tolerance, int blur, int taa,
             uint32_t seed, float* d_out, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!mesh || !d_norm || !d_image || !d_out) return fail(c, TINA_ERR_INVALID, "tina_ssr: buffer is NULL");
  if (mesh->nfaces < 0) return fail(c, TINA_ERR_INVALID, "nfaces must be >= 0");
  if (nsamples < 1 || nsteps < 1 || blur < 1) return fail(c, TINA_ERR_INVALID, "tina_ssr: nsamples, nsteps, blur >= 1");
  int rc = check_material(c, mat);
  if (rc) return rc;
  if (mesh->nfaces == 0) return TINA_OK;
  ShadeArgs a;
  if ((rc = make_shade_args(c, mesh, mat, nullptr, nullptr, nullptr, 0, a))) return rc;
  if ((rc = flush_clear(c, (hipStream_t)stream))) return rc;
  SsrArgs sa;
  sa.norm = d_norm;
  sa.coor = d_coor;
  sa.image = d_image;
  sa.out = d_out;
  sa.nsamples = nsamples;
  sa.nsteps = nsteps;
  sa.blur = blur;
  sa.taa = taa ? 1 : 0;
  sa.stepsize = stepsize;
  sa.tolerance = tolerance;
  sa.seed = seed;
  const long npix = (long)c->W * c->H;
  hipLaunchKernelGGL(k_ssr, dim3(grid_for(npix, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, a,
                     (const uint64_t*)c->keys, sa);
  TINA_LAUNCH(c, "k_ssr");
  return TINA_OK;
}

int tina_ssr_apply(float* d_img, int width, int height, const float* d_ssr, int blur, int taa, void* stream) {
  if (!d_img || !d_ssr) return fail(nullptr, TINA_ERR_INVALID, "tina_ssr_apply: buffer is NULL");
  if (width < 1 || height < 1 || blur < 1) return fail(nullptr, TINA_ERR_INVALID, "tina_ssr_apply: bad size");
  const long n = (long)width * height;
  hipLaunchKernelGGL(k_ssr_apply, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, d_img, d_ssr,
                     width, height, blur, taa ? 1 : 0);
  return hip_check(nullptr, hipGetLastError(), "k_ssr_apply");
}

int tina_shade(tina_ctx* c, const tina_mesh* mesh, const tina_material* mat, const tina_lights* lights,
               const tina_ibl* ibl, float* d_img, const float bg[3], uint32_t flags, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!mesh || !d_img || !mesh->verts) return fail(c, TINA_ERR_INVALID, "mesh, verts or image is NULL");
  if (mesh->nfaces < 0) return fail(c, TINA_ERR_INVALID, "nfaces must be >= 0");
  if ((flags & TINA_SMOOTH) && !mesh->norms) return fail(c, TINA_ERR_INVALID, "TINA_SMOOTH without normals");
  if ((flags & TINA_TEX) && !mesh->coors) return fail(c, TINA_ERR_INVALID, "TINA_TEX without texcoords");
  if ((flags & TINA_VCOLOR) && !mesh->colors) return fail(c, TINA_ERR_INVALID, "TINA_VCOLOR without colours");
  if ((flags & TINA_PARTICLES) && (flags & (TINA_SMOOTH | TINA_TEX)))
    return fail(c, TINA_ERR_INVALID, "particles take neither TINA_SMOOTH nor TINA_TEX");
  if ((flags & TINA_PARTICLES) && mesh->nfaces > 0 && !mesh->norms)
    return fail(c, TINA_ERR_INVALID, "TINA_PARTICLES needs the radii in mesh->norms");
  int rc = check_material(c, mat);
  if (rc) return rc;
  if (!lights && !(flags & TINA_NOLIGHT)) return fail(c, TINA_ERR_INVALID, "lights is NULL");
  ShadeArgs a;
  rc = make_shade_args(c, mesh, mat, lights, ibl, bg, flags, a);
  if (rc) return rc;
  if ((rc = flush_clear(c, (hipStream_t)stream))) return rc;
  const long npix = (long)c->W * c->H;
  const bool ibl_mode = (flags & TINA_IBL) != 0;
  const int prog = (flags & TINA_PARTICLES) ? (int)PROG_PARTICLE : classify_program(a.mat, ibl_mode);
  dispatch_shade(prog, ibl_mode, grid_for(npix, kShadePix, 0x7fffffff),
                 (hipStream_t)stream, a, c->keys, d_img);
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

int tina_bloom(float* d_img, int width, int height, float thresh, float factor, float scale, const float* d_gwei,
               int radius, float* d_half, float* d_tmp, void* stream) {
  if (!d_img || !d_gwei || !d_half || !d_tmp) return fail(nullptr, TINA_ERR_INVALID, "tina_bloom: buffer is NULL");
  if (width < 2 || height < 2 || radius < 0)
    return fail(nullptr, TINA_ERR_INVALID, "tina_bloom: the image needs 2x2 pixels and radius >= 0");
  const int hw = width / 2, hh = height / 2;
  hipStream_t s = (hipStream_t)stream;
  const long nh = (long)hw * hh, n = (long)width * height;
  hipLaunchKernelGGL(k_bloom_down, dim3(grid_for(nh, kBlock, 4096)), dim3(kBlock), 0, s, d_img, width, d_half, hw, hh,
                     thresh, scale, factor);
  // the LDS-staged passes when their tiles fit 64 KB (radius up to ~1300 along x, ~150 along y)
  const size_t lds_x = (size_t)(kBlurRow + 2 * radius) * 12;
  const size_t lds_y32 = (size_t)(kBlurRows + 2 * radius) * 32 * 12, lds_y16 = lds_y32 / 2;
  if (lds_x <= 65536)
    hipLaunchKernelGGL(k_bloom_blur_x, dim3((hw + kBlurRow - 1) / kBlurRow, hh), dim3(kBlock), lds_x, s, d_half,
                       d_tmp, hw, hh, d_gwei, radius);
  else
    hipLaunchKernelGGL(k_bloom_blur, dim3(grid_for(nh, kBlock, 4096)), dim3(kBlock), 0, s, d_half, d_tmp, hw, hh,
                       d_gwei, radius, 0);
  const unsigned gy = (unsigned)((hh + kBlurRows - 1) / kBlurRows);
  if (lds_y32 <= 65536)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bloom_blur_y<32>), dim3((hw + 31) / 32, gy), dim3(kBlock), lds_y32, s, d_tmp,
                       d_half, hw, hh, d_gwei, radius);
  else if (lds_y16 <= 65536)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bloom_blur_y<16>), dim3((hw + 15) / 16, gy), dim3(kBlock), lds_y16, s, d_tmp,
                       d_half, hw, hh, d_gwei, radius);
  else
    hipLaunchKernelGGL(k_bloom_blur, dim3(grid_for(nh, kBlock, 4096)), dim3(kBlock), 0, s, d_tmp, d_half, hw, hh,
                       d_gwei, radius, 1);
  hipLaunchKernelGGL(k_bloom_add, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, s, d_img, width, height, d_half,
                     hw, hh);
  return hip_check(nullptr, hipGetLastError(), "k_bloom");
}

int tina_fxaa(float* d_img, int width, int height, float abs_thresh, float rel_thresh, float factor, float* d_lumi,
              float* d_copy, void* stream) {
  if (!d_img || !d_lumi || !d_copy) return fail(nullptr, TINA_ERR_INVALID, "tina_fxaa: buffer is NULL");
  if (width < 1 || height < 1) return fail(nullptr, TINA_ERR_INVALID, "tina_fxaa: empty image");
  const long n = (long)width * height;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fxaa_lumi, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, s, d_img, d_lumi, d_copy, n);
  hipLaunchKernelGGL(k_fxaa, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, s, d_img, d_lumi, d_copy, width,
                     height, abs_thresh, rel_thresh, factor);
  return hip_check(nullptr, hipGetLastError(), "k_fxaa");
}

int tina_accumulate(float* d_acc, const float* d_src, int64_t n, int count, void* stream) {
  if (!d_acc || !d_src || n < 0) return fail(nullptr, TINA_ERR_INVALID, "tina_accumulate: bad buffer or size");
  if (count < 1) return fail(nullptr, TINA_ERR_INVALID, "tina_accumulate: count counts from 1");
  if (n == 0) return TINA_OK;
  const float inv = 1.0f / (float)count;
  hipLaunchKernelGGL(k_accumulate, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, d_acc, d_src,
                     (long)n, inv);
  return hip_check(nullptr, hipGetLastError(), "k_accumulate");
}

int tina_bake_ibl(const float* d_env, int ew, int eh, float* d_out, int ow, int oh, float roughness, int nsamples,
                  void* stream) {
  if (!d_env || !d_out) return fail(nullptr, TINA_ERR_INVALID, "tina_bake_ibl: map pointer is NULL");
  if (ew < 2 || eh < 2 || ow < 2 || oh < 2 || ew > 65536 || eh > 65536 || ow > 65536 || oh > 65536)
    return fail(nullptr, TINA_ERR_INVALID, "tina_bake_ibl: map sizes must lie in [2, 65536]");
  if (nsamples < 1) return fail(nullptr, TINA_ERR_INVALID, "tina_bake_ibl: nsamples must be positive");
  if (!(roughness == roughness)) return fail(nullptr, TINA_ERR_INVALID, "tina_bake_ibl: roughness is NaN");
  const double r = (double)roughness;
  const double alpha2 = roughness < 0.0f ? -1.0 : fmax(0.0, r * r);
  const long n = (long)ow * oh;
  hipLaunchKernelGGL(k_bake_ibl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d_env, ew, eh,
                     d_out, ow, oh, alpha2, nsamples);
  return hip_check(nullptr, hipGetLastError(), "k_bake_ibl");
}

// ---- learned shading ----------------------------------------------------------

static int check_sun(tina_ctx* c, const tina_lights* lights) {
  if (!lights || lights->n < 1 || lights->n > TINA_MAX_LIGHTS)
    return fail(c, TINA_ERR_INVALID, "the surrogate needs the lighting with its directional light 0");
  if (lights->dirs[0][3] != 0.0f) return fail(c, TINA_ERR_INVALID, "light 0 of the surrogate must be directional");
  return TINA_OK;
}

int tina_nn_samples(const tina_lights* lights, const tina_ibl* ibl, uint32_t seed, int64_t nsamples, float r_lo,
                    float r_hi, float p_metal, float specular, float* d_x, float* d_y, void* stream) {
  int rc = check_sun(nullptr, lights);
  if (rc) return rc;
  if ((rc = check_ibl(nullptr, ibl))) return rc;
  if (!d_x || !d_y) return fail(nullptr, TINA_ERR_INVALID, "tina_nn_samples: output is NULL");
  if (nsamples < 0) return fail(nullptr, TINA_ERR_INVALID, "tina_nn_samples: nsamples < 0");
  if (!(r_lo > 0.0f && r_lo <= r_hi && r_hi <= 1.0f))
    return fail(nullptr, TINA_ERR_INVALID, "tina_nn_samples: roughness range must lie in (0, 1]");
  if (nsamples == 0) return TINA_OK;
  ShadeArgs a;
  memset(&a, 0, sizeof(a));
  a.lights = *lights;
  a.ibl = *ibl;
  a.flags = TINA_IBL;
  hipLaunchKernelGGL(k_nn_samples, dim3((unsigned)((nsamples + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     seed, (long)nsamples, r_lo, r_hi, p_metal, specular, d_x, d_y);
  return hip_check(nullptr, hipGetLastError(), "k_nn_samples");
}

int tina_nn_analytic_rows(const tina_lights* lights, const tina_ibl* ibl, const float* d_x, int64_t n, float specular,
                          float* d_out, void* stream) {
  int rc = check_sun(nullptr, lights);
  if (rc) return rc;
  if ((rc = check_ibl(nullptr, ibl))) return rc;
  if (n < 0) return fail(nullptr, TINA_ERR_INVALID, "tina_nn_analytic_rows: n < 0");
  if (n == 0) return TINA_OK;
  if (!d_x || !d_out) return fail(nullptr, TINA_ERR_INVALID, "tina_nn_analytic_rows: buffer is NULL");
  ShadeArgs a;
  memset(&a, 0, sizeof(a));
  a.lights = *lights;
  a.ibl = *ibl;
  a.flags = TINA_IBL;
  hipLaunchKernelGGL(k_nn_analytic_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     d_x, (long)n, specular, d_out);
  return hip_check(nullptr, hipGetLastError(), "k_nn_analytic_rows");
}

int tina_mlp_train_step(const float* d_x, const float* d_y, int64_t first, int64_t nsamples, int batch,
                        float* d_params, float* d_m1, float* d_m2, float* d_partial, float lr, int step,
                        float* d_grad, void* stream) {
  if (!d_x || !d_y || !d_params || !d_m1 || !d_m2 || !d_partial)
    return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_train_step: buffer is NULL");
  if (batch < 1 || first < 0 || first + batch > nsamples)
    return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_train_step: the batch must lie inside the samples");
  if (step < 1 || !(lr >= 0.0f)) return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_train_step: step >= 1, lr >= 0");
  const int nblocks = (batch + kNnBatchPerBlock - 1) / kNnBatchPerBlock;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(kGradBlock), 0, (hipStream_t)stream, d_params, d_x, d_y, (long)first,
                     (long)(first + batch), batch, d_partial);
  int rc = hip_check(nullptr, hipGetLastError(), "k_mlp_grad");
  if (rc) return rc;
  const double b1 = 0.9, b2 = 0.999;
  const float c1 = (float)(1.0 / (1.0 - pow(b1, step))), c2 = (float)(1.0 / (1.0 - pow(b2, step)));
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNnParams + 63) / 64), dim3(64 * kAdamGroups), 0, (hipStream_t)stream, d_params, d_m1,
                     d_m2, d_partial, nblocks, lr, (float)b1, (float)b2, 1e-8f, c1, c2, d_grad);
  return hip_check(nullptr, hipGetLastError(), "k_mlp_adam");
}

}  // extern "C"

// ---- the training step without a workspace argument ---------------------------
//
// The gradient rows of k_mlp_grad live in a file-scope device array instead of a
// caller's workspace: kStaticRows rows of kNnParams floats (9.6 MB on every GPU
// the library runs on, held until the process ends).  In host code `scratch`
// names a host-side stand-in; the kernels get the device address that
// hipGetSymbolAddress returns, taken once per device on the entry point's first
// call there.  Every row a step uses is written in full by k_mlp_grad before
// k_mlp_adam reads it, so nothing needs clearing; calls that overlap on one GPU
// share the array and race.
namespace {
constexpr int kStaticRows = 256;  // batch <= 256 * kNnBatchPerBlock = 16384 samples
static __device__ __align__(256) float scratch[(size_t)kStaticRows * kNnParams];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

extern "C" {

int tina_mlp_train_step_static(const float* d_x, const float* d_y, int64_t first, int64_t nsamples, int batch,
                               float* d_params, float* d_m1, float* d_m2, float lr, int step, float* d_grad,
                               void* stream) {
  if (batch > kStaticRows * kNnBatchPerBlock)
    return fail(nullptr, TINA_ERR_CAPACITY, "tina_mlp_train_step_static: batch above " +
                                                std::to_string(kStaticRows * kNnBatchPerBlock) + " samples");
  int dev = 0;
  if (hip_check(nullptr, hipGetDevice(&dev), "hipGetDevice")) return TINA_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, TINA_ERR_CAPACITY, "device index above 63");
  float* rows = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      if (hip_check(nullptr, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
        return TINA_ERR_HIP;
      g_scratch_ptr[dev] = (float*)p;
    }
    rows = g_scratch_ptr[dev];
  }
  return tina_mlp_train_step(d_x, d_y, first, nsamples, batch, d_params, d_m1, d_m2, rows, lr, step, d_grad,
                             stream);
}

int tina_mlp_pack(const float* d_params, void* d_packed, void* stream) {
  if (!d_params || !d_packed) return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_pack: buffer is NULL");
  const int n = kNnFrags * 64 * 8 + 3 * 64 + 4;
  hipLaunchKernelGGL(k_mlp_pack, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_params,
                     (__bf16*)d_packed);
  return hip_check(nullptr, hipGetLastError(), "k_mlp_pack");
}

int tina_mlp_inference(const void* d_packed, const float* d_x, int64_t n, float* d_out, void* stream) {
  if (n < 0) return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_inference: n < 0");
  if (!d_packed) return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_inference: network is NULL");
  if (n == 0) return TINA_OK;
  if (!d_x || !d_out) return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_inference: buffer is NULL");
  if (((uintptr_t)d_x | (uintptr_t)d_out) & 15)
    return fail(nullptr, TINA_ERR_INVALID, "tina_mlp_inference: rows must be 16-byte aligned");
  static const int rows_grid = resident_grid((const void*)k_mlp_rows, kBlock, 768);
  hipLaunchKernelGGL(k_mlp_rows, dim3(grid_for(n, kBlock, (unsigned)rows_grid)), dim3(kBlock), 0, (hipStream_t)stream,
                     (const __bf16*)d_packed, d_x, (long)n, d_out);
  return hip_check(nullptr, hipGetLastError(), "k_mlp_rows");
}

int tina_shade_nn(tina_ctx* c, const tina_mesh* mesh, const tina_material* mat, const tina_lights* lights,
                  const tina_ibl* ibl, const void* d_packed, float* d_img, const float bg[3], uint32_t flags,
                  float threshold, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!mesh || !d_img || !mesh->verts || !d_packed)
    return fail(c, TINA_ERR_INVALID, "mesh, verts, image or network is NULL");
  if (mesh->nfaces < 0) return fail(c, TINA_ERR_INVALID, "nfaces must be >= 0");
  if ((flags & TINA_SMOOTH) && !mesh->norms) return fail(c, TINA_ERR_INVALID, "TINA_SMOOTH without normals");
  if ((flags & TINA_VCOLOR) && !mesh->colors) return fail(c, TINA_ERR_INVALID, "TINA_VCOLOR without colours");
  if ((flags & (TINA_IBL | TINA_TONEMAP)) != (TINA_IBL | TINA_TONEMAP) ||
      (flags & (TINA_NOLIGHT | TINA_TEX | TINA_PARTICLES)))
    return fail(c, TINA_ERR_INVALID, "the surrogate shades triangle meshes, IBL + lights with ACES, without textures");
  int rc = check_material(c, mat);
  if (rc) return rc;
  if ((rc = check_sun(c, lights))) return rc;
  if ((rc = check_ibl(c, ibl))) return rc;
  if (classify_program(*mat, true) != PROG_PBR || mat->params[mat->ops[2].b].kind != TINA_P_FRESNEL ||
      mat->ops[2].b != mat->ops[3].a)
    return fail(c, TINA_ERR_INVALID, "the surrogate replaces the PBR material (CookTorrance f0 = FresnelFactor)");
  ShadeArgs a;
  rc = make_shade_args(c, mesh, mat, lights, ibl, bg, flags, a);
  if (rc) return rc;
  if ((rc = flush_clear(c, (hipStream_t)stream))) return rc;
  const long npix = (long)c->W * c->H;
  hipStream_t s = (hipStream_t)stream;
  uint32_t* fbn = c->fbn;
  hipLaunchKernelGGL(k_shade_nn, dim3(grid_for(npix, kNnShadePix, (unsigned)c->nn_grid)), dim3(kBlock), 0, s, a,
                     c->keys, (const __bf16*)d_packed, threshold, c->fbq, c->fb_segcap, fbn, d_img);
  TINA_LAUNCH(c, "k_shade_nn");
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade_list<PROG_PBR, true>), dim3(grid_for(npix, kBlock, 512)), dim3(kBlock), 0,
                     s, a, c->keys, c->fbq, c->fb_segcap, fbn, d_img);
  TINA_LAUNCH(c, "k_shade_list");
  return TINA_OK;
}

int tina_nn_fallback_count(tina_ctx* c, uint32_t* d_count, void* stream) {
  if (!c || !d_count) return fail(c, TINA_ERR_INVALID, "ctx or out is NULL");
  TINA_HIP(c, hipMemcpyAsync(d_count, c->fbn + kFbSeg * kFbLine, sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return TINA_OK;
}

// ---- differentiable shading ------------------------------------------------------

static int make_diff_args(tina_ctx* c, const tina_dmesh* m, const tina_lights* lights, DiffArgs& A) {
  if (!m || !lights) return fail(c, TINA_ERR_INVALID, "mesh or lights is NULL");
  if (!m->verts || !m->faces || !m->normals || !m->colors || !m->rough || !m->metal || !m->intensity)
    return fail(c, TINA_ERR_INVALID, "a differentiable-mesh array is NULL");
  if (m->nfaces < 0 || m->nverts < 0) return fail(c, TINA_ERR_INVALID, "negative mesh size");
  if (lights->n < 0 || lights->n > TINA_MAX_LIGHTS) return fail(c, TINA_ERR_CAPACITY, "light count out of range");
  memset(&A, 0, sizeof(A));
  A.cam = make_cam(c);
  A.soup = m->verts;
  A.faces = m->faces;
  A.nrm = m->normals;
  A.col = m->colors;
  A.rough = m->rough;
  A.metal = m->metal;
  A.inten = m->intensity;
  A.nfaces = m->nfaces;
  A.id_base = m->id_base;
  A.nlights = lights->n;
  for (int l = 0; l < lights->n; l++) {
    if (lights->dirs[l][3] != 1.0f) return fail(c, TINA_ERR_INVALID, "differentiable shading takes point lights");
    for (int k = 0; k < 3; k++) {
      A.lpos[l][k] = lights->dirs[l][k];
      A.lcol[l][k] = lights->colors[l][k];
    }
  }
  for (int k = 0; k < 3; k++) A.amb[k] = m->ambient[k];
  A.f0_dielectric = 0.16f * (m->specular * m->specular);
  return TINA_OK;
}

int tina_dshade_forward(tina_ctx* c, const tina_dmesh* m, const tina_lights* lights, float* d_img, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!d_img) return fail(c, TINA_ERR_INVALID, "image is NULL");
  DiffArgs A;
  int rc = make_diff_args(c, m, lights, A);
  if (rc) return rc;
  if ((rc = flush_clear(c, (hipStream_t)stream))) return rc;
  const long npix = (long)c->W * c->H;
  hipLaunchKernelGGL(k_dshade_fwd, dim3(grid_for(npix, kBlock, 0x7fffffff)), dim3(kBlock), 0, (hipStream_t)stream,
                     A, c->keys, d_img);
  TINA_LAUNCH(c, "k_dshade_fwd");
  return TINA_OK;
}

int tina_dshade_backward(tina_ctx* c, const tina_dmesh* m, const tina_lights* lights, const float* d_gimg,
                         const tina_dgrads* g, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!d_gimg || !g || !g->colors || !g->normals || !g->rough || !g->metal || !g->intensity)
    return fail(c, TINA_ERR_INVALID, "gradient buffer is NULL");
  DiffArgs A;
  int rc = make_diff_args(c, m, lights, A);
  if (rc) return rc;
  if ((rc = flush_clear(c, (hipStream_t)stream))) return rc;
  const long npix = (long)c->W * c->H;
  hipLaunchKernelGGL(k_dshade_bwd, dim3(grid_for(npix, kBlock, 0x7fffffff)), dim3(kBlock), 0, (hipStream_t)stream,
                     A, c->keys, d_gimg, (float*)g->colors, (float*)g->normals, (float*)g->rough, (float*)g->metal,
                     (float*)g->intensity);
  TINA_LAUNCH(c, "k_dshade_bwd");
  return TINA_OK;
}

// ---- batched multi-view rendering ---------------------------------------------------

}  // extern "C"

struct tina_batch {
  int n = 0;
  tina_batch_tuning tune{};
  int prog = 0;
  long total_pix = 0;
  int max_ntiles = 0, max_groups = 0;
  uint64_t* keys = nullptr;           // all scenes' keys, back to back
  std::vector<void*> allocs;          // per-scene raster scratch
  BinJob* d_jobs = nullptr;           // [n]
  BatchShade* d_shade = nullptr;      // [n]
  int* d_prefix = nullptr;            // [3][n + 1]: binning groups, tiles, shade chunks
  int totals[3] = {0, 0, 0};
  std::vector<long> key_off;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

static void batch_free(tina_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void* p : b->allocs) (void)hipFree(p);
  for (void* p : {(void*)b->keys, (void*)b->d_jobs, (void*)b->d_shade, (void*)b->d_prefix})
    if (p) (void)hipFree(p);
  delete b;
}

int tina_batch_create(int nscenes, const tina_scene_desc* scenes, const tina_batch_tuning* tune, tina_batch** out) {
  if (!out || !scenes || !tune) return fail(nullptr, TINA_ERR_INVALID, "tina_batch_create: NULL argument");
  *out = nullptr;
  if (nscenes < 1 || nscenes > kBlock) return fail(nullptr, TINA_ERR_INVALID, "a batch holds 1 to 256 scenes");
  const int bf = tune->bin_faces;
  if (!(bf == 1024 || bf == 2048 || bf == 4096 || bf == 8192))
    return fail(nullptr, TINA_ERR_INVALID, "bin_faces must be 1024, 2048, 4096 or 8192");
  if (tune->max_tiles < 1 || tune->max_tiles > 256) return fail(nullptr, TINA_ERR_INVALID, "max_tiles out of [1, 256]");
  if (tune->grid_bin < 0 || tune->grid_tile < 0 || tune->grid_shade < 0)
    return fail(nullptr, TINA_ERR_INVALID, "grid sizes must be >= 0");
  tina_batch* b = new (std::nothrow) tina_batch();
  if (!b) return fail(nullptr, TINA_ERR_ALLOC, "host allocation failed");
  b->n = nscenes;
  b->tune = *tune;
  std::vector<BinJob> jobs(nscenes);
  std::vector<BatchShade> sh(nscenes);
  std::vector<int> prefix(3 * (nscenes + 1), 0);
  b->key_off.resize(nscenes);
  for (int i = 0; i < nscenes; i++) {
    const tina_scene_desc& d = scenes[i];
    if (d.width <= 0 || d.height <= 0 || d.width >= 32768 || d.height >= 32768 || d.nfaces < 0 || !d.mesh.verts ||
        !d.img || d.mesh.nfaces != d.nfaces || d.mesh.id_base != 0 ||
        (d.shade_flags & (TINA_PARTICLES | TINA_IBL | TINA_NOLIGHT))) {
      batch_free(b);
      return fail(nullptr, TINA_ERR_INVALID, "scene " + std::to_string(i) + ": malformed descriptor");
    }
    int rc = check_material(nullptr, &d.mat);
    if (rc) {
      batch_free(b);
      return rc;
    }
    const int prog = classify_program(d.mat, false);
    if (i == 0) b->prog = prog;
    if (prog != b->prog) {
      batch_free(b);
      return fail(nullptr, TINA_ERR_INVALID, "every scene of a batch needs the same material program");
    }
    const int tx = (d.width + kTile - 1) / kTile, ty = (d.height + kTile - 1) / kTile;
    const int ntiles = tx * ty;
    const int ngroups = (d.nfaces + bf - 1) / bf;
    if ((ntiles + bf) * 4 > kLdsBudget || ngroups > kMaxBinGroups ||
        (long)d.nfaces * tune->max_tiles > 0xFFFFFFFFL) {
      batch_free(b);
      return fail(nullptr, TINA_ERR_CAPACITY, "scene " + std::to_string(i) + ": screen or face count too large for the binned path");
    }
    b->max_ntiles = std::max(b->max_ntiles, ntiles);
    b->max_groups = std::max(b->max_groups, ngroups);
    b->key_off[i] = b->total_pix;
    b->total_pix += (long)d.width * d.height;
    const size_t T = (size_t)std::max(d.nfaces, 1);
    void *co = nullptr, *cc = nullptr, *en = nullptr, *rc_ = nullptr, *gq = nullptr, *gi = nullptr, *gn = nullptr,
         *cu = nullptr;
    bool ok = hipMalloc(&co, 4 * (size_t)ntiles * std::max(ngroups, 1)) == hipSuccess &&
              hipMalloc(&cc, 4 * (size_t)ntiles * std::max(ngroups, 1)) == hipSuccess &&
              hipMalloc(&en, 4 * T * (size_t)tune->max_tiles) == hipSuccess &&
              hipMalloc(&rc_, sizeof(FaceRec) * T) == hipSuccess && hipMalloc(&gq, sizeof(BigRec) * T) == hipSuccess &&
              hipMalloc(&gi, sizeof(uint32_t) * T) == hipSuccess &&
              hipMalloc(&gn, sizeof(uint32_t) * (size_t)std::max(ngroups, 1)) == hipSuccess &&
              hipMalloc(&cu, sizeof(uint32_t)) == hipSuccess;
    for (void* p : {co, cc, en, rc_, gq, gi, gn, cu})
      if (p) b->allocs.push_back(p);
    if (!ok || hipMemset(cu, 0, sizeof(uint32_t)) != hipSuccess) {
      (void)hipGetLastError();
      batch_free(b);
      return fail(nullptr, TINA_ERR_ALLOC, "hipMalloc of the batch scratch failed");
    }
    BinJob& j = jobs[i];
    j.verts = (const float*)(uintptr_t)d.mesh.verts;
    j.keys = nullptr;  // set below, once the key buffer exists
    j.chunk_off = (uint32_t*)co;
    j.chunk_cnt = (uint32_t*)cc;
    j.entries = (uint32_t*)en;
    j.recs = (FaceRec*)rc_;
    j.cursor = (uint32_t*)cu;
    j.giantq = (BigRec*)gq;
    j.gids = (uint32_t*)gi;
    j.gcnt = (uint32_t*)gn;
    memcpy(j.cam.W2V, d.cam.W2V, sizeof(j.cam.W2V));
    memcpy(j.cam.V2W, d.cam.V2W, sizeof(j.cam.V2W));
    j.cam.bias[0] = d.cam.bias[0];
    j.cam.bias[1] = d.cam.bias[1];
    j.cam.W = d.width;
    j.cam.H = d.height;
    j.T = d.nfaces;
    j.id_base = 0;
    j.flags = d.raster_flags;
    j.tiles_x = tx;
    j.ntiles = ntiles;
    j.ngroups = ngroups;
    j.bin_faces = bf;
    j.max_tiles = tune->max_tiles;
    j.load_keys = 0;  // k_btile writes every tile of every scene: no clear pass
    ShadeArgs& a = sh[i].a;
    memset(&a, 0, sizeof(a));
    a.mesh = d.mesh;
    a.mat = d.mat;
    a.lights = d.lights;
    a.cam = j.cam;
    for (int k = 0; k < 3; k++) a.bg[k] = d.bg[k];
    a.flags = d.shade_flags;
    if (d.lights.n < 0 || d.lights.n > TINA_MAX_LIGHTS) {
      batch_free(b);
      return fail(nullptr, TINA_ERR_CAPACITY, "light count out of range");
    }
    sh[i].img = (float*)(uintptr_t)d.img;
    prefix[0 * (nscenes + 1) + i + 1] = prefix[0 * (nscenes + 1) + i] + ngroups;
    prefix[1 * (nscenes + 1) + i + 1] = prefix[1 * (nscenes + 1) + i] + ntiles;
    prefix[2 * (nscenes + 1) + i + 1] =
        prefix[2 * (nscenes + 1) + i] + (int)(((long)d.width * d.height + kBatchShadePix - 1) / kBatchShadePix);
  }
  for (int k = 0; k < 3; k++) b->totals[k] = prefix[k * (nscenes + 1) + nscenes];
  if (hipMalloc(&b->keys, sizeof(uint64_t) * (size_t)b->total_pix) != hipSuccess ||
      hipMalloc(&b->d_jobs, sizeof(BinJob) * nscenes) != hipSuccess ||
      hipMalloc(&b->d_shade, sizeof(BatchShade) * nscenes) != hipSuccess ||
      hipMalloc(&b->d_prefix, sizeof(int) * prefix.size()) != hipSuccess) {
    (void)hipGetLastError();
    batch_free(b);
    return fail(nullptr, TINA_ERR_ALLOC, "hipMalloc of the batch tables failed");
  }
  for (int i = 0; i < nscenes; i++) {
    jobs[i].keys = b->keys + b->key_off[i];
    sh[i].keys = b->keys + b->key_off[i];
  }
  if (hipMemcpy(b->d_jobs, jobs.data(), sizeof(BinJob) * nscenes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_shade, sh.data(), sizeof(BatchShade) * nscenes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_prefix, prefix.data(), sizeof(int) * prefix.size(), hipMemcpyHostToDevice) != hipSuccess) {
    batch_free(b);
    return fail(nullptr, TINA_ERR_HIP, "upload of the batch tables failed");
  }
  *out = b;
  return TINA_OK;
}

int tina_batch_destroy(tina_batch* b) {
  batch_free(b);
  return TINA_OK;
}

static unsigned batch_grid(int want, int items) {
  const int g = want > 0 ? std::min(want, items) : items;
  return (unsigned)std::max(g, 1);
}

int tina_batch_render(tina_batch* b, void* stream) {
  if (!b) return fail(nullptr, TINA_ERR_INVALID, "batch is NULL");
  hipStream_t s = (hipStream_t)stream;
  const int n = b->n;
  int rc;
  allow_big_lds();
  hipLaunchKernelGGL(k_bbin, dim3(batch_grid(b->tune.grid_bin, b->totals[0])), dim3(kBinBlock),
                     (size_t)(b->max_ntiles + b->tune.bin_faces) * 4, s, b->d_jobs, b->d_prefix, n);
  if ((rc = hip_check(nullptr, hipGetLastError(), "k_bbin"))) return rc;
  // every tile of every scene starts from empty keys and is written whole: no clear pass
  hipLaunchKernelGGL(k_btile, dim3(batch_grid(b->tune.grid_tile, b->totals[1])), dim3(kTileBlock),
                     tile_lds_bytes(b->max_groups), s, b->d_jobs, b->d_prefix + (n + 1), n);
  if ((rc = hip_check(nullptr, hipGetLastError(), "k_btile"))) return rc;
  const unsigned gs = batch_grid(b->tune.grid_shade, b->totals[2]);
  const int* pre = b->d_prefix + 2 * (n + 1);
  switch (b->prog) {
    case PROG_DIFFUSE:
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bshade<PROG_DIFFUSE, false>), dim3(gs), dim3(kBlock), 0, s, b->d_shade, pre, n);
      break;
    case PROG_CLASSIC:
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bshade<PROG_CLASSIC, false>), dim3(gs), dim3(kBlock), 0, s, b->d_shade, pre, n);
      break;
    case PROG_BLINN:
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bshade<PROG_BLINN, false>), dim3(gs), dim3(kBlock), 0, s, b->d_shade, pre, n);
      break;
    case PROG_PBR:
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bshade<PROG_PBR, false>), dim3(gs), dim3(kBlock), 0, s, b->d_shade, pre, n);
      break;
    default:
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bshade<PROG_GENERIC, false>), dim3(gs), dim3(kBlock), 0, s, b->d_shade, pre, n);
      break;
  }
  return hip_check(nullptr, hipGetLastError(), "k_bshade");
}

int tina_batch_capture(tina_batch* b, void* stream) {
  if (!b) return fail(nullptr, TINA_ERR_INVALID, "batch is NULL");
  if (b->exec) return TINA_OK;
  hipStream_t s = (hipStream_t)stream;
  int rc = hip_check(nullptr, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (rc) return rc;
  const int rrc = tina_batch_render(b, stream);
  hipGraph_t g = nullptr;
  rc = hip_check(nullptr, hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  if (rrc) {
    if (g) (void)hipGraphDestroy(g);
    return rrc;
  }
  if (rc) return rc;
  b->graph = g;
  return hip_check(nullptr, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
}

int tina_batch_launch(tina_batch* b, void* stream) {
  if (!b || !b->exec) return fail(nullptr, TINA_ERR_INVALID, "batch is NULL or not captured");
  return hip_check(nullptr, hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch");
}

int tina_batch_keys(tina_batch* b, int scene, uint64_t** d_keys) {
  if (!b || !d_keys || scene < 0 || scene >= b->n) return fail(nullptr, TINA_ERR_INVALID, "bad batch or scene");
  *d_keys = b->keys + b->key_off[scene];
  return TINA_OK;
}

int tina_resolve_keys(const uint64_t* d_keys, int64_t n, int32_t* d_id, int32_t* d_depth, void* stream) {
  if (!d_keys || n < 0) return fail(nullptr, TINA_ERR_INVALID, "tina_resolve_keys: keys NULL or n < 0");
  if (n == 0 || (!d_id && !d_depth)) return TINA_OK;
  hipLaunchKernelGGL(k_resolve, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, d_keys, d_id,
                     d_depth, (long)n);
  return hip_check(nullptr, hipGetLastError(), "k_resolve");
}

int tina_resolve(tina_ctx* c, int32_t* d_id, int32_t* d_depth, void* stream) {
  if (!c) return fail(nullptr, TINA_ERR_INVALID, "ctx is NULL");
  if (!d_id && !d_depth) return TINA_OK;
  int rc = flush_clear(c, (hipStream_t)stream);
  if (rc) return rc;
  const long n = (long)c->W * c->H;
  hipLaunchKernelGGL(k_resolve, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, c->keys,
                     d_id, d_depth, n);
  TINA_LAUNCH(c, "k_resolve");
  return TINA_OK;
}

int tina_selftest_div(int mode, uint64_t npairs, uint32_t seed, uint64_t* d_counts, void* stream) {
  if (!d_counts || mode < 0 || mode > 2) return fail(nullptr, TINA_ERR_INVALID, "tina_selftest_div: bad mode or NULL");
  if (npairs == 0) return TINA_OK;
  hipLaunchKernelGGL(k_selftest_div, dim3(4096), dim3(kBlock), 0, (hipStream_t)stream, mode, npairs, seed,
                     (unsigned long long*)d_counts);
  return hip_check(nullptr, hipGetLastError(), "k_selftest_div");
}

}  // extern "C"
