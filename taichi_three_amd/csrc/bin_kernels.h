// bin_kernels.h -- binned, tiled coverage: the main render_occup path
// (tina/core/triangle.py:88-131).
//
//   k_bin    one 1024-lane workgroup per bin_faces consecutive faces.  Each face is
//            set up once (cull, clip, bbox, edge coefficients; triangle.py:93-113)
//            and its setup record stored; it is counted into every kTile x kTile
//            screen tile its bbox touches, in an LDS histogram.  One global atomic
//            reserves the workgroup's block of bin entries; the workgroup writes,
//            per tile, its (offset, count) chunk record and scatters its face ids
//            into its block.  Faces whose bbox touches more than kBinMaxTiles
//            tiles go to the giant queue instead.
//   k_tile   one workgroup per screen tile.  The tile's chunks from all binning
//            workgroups form one virtual list (prefix sum of their counts in LDS).
//            The tile's keys are loaded into LDS once (or start empty when the
//            frame's clear is deferred to this call); each wave then pulls 64 faces
//            at a time from the list, reads their setup records, clips their bboxes
//            to the tile and rasterises them with the wave-cooperative fragment
//            expansion (the bbox pixels of the wave's faces laid end to end, one
//            pixel per lane per step), reducing every covered pixel with an LDS
//            64-bit atomicMin.  Waves run independently (no workgroup barrier in
//            the loop).  The giant faces follow the binned ones in every tile's
//            list (each tile takes the part of each giant inside it).  The keys
//            go back to HBM once, coalesced.
//
// No global atomic touches the key buffer on this path: the 64-bit atomic rate
// to HBM-resident lines (about 27 G/s measured on MI355X) would bound the pass.
#pragma once
#include "raster_atomic.h"

namespace tina {

constexpr int kTileLog = 4;
// a queued fragment packs its face slot in the wave (6 bits) and its tile pixel in 16 bits
static_assert(6 + 2 * kTileLog <= 16, "fragment code does not fit 16 bits");
// a wave's fragment offsets (64 faces of at most kTile^2 pixels) fit 14 bits
static_assert(64 << (2 * kTileLog) <= (1 << 14), "fragment offset does not fit 14 bits");
constexpr int kTile = 1 << kTileLog;  // 16 x 16 pixel tiles, 2 KB of keys in LDS
constexpr int kBinBlock = 1024;       // lanes of a binning workgroup
constexpr int kTileBlock = 256;       // lanes of a tile workgroup (128: k_tile 91 us, 512: 97 us against 71 at 1080p/2M)
constexpr int kBinMaxTiles = 64;      // bbox tiles above which a face is rasterised as a giant
// binning workgroups a k_tile workgroup can index: its dynamic LDS of g + 1
// words and g half-words (tile_lds_bytes) plus ~21 KB of static LDS stay within
// a 64 KB workgroup allocation (7,168 groups: 64,238 bytes)
constexpr int kMaxBinGroups = 7168;
// bytes of dynamic LDS a binning workgroup may take (its static LDS is 72 bytes):
// a gfx950 CU has 160 KB of LDS and one workgroup may use all of it, so screens
// up to ~40k tiles (3840x2160 is 32,400) stay on the binned path
constexpr int kLdsBudget = 160 * 1024 - 1024;

// Faces per binning workgroup for a screen of ntiles tiles: the largest of 8192,
// 4096, 2048 and 1024 whose LDS tables (ntiles + faces words) fit the budget,
// or 0.  Larger groups mean fewer (group, tile) chunk records for k_tile to read.
inline int bin_faces_for(long ntiles) {
  for (int f = 8192; f >= 1024; f >>= 1)
    if ((ntiles + f) * 4 <= kLdsBudget) return f;
  return 0;
}

// Faces per binning group of a single raster call of nfaces faces on ncu CUs:
// equal groups, a whole number of rounds of one group per CU, so that no CU
// waits while others finish a partial last round, and at least kBinMinFaces
// faces each: a group of up to 1,024 faces sets them up in one pass of its
// lanes, and every group clears, scans and writes the whole tile table whatever
// its face count.  At 1080p: 2 M faces take 256 groups of 7,813 faces instead
// of 245 of 8,192; 1.1 M, 256 of 4,297 instead of 135 of 8,192; 70 k to 900 k
// faces, 256 groups instead of 9 to 110 groups of 8,192 after a separate set-up
// pass of one face a lane (1.5-4 us faster for the two raster kernels).
// Minimum group sizes at 1080p, raster kernels for 1 k / 4 k / 20 k / 65 k /
// 200 k faces: 32 faces 25.4 / 28.7 / 34.1 / 36.3 / 44.7 us, 512 faces
// 24.9 / 27.2 / 29.8 / 33.3 / 44.7, 1,024 faces - / 27.9 / 30.8 / 32.9 / 43.6,
// 2,048 faces (two set-up passes a lane) 32.7 / 35.5 / 37.5 / 45.1.
constexpr int kBinMinFaces = 512;
inline int bin_faces_call(long ntiles, long nfaces, int ncu) {
  const int bf = bin_faces_for(ntiles);
  if (bf == 0 || ncu <= 0) return bf;
  const long rounds = (nfaces + (long)ncu * bf - 1) / ((long)ncu * bf);
  const long eq = rounds > 0 ? (nfaces + rounds * ncu - 1) / (rounds * ncu) : bf;
  return (int)(eq < kBinMinFaces ? kBinMinFaces : eq < bf ? eq : bf);
}

// Binning groups to allocate for calls of up to cap faces: the most any call
// of at most cap faces can use under bin_faces_call (both bounds below grow
// with the face count).
inline long bin_groups_cap(long ntiles, long cap, int ncu) {
  const int bf = bin_faces_for(ntiles);
  if (bf == 0) return 0;
  if (ncu <= 0) return (cap + bf - 1) / bf;
  const long rounds = (cap + (long)ncu * bf - 1) / ((long)ncu * bf);
  const long small = (cap + kBinMinFaces - 1) / kBinMinFaces;
  return rounds * ncu < small ? rounds * ncu : small;
}

// Setup record of a binned face (64 B, one cache line): the state the coverage
// loop needs (triangle.py:127-131 plus the view-space z of A, B, C) and its bbox.
struct FaceRec {
  float4 a;  // bx, by, cx, cy
  float4 b;  // bcnx, bcny, canx, cany
  float4 c;  // ws0, ws1, ws2, Az
  float4 d;  // Bz, Cz, botx | boty << 16, topx | topy << 16
};

// One raster call on the binned path: its faces, screen, key buffer and the
// scratch the passes share.  bin_faces (faces per binning workgroup) and
// max_tiles (bbox tiles above which a face is rasterised as a giant) are the
// tunable bin sizes; a single raster call takes bin_faces_call(ntiles, T, CUs)
// and kBinMaxTiles.
struct BinJob {
  const float* verts;
  uint64_t* keys;
  uint32_t* chunk_off;
  uint32_t* chunk_cnt;
  uint32_t* entries;
  FaceRec* recs;
  // Every call leaves the state the next one starts from, so a call captured in
  // a graph replays as it ran: the entry cursor (binning groups reserve their
  // entries with one atomic each) is emptied by the tile pass, which does not
  // read it, and each binning group lists its giant faces at
  // gids[g * bin_faces, + gcnt[g]) instead of appending them to a shared queue.
  uint32_t* cursor; // entries reserved so far by the call's binning groups
  BigRec* giantq;   // [T] the set-up of each giant face, at its face index
  uint32_t* gids;   // [T] giant face ids by group (see above)
  uint32_t* gcnt;   // [ngroups] giant faces of each binning group
  Cam cam;
  int T;
  uint32_t id_base, flags;
  int tiles_x, ntiles, ngroups;
  int bin_faces, max_tiles;
  int load_keys;   // 1: k_tile merges into the keys as they stand; 0: the keys are
                   // cleared (a clear deferred to this call), so k_tile starts every
                   // tile from empty keys and writes all of them, empty tiles included
};

// The tiles a face's bbox touches: x0 (12 bits), y0 (12 bits), x span - 1 (4 bits),
// y span - 1 (4 bits).  0xFFFFFFFF: nothing to bin.
__device__ __forceinline__ uint32_t pack_span(int tx0, int ty0, int nx, int ny) {
  return (uint32_t)tx0 | ((uint32_t)ty0 << 12) | ((uint32_t)(nx - 1) << 24) | ((uint32_t)(ny - 1) << 28);
}

// Block-wide exclusive scan of n values v(i) into out(i); returns the total.
// Up to NT values: one per lane.  More: each lane takes a run of consecutive
// values, sums it, one block scan of the lane sums gives each run its offset,
// and the lane walks its run again (v is read twice) -- one block scan instead
// of one per NT values (8,160 tiles with 1,024 lanes: 6.6 -> ~1 us).
// wsum: LDS scratch of NT/64 words; carry: one LDS word.
template <int NT, class Get, class Put>
__device__ __forceinline__ uint32_t block_exclusive_scan(int n, Get v, Put out, uint32_t* wsum, uint32_t* carry) {
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const int per = (n + NT - 1) / NT;  // values per lane
  const int i0 = tid * per;
  uint32_t local = 0, x0 = 0;
  if (per <= 1) {
    x0 = tid < n ? v(tid) : 0u;
    local = x0;
  } else {
    for (int k = 0; k < per; k++)
      if (i0 + k < n) local += v(i0 + k);
  }
  const uint32_t incl = (uint32_t)wave_inclusive_scan((int)local, lane);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t wofs = 0, tot = 0;
  for (int w = 0; w < NT / 64; w++) {
    if (w < wave) wofs += wsum[w];
    tot += wsum[w];
  }
  uint32_t run = wofs + incl - local;
  if (per <= 1) {
    if (tid < n) out(tid, run);
  } else {
    for (int k = 0; k < per; k++) {
      const int i = i0 + k;
      if (i < n) {
        const uint32_t x = v(i);
        out(i, run);
        run += x;
      }
    }
  }
  if (tid == 0) *carry = tot;
  __syncthreads();
  return tot;
}

// In-place exclusive scan of the LDS array a[0, n) (16-byte aligned); returns
// the total.  As above, but each lane's run is a multiple of four values long
// and is read and written as 16-byte vectors: single words at a stride of the
// run length would put eight lanes on each bank.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan_lds(uint32_t* a, int n, uint32_t* wsum, uint32_t* carry) {
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const int per = (((n + NT - 1) / NT) + 3) & ~3;  // values per lane
  const int i0 = tid * per;
  uint32_t local = 0;
  for (int k = 0; k < per; k += 4) {
    const int i = i0 + k;
    if (i + 3 < n) {
      const uint4 q = *(const uint4*)(a + i);
      local += q.x + q.y + q.z + q.w;
    } else {
      for (int m = 0; m < 4; m++)
        if (i + m < n) local += a[i + m];
    }
  }
  const uint32_t incl = (uint32_t)wave_inclusive_scan((int)local, lane);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t wofs = 0, tot = 0;
  for (int w = 0; w < NT / 64; w++) {
    if (w < wave) wofs += wsum[w];
    tot += wsum[w];
  }
  uint32_t run = wofs + incl - local;
  for (int k = 0; k < per; k += 4) {
    const int i = i0 + k;
    if (i + 3 < n) {
      const uint4 q = *(const uint4*)(a + i);
      uint4 o;
      o.x = run; run += q.x;
      o.y = run; run += q.y;
      o.z = run; run += q.z;
      o.w = run; run += q.w;
      *(uint4*)(a + i) = o;
    } else {
      for (int m = 0; m < 4; m++)
        if (i + m < n) {
          const uint32_t x = a[i + m];
          a[i + m] = run;
          run += x;
        }
    }
  }
  if (tid == 0) *carry = tot;
  __syncthreads();
  return tot;
}

// Shrinks the pixel bbox of a well-conditioned face to the pixels that can
// survive the cheap rejection (and hence the coverage test); false when none can.
//
// The face's set-up values define three affine functions of the pixel centre P,
// evaluated exactly: a(P) = (P.x - bx) bcny - (P.y - by) bcnx, b(P) likewise for
// edge CA, and 1 - a - b.  The cheap test computes them in binary32 and rejects
// when one is below -2**-40; its rounding error over the reference bbox is at
// most e_a = 8u (|P.x - bx| |bcny| + |P.y - by| |bcnx| + the rounding of P), u =
// 2**-24 (and e_b likewise, e_t = e_a + e_b + 4u (1 + |a| + |b|)).  So every
// survivor lies in the triangle a >= -e_a', b >= -e_b', 1 - a - b >= -e_t' (e' =
// e + 2**-40), whose vertices are solved here in float64 (the products of two
// binary32 values are exact; the rest carries a relative error near 2**-50, and
// the determinant is used only when it keeps 20 of its bits).  The pixels whose
// centres fall inside that triangle's bbox, widened by 2**-30 of the magnitudes
// involved, are kept.  This is the set the cheap test could accept, not the
// triangle's hull: a face whose rounded coefficients describe a slightly
// different triangle keeps the pixels of that one.
__device__ __forceinline__ bool tight_bbox(const Cam& cam, FaceSetup& s) {
  const double u = 5.9604644775390625e-8, tiny = 9.094947017729282e-13;  // 2**-24, 2**-40
  const double bias_x = cam.bias[0], bias_y = cam.bias[1];
  const double px0 = s.botx + bias_x, px1 = s.topx + bias_x, py0 = s.boty + bias_y, py1 = s.topy + bias_y;
  const double bx = s.bx, by = s.by, cx = s.cx, cy = s.cy;
  const double bcnx = s.bcnx, bcny = s.bcny, canx = s.canx, cany = s.cany;
  // a NaN would drop out of the min / max below: keep the reference bbox
  if (!(fabs(bx) + fabs(by) + fabs(cx) + fabs(cy) + fabs(bcnx) + fabs(bcny) + fabs(canx) + fabs(cany) < 1e300))
    return true;
  // |P - vertex| over the bbox, plus the rounding of P itself
  const double rx = fmax(fabs(px0), fabs(px1)) * u, ry = fmax(fabs(py0), fabs(py1)) * u;
  const double dbx = fmax(fabs(px0 - bx), fabs(px1 - bx)) + rx + 1.0, dby = fmax(fabs(py0 - by), fabs(py1 - by)) + ry + 1.0;
  const double dcx = fmax(fabs(px0 - cx), fabs(px1 - cx)) + rx + 1.0, dcy = fmax(fabs(py0 - cy), fabs(py1 - cy)) + ry + 1.0;
  const double m_a = dbx * fabs(bcny) + dby * fabs(bcnx), m_b = dcx * fabs(cany) + dcy * fabs(canx);
  const double e_a = tiny + 8.0 * u * m_a, e_b = tiny + 8.0 * u * m_b;
  const double e_t = tiny + e_a + e_b + 4.0 * u * (1.0 + m_a + m_b);
  const double det = bcnx * cany - bcny * canx;
  const double dterm = fabs(bcnx * cany) + fabs(bcny * canx);
  if (!(fabs(det) > dterm * 9.5367431640625e-7)) return true;  // 2**-20: keep the reference bbox
  const double inv = 1.0 / det;
  const double A0 = bx * bcny - by * bcnx, B0 = cx * cany - cy * canx;
  // P(a, b) solves (P.x - bx) bcny - (P.y - by) bcnx = a, (P.x - cx) cany - (P.y - cy) canx = b
  const double av[3] = {-e_a, 1.0 + e_t + e_b, -e_a};
  const double bv[3] = {-e_b, -e_b, 1.0 + e_t + e_a};
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300, mag = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double ta = av[k] + A0, tb = bv[k] + B0;
    const double X = (bcnx * tb - canx * ta) * inv, Y = (bcny * tb - cany * ta) * inv;
    xmin = fmin(xmin, X); xmax = fmax(xmax, X);
    ymin = fmin(ymin, Y); ymax = fmax(ymax, Y);
    mag = fmax(mag, (fabs(bcnx * tb) + fabs(canx * ta) + fabs(bcny * tb) + fabs(cany * ta)) * fabs(inv) +
                        fabs(X) + fabs(Y));
  }
  const double d = 9.313225746154785e-10 * (mag + 1.0);  // 2**-30
  // keep pixel x when its centre x + bias lies in [xmin - d, xmax + d]
  const double lx = fmax(ceil(xmin - d - bias_x), (double)s.botx), hx = fmin(floor(xmax + d - bias_x), (double)s.topx);
  const double ly = fmax(ceil(ymin - d - bias_y), (double)s.boty), hy = fmin(floor(ymax + d - bias_y), (double)s.topy);
  if (lx > hx || ly > hy) return false;
  s.botx = (int32_t)lx; s.topx = (int32_t)hx;
  s.boty = (int32_t)ly; s.topy = (int32_t)hy;
  return true;
}

// Set-up of face f (cull, clip, bbox, edges; triangle.py:93-113) and its bbox
// shrunk to the pixels the coverage test can accept: the setup record of a
// binned face is stored and its tile span returned (0xFFFFFFFF: no pixel);
// `giant` is set for a face over more than max_tiles tiles (not binned).
__device__ __forceinline__ uint32_t face_setup_record(const BinJob& j, long f, const float (&v)[9], FaceSetup& s,
                                                      bool& giant) {
  giant = false;
  if (!(f < j.T && setup_face(j.cam, v, j.flags, s) && s.topx >= s.botx && s.topy >= s.boty &&
        (!face_well_conditioned(s) || tight_bbox(j.cam, s))))
    return 0xFFFFFFFFu;
  const int tx0 = s.botx >> kTileLog, tx1 = s.topx >> kTileLog;
  const int ty0 = s.boty >> kTileLog, ty1 = s.topy >> kTileLog;
  const int nx = tx1 - tx0 + 1, ny = ty1 - ty0 + 1;
  if (nx * ny > j.max_tiles || nx > 16 || ny > 16) {
    giant = true;
    return 0xFFFFFFFFu;
  }
  FaceRec r;
  r.a = make_float4(s.bx, s.by, s.cx, s.cy);
  r.b = make_float4(s.bcnx, s.bcny, s.canx, s.cany);
  r.c = make_float4(s.ws0, s.ws1, s.ws2, s.Az);
  // bbox corners (16 bits each; screen < 32768 px) and bit 31: well-conditioned
  r.d = make_float4(s.Bz, s.Cz,
                    __int_as_float(s.botx | (s.boty << 16) | (face_well_conditioned(s) ? (int)0x80000000 : 0)),
                    __int_as_float(s.topx | (s.topy << 16)));
  j.recs[f] = r;
  return pack_span(tx0, ty0, nx, ny);
}

// The set-up of giant face f, at its face index (k_tile reads it through gids).
__device__ __forceinline__ void store_giant(const BinJob& j, long f, const FaceSetup& s) {
  BigRec r;
  r.s = s;
  r.id = j.id_base + (uint32_t)f;
  r.pad = 0;
  j.giantq[f] = r;
}

// The tiles of span sp counted into the group's LDS histogram.
__device__ __forceinline__ void count_span(uint32_t* cnt, int tiles_x, uint32_t sp) {
  if (sp == 0xFFFFFFFFu) return;
  const int tx0 = sp & 0xFFF, ty0 = (sp >> 12) & 0xFFF, nx = ((sp >> 24) & 0xF) + 1, ny = (sp >> 28) + 1;
  for (int y = ty0; y < ty0 + ny; y++)
    for (int x = tx0; x < tx0 + nx; x++) atomicAdd(&cnt[y * tiles_x + x], 1u);
}

// Binning workgroup g of job j (the body of k_bin).  Needs (ntiles + bin_faces)
// words of dynamic LDS.
__device__ __forceinline__ void bin_group(const BinJob& j, int g) {
  extern __shared__ __align__(16) uint32_t lds[];
  const int ntiles = j.ntiles, tiles_x = j.tiles_x, nbf = j.bin_faces;
  const Cam& cam = j.cam;
  uint32_t* cnt = lds;            // [ntiles] counts, then running offsets
  uint32_t* span = lds + ntiles;  // [bin_faces] per-face tile span
  __shared__ uint32_t s_base, s_carry, s_gcnt, s_wsum[kBinBlock / 64];
  const int tid = threadIdx.x;
  const long f0 = (long)g * nbf;
  uint32_t* __restrict__ chunk_off = j.chunk_off;
  uint32_t* __restrict__ chunk_cnt = j.chunk_cnt;
  uint32_t* __restrict__ entries = j.entries;
  uint32_t* __restrict__ gids = j.gids + f0;
  for (int t = tid; t < ntiles; t += kBinBlock) cnt[t] = 0;
  if (tid == 0) s_gcnt = 0;
  __syncthreads();

  // The positions of a lane's faces are fetched two faces ahead of their set-up,
  // so that the loads overlap the arithmetic of the faces before them.
  const float* __restrict__ verts = j.verts;
  auto fetch = [&](int i, float (&v)[9]) {
    const long f = f0 + i;
    if (i < nbf && f < j.T) {
#pragma unroll
      for (int k = 0; k < 9; k++) v[k] = verts[f * 9 + k];
    }
  };
  float p0[9], p1[9];
  fetch(tid, p0);
  fetch(tid + kBinBlock, p1);
  for (int i = tid; i < nbf; i += kBinBlock) {
    const long f = f0 + i;
    float v[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      v[k] = p0[k];
      p0[k] = p1[k];
    }
    fetch(i + 2 * kBinBlock, p1);
    FaceSetup s;
    bool giant;
    const uint32_t sp = face_setup_record(j, f, v, s, giant);
    count_span(cnt, tiles_x, sp);
    span[i] = sp;
    if (giant) {
      store_giant(j, f, s);
      gids[atomicAdd(&s_gcnt, 1u)] = (uint32_t)f;
    }
  }
  __syncthreads();
  if (tid == 0) j.gcnt[g] = s_gcnt;

  for (int t = tid; t < ntiles; t += kBinBlock) chunk_cnt[(size_t)g * ntiles + t] = cnt[t];
  const uint32_t total = block_exclusive_scan_lds<kBinBlock>(cnt, ntiles, s_wsum, &s_carry);
  if (tid == 0) s_base = atomicAdd(j.cursor, total);
  __syncthreads();
  const uint32_t base = s_base;
  for (int t = tid; t < ntiles; t += kBinBlock) chunk_off[(size_t)g * ntiles + t] = base + cnt[t];
  __syncthreads();

  // scatter the face ids (order within a tile is irrelevant: the reduction is a min)
  for (int i = tid; i < nbf; i += kBinBlock) {
    const uint32_t sp = span[i];
    if (sp == 0xFFFFFFFFu) continue;
    const int tx0 = sp & 0xFFF, ty0 = (sp >> 12) & 0xFFF, nx = ((sp >> 24) & 0xF) + 1, ny = (sp >> 28) + 1;
    for (int y = ty0; y < ty0 + ny; y++)
      for (int x = tx0; x < tx0 + nx; x++) {
        const uint32_t pos = atomicAdd(&cnt[y * tiles_x + x], 1u);
        entries[base + pos] = (uint32_t)(f0 + i);
      }
  }
}

// Dynamic LDS of tile_body for a job of ngroups binning groups.
inline size_t tile_lds_bytes(int ngroups) { return (size_t)(ngroups + 1) * 4 + ((size_t)ngroups * 2 + 3) / 4 * 4; }

// Makes this lane's LDS writes visible to the rest of its wave before they are read.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Tile t of job j (the body of k_tile).  Needs tile_lds_bytes(ngroups) of dynamic LDS.
__device__ __forceinline__ void tile_body(const BinJob& j, int t) {
  // per binning group: its faces for this tile, the chunk's binned faces then the
  // group's giant faces (each tile takes the part of every giant inside it)
  extern __shared__ uint32_t pre[];  // [ngroups + 1] exclusive prefix of those counts
  // [ngroups] the groups' giant counts (at most bin_faces each); 16 bits, as the
  // tags above, keep seven workgroups per CU up to 336 groups
  uint16_t* const gc = (uint16_t*)(pre + j.ngroups + 1);
  __shared__ unsigned long long tk[kTile * kTile];
  __shared__ float4 R0[kTileBlock], R1[kTileBlock], R2[kTileBlock], R3[kTileBlock];
  // per wave, per live face: its first fragment pair (bits 0-13) and ceil(2**16 / pairs a row) (bits 14-30)
  __shared__ uint32_t Lstart[kTileBlock];
  // per wave: window tag (a window start below 2**13) at the position of a face's last fragment pair
  __shared__ int16_t Mark[kTileBlock];
  __shared__ uint16_t Q[kTileBlock * 2];  // per wave: up to 127 pending fragments (face slot | lx | ly)
  __shared__ uint32_t s_wsum[kTileBlock / 64], s_carry, s_next;
  const int tid = threadIdx.x, lane = __lane_id(), wbase = tid & ~63;
  const Cam& cam = j.cam;
  // read once: the kernel argument block is not provably unchanged across the
  // loops' stores, so uses of cam.bias inside them reload it
  const float bias_x = cam.bias[0], bias_y = cam.bias[1];
  const int tiles_x = j.tiles_x, ngroups = j.ngroups;
  const uint32_t id_base = j.id_base;
  const uint32_t* __restrict__ chunk_off = j.chunk_off;
  const uint32_t* __restrict__ chunk_cnt = j.chunk_cnt;
  const uint32_t* __restrict__ entries = j.entries;
  const FaceRec* __restrict__ recs = j.recs;
  uint64_t* __restrict__ keys = j.keys;
  const int tx = t % tiles_x, ty = t / tiles_x;
  const int ox = tx << kTileLog, oy = ty << kTileLog;
  const int ntiles = j.ntiles;  // chunk tables are group-major: [group][tile]
  // (Staging the chunks' entry offsets in LDS with the counts, to take one
  // dependent read out of the face loop, measured no faster.)
  const uint32_t total = block_exclusive_scan<kTileBlock>(
      ngroups,
      [&](int c) {
        const uint32_t ng = j.gcnt[c];
        gc[c] = (uint16_t)ng;
        return chunk_cnt[(size_t)c * ntiles + t] + ng;
      },
      [&](int c, uint32_t ofs) { pre[c] = ofs; }, s_wsum, &s_carry);
  const int tw = min(kTile, cam.W - ox), th = min(kTile, cam.H - oy);
  if (total == 0) {
    if (!j.load_keys)  // the deferred clear of a tile no face touches
      for (int i = tid; i < kTile * kTile; i += kTileBlock) {
        const int lx = i & (kTile - 1), ly = i >> kTileLog;
        if (lx < tw && ly < th) keys[(size_t)(oy + ly) * cam.W + ox + lx] = empty_key();
      }
    return;
  }
  if (tid == 0) {
    pre[ngroups] = total;
    s_next = 0;
  }
  for (int i = tid; i < kTile * kTile; i += kTileBlock) {
    const int lx = i & (kTile - 1), ly = i >> kTileLog;
    tk[i] = (j.load_keys && lx < tw && ly < th) ? keys[(size_t)(oy + ly) * cam.W + ox + lx] : empty_key();
  }
  __syncthreads();

  for (;;) {
    uint32_t b0 = 0;
    if (lane == 0) b0 = atomicAdd(&s_next, 64u);
    b0 = __shfl(b0, 0, 64);
    if (b0 >= total) break;
    const uint32_t v = b0 + lane;
    int area = 0;
    float4 ra, rb, rc, rd;
    if (v < total) {
      int lo = 0, hi = ngroups;  // pre[lo] <= v < pre[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= v) lo = mid; else hi = mid;
      }
      const uint32_t i = v - pre[lo], nb = pre[lo + 1] - pre[lo] - gc[lo];
      if (i < nb) {
        const uint32_t fid = entries[chunk_off[(size_t)lo * ntiles + t] + i];
        const FaceRec r = recs[fid];
        const uint32_t lo16 = (uint32_t)__float_as_int(r.d.z), hi16 = (uint32_t)__float_as_int(r.d.w);
        const int wellc = (int)(lo16 >> 31);
        const int bx0 = max((int)(lo16 & 0xFFFF), ox), by0 = max((int)((lo16 >> 16) & 0x7FFF), oy);
        const int bx1 = min((int)(hi16 & 0xFFFF), ox + tw - 1), by1 = min((int)(hi16 >> 16), oy + th - 1);
        const int w = bx1 - bx0 + 1, h = by1 - by0 + 1;
        if (w > 0 && h > 0) {
          area = ((w + 1) >> 1) * h;
          ra = r.a;
          rb = r.b;
          rc = r.c;
          // Bz, Cz, clipped origin (x | y << 8), width, well-conditioned bit; face id
          rd = make_float4(r.d.x, r.d.y,
                           __int_as_float((bx0 - ox) | ((by0 - oy) << 8) | (w << 16) | (wellc << 24)),
                           __int_as_float((int)(id_base + fid)));
        }
      } else {
        const BigRec& g = j.giantq[j.gids[(size_t)lo * j.bin_faces + (i - nb)]];
        const FaceSetup& gs = g.s;
        const int bx0 = max(gs.botx, ox), by0 = max(gs.boty, oy);
        const int bx1 = min(gs.topx, ox + tw - 1), by1 = min(gs.topy, oy + th - 1);
        const int w = bx1 - bx0 + 1, h = by1 - by0 + 1;
        if (w > 0 && h > 0) {
          area = ((w + 1) >> 1) * h;
          ra = make_float4(gs.bx, gs.by, gs.cx, gs.cy);
          rb = make_float4(gs.bcnx, gs.bcny, gs.canx, gs.cany);
          rc = make_float4(gs.ws0, gs.ws1, gs.ws2, gs.Az);
          rd = make_float4(gs.Bz, gs.Cz,
                           __int_as_float((bx0 - ox) | ((by0 - oy) << 8) | (w << 16) |
                                          ((face_well_conditioned(gs) ? 1 : 0) << 24)),
                           __int_as_float((int)g.id));
        }
      }
    }
    // A lane takes two horizontally adjacent pixels of a face's clipped bbox (a
    // fragment pair; the second is absent at the end of an odd row), so the
    // face lookup below is paid once for both.  The faces with pixels in the
    // tile take consecutive slots, so their pair ranges [start, end) tile
    // [0, wtotal) with no empty range.
    const unsigned long long live = __ballot(area > 0);
    const int end = wave_inclusive_scan(area, lane);
    const int wtotal = __shfl(end, 63, 64);
    if (area > 0) {
      const int slot = wbase + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(live >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)live, 0u));
      R0[slot] = ra;
      R1[slot] = rb;
      R2[slot] = rc;
      R3[slot] = rd;
      const uint32_t pw = (((uint32_t)(__float_as_int(rd.z) >> 16) & 0xFF) + 1u) >> 1;  // pairs a row
      Lstart[slot] = (uint32_t)(end - area) | (((65536u + pw - 1u) / pw) << 14);
    }
    Mark[tid] = -1;
    wave_sync_lds();
    // Cheap pass over every fragment pair.  Pair g belongs to the first face whose
    // last pair is >= g: per window of 64 pairs, the faces whose last pair falls
    // inside it tag that position, and a ballot of the tags gives each lane its
    // face as (faces ended before the window) + (tags below it).  Each pixel then
    // evaluates the unscaled edge values and is dropped when a well-conditioned
    // face rejects it.  Survivors (about two in five) are compacted into the
    // wave's LDS queue and the exact test runs on full waves of them.
    uint16_t* const q = Q + 2 * wbase;
    int qn = 0;     // uniform: pending fragments in the queue
    int kbase = 0;  // uniform: faces whose fragments all precede the window
    auto exact = [&](int e) {  // exact test + depth + LDS min for queue entry e
      const int code = q[e];
      const int j = wbase + (code >> (2 * kTileLog)), lx = code & (kTile - 1), ly = (code >> kTileLog) & (kTile - 1);
      const float4 r0 = R0[j], r1 = R1[j], r2 = R2[j], r3 = R3[j];
      FaceSetup s;
      s.bx = r0.x; s.by = r0.y; s.cx = r0.z; s.cy = r0.w;
      s.bcnx = r1.x; s.bcny = r1.y; s.canx = r1.z; s.cany = r1.w;
      s.ws0 = r2.x; s.ws1 = r2.y; s.ws2 = r2.z; s.Az = r2.w;
      s.Bz = r3.x; s.Cz = r3.y;
      float w0, w1, w2;
      // the reference's weights and test as written: nearly every full wave of
      // survivors holds a covered pixel, so the divisions run anyway and the
      // sign-based early rejection of pixel_cover_fast would only add work
      if (pixel_weights(s, (float)(ox + lx) + bias_x, (float)(oy + ly) + bias_y, w0, w1, w2)) {
        const uint64_t key = pack_key(pixel_depth(s, w0, w1, w2), (uint32_t)__float_as_int(r3.w));
        atomicMin(&tk[(ly << kTileLog) + lx], (unsigned long long)key);
      }
    };
    for (int g0 = 0; g0 < wtotal; g0 += 64) {
      if (area > 0 && end - 1 >= g0 && end - 1 < g0 + 64) Mark[wbase + end - 1 - g0] = (int16_t)g0;
      wave_sync_lds();
      const unsigned long long ends = __ballot(Mark[tid] == g0);
      const int gI = g0 + lane;
      bool keep0 = false, keep1 = false;
      int code = 0;
      if (gI < wtotal) {
        const int k = kbase + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ends >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)ends, 0u));
        const int j = wbase + k;
        const uint32_t ls = Lstart[j];
        const int packed = __float_as_int(R3[j].z);
        const int lx0 = packed & 0xFF, ly0 = (packed >> 8) & 0xFF, wj = (packed >> 16) & 0xFF;
        const int pw = (wj + 1) >> 1;
        const int li = gI - (int)(ls & 0x3FFFu);
        // li / pw as (li * ceil(2**16 / pw)) >> 16: the multiplier exceeds 2**16 / pw by
        // less than 1, so the quotient gains less than li / 2**16 < 1/512, while li / pw
        // sits at least 1/pw >= 2/kTile below the next integer (li < kTile^2 / 2, pw <= kTile / 2)
        const int dy = (int)(__umul24((uint32_t)li, ls >> 14) >> 16);
        const int dx = 2 * (li - (int)__umul24((uint32_t)dy, (uint32_t)pw));
        const int lx = lx0 + dx, ly = ly0 + dy;
        const float4 r0 = R0[j], r1 = R1[j];
        FaceSetup s;
        s.bx = r0.x; s.by = r0.y; s.cx = r0.z; s.cy = r0.w;
        s.bcnx = r1.x; s.bcny = r1.y; s.canx = r1.z; s.cany = r1.w;
        const bool wellc = (packed >> 24) != 0;
        const float posy = (float)(oy + ly) + bias_y;
        keep0 = !(wellc && pixel_rejects(s, (float)(ox + lx) + bias_x, posy));
        keep1 = dx + 1 < wj && !(wellc && pixel_rejects(s, (float)(ox + lx + 1) + bias_x, posy));
        code = (k << (2 * kTileLog)) | (ly << kTileLog) | lx;
      }
      kbase += __popcll(ends);
      // the first pixels of the pairs, then the second ones: each push adds at
      // most 64 to at most 63 pending, and a full wave is drained after each
      for (int half = 0; half < 2; half++) {
        const bool keep = half ? keep1 : keep0;
        const unsigned long long m = __ballot(keep);
        if (keep)
          q[qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
              (uint16_t)(code + half);
        qn += __popcll(m);
        wave_sync_lds();
        if (qn >= 64) {  // a full wave of survivors
          exact(lane);
          const int rest = qn - 64;  // <= 63
          const int moved = lane < rest ? q[64 + lane] : 0;
          wave_sync_lds();
          if (lane < rest) q[lane] = (uint16_t)moved;
          qn = rest;
          wave_sync_lds();
        }
      }
    }
    if (lane < qn) exact(lane);
    wave_sync_lds();
  }
  __syncthreads();
  for (int i = tid; i < kTile * kTile; i += kTileBlock) {
    const int lx = i & (kTile - 1), ly = i >> kTileLog;
    if (lx < tw && ly < th) keys[(size_t)(oy + ly) * cam.W + ox + lx] = tk[i];
  }
}

// ---- kernels: one raster call -------------------------------------------------
//
// The job arrives by value as the first kernel argument and is read in place in
// the kernarg segment (naming the by-value parameter would copy it to scratch).
#if defined(__HIP_DEVICE_COMPILE__)
#define TINA_KERNARG_JOB(T, name, byval) \
  (void)byval;                            \
  const T& name = *(const T*)__builtin_amdgcn_kernarg_segment_ptr()
#else
#define TINA_KERNARG_JOB(T, name, byval) const T& name = byval
#endif

__global__ void __launch_bounds__(kBinBlock) k_bin(const BinJob jv) {
  TINA_KERNARG_JOB(BinJob, j, jv);
  bin_group(j, blockIdx.x);
}

// Workgroup b of n -> tile.  The dispatcher deals workgroups round-robin over
// the 8 XCDs (each with its own L2).  The 32 tiles whose chunk-table words share
// a 128-byte line go to one XCD, so a table line is fetched once rather than by
// every XCD; the 32-tile strips of an XCD are spread over the whole screen, so
// the uneven per-tile work stays balanced.  (One contiguous band per XCD cuts the
// fetches further but measured 21% slower: the bands over the dense centre
// finish last.)  The tail beyond a multiple of 256 keeps the identity order.
__device__ __forceinline__ long strip_item(long b, long n) {
  const long full = (n / 256) * 256;
  if (b >= full) return b;
  return 32 * (8 * (b / 256) + b % 8) + (b / 8) % 32;
}

__global__ void __launch_bounds__(kTileBlock) k_tile(const BinJob jv) {
  TINA_KERNARG_JOB(BinJob, j, jv);
  // the binning groups have all reserved their entries (stream order) and no tile
  // reads the cursor: empty it for the next call
  if (blockIdx.x == 0 && threadIdx.x == 0) *j.cursor = 0u;
  tile_body(j, (int)strip_item(blockIdx.x, gridDim.x));
}

// ---- kernels: a batch of raster calls (independent scenes) ---------------------
//
// Persistent workgroups walk the batch's work items (binning groups, tiles) in
// order; item i belongs to the job s with prefix[s] <= i < prefix[s + 1].  The
// jobs live in device memory.

__device__ __forceinline__ int batch_job(const int* __restrict__ prefix, int njobs, long item) {
  int lo = 0, hi = njobs;  // prefix[lo] <= item < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kBinBlock)
k_bbin(const BinJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs) {
  const long total = prefix[njobs];
  for (long item = blockIdx.x; item < total; item += gridDim.x) {
    const int s = batch_job(prefix, njobs, item);
    bin_group(jobs[s], (int)(item - prefix[s]));
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kTileBlock)
k_btile(const BinJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs) {
  // every scene's binning is done and no tile reads the cursors: empty them
  if (blockIdx.x == 0)
    for (int s = threadIdx.x; s < njobs; s += kTileBlock) *jobs[s].cursor = 0u;
  const long total = prefix[njobs];
  for (long b = blockIdx.x; b < total; b += gridDim.x) {
    const long sweep = b - blockIdx.x, width = min((long)gridDim.x, total - sweep);
    const long item = sweep + strip_item(b - sweep, width);
    const int s = batch_job(prefix, njobs, item);
    tile_body(jobs[s], (int)(item - prefix[s]));
    __syncthreads();
  }
}

}  // namespace tina
