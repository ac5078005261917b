// mesh_kernels.hip -- the device side of the mesh nodes upstream of set_object
// (tina/mesh/trans.py, tina/mesh/norm.py), behind the C ABI of include/tina_hip.h.
//
//   k_mesh_transform   MeshTransform.get_face_verts / get_face_norms
//                      (trans.py:29-40): mapply_pos of every position
//                      (tina/common.py:169-183) and trans_normal @ normal
//   k_face_normals     normalized((b - a) x (c - a)) per face (norm.py:7-14, 41)
//   k_vertex_offsets   vertex -> incidence list offsets from the incidences
//                      sorted by vertex (a stable radix sort keeps each list in
//                      ascending face and corner order)
//   k_smooth_normals   MeshSmoothNormal._update_normal (norm.py:35-47): per
//                      vertex, the face normals summed in the order the
//                      reference's serial loop adds them, then normalized
//
// Every float expression keeps the reference's grouping; the file is compiled
// with -ffp-contract=off like the rest of the library, so results are
// bit-identical to the binary32 evaluation of the reference's loops.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <stdint.h>
#include <string>

#include "../../include/tina_hip.h"

namespace tina {
void set_global_error(const std::string& msg);  // tina_kernels.hip: read back by tina_last_error(NULL)
}

namespace {

constexpr int kBlk = 256;

struct Mat4 {
  float m[16];
};
struct Mat3 {
  float m[9];
};

// mapply_pos (tina/common.py:169-183): row r starts at mat[r][3] * 1 and adds
// mat[r][j] * pos[j] for j = 0, 1, 2; the result is divided by the w row.
__global__ void __launch_bounds__(kBlk)
k_mesh_transform(const Mat4 T, const Mat3 N, const float* __restrict__ vin, float* __restrict__ vout, long nv,
                 const float* __restrict__ nin, float* __restrict__ nout, long nn) {
  const long stride = (long)gridDim.x * kBlk;
  const float* M = T.m;
  for (long i = (long)blockIdx.x * kBlk + threadIdx.x; i < nv; i += stride) {
    const float p0 = vin[3 * i], p1 = vin[3 * i + 1], p2 = vin[3 * i + 2];
    float r[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float a = M[4 * k + 3] * 1.0f;
      a = a + M[4 * k + 0] * p0;
      a = a + M[4 * k + 1] * p1;
      a = a + M[4 * k + 2] * p2;
      r[k] = a;
    }
    float w = M[15] * 1.0f;
    w = w + M[12] * p0;
    w = w + M[13] * p1;
    w = w + M[14] * p2;
    vout[3 * i] = r[0] / w;
    vout[3 * i + 1] = r[1] / w;
    vout[3 * i + 2] = r[2] / w;
  }
  // Matrix @ Vector: entry r = (M[r][0] n0 + M[r][1] n1) + M[r][2] n2
  for (long i = (long)blockIdx.x * kBlk + threadIdx.x; i < nn; i += stride) {
    const float n0 = nin[3 * i], n1 = nin[3 * i + 1], n2 = nin[3 * i + 2];
#pragma unroll
    for (int k = 0; k < 3; k++) nout[3 * i + k] = ((N.m[3 * k] * n0) + (N.m[3 * k + 1] * n1)) + (N.m[3 * k + 2] * n2);
  }
}

// (b - a).cross(c - a).normalized(): Taichi's cross, then v * (1 / sqrt((x*x + y*y) + z*z)).
__device__ __forceinline__ void face_normal(const float* v, float* n) {
  const float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
  const float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
  const float cx = e1y * e2z - e1z * e2y;
  const float cy = e1z * e2x - e1x * e2z;
  const float cz = e1x * e2y - e1y * e2x;
  const float inv = 1.0f / sqrtf(((cx * cx) + (cy * cy)) + (cz * cz));
  n[0] = inv * cx;
  n[1] = inv * cy;
  n[2] = inv * cz;
}

__global__ void __launch_bounds__(kBlk)
k_face_normals(const float* __restrict__ soup, long nf, int corners, float* __restrict__ out) {
  const long stride = (long)gridDim.x * kBlk;
  for (long f = (long)blockIdx.x * kBlk + threadIdx.x; f < nf; f += stride) {
    float v[9], n[3];
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = soup[9 * f + k];
    face_normal(v, n);
    for (int c = 0; c < corners; c++) {
      out[3 * (corners * f + c)] = n[0];
      out[3 * (corners * f + c) + 1] = n[1];
      out[3 * (corners * f + c) + 2] = n[2];
    }
  }
}

// keys[i] = vertex of incidence i (3 f + corner), vals[i] = i
__global__ void __launch_bounds__(kBlk)
k_incidences(const int32_t* __restrict__ faces, long ninc, int nverts, int32_t* __restrict__ keys,
             int32_t* __restrict__ vals, uint32_t* __restrict__ bad) {
  const long stride = (long)gridDim.x * kBlk;
  for (long i = (long)blockIdx.x * kBlk + threadIdx.x; i < ninc; i += stride) {
    const int32_t v = faces[i];
    if (v < 0 || v >= nverts) {
      atomicOr(bad, 1u);
      keys[i] = 0;
    } else {
      keys[i] = v;
    }
    vals[i] = (int32_t)i;
  }
}

// offsets[v] = first position of vertex v in the sorted keys (lower bound);
// offsets[nverts] = ninc.
__global__ void __launch_bounds__(kBlk)
k_vertex_offsets(const int32_t* __restrict__ sorted, long ninc, int nverts, int32_t* __restrict__ offsets) {
  const long stride = (long)gridDim.x * kBlk;
  for (long v = (long)blockIdx.x * kBlk + threadIdx.x; v <= nverts; v += stride) {
    long lo = 0, hi = ninc;  // first i with sorted[i] >= v
    while (lo < hi) {
      const long mid = (lo + hi) >> 1;
      if (sorted[mid] < v) lo = mid + 1; else hi = mid;
    }
    offsets[v] = (int32_t)lo;
  }
}

// norm[v] = 0; for each face n ascending: norm[i] += nrm, norm[j] += nrm,
// norm[k] += nrm; then norm[v] = norm[v].normalized()  (norm.py:37-47).
__global__ void __launch_bounds__(kBlk)
k_smooth_normals(const float* __restrict__ fnrm, const int32_t* __restrict__ offsets, const int32_t* __restrict__ list,
                 int nverts, float* __restrict__ out) {
  const long stride = (long)gridDim.x * kBlk;
  for (long v = (long)blockIdx.x * kBlk + threadIdx.x; v < nverts; v += stride) {
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int q = offsets[v]; q < offsets[v + 1]; q++) {
      const long f = list[q] / 3;
      sx = sx + fnrm[3 * f];
      sy = sy + fnrm[3 * f + 1];
      sz = sz + fnrm[3 * f + 2];
    }
    const float inv = 1.0f / sqrtf(((sx * sx) + (sy * sy)) + (sz * sz));
    out[3 * v] = inv * sx;
    out[3 * v + 1] = inv * sy;
    out[3 * v + 2] = inv * sz;
  }
}

unsigned grid_of(long n) {
  long g = (n + kBlk - 1) / kBlk;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (unsigned)g;
}

int fail(int code, const std::string& msg) {
  tina::set_global_error(msg);
  return code;
}

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return TINA_OK;
  return fail(TINA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int bits_for(int nverts) {
  int b = 1;
  while (b < 31 && (1L << b) < (long)nverts) b++;
  return b;
}

}  // namespace

extern "C" {

int tina_mesh_transform(const float trans[16], const float trans_normal[9], const float* d_verts, float* d_out_verts,
                        int64_t nverts, const float* d_norms, float* d_out_norms, int64_t nnorms, void* stream) {
  if (!trans) return fail(TINA_ERR_INVALID, "tina_mesh_transform: trans is NULL");
  if (nverts < 0 || nnorms < 0) return fail(TINA_ERR_INVALID, "tina_mesh_transform: negative count");
  if (nverts > 0 && (!d_verts || !d_out_verts)) return fail(TINA_ERR_INVALID, "tina_mesh_transform: positions NULL");
  if (nnorms > 0 && (!d_norms || !d_out_norms || !trans_normal))
    return fail(TINA_ERR_INVALID, "tina_mesh_transform: normals or trans_normal NULL");
  if (nverts == 0 && nnorms == 0) return TINA_OK;
  Mat4 T;
  Mat3 N;
  for (int i = 0; i < 16; i++) T.m[i] = trans[i];
  for (int i = 0; i < 9; i++) N.m[i] = trans_normal ? trans_normal[i] : 0.0f;
  hipLaunchKernelGGL(k_mesh_transform, dim3(grid_of(nverts > nnorms ? nverts : nnorms)), dim3(kBlk), 0,
                     (hipStream_t)stream, T, N, d_verts, d_out_verts, (long)nverts, d_norms, d_out_norms, (long)nnorms);
  return hip_status(hipGetLastError(), "k_mesh_transform");
}

int tina_face_normals(const float* d_soup, int64_t nfaces, int corners, float* d_out, void* stream) {
  if (nfaces < 0 || !(corners == 1 || corners == 3)) return fail(TINA_ERR_INVALID, "tina_face_normals: bad size");
  if (nfaces == 0) return TINA_OK;
  if (!d_soup || !d_out) return fail(TINA_ERR_INVALID, "tina_face_normals: buffer is NULL");
  hipLaunchKernelGGL(k_face_normals, dim3(grid_of(nfaces)), dim3(kBlk), 0, (hipStream_t)stream, d_soup, (long)nfaces,
                     corners, d_out);
  return hip_status(hipGetLastError(), "k_face_normals");
}

int tina_vertex_incidence(const int32_t* d_faces, int64_t nfaces, int nverts, int32_t* d_offsets, int32_t* d_list,
                          void* d_temp, size_t* temp_bytes, void* stream) {
  if (!temp_bytes || nfaces < 0 || nverts < 0 || 3 * nfaces > 0x7fffffffL)
    return fail(TINA_ERR_INVALID, "tina_vertex_incidence: bad size or temp_bytes NULL");
  const long ninc = 3 * (long)nfaces;
  // temp layout: keys in, keys out, vals in, the bad-index flag, then the sort's own storage
  const size_t lists = (size_t)4 * (size_t)(ninc > 0 ? ninc : 1);
  const size_t head = 3 * ((lists + 255) / 256 * 256) + 256;
  size_t sort_bytes = 0;
  if (hip_status(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                                                    (const int32_t*)nullptr, (int32_t*)nullptr, (int)ninc, 0,
                                                    bits_for(nverts), (hipStream_t)stream),
                 "hipcub::DeviceRadixSort::SortPairs (size query)"))
    return TINA_ERR_HIP;
  const size_t need = head + sort_bytes;
  if (!d_temp) {
    *temp_bytes = need;
    return TINA_OK;
  }
  if (*temp_bytes < need) return fail(TINA_ERR_CAPACITY, "tina_vertex_incidence: temp storage too small");
  if (!d_offsets || (ninc > 0 && (!d_faces || !d_list)))
    return fail(TINA_ERR_INVALID, "tina_vertex_incidence: buffer is NULL");
  hipStream_t s = (hipStream_t)stream;
  char* t = (char*)d_temp;
  const size_t seg = (lists + 255) / 256 * 256;
  int32_t* kin = (int32_t*)t;
  int32_t* kout = (int32_t*)(t + seg);
  int32_t* vin = (int32_t*)(t + 2 * seg);
  uint32_t* bad = (uint32_t*)(t + 3 * seg);
  void* sort_tmp = t + head;
  uint32_t h_bad = 0;
  if (hip_status(hipMemcpyAsync(bad, &h_bad, 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync")) return TINA_ERR_HIP;
  if (ninc > 0) {
    hipLaunchKernelGGL(k_incidences, dim3(grid_of(ninc)), dim3(kBlk), 0, s, d_faces, ninc, nverts, kin, vin, bad);
    if (hip_status(hipGetLastError(), "k_incidences")) return TINA_ERR_HIP;
    size_t sb = sort_bytes;
    if (hip_status(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sb, kin, kout, vin, d_list, (int)ninc, 0,
                                                      bits_for(nverts), s),
                   "hipcub::DeviceRadixSort::SortPairs"))
      return TINA_ERR_HIP;
  }
  hipLaunchKernelGGL(k_vertex_offsets, dim3(grid_of((long)nverts + 1)), dim3(kBlk), 0, s, kout, ninc, nverts,
                     d_offsets);
  if (hip_status(hipGetLastError(), "k_vertex_offsets")) return TINA_ERR_HIP;
  if (hip_status(hipMemcpyAsync(&h_bad, bad, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync") ||
      hip_status(hipStreamSynchronize(s), "hipStreamSynchronize"))
    return TINA_ERR_HIP;
  if (h_bad) return fail(TINA_ERR_INVALID, "tina_vertex_incidence: a face index lies outside [0, nverts)");
  return TINA_OK;
}

int tina_smooth_normals(const float* d_face_nrm, const int32_t* d_offsets, const int32_t* d_list, int nverts,
                        float* d_out, void* stream) {
  if (nverts < 0) return fail(TINA_ERR_INVALID, "tina_smooth_normals: nverts < 0");
  if (nverts == 0) return TINA_OK;
  if (!d_offsets || !d_out) return fail(TINA_ERR_INVALID, "tina_smooth_normals: buffer is NULL");
  hipLaunchKernelGGL(k_smooth_normals, dim3(grid_of(nverts)), dim3(kBlk), 0, (hipStream_t)stream, d_face_nrm,
                     d_offsets, d_list, nverts, d_out);
  return hip_status(hipGetLastError(), "k_smooth_normals");
}

}  // extern "C"
