/*
 * tina_oracle.h -- CPU restatement of Tina's rasterise-and-shade path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the HIP path is compared with: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (taichi_three_amd + libtina_hip.so) never links or calls it.
 *
 * It restates, in IEEE binary32 with no FMA contraction (built with
 * -ffp-contract=off) and serial per-face order, the Taichi kernels of the
 * reference:
 *   orc_raster        tina/core/triangle.py:88-131 (render_occup) with the serial
 *                     update rule "depth < depth_buf -> take it", faces ascending
 *   orc_shade_f32/f64 tina/core/triangle.py:133-153 (render_color, interpolate :31-49),
 *                     tina/core/shader.py:81-93,112-131, tina/core/lighting.py:84-98,
 *                     tina/matr/material.py (Lambert, Phong, CookTorrance, Mix,
 *                     Scale, Add, FresnelFactor), tina/matr/nodes.py:42-111,
 *                     tina/postp/tonemap.py + tina/advans.py:31-34 (ACES)
 *   orc_ibl_*         tina/core/lighting.py:4-21, material.py:251-320,407-442,
 *                     tina/skybox.py:23-145, tina/common.py:140-149 (bilerp)
 *   orc_raster_particles  tina/core/particle.py:77-116 (render_occup), particles ascending
 *   orc_raster_wires  tina/core/wireframe.py:49-95 (draw_line steps), wires ascending
 *   orc_raster_omp    the same coverage rule, parallel over faces with a 64-bit
 *                     compare-and-swap min (the CPU baseline; same result)
 *
 * Pinning status of each function is listed in oracle/README.md.
 */
#ifndef TINA_ORACLE_H
#define TINA_ORACLE_H
#include <stdint.h>
#include "../include/tina_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Serial coverage.  depth/occup are [H][W] i32, in/out (clear with orc_clear). */
void orc_clear(int W, int H, int32_t *depth, int32_t *occup);
void orc_raster(int W, int H, const tina_camera *cam, const float *verts, int T,
                uint32_t id_base, uint32_t flags, int32_t *depth, int32_t *occup);

/* ParticleRaster.render_occup (tina/core/particle.py:77-116), serial over
 * particles: pos [n][3], radius [n]; flags: TINA_CLIP. */
void orc_raster_particles(int W, int H, const tina_camera *cam, const float *pos, const float *radius, int n,
                          uint32_t id_base, uint32_t flags, int32_t *depth, int32_t *occup);

/* WireframeRaster's depth pass (tina/core/wireframe.py:49-95), serial over wires
 * and over the steps of draw_line: verts [n][2][3]; flags: TINA_CLIP. */
void orc_raster_wires(int W, int H, const tina_camera *cam, const float *verts, int n, uint32_t id_base,
                      uint32_t flags, int32_t *depth, int32_t *occup);

/* Parallel coverage into packed keys ([H][W] u64), nthreads OpenMP threads. */
void orc_clear_keys(int W, int H, uint64_t *keys, int nthreads);
void orc_raster_omp(int W, int H, const tina_camera *cam, const float *verts, int T,
                    uint32_t id_base, uint32_t flags, uint64_t *keys, int nthreads);
void orc_unpack_keys(int W, int H, const uint64_t *keys, int32_t *depth, int32_t *occup);

/* Host-side mesh description (pointers are host addresses cast to uint64). */
/* Shading of the pixels whose id lies in [mesh->id_base, +nfaces).  img [H][W][3].
 * f32: op-exact binary32; f64: the same formulas in binary64 (written as f64 too). */
void orc_shade_f32(int W, int H, const tina_camera *cam, const tina_mesh *mesh,
                   const tina_material *mat, const tina_lights *lights,
                   const tina_ibl *ibl, const int32_t *occup, float *img,
                   const float bg[3], uint32_t flags, int nthreads);
void orc_shade_f64(int W, int H, const tina_camera *cam, const tina_mesh *mesh,
                   const tina_material *mat, const tina_lights *lights,
                   const tina_ibl *ibl, const int32_t *occup, double *img,
                   const double bg[3], uint32_t flags, int nthreads);

/* ACES (tina/advans.py:31-34), in place over n RGB pixels. */
void orc_aces_f32(float *img, int npixels);
void orc_aces_f64(double *img, int npixels);

/* Single-sample entry points used by the closed-form pinning tests and by the
 * neural-surrogate training-set generator: shade one (n, view, pos) with the
 * material and lights, no tone mapping.  params_override (may be NULL) gives
 * the per-sample value of each parameter slot as float3. */
void orc_shade_point_f64(const tina_material *mat, const tina_lights *lights,
                         const tina_ibl *ibl, uint32_t flags, const double pos[3],
                         const double nrm[3], const double viewdir[3],
                         const double *slot_values /* [nparams][3] or NULL */,
                         double out[3]);
void orc_shade_point_f32(const tina_material *mat, const tina_lights *lights,
                         const tina_ibl *ibl, uint32_t flags, const float pos[3],
                         const float nrm[3], const float viewdir[3],
                         const float *slot_values, float out[3]);

/* IBL helpers: Skybox.sample (tina/skybox.py:138-145) and the deterministic
 * Hammersley-quadrature bakes that stand in for the ti.random bakes
 * (material.py:283-299, 414-426). env/out are [w][h][3] (Skybox.img layout). */
void orc_skybox_sample_f64(const float *img, int w, int h, const double dir[3], double out[3]);
void orc_bake_lambert(const float *env, int ew, int eh, float *out, int ow, int oh, int nsamples, int nthreads);
void orc_bake_ggx(const float *env, int ew, int eh, float *out, int ow, int oh,
                  float roughness, int nsamples, int nthreads);

/* Mesh nodes upstream of set_object (binary32, serial):
 *   orc_mesh_transform  MeshTransform.get_face_verts / get_face_norms
 *                       (tina/mesh/trans.py:29-40): mapply_pos (tina/common.py:169-183)
 *                       of every position, trans_normal @ n of every normal
 *   orc_face_normals    normalized((b - a) x (c - a)) per face (tina/mesh/norm.py:7-14)
 *   orc_smooth_normals  MeshSmoothNormal._update_normal (norm.py:35-47) with the
 *                       faces visited in ascending order; faces [nf][3] vertex
 *                       indices, positions taken from the face soup [nf][3][3] */
/* SSR.render / SSR.apply (tina/postp/ssr.py:31-103), binary32, for the pixels
 * whose id (occup) lies in the mesh's range; depth/occup [H][W] i32, norm
 * [H][W][3], coor [H][W][2] or NULL, image [H][W][3]; out [H][W][4].  The
 * non-TAA mode: WangHashRNG(P % blur).  orc_ssr_apply: the blur x blur box at
 * offset -blur / 2 (clamped), img = img * (1 - a) + rgb. */
void orc_ssr(int W, int H, const tina_camera *cam, const tina_mesh *mesh, const tina_material *mat,
             const int32_t *depth, const int32_t *occup, const float *norm, const float *coor, const float *image,
             int nsamples, int nsteps, float stepsize, float tolerance, int blur, float *out);
void orc_ssr_apply(int W, int H, float *img, const float *buf, int blur);

void orc_mesh_transform(const float trans[16], const float trans_normal[9], const float *verts, float *out_verts,
                        long nverts, const float *norms, float *out_norms, long nnorms);
void orc_face_normals(const float *soup, long nfaces, float *out);
void orc_smooth_normals(const float *soup, const int32_t *faces, long nfaces, int nverts, float *out);

#ifdef __cplusplus
}
#endif
#endif
