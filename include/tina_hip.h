/*
 * tina_hip.h -- C ABI of the MI355X rasterise-and-shade library (libtina_hip.so).
 *
 * The reference (taichi_three / Tina) has no FFI: its "operator API" is a set of
 * duck-typed Python protocols that Taichi inlines at JIT time.  Each entry point
 * below replaces one Taichi kernel of that path; the comment on each names the
 * reference function it stands in for (paths relative to the reference root).
 *
 * Conventions
 *   - Every pointer named "d_*" is a device (HBM) pointer; everything else is host.
 *   - Screen buffers are row-major [y][x] with y = 0 at the bottom row, i.e. the
 *     transpose of Taichi's [x, y] fields (tina/core/engine.py:9-11).
 *   - Triangle soups are [T][3][3] float32 (face, corner, xyz), the layout of
 *     TriangleRaster.verts (tina/core/triangle.py:18).
 *   - Every call returns TINA_OK (0) or an error code; tina_last_error() gives the
 *     message.  No C++ exception crosses this boundary.  Calls only enqueue work on
 *     the given stream (a hipStream_t passed as void*; NULL = default stream).
 *   - A context is not thread-safe.  All device scratch is allocated by
 *     tina_ctx_create / tina_ctx_resize, never inside the per-frame calls.
 */
#ifndef TINA_HIP_H
#define TINA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define TINA_OK            0
#define TINA_ERR_INVALID   1   /* bad argument (null pointer, bad size, bad enum) */
#define TINA_ERR_CAPACITY  2   /* more faces / lights / ops than the context holds */
#define TINA_ERR_HIP       3   /* a HIP runtime call failed */
#define TINA_ERR_ALLOC     4   /* device allocation failed (or exceeds the set limit) */

/* ---- raster flags (TriangleRaster options, tina/core/triangle.py:6-7) ---- */
#define TINA_CULL   1u   /* drop faces with facing <= 0 (triangle.py:95-98) */
#define TINA_CLIP   2u   /* drop faces whose 3 vertices are all outside NDC (triangle.py:100-104) */

/* ---- shade flags --------------------------------------------------------- */
#define TINA_SMOOTH   1u    /* interpolate per-vertex normals (triangle.py:36-38) */
#define TINA_TEX      2u    /* interpolate per-vertex texcoords (triangle.py:43-46) */
#define TINA_VCOLOR   4u    /* interpolate per-vertex colours for Input('color') (extension) */
#define TINA_TONEMAP  8u    /* apply ACES to written pixels (tina/postp/tonemap.py:9-12) */
#define TINA_BG      16u    /* write the background colour where no face won */
#define TINA_IBL     32u    /* SkyboxLighting instead of Lighting (tina/core/lighting.py:4-21) */
#define TINA_NOLIGHT 64u    /* skip direct lights (IBL only) */
#define TINA_PARTICLES 128u /* the mesh holds particles: verts = centres [n][3], norms = radii [n],
                               colors = per-particle colours [n][3] (ParticleRaster.render_color,
                               tina/core/particle.py:119-148) */

/* ---- limits -------------------------------------------------------------- */
#define TINA_MAX_LIGHTS  16   /* Lighting(maxlights=16), tina/core/lighting.py:25 */
#define TINA_MAX_PARAMS  12
#define TINA_MAX_OPS     16
#define TINA_MAX_STACK    4

/* Camera: Engine.W2V / V2W / bias (tina/core/engine.py:14-17, 72-76).
 * Matrices are row-major float32, M[i][j] = m[4*i + j]. */
typedef struct tina_camera {
    float W2V[16];
    float V2W[16];
    float bias[2];
    float pad[2];
} tina_camera;

/* Lighting fields (tina/core/lighting.py:26-29): light_dirs vec4 (w = 0 for a
 * directional light, 1 for a point light), light_colors vec3, ambient vec3. */
typedef struct tina_lights {
    int32_t n;
    int32_t pad[3];
    float dirs[TINA_MAX_LIGHTS][4];
    float colors[TINA_MAX_LIGHTS][4];
    float ambient[4];
} tina_lights;

/* Material parameter sources: the lowered form of tina.matr.nodes
 * (tina/matr/nodes.py:42-111) plus FresnelFactor (tina/matr/material.py:68-83). */
enum tina_param_kind {
    TINA_P_CONST    = 0,  /* Const: value[0..2]                                  */
    TINA_P_PARAM    = 1,  /* Param: device scalar/vec3 at ptr (dim 1 or 3)        */
    TINA_P_COLOR    = 2,  /* Input('color'): per-vertex colour or (1,1,1)         */
    TINA_P_POS      = 3,  /* Input('pos')                                         */
    TINA_P_NORMAL   = 4,  /* Input('normal')                                      */
    TINA_P_TEXCOORD = 5,  /* Input('texcoord') (z = 0)                            */
    TINA_P_TEXTURE  = 6,  /* Texture: bilerp(tex[w][h][c], slot(a).xy * (shape-1)) */
    TINA_P_FACE     = 7,  /* per-face array at ptr, dim 1 or 3 (extension)        */
    TINA_P_FRESNEL  = 8   /* FresnelFactor(metallic=slot a, albedo=slot b, specular=slot c) */
};

typedef struct tina_param {
    int32_t kind;
    int32_t dim;          /* 1 (scalar, broadcast) or 3 */
    int32_t a, b, c;      /* operand slots (FRESNEL, TEXTURE) */
    int32_t tex_w, tex_h, tex_c;
    float value[4];       /* CONST */
    uint64_t ptr;         /* device pointer: PARAM / TEXTURE / FACE */
} tina_param;

/* Material program: post-order walk of the material tree (tina/matr/material.py).
 * Leaves push a lobe value, combinators pop and push. */
enum tina_op_kind {
    TINA_OP_LAMBERT      = 0,  /* Lambert (material.py:387-442)                     */
    TINA_OP_PHONG        = 1,  /* Phong(shineness=slot a) (material.py:445-457)     */
    TINA_OP_BLINNPHONG   = 2,  /* BlinnPhong(shineness=slot a) (extension)          */
    TINA_OP_COOKTORRANCE = 3,  /* CookTorrance(roughness=a, fresnel=b) (:244-362)   */
    TINA_OP_EMISSION     = 4,  /* Emission (material.py:636-657)                    */
    TINA_OP_SCALE        = 5,  /* ScaleMaterial(factor=slot a) (material.py:148-191) */
    TINA_OP_MIX          = 6,  /* MixMaterial(factor=slot a) (material.py:86-145)   */
    TINA_OP_ADD          = 7   /* AddMaterial (material.py:194-238)                 */
};

typedef struct tina_op {
    int32_t op, a, b, pad;
} tina_op;

typedef struct tina_material {
    int32_t nparams, nops;
    int32_t pad[2];
    tina_param params[TINA_MAX_PARAMS];
    tina_op ops[TINA_MAX_OPS];
} tina_material;

/* Per-object face data bound by TriangleRaster.set_object (triangle.py:71-86).
 * All pointers are device pointers; optional ones may be 0. */
typedef struct tina_mesh {
    uint64_t verts;    /* [T][3][3] f32 */
    uint64_t norms;    /* [T][3][3] f32 (TINA_SMOOTH) */
    uint64_t coors;    /* [T][3][2] f32 (TINA_TEX)    */
    uint64_t colors;   /* [T][3][3] f32 (TINA_VCOLOR) */
    int32_t nfaces;
    uint32_t id_base;  /* global id of face 0: earlier objects own lower ids */
} tina_mesh;

/* Image-based lighting tables (SkyboxLighting, tina/core/lighting.py:4-21):
 * equirect maps [w][h][3] f32 (x-major, indexed like Skybox.img[x, y],
 * tina/skybox.py:39-145).  spec[0] is the environment itself (also the
 * background); spec[1..6] are the prefiltered levels of CookTorrance.rough_levels.
 * rough_lo/hi/step are the f32 values of (rough - step, rough, step) for the six
 * levels, computed in f64 on the host as Python does (material.py:251-253). */
#define TINA_IBL_LEVELS 7
typedef struct tina_ibl {
    uint64_t diff;                    /* Lambert irradiance map          */
    int32_t diff_w, diff_h;
    uint64_t spec[TINA_IBL_LEVELS];   /* env + 6 GGX-prefiltered levels  */
    int32_t spec_w[TINA_IBL_LEVELS], spec_h[TINA_IBL_LEVELS];
    float lut[3][3][2];               /* LUT texels [0..2]x[0..2] (x = VoH, y = roughness, [A,B]):
                                         bilerp at coordinates in [0,1]^2 reads only these (material.py:275) */
    float rough_lo[6], rough_hi[6], rough_step[6];
} tina_ibl;

typedef struct tina_ctx tina_ctx;

/* ---- context ------------------------------------------------------------- */

/* Engine(res_x, res_y) + TriangleRaster(maxfaces) storage (engine.py:6-28,
 * triangle.py:6-29).  Allocates the packed depth|id buffer and raster scratch. */
int tina_ctx_create(int width, int height, int max_faces, tina_ctx **out);
int tina_ctx_destroy(tina_ctx *ctx);
/* Grows the face capacity (TriangleRaster maxfaces) without touching the keys. */
int tina_ctx_reserve(tina_ctx *ctx, int max_faces);
/* Message of the last failed call on ctx (or of the last failed create when ctx is NULL). */
const char *tina_last_error(const tina_ctx *ctx);
/* Library version, for the Python loader's sanity check. */
int tina_version(void);

/* ---- per-frame path ------------------------------------------------------ */

/* Engine.set_camera / bias (engine.py:72-76, :17) -- host copy, passed by value. */
int tina_set_camera(tina_ctx *ctx, const tina_camera *cam);

/* Engine.clear_depth (engine.py:67-70) + TriangleRaster.occup = -1 (triangle.py:90-91):
 * every packed key becomes (depth 2**30, no face).  The clear is recorded and
 * applied by the next call that touches the keys, on that call's stream (the
 * binned tina_raster folds it into its tile pass, which starts every tile from
 * empty keys and writes it whole); tina_keys applies it on this call's stream. */
int tina_clear(tina_ctx *ctx, void *stream);

/* TriangleRaster.render_occup (triangle.py:88-131): setup, cull, clip, coverage,
 * depth and the depth|id min for faces [0, T) of d_verts with ids id_base + f. */
int tina_raster(tina_ctx *ctx, const float *d_verts, int nfaces, uint32_t id_base,
                uint32_t flags, void *stream);

/* ParticleRaster.render_occup (tina/core/particle.py:77-116): the disc of each
 * particle i in [0, n) (centre d_pos[i] [n][3], radius d_radius[i]) enters the
 * depth|id min with depth int(z_view * 2**30) and id id_base + i.  flags:
 * TINA_CLIP drops particles whose view-space z lies outside [-1, 1]
 * (particle.py:85-87).  Shade the result with tina_shade and TINA_PARTICLES. */
int tina_raster_particles(tina_ctx *ctx, const float *d_pos, const float *d_radius, int n,
                          uint32_t id_base, uint32_t flags, void *stream);

/* WireframeRaster.render_color's coverage (tina/core/wireframe.py:49-95): the
 * samples draw_line steps along the screen segment of each wire i in [0, n)
 * (d_verts [n][2][3], the two end points) enter the depth|id min with the
 * perspective-interpolated depth and id id_base + i.  flags: TINA_CLIP drops a
 * wire whose two end points both leave the view cube (wireframe.py:75-78). */
int tina_raster_wires(tina_ctx *ctx, const float *d_verts, int n, uint32_t id_base, uint32_t flags,
                      void *stream);

/* The colour pass of WireframeRaster (wireframe.py:92-95, Shader.blend_color
 * with factor 1): every pixel whose winning id lies in [id_base, id_base + n)
 * takes color.  With TINA_BG the pixels no id won take bg; TINA_TONEMAP maps
 * what is written through ACES.  Other flags are rejected. */
int tina_paint_ids(tina_ctx *ctx, uint32_t id_base, int n, const float color[3], const float bg[3],
                   uint32_t flags, float *d_img, void *stream);

/* NormalShader (tina/core/shader.py:43-46): writes the shading normal of every
 * pixel whose winning id lies in [mesh->id_base, +mesh->nfaces) into d_norm
 * ([H][W][3] f32), as the colour pass computes it.  flags: TINA_SMOOTH
 * (interpolated vertex normals) or TINA_PARTICLES (the sphere normal). */
int tina_normals(tina_ctx *ctx, const tina_mesh *mesh, uint32_t flags, float *d_norm, void *stream);

/* The auxiliary shaders of tina/core/shader.py:21-109: for every pixel whose
 * winning id lies in [mesh->id_base, +mesh->nfaces) writes `channels` (1..3)
 * floats per pixel into d_out ([H][W][channels]): mode 0 NormalShader,
 * 1 ViewNormalShader, 2 TexcoordShader, 3 ColorShader, 4 PositionShader,
 * 5 DepthShader, 6 ConstShader (value), 7 ViewdirShader, 8 SimpleShader,
 * 9 ChessboardShader (size).  flags: TINA_SMOOTH, TINA_TEX, TINA_VCOLOR or
 * TINA_PARTICLES select the fragment as tina_shade does. */
int tina_gbuffer(tina_ctx *ctx, const tina_mesh *mesh, uint32_t flags, int mode, int channels,
                 const float value[3], float size, float *d_out, void *stream);

/* SSR.render (tina/postp/ssr.py:51-103) for the pixels whose winning id lies in
 * [mesh->id_base, +mesh->nfaces): per pixel, unless its normal in d_norm
 * ([H][W][3]) is zero (then 0), nsamples directions drawn from the material's
 * sample() (Lambert, Phong, CookTorrance, Mix, Scale, Add, Emission;
 * material.py) at the view position of the depth buffer, each marched nsteps
 * steps of stepsize / (sqrt(1 - (d.v)^2) nsteps); a step within `tolerance` view
 * steps behind the depth buffer takes bilerp(d_image) times the sample weight.
 * Writes (rgb sum, hits) / nsamples to d_out ([H][W][4]).  d_coor: [H][W][2]
 * texcoords (TexcoordShader) or NULL.  taa = 0 draws the reference's
 * WangHashRNG(P % blur) stream; taa = 1 a stream hashed from the pixel and seed. */
int tina_ssr(tina_ctx *ctx, const tina_mesh *mesh, const tina_material *mat, const float *d_norm,
             const float *d_coor, const float *d_image, int nsamples, int nsteps, float stepsize, float tolerance,
             int blur, int taa, uint32_t seed, float *d_out, void *stream);

/* SSR.apply (ssr.py:31-45): res = the blur x blur mean of d_ssr at offset
 * -blur / 2 (reads clamped to the screen; the pixel itself with taa = 1);
 * d_img = d_img * (1 - res.w) + res.xyz. */
int tina_ssr_apply(float *d_img, int width, int height, const float *d_ssr, int blur, int taa, void *stream);

/* SSAO.render (tina/postp/ssao.py:61-96): ambient occlusion per pixel from the
 * depth buffer and d_norm, with nsamples hemisphere samples d_samples [n][3]
 * turned by the nrot x nrot tile d_rot [nrot][nrot][2] (cos, sin); writes
 * clamp(factor * (occ / n - thresh), 0, 1) to d_ao ([H][W] f32). */
int tina_ssao(tina_ctx *ctx, const float *d_norm, const float *d_samples, int nsamples, const float *d_rot,
              int nrot, float radius, float thresh, float factor, float *d_ao, void *stream);

/* SSAO.apply (ssao.py:37-48): d_img ([H][W][3]) times 1 - (d_ao averaged over
 * an nrot x nrot box, box = 1) or times 1 - d_ao (box = 0, the TAA mode). */
int tina_ssao_apply(float *d_img, int width, int height, const float *d_ao, int nrot, int box, void *stream);

/* TriangleRaster.render_color + Shader/Lighting/material (triangle.py:133-153,
 * shader.py:112-131, lighting.py:84-98): shades the pixels whose winning id lies
 * in [mesh->id_base, mesh->id_base + mesh->nfaces) into d_img ([H][W][3] f32).
 * With TINA_BG also writes bg to empty pixels; with TINA_TONEMAP applies ACES. */
int tina_shade(tina_ctx *ctx, const tina_mesh *mesh, const tina_material *mat,
               const tina_lights *lights, const tina_ibl *ibl, float *d_img,
               const float bg[3], uint32_t flags, void *stream);

/* Scene.image.fill(bgcolor) (tina/scene/raster.py:176), ACES-mapped with TINA_TONEMAP. */
int tina_fill(tina_ctx *ctx, float *d_img, const float bg[3], uint32_t flags, void *stream);

/* ToneMapping.apply (tina/postp/tonemap.py:9-12): in-place ACES over n RGB pixels. */
int tina_tonemap_aces(float *d_img, int npixels, void *stream);

/* Engine.depth and TriangleRaster.occup / ProbeShader.elmid readback
 * (engine.py:9, triangle.py:16, tina/probe.py:17-20): unpacks the keys into
 * d_id ([H][W] i32, winning global id or -1) and d_depth ([H][W] i32, 2**30 when
 * empty).  Either output may be NULL. */
int tina_resolve(tina_ctx *ctx, int32_t *d_id, int32_t *d_depth, void *stream);

/* Device pointer of the packed u64 key buffer ([H][W]) for zero-copy consumers. */
int tina_keys(tina_ctx *ctx, uint64_t **d_keys);

/* ---- one-time IBL bakes -------------------------------------------------- */

/* One prefiltered equirect level of the IBL tables from the environment map
 * d_env ([ew][eh][3] f32) into d_out ([ow][oh][3] f32):
 * Lambert.cook_for_ibl (tina/matr/material.py:407-438) when roughness < 0,
 * CookTorrance.cook_for_ibl at that roughness (material.py:279-320) otherwise.
 * nsamples Hammersley points replace ti.random; maps need w, h >= 2.
 * Errors are reported through tina_last_error(NULL). */
int tina_bake_ibl(const float *d_env, int ew, int eh, float *d_out, int ow, int oh,
                  float roughness, int nsamples, void *stream);

/* ---- learned shading surrogate ------------------------------------------- */
/* No counterpart in the reference: a 12-64-64-64-4 ReLU network trained on the
 * analytic PBR material (tina/matr/material.py:322-362, 701-706) under
 * SkyboxLighting plus one directional light (tina/core/lighting.py:4-21, 84-98).
 * Inputs per sample: n, R = 2 (n.V) n - V, R.L (L = light 0), log(roughness),
 * metallic, albedo.  Outputs: the ACES-mapped colour and an estimate of its
 * largest channel error.  Parameters are TINA_NN_PARAMS floats: W1[64][12] b1[64]
 * W2[64][64] b2[64] W3[64][64] b3[64] W4[4][64] b4[4].  Errors are reported
 * through tina_last_error(NULL) by the context-free calls. */
#define TINA_NN_PARAMS 9412
#define TINA_NN_PACK_BYTES (22 * 64 * 16 + 196 * 4)   /* bf16 MFMA fragments + f32 biases */
#define TINA_NN_BLOCK 64                               /* samples per gradient row */

/* Training pairs: d_x [n][12] features, d_y [n][4] (ACES colour, 0) from the
 * analytic shader at hashed inputs of `seed`: n uniform on the sphere, V uniform
 * on its hemisphere, roughness U[r_lo, r_hi], metallic Bernoulli(p_metal)
 * (uniform if p_metal < 0), albedo U[0,1]^3, FresnelFactor specular `specular`. */
int tina_nn_samples(const tina_lights *lights, const tina_ibl *ibl, uint32_t seed, int64_t nsamples,
                    float r_lo, float r_hi, float p_metal, float specular, float *d_x, float *d_y,
                    void *stream);

/* One Adam step (b1 0.9, b2 0.999, eps 1e-8; step counts from 1) on samples
 * [first, first + batch) of d_x/d_y.  d_partial holds ceil(batch / TINA_NN_BLOCK)
 * rows of TINA_NN_PARAMS floats (workspace).  d_grad (optional) receives the
 * batch gradient; lr = 0 computes it without updating anything. */
int tina_mlp_train_step(const float *d_x, const float *d_y, int64_t first, int64_t nsamples, int batch,
                        float *d_params, float *d_m1, float *d_m2, float *d_partial, float lr, int step,
                        float *d_grad, void *stream);

/* No counterpart in the reference: tina_mlp_train_step for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The per-workgroup gradient rows live in a static
 * device array of the library (TINA_MLP_STATIC_MAX_BATCH / 64 rows of
 * TINA_NN_PARAMS floats, 9.6 MB per GPU, held until the process ends) instead of
 * a caller workspace; a batch above TINA_MLP_STATIC_MAX_BATCH samples is
 * rejected with TINA_ERR_CAPACITY.  Allocates nothing.  Every call on one GPU
 * shares that array: calls that overlap on one device race. */
#define TINA_MLP_STATIC_MAX_BATCH 16384
int tina_mlp_train_step_static(const float *d_x, const float *d_y, int64_t first, int64_t nsamples, int batch,
                               float *d_params, float *d_m1, float *d_m2, float lr, int step, float *d_grad,
                               void *stream);

/* Blooming.apply (tina/postp/blooming.py:45-68) on d_img [height][width][3]:
 * the bright pass factor * (1 - 1 / (1 + scale * max(0, c - thresh))) averaged
 * over 2x2 blocks into d_half, blurred along x into d_tmp and along y back into
 * d_half with the radius + 1 weights d_gwei (taps clamped at the borders), then
 * d_img += bilerp(d_half, (x / 2, y / 2)).  d_half, d_tmp: [height/2][width/2][3]. */
int tina_bloom(float *d_img, int width, int height, float thresh, float factor, float scale,
               const float *d_gwei, int radius, float *d_half, float *d_tmp, void *stream);

/* FXAA.apply (tina/postp/fxaa.py:26-68) on d_img [height][width][3] in place;
 * d_lumi [height][width] and d_copy [height][width][3] are scratch.  Neighbours
 * past the border clamp to it. */
int tina_fxaa(float *d_img, int width, int height, float abs_thresh, float rel_thresh, float factor,
              float *d_lumi, float *d_copy, void *stream);

/* Accumator.update (tina/util/accumator.py:17-23): with inv = 1 / count (count
 * counts the frames so far, this one included), d_acc[i] = d_acc[i] * (1 - inv)
 * + d_src[i] * inv for i < n floats. */
int tina_accumulate(float *d_acc, const float *d_src, int64_t n, int count, void *stream);

/* The analytic PBR + IBL + light-0 shader (ACES) on feature rows d_x [n][12] ->
 * d_out [n][4]: the evaluation the network replaces, on the network's inputs
 * (V is recovered as 2 (n.R) n - R, roughness as exp(x7)). */
int tina_nn_analytic_rows(const tina_lights *lights, const tina_ibl *ibl, const float *d_x, int64_t n,
                          float specular, float *d_out, void *stream);

/* float32 parameters -> the packed bf16 inference weights (TINA_NN_PACK_BYTES). */
int tina_mlp_pack(const float *d_params, void *d_packed, void *stream);

/* The packed network on d_x [n][12] -> d_out [n][4] (max(0, colour), error);
 * both buffers 16-byte aligned. */
int tina_mlp_inference(const void *d_packed, const float *d_x, int64_t n, float *d_out, void *stream);

/* tina_shade for the PBR material under IBL (TINA_IBL | TINA_TONEMAP required)
 * through the network: pixels whose error estimate exceeds `threshold` are
 * queued and shaded by the analytic kernel in the same call. */
int tina_shade_nn(tina_ctx *ctx, const tina_mesh *mesh, const tina_material *mat,
                  const tina_lights *lights, const tina_ibl *ibl, const void *d_packed, float *d_img,
                  const float bg[3], uint32_t flags, float threshold, void *stream);

/* Copies (on the stream) the number of pixels the last tina_shade_nn of ctx
 * shaded analytically into d_count (one uint32 of device memory). */
int tina_nn_fallback_count(tina_ctx *ctx, uint32_t *d_count, void *stream);

/* ---- differentiable shading ------------------------------------------------ */
/* No counterpart in the reference (Tina has no gradients): the PBR material
 * (tina/matr/material.py:322-362, 387-396, 701-706) under point lights
 * (tina/core/lighting.py:84-98) with albedo, smooth normal, roughness and
 * metallic per vertex of an indexed mesh and one scalar intensity per light
 * (light l's colour is intensity[l] * lights->colors[l]).  Coverage comes from
 * the packed buffer of the last tina_raster of `verts`, held fixed. */
typedef struct tina_dmesh {
    uint64_t verts;     /* [F][3][3] f32: the positions tina_raster was given   */
    uint64_t faces;     /* [F][3] i32 vertex indices                            */
    uint64_t normals;   /* [V][3] f32 per-vertex normals (interpolated, then normalised) */
    uint64_t colors;    /* [V][3] f32 albedo                                    */
    uint64_t rough;     /* [V] f32                                              */
    uint64_t metal;     /* [V] f32                                              */
    uint64_t intensity; /* [lights->n] f32                                      */
    int32_t nfaces, nverts;
    uint32_t id_base;
    float ambient[3];
    float specular;     /* FresnelFactor specular (0.5 in the PBR preset)       */
} tina_dmesh;

/* Gradient outputs, accumulated (+=) into caller-zeroed device arrays shaped
 * like the inputs. */
typedef struct tina_dgrads {
    uint64_t colors, normals, rough, metal, intensity;
} tina_dgrads;

/* Colour of every pixel covered by this mesh into d_img ([H][W][3]); empty
 * pixels become 0, other objects' pixels are left alone. */
int tina_dshade_forward(tina_ctx *ctx, const tina_dmesh *mesh, const tina_lights *lights, float *d_img,
                        void *stream);

/* Adjoint of tina_dshade_forward: d_gimg ([H][W][3]) is dLoss/d(image); the
 * face weights are recomputed from the id buffer. */
int tina_dshade_backward(tina_ctx *ctx, const tina_dmesh *mesh, const tina_lights *lights,
                         const float *d_gimg, const tina_dgrads *grads, void *stream);

/* ---- batched multi-view rendering ------------------------------------------ */
/* No counterpart in the reference (one Scene renders one view at a time,
 * tina/scene/raster.py:168-214): a batch of independent scenes of any sizes is
 * cleared, binned, rasterised and shaded by five persistent kernels that walk
 * the work items of every scene, and the five launches can be captured once
 * into a HIP graph and replayed with one launch.  Every scene of a batch uses
 * the same material program (the per-scene parameters may differ). */
typedef struct tina_scene_desc {
    int32_t width, height;
    int32_t nfaces;
    uint32_t raster_flags;   /* TINA_CULL | TINA_CLIP                            */
    uint32_t shade_flags;    /* TINA_SMOOTH | TINA_VCOLOR | TINA_TONEMAP | TINA_BG */
    int32_t pad0;
    uint64_t img;            /* [H][W][3] f32 output                              */
    tina_camera cam;
    tina_mesh mesh;          /* mesh.verts: [F][3][3] f32 positions; id_base 0    */
    tina_material mat;
    tina_lights lights;
    float bg[3];
    float pad1;
} tina_scene_desc;

/* The autotuned bin sizes and launch shapes. */
typedef struct tina_batch_tuning {
    int32_t bin_faces;   /* faces per binning workgroup: 1024, 2048, 4096 or 8192      */
    int32_t max_tiles;   /* bbox tiles (16 x 16 px) above which a face is not binned: every
                            tile of its scene takes its part from the giant queue      */
    int32_t grid_bin, grid_tile, grid_shade;  /* persistent workgroups (0: one per work item) */
} tina_batch_tuning;

typedef struct tina_batch tina_batch;

/* Allocates the batch's key buffers and raster scratch and uploads the scene
 * table.  The descriptors' device pointers must stay valid for its lifetime. */
int tina_batch_create(int nscenes, const tina_scene_desc *scenes, const tina_batch_tuning *tune,
                      tina_batch **out);
int tina_batch_destroy(tina_batch *batch);
/* One frame of every scene: five launches on the stream. */
int tina_batch_render(tina_batch *batch, void *stream);
/* Captures tina_batch_render into a HIP graph (once) ... */
int tina_batch_capture(tina_batch *batch, void *stream);
/* ... and replays it with one graph launch. */
int tina_batch_launch(tina_batch *batch, void *stream);
/* Device pointer of scene i's packed depth|id buffer ([H][W] u64). */
int tina_batch_keys(tina_batch *batch, int scene, uint64_t **d_keys);
/* tina_resolve over any n packed keys (a batch scene's, for instance). */
int tina_resolve_keys(const uint64_t *d_keys, int64_t n, int32_t *d_id, int32_t *d_depth, void *stream);

/* ---- mesh nodes ---------------------------------------------------------------- */
/* The per-frame getters of the edit nodes that set_object inlines in the
 * reference, as device passes over resident buffers.  Context-free: errors are
 * read with tina_last_error(NULL). */

/* MeshTransform.get_face_verts / get_face_norms (tina/mesh/trans.py:29-40):
 * d_out_verts[i] = mapply_pos(trans, d_verts[i]) for nverts points, with the
 * op order of tina/common.py:169-183 (row r: trans[r][3] * 1, then + trans[r][j]
 * * p[j] for j = 0, 1, 2; divided by the w row); d_out_norms[i] = trans_normal @
 * d_norms[i] (entry r: (M[r][0] n0 + M[r][1] n1) + M[r][2] n2) for nnorms normals.
 * trans: 16 floats row-major; trans_normal: 9 floats row-major (the reference's
 * trans_normal field).  Points and normals are [n][3] f32. */
int tina_mesh_transform(const float trans[16], const float trans_normal[9], const float *d_verts,
                        float *d_out_verts, int64_t nverts, const float *d_norms, float *d_out_norms,
                        int64_t nnorms, void *stream);

/* MeshFlatNormal.get_face_norms (tina/mesh/norm.py:7-14): normalized((b - a) x
 * (c - a)) of each face of d_soup ([nfaces][3][3] f32), written once per face
 * (corners = 1, [nfaces][3]) or to its three corners (corners = 3, [nfaces][3][3]). */
int tina_face_normals(const float *d_soup, int64_t nfaces, int corners, float *d_out, void *stream);

/* The vertex -> incidence lists MeshSmoothNormal sums over: d_list holds the
 * incidences 3 f + c (face f, corner c) of d_faces ([nfaces][3] i32 vertex
 * indices) grouped by vertex, each group in ascending order; vertex v's group is
 * d_list[d_offsets[v] .. d_offsets[v + 1]).  d_offsets: nverts + 1 i32; d_list:
 * 3 nfaces i32.  With d_temp NULL, stores the workspace size in *temp_bytes;
 * otherwise d_temp holds *temp_bytes bytes.  Synchronises the stream (an index
 * outside [0, nverts) is reported as TINA_ERR_INVALID).  A set-up call, once per
 * topology. */
int tina_vertex_incidence(const int32_t *d_faces, int64_t nfaces, int nverts, int32_t *d_offsets,
                          int32_t *d_list, void *d_temp, size_t *temp_bytes, void *stream);

/* MeshSmoothNormal._update_normal (tina/mesh/norm.py:35-47): d_out[v] = the sum
 * of d_face_nrm[f] ([nfaces][3], tina_face_normals with corners = 1) over the
 * incidences of v in list order -- the order the reference's loop adds them run
 * serially, faces ascending -- then normalized.  d_out: [nverts][3] f32. */
int tina_smooth_normals(const float *d_face_nrm, const int32_t *d_offsets, const int32_t *d_list, int nverts,
                        float *d_out, void *stream);

/* ---- diagnostics ------------------------------------------------------------ */
/* No counterpart in the reference.  Device self-test of the shared-divisor
 * division the kernels use wherever the reference divides (every quotient must
 * equal the correctly rounded a / d bit for bit): runs npairs operand pairs of
 * `mode` (0: every numerator mantissa in [1, 2) against hashed divisors in
 * [1, 2), 2^23 pairs per divisor; 1: random bit patterns over the whole binary32
 * range; 2: random operands with exponents in [-70, 70]) and adds to d_counts[0]
 * the pairs that took the shared-divisor path and to d_counts[1] the pairs whose
 * quotient differs from the compiler's division.  d_counts: 2 zeroed u64. */
int tina_selftest_div(int mode, uint64_t npairs, uint32_t seed, uint64_t *d_counts, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TINA_HIP_H */
