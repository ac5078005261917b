/*
 * tina_oracle.c -- TEST INFRASTRUCTURE: the CPU checker of the HIP path.
 * Coverage part (orc_raster*).  Shading lives in shade_impl.h, instantiated
 * for float and double at the bottom of this file.
 *
 * Coverage restates tina/core/triangle.py:88-131 step by step (the comments give
 * the line each step follows).  All arithmetic is binary32, evaluated left to
 * right as the Python source groups it, with no FMA (-ffp-contract=off).
 * Conversions float -> int saturate (NaN -> 0), the behaviour of the AMD
 * conversion instruction and of Taichi's CUDA backend; Taichi's x86 backend
 * would give INT_MIN out of range, which no workload here reaches.
 */
#include "tina_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define MAXDEPTH 1073741824 /* Engine.maxdepth = 2**30, tina/core/engine.py:12 */

static inline int32_t f2i_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}

/* mapply(mat, pos, 1) (tina/common.py:169-177): r_i = M[i][3]*1 then += M[i][j]*p_j
 * for j = 0,1,2; w likewise from row 3. */
static inline void mapply1(const float *M, const float *p, float r[3], float *w)
{
    for (int i = 0; i < 3; i++) {
        float acc = M[4 * i + 3] * 1.0f;
        acc = acc + M[4 * i + 0] * p[0];
        acc = acc + M[4 * i + 1] * p[1];
        acc = acc + M[4 * i + 2] * p[2];
        r[i] = acc;
    }
    float ww = M[15] * 1.0f;
    ww = ww + M[12] * p[0];
    ww = ww + M[13] * p[1];
    ww = ww + M[14] * p[2];
    *w = ww;
}

typedef struct {
    float Av[3], Bv[3], Cv[3];
    float b[2], c[2];
    float bcn[2], can[2], wsc[3];
    int32_t bot[2], top[2];
} face_setup;

/* Per-face part of render_occup (triangle.py:93-113).  Returns 0 when the face
 * is culled or clipped. */
static int setup_face(const tina_camera *cam, int W, int H, const float *v,
                      uint32_t flags, face_setup *s)
{
    const float *M = cam->W2V;
    float r[3][3], w[3];
    for (int k = 0; k < 3; k++) mapply1(M, v + 3 * k, r[k], &w[k]);
    float *V[3] = {s->Av, s->Bv, s->Cv};
    for (int k = 0; k < 3; k++)                       /* mapply_pos: res / rew */
        for (int i = 0; i < 3; i++) V[k][i] = r[k][i] / w[k];

    /* facing = (Bv.xy - Av.xy).cross(Cv.xy - Av.xy)   (:95) */
    float facing = (s->Bv[0] - s->Av[0]) * (s->Cv[1] - s->Av[1])
                 - (s->Bv[1] - s->Av[1]) * (s->Cv[0] - s->Av[0]);
    if ((flags & TINA_CULL) && facing <= 0.0f) return 0;          /* :96-98 */

    if (flags & TINA_CLIP) {                                       /* :100-104 */
        int outside = 0;
        for (int k = 0; k < 3; k++) {
            int inside = 1;
            for (int i = 0; i < 3; i++)
                inside &= (-1.0f <= V[k][i]) & (V[k][i] <= 1.0f);
            outside += !inside;
        }
        if (outside == 3) return 0;
    }

    /* to_viewport: (p.xy * 0.5 + 0.5) * res   (engine.py:59-61) */
    float a[2], b[2], c[2];
    const float res[2] = {(float)W, (float)H};
    for (int i = 0; i < 2; i++) {
        a[i] = (s->Av[i] * 0.5f + 0.5f) * res[i];
        b[i] = (s->Bv[i] * 0.5f + 0.5f) * res[i];
        c[i] = (s->Cv[i] * 0.5f + 0.5f) * res[i];
    }
    /* bot, top = ifloor(min(a,b,c)), iceil(max(a,b,c)); clamp to [0, res-1] (:108-109) */
    const int32_t lim[2] = {W - 1, H - 1};
    for (int i = 0; i < 2; i++) {
        float lo = fminf(a[i], fminf(b[i], c[i]));
        float hi = fmaxf(a[i], fmaxf(b[i], c[i]));
        int32_t bo = f2i_sat(floorf(lo));
        int32_t to = f2i_sat(ceilf(hi));
        s->bot[i] = bo > 0 ? bo : 0;
        s->top[i] = to < lim[i] ? to : lim[i];
    }
    /* n = (b - a).cross(c - a); bcn = (b - c) / n; can = (c - a) / n   (:110-112) */
    float n = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
    s->bcn[0] = (b[0] - c[0]) / n;
    s->bcn[1] = (b[1] - c[1]) / n;
    s->can[0] = (c[0] - a[0]) / n;
    s->can[1] = (c[1] - a[1]) / n;
    s->b[0] = b[0]; s->b[1] = b[1];
    s->c[0] = c[0]; s->c[1] = c[1];
    /* wscale = 1 / (w_A, w_B, w_C)   (:113) */
    for (int k = 0; k < 3; k++) s->wsc[k] = 1.0f / w[k];
    return 1;
}

/* Per-pixel weights (:115-120) and the coverage test (:120). */
static inline int pixel_cover(const face_setup *s, int32_t px, int32_t py,
                              const float bias[2], float wei[3])
{
    float posx = (float)px + bias[0];
    float posy = (float)py + bias[1];
    float w_bc = (posx - s->b[0]) * s->bcn[1] - (posy - s->b[1]) * s->bcn[0];
    float w_ca = (posx - s->c[0]) * s->can[1] - (posy - s->c[1]) * s->can[0];
    float t = (1.0f - w_bc) - w_ca;
    float x = w_bc * s->wsc[0];
    float y = w_ca * s->wsc[1];
    float z = t * s->wsc[2];
    float sum = (x + y) + z;
    wei[0] = x / sum;
    wei[1] = y / sum;
    wei[2] = z / sum;
    return (wei[0] >= 0.0f) & (wei[1] >= 0.0f) & (wei[2] >= 0.0f);
}

/* depth = int((wei.x*Av.z + wei.y*Bv.z + wei.z*Cv.z) * 2**30)   (:121-122) */
static inline int32_t pixel_depth(const face_setup *s, const float wei[3])
{
    float d = ((wei[0] * s->Av[2]) + (wei[1] * s->Bv[2])) + (wei[2] * s->Cv[2]);
    return f2i_sat(d * 1073741824.0f);
}

void orc_clear(int W, int H, int32_t *depth, int32_t *occup)
{
    for (long i = 0; i < (long)W * H; i++) {
        depth[i] = MAXDEPTH;       /* engine.py:67-70 */
        occup[i] = -1;             /* triangle.py:90-91 */
    }
}

void orc_raster(int W, int H, const tina_camera *cam, const float *verts, int T,
                uint32_t id_base, uint32_t flags, int32_t *depth, int32_t *occup)
{
    for (int f = 0; f < T; f++) {
        face_setup s;
        if (!setup_face(cam, W, H, verts + 9L * f, flags, &s)) continue;
        for (int32_t px = s.bot[0]; px <= s.top[0]; px++)
            for (int32_t py = s.bot[1]; py <= s.top[1]; py++) {
                float wei[3];
                if (!pixel_cover(&s, px, py, cam->bias, wei)) continue;
                int32_t d = pixel_depth(&s, wei);
                long P = (long)py * W + px;
                if (d < depth[P]) {          /* serial form of atomic_min + occup (:123-125) */
                    depth[P] = d;
                    occup[P] = (int32_t)(id_base + (uint32_t)f);
                }
            }
    }
}

/* ---- particles: tina/core/particle.py:77-116 (render_occup) ---------------- */

/* mapply_pos(M, p) = (r / w) per component (tina/common.py:180-183) */
static inline void mapply_pos3(const float *M, const float p[3], float out[3])
{
    float r[3], w;
    mapply1(M, p, r, &w);
    for (int i = 0; i < 3; i++) out[i] = r[i] / w;
}

/* mapply_dir(M, e_axis).normalized() (common.py:186-189; Matrix.normalized):
 * mapply with weight 0, then v * (1 / sqrt(v.v)) */
static inline void unit_dir(const float *M, int axis, float out[3])
{
    const float e[3] = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
    for (int i = 0; i < 3; i++) {
        float acc = M[4 * i + 3] * 0.0f;
        for (int j = 0; j < 3; j++) acc = acc + M[4 * i + j] * e[j];
        out[i] = acc;
    }
    float inv = 1.0f / sqrtf(((out[0] * out[0]) + (out[1] * out[1])) + (out[2] * out[2]));
    for (int i = 0; i < 3; i++) out[i] = inv * out[i];
}

void orc_raster_particles(int W, int H, const tina_camera *cam, const float *pos, const float *radius, int n,
                          uint32_t id_base, uint32_t flags, int32_t *depth, int32_t *occup)
{
    const float res[2] = {(float)W, (float)H};
    float DX[3], DY[3];
    unit_dir(cam->V2W, 0, DX);                                           /* :89 */
    unit_dir(cam->V2W, 1, DY);                                           /* :90 */
    for (int f = 0; f < n; f++) {
        const float *Al = pos + 3L * f;
        const float Rl = radius[f];
        float Av[3];
        mapply_pos3(cam->W2V, Al, Av);                                   /* :84 */
        if ((flags & TINA_CLIP) && !(-1.0f <= Av[2] && Av[2] <= 1.0f)) continue;  /* :85-87 */
        /* Rv = (to_viewspace(Al + DX*Rl).x - Av.x, to_viewspace(Al + DY*Rl).y - Av.y)  (:91-93) */
        float px_[3], py_[3], tx[3], ty[3];
        for (int i = 0; i < 3; i++) {
            px_[i] = Al[i] + DX[i] * Rl;
            py_[i] = Al[i] + DY[i] * Rl;
        }
        mapply_pos3(cam->W2V, px_, tx);
        mapply_pos3(cam->W2V, py_, ty);
        const float Rv[2] = {tx[0] - Av[0], ty[1] - Av[1]};
        /* the four points Av -+ (Rv.x, 0, 0), Av -+ (0, Rv.y, 0) in viewport units (:94-101);
         * bot = ifloor(min(b0, b2)), top = iceil(max(b1, b3)), clamped (:103-104) */
        int32_t bot[2], top[2];
        for (int i = 0; i < 2; i++) {
            /* along axis i one of the two points is moved by -+Rv[i], the other is not */
            const float lo_moved = ((Av[i] - Rv[i]) * 0.5f + 0.5f) * res[i];
            const float lo_still = ((Av[i] - 0.0f) * 0.5f + 0.5f) * res[i];
            const float hi_moved = ((Av[i] + Rv[i]) * 0.5f + 0.5f) * res[i];
            const float hi_still = ((Av[i] + 0.0f) * 0.5f + 0.5f) * res[i];
            const int32_t bo = f2i_sat(floorf(fminf(lo_moved, lo_still)));
            const int32_t to = f2i_sat(ceilf(fmaxf(hi_moved, hi_still)));
            bot[i] = bo > 0 ? bo : 0;
            top[i] = to < (i == 0 ? W - 1 : H - 1) ? to : (i == 0 ? W - 1 : H - 1);
        }
        const float r2 = Rl * Rl;
        const int32_t d = f2i_sat(Av[2] * 1073741824.0f);              /* :112-113 */
        for (int32_t X = bot[0]; X <= top[0]; X++)
            for (int32_t Y = bot[1]; Y <= top[1]; Y++) {
                /* Pl = from_viewspace((P + bias) / res * 2 - 1, Av.z)   (:106-108) */
                const float Pv[3] = {((((float)X + cam->bias[0]) / res[0]) * 2.0f) - 1.0f,
                                     ((((float)Y + cam->bias[1]) / res[1]) * 2.0f) - 1.0f, Av[2]};
                float Pl[3];
                mapply_pos3(cam->V2W, Pv, Pl);
                const float e0 = Pl[0] - Al[0], e1 = Pl[1] - Al[1], e2 = Pl[2] - Al[2];
                if ((((e0 * e0) + (e1 * e1)) + (e2 * e2)) > r2) continue;   /* :109-110 */
                const long P = (long)Y * W + X;
                if (d < depth[P]) {          /* serial form of atomic_min + occup (:114-116) */
                    depth[P] = d;
                    occup[P] = (int32_t)(id_base + (uint32_t)f);
                }
            }
    }
}

/* ---- wires: tina/core/wireframe.py:49-95 (draw_line + render_color's depth test) ---- */

static inline int in_cube(const float v[3])
{
    int ok = 1;
    for (int i = 0; i < 3; i++) ok &= (-1.0f <= v[i]) & (v[i] <= 1.0f);
    return ok;
}

void orc_raster_wires(int W, int H, const tina_camera *cam, const float *verts, int n, uint32_t id_base,
                      uint32_t flags, int32_t *depth, int32_t *occup)
{
    const float res[2] = {(float)W, (float)H};
    for (int f = 0; f < n; f++) {
        const float *Al = verts + 6L * f, *Bl = verts + 6L * f + 3;
        float r[2][3], w[2], Av[3], Bv[3];
        mapply1(cam->W2V, Al, r[0], &w[0]);
        mapply1(cam->W2V, Bl, r[1], &w[1]);
        for (int i = 0; i < 3; i++) { Av[i] = r[0][i] / w[0]; Bv[i] = r[1][i] / w[1]; }
        if ((flags & TINA_CLIP) && !in_cube(Av) && !in_cube(Bv)) continue;    /* :75-78 */
        float a[2], b[2];
        for (int i = 0; i < 2; i++) {                                       /* :80 */
            a[i] = (Av[i] * 0.5f + 0.5f) * res[i];
            b[i] = (Bv[i] * 0.5f + 0.5f) * res[i];
        }
        const float ws[2] = {1.0f / w[0], 1.0f / w[1]};                     /* :83 */
        /* draw_line(a, b) (:50-64) */
        const float dlt[2] = {b[0] - a[0], b[1] - a[1]};
        const float adlt[2] = {fabsf(dlt[0]), fabsf(dlt[1])};
        float k[2];
        int32_t siz;
        if (adlt[0] >= adlt[1]) {
            k[0] = dlt[0] >= 0.0f ? 1.0f : -1.0f;
            k[1] = (k[0] * dlt[1]) / dlt[0];
            siz = f2i_sat(adlt[0]);
        } else {
            k[1] = dlt[1] >= 0.0f ? 1.0f : -1.0f;
            k[0] = (k[1] * dlt[0]) / dlt[1];
            siz = f2i_sat(adlt[1]);
        }
        for (long i = 0; i <= (long)siz; i++) {
            const float fi = (float)i;
            const float pos[2] = {(a[0] + k[0] * fi) + cam->bias[0], (a[1] + k[1] * fi) + cam->bias[1]};
            const float cor = fi / (float)siz;
            const int32_t X = f2i_sat(floorf(pos[0])), Y = f2i_sat(floorf(pos[1]));
            if (!(0 <= X && X < W && 0 <= Y && Y < H)) continue;            /* :87 */
            float w0 = (1.0f - cor) * ws[0], w1 = cor * ws[1];              /* :88-89 */
            const float sum = w0 + w1;
            w0 = w0 / sum;
            w1 = w1 / sum;
            const int32_t d = f2i_sat(((w0 * Av[2]) + (w1 * Bv[2])) * 1073741824.0f);   /* :90-91 */
            const long P = (long)Y * W + X;
            if (d < depth[P]) {          /* serial form of atomic_min (:92-93) */
                depth[P] = d;
                occup[P] = (int32_t)(id_base + (uint32_t)f);
            }
        }
    }
}

static inline uint64_t pack_key(int32_t depth, uint32_t id)
{
    return ((uint64_t)((uint32_t)depth ^ 0x80000000u) << 32) | (uint64_t)id;
}

void orc_clear_keys(int W, int H, uint64_t *keys, int nthreads)
{
    uint64_t k = pack_key(MAXDEPTH, 0);
    long n = (long)W * H;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long i = 0; i < n; i++) keys[i] = k;
}

void orc_raster_omp(int W, int H, const tina_camera *cam, const float *verts, int T,
                    uint32_t id_base, uint32_t flags, uint64_t *keys, int nthreads)
{
    /* One task per face, each scanning its bbox serially: the reference's
     * launch shape (triangle.py:92-125), with the race removed by a 64-bit
     * compare-and-swap min on the packed (depth, id) key. */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (int f = 0; f < T; f++) {
        face_setup s;
        if (!setup_face(cam, W, H, verts + 9L * f, flags, &s)) continue;
        uint32_t id = id_base + (uint32_t)f;
        for (int32_t px = s.bot[0]; px <= s.top[0]; px++)
            for (int32_t py = s.bot[1]; py <= s.top[1]; py++) {
                float wei[3];
                if (!pixel_cover(&s, px, py, cam->bias, wei)) continue;
                uint64_t k = pack_key(pixel_depth(&s, wei), id);
                uint64_t *dst = keys + (long)py * W + px;
                uint64_t cur = __atomic_load_n(dst, __ATOMIC_RELAXED);
                while (k < cur &&
                       !__atomic_compare_exchange_n(dst, &cur, k, 1, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED)) {
                }
            }
    }
}

void orc_unpack_keys(int W, int H, const uint64_t *keys, int32_t *depth, int32_t *occup)
{
    for (long i = 0; i < (long)W * H; i++) {
        int32_t d = (int32_t)((uint32_t)(keys[i] >> 32) ^ 0x80000000u);
        if (depth) depth[i] = d;
        if (occup) occup[i] = d >= MAXDEPTH ? -1 : (int32_t)(uint32_t)keys[i];
    }
}

/* ---------------------------------------------------------------- shading */
#define REAL float
#define RSUF(name) name##_f32
#define RC(x) ((float)(x))
#include "shade_impl.h"
#undef REAL
#undef RSUF
#undef RC

#define REAL double
#define RSUF(name) name##_f64
#define RC(x) ((double)(x))
#include "shade_impl.h"
#undef REAL
#undef RSUF
#undef RC

/* ---------------------------------------------------------------- IBL bakes */


void orc_skybox_sample_f64(const float *img, int w, int h, const double dir[3], double out[3])
{
    vec_f64 r = skybox_sample_f64(img, w, h, v3_f64(dir[0], dir[1], dir[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* Van der Corput radical inverse in base 2: the second Hammersley coordinate. */
static double radical_inverse2(uint32_t i)
{
    i = (i << 16) | (i >> 16);
    i = ((i & 0x55555555u) << 1) | ((i & 0xAAAAAAAAu) >> 1);
    i = ((i & 0x33333333u) << 2) | ((i & 0xCCCCCCCCu) >> 2);
    i = ((i & 0x0F0F0F0Fu) << 4) | ((i & 0xF0F0F0F0u) >> 4);
    i = ((i & 0x00FF00FFu) << 8) | ((i & 0xFF00FF00u) >> 8);
    return (double)i * 2.3283064365386963e-10;
}

/* Skybox.unmapcoor (tina/skybox.py:131-136) + unspheremap (:31-35) + spherical
 * (tina/advans.py:104-108). */
static vec_f64 unmapcoor_f64(int ix, int iy, int w, int h)
{
    const double tau = 6.283185307179586;
    double v = (double)ix / (double)(w - 1), u = (double)iy / (double)(h - 1);
    double hh = u * 2.0 - 1.0;
    double r = sqrt(fmax(0.0, 1.0 - hh * hh));
    double dx = r * cos(v * tau), dy = r * sin(v * tau), dz = hh;
    return v3_f64(dx, dz, -dy);
}

/* tangentspace(nrm) @ spherical(u, v) (tina/advans.py:96-108) */
static vec_f64 tangent_dir_f64(vec_f64 nrm, double u, double v)
{
    const double tau = 6.283185307179586;
    vec_f64 up = v3_f64(233.0, 666.0, 512.0);
    vec_f64 bitan = vnormalized_f64(vcross_f64(nrm, up));
    vec_f64 tan = vcross_f64(bitan, nrm);
    double r = sqrt(fmax(0.0, 1.0 - u * u));
    double sx = r * cos(v * tau), sy = r * sin(v * tau), sz = u;
    return vadd_f64(vadd_f64(vsmul_f64(sx, tan), vsmul_f64(sy, bitan)), vsmul_f64(sz, nrm));
}

/* Lambert.cook_for_ibl bake (material.py:414-426) with Hammersley points in place
 * of ti.random: res += env(odir) * u; dem += u; out = res / dem. */
void orc_bake_lambert(const float *env, int ew, int eh, float *out, int ow, int oh, int nsamples, int nthreads)
{
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4) collapse(2)
    for (int ix = 0; ix < ow; ix++)
        for (int iy = 0; iy < oh; iy++) {
            vec_f64 dir = unmapcoor_f64(ix, iy, ow, oh);
            vec_f64 res = vs_f64(0.0);
            double dem = 0.0;
            for (int s = 0; s < nsamples; s++) {
                double u = ((double)s + 0.5) / (double)nsamples, v = radical_inverse2((uint32_t)s);
                vec_f64 odir = tangent_dir_f64(dir, u, v);
                vec_f64 wei = skybox_sample_f64(env, ew, eh, odir);
                res = vadd_f64(res, vsmul_f64(u, wei));
                dem += u;
            }
            float *o = out + ((long)ix * oh + iy) * 3;
            o[0] = (float)(res.x / dem); o[1] = (float)(res.y / dem); o[2] = (float)(res.z / dem);
        }
}

/* CookTorrance.cook_for_ibl bake (material.py:283-299): GGX-distributed cosines
 * u = sqrt((1 - u) / (1 - u * (1 - alpha2))), weights u. */
void orc_bake_ggx(const float *env, int ew, int eh, float *out, int ow, int oh,
                  float roughness, int nsamples, int nthreads)
{
    double alpha2 = fmax(0.0, (double)roughness * (double)roughness);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4) collapse(2)
    for (int ix = 0; ix < ow; ix++)
        for (int iy = 0; iy < oh; iy++) {
            vec_f64 dir = unmapcoor_f64(ix, iy, ow, oh);
            vec_f64 res = vs_f64(0.0);
            double dem = 0.0;
            for (int s = 0; s < nsamples; s++) {
                double u0 = ((double)s + 0.5) / (double)nsamples, v = radical_inverse2((uint32_t)s);
                double u = sqrt((1.0 - u0) / (1.0 - u0 * (1.0 - alpha2)));
                vec_f64 odir = tangent_dir_f64(dir, u, v);
                vec_f64 wei = skybox_sample_f64(env, ew, eh, odir);
                res = vadd_f64(res, vsmul_f64(u, wei));
                dem += u;
            }
            float *o = out + ((long)ix * oh + iy) * 3;
            o[0] = (float)(res.x / dem); o[1] = (float)(res.y / dem); o[2] = (float)(res.z / dem);
        }
}

/* ---------------------------------------------------------------- SSR */

/* WangHashRNG (tina/random.py:17-64) */
static uint32_t wang_noise(uint32_t v)
{
    v = (v ^ 61u) ^ (v >> 16);
    v *= 9u;
    v ^= v << 4;
    v *= 0x27d4eb2du;
    v ^= v >> 15;
    return v;
}
static float wang_random(uint32_t *seed)
{
    uint32_t u = wang_noise(*seed) >> 1;
    *seed += 1u;
    return (float)u * 4.656612873077393e-10f;    /* u * (2 / 2**32) */
}
static uint32_t wang_random_int(uint32_t *seed)
{
    uint32_t u = wang_noise(*seed);
    *seed += 1u;
    return u;
}

/* spherical(h, p) (tina/advans.py:105-108); cos and sin in binary64, rounded */
static vec_f32 spherical_f32(float h, float p)
{
    float a = p * 6.2831855f;
    float c = (float)cos((double)a), sn = (float)sin((double)a);
    float r = sqrtf(fmaxf(0.0f, 1.0f - h * h));
    return v3_f32(r * c, r * sn, h);
}

/* tangentspace(n) @ d (advans.py:97-100): columns tan, bitan, n */
static vec_f32 tangent_apply_f32(vec_f32 n, vec_f32 d)
{
    vec_f32 bitan = vnormalized_f32(vcross_f32(n, v3_f32(233.0f, 666.0f, 512.0f)));
    vec_f32 tan = vcross_f32(bitan, n);
    return v3_f32(((tan.x * d.x) + (bitan.x * d.y)) + (n.x * d.z), ((tan.y * d.x) + (bitan.y * d.y)) + (n.y * d.z),
                  ((tan.z * d.x) + (bitan.z * d.y)) + (n.z * d.z));
}

/* material.sample(idir, nrm, 1, rng) of the program (material.py:25-32, 124-138,
 * 181-184, 228-238, 365-384, 398-405, 467-477), walked recursively. */
static void sample_node(const tina_material *m, const vec_f32 *pv, int node, const int *size, vec_f32 idir,
                        vec_f32 nrm, uint32_t *rng, vec_f32 *odir, vec_f32 *wei)
{
    const tina_op *o = &m->ops[node];
    if (o->op == TINA_OP_SCALE) {
        sample_node(m, pv, node - 1, size, idir, nrm, rng, odir, wei);
        *wei = vmul_f32(*wei, pv[o->a]);
        return;
    }
    if (o->op == TINA_OP_MIX) {
        vec_f32 fac = pv[o->a];
        int scalar = m->params[o->a].dim == 1;
        float f = scalar ? fac.x : ((fac.x + fac.y) + fac.z) / 3.0f;
        if (f != 0.0f && f != 1.0f) {
            float t = (f - 0.0f) / (1.0f - 0.0f);
            t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
            float sm = (t * t) * (3.0f - 2.0f * t);
            f = 0.12f * (1.0f - sm) + 0.88f * sm;
        }
        if (wang_random(rng) < f) {
            sample_node(m, pv, node - 1, size, idir, nrm, rng, odir, wei);
            vec_f32 k = scalar ? vs_f32(fac.x / f) : vdivs_f32(fac, f);
            *wei = vmul_f32(*wei, k);
        } else {
            sample_node(m, pv, node - 1 - size[node - 1], size, idir, nrm, rng, odir, wei);
            vec_f32 k = scalar ? vs_f32((1.0f - fac.x) / (1.0f - f)) : vdivs_f32(vsub_f32(vs_f32(1.0f), fac), 1.0f - f);
            *wei = vmul_f32(*wei, k);
        }
        return;
    }
    if (o->op == TINA_OP_ADD) {
        int first = (wang_random_int(rng) % 2u) == 0u;
        sample_node(m, pv, first ? node - 1 - size[node - 1] : node - 1, size, idir, nrm, rng, odir, wei);
        *wei = vmul_f32(*wei, vs_f32(2.0f));
        return;
    }
    *odir = idir;
    *wei = vs_f32(0.0f);
    if (o->op == TINA_OP_LAMBERT) {
        float u = wang_random(rng);
        float v = wang_random(rng);
        u = sqrtf(u);
        *odir = vnormalized_f32(tangent_apply_f32(nrm, spherical_f32(u, v)));
        *wei = vs_f32(1.0f);
    } else if (o->op == TINA_OP_COOKTORRANCE) {
        float roughness = pv[o->a].x;
        vec_f32 f0 = pv[o->b];
        float alpha2 = fmaxf(0.0f, roughness * roughness);
        float u = wang_random(rng);
        float v = wang_random(rng);
        u = sqrtf((1.0f - u) / (1.0f - u * (1.0f - alpha2)));
        vec_f32 rdir = reflect_f32(vneg_f32(idir), nrm);
        vec_f32 od = tangent_apply_f32(rdir, spherical_f32(u, v));
        /* sub_brdf(nrm, idir, od): the Fresnel and visibility terms */
        const float EPS = 1e-10f, ONE_M_EPS = (float)(1.0 - 1e-10);
        vec_f32 half = vnormalized_f32(vadd_f32(idir, od));
        float NoL = fmaxf(EPS, vdot_f32(idir, nrm));
        float NoV = fmaxf(EPS, vdot_f32(od, nrm));
        float VoH = fminf(ONE_M_EPS, fmaxf(EPS, vdot_f32(half, od)));
        float a2 = fmaxf(1e-6f, roughness * roughness);
        float k = a2 / 2.0f;
        float vdf = 1.0f / (((NoV * k) + 1.0f) - k);
        vdf = vdf * (1.0f / (((NoL * k) + 1.0f) - k));
        vdf = vdf / ((1.0f * (1.0f - a2)) + ((float)(4.0 * 3.141592653589793) * a2));
        float x1 = 1.0f - VoH, x2 = x1 * x1, p5 = (x2 * x2) * x1;
        vec_f32 fdf = vadd_f32(f0, vsmul_f32(p5, vsub_f32(vs_f32(1.0f), f0)));
        if (vdot_f32(od, nrm) < 0.0f) {
            od = vneg_f32(od);
            fdf = vs_f32(0.0f);
        }
        *odir = od;
        *wei = vsmul_f32(vdf, fdf);
    } else if (o->op == TINA_OP_PHONG || o->op == TINA_OP_BLINNPHONG) {
        float mm = pv[o->a].x;
        float u = wang_random(rng);
        float v = wang_random(rng);
        u = (float)pow((double)u, (double)(1.0f / (mm + 1.0f)));
        vec_f32 rdir = reflect_f32(vneg_f32(idir), nrm);
        vec_f32 od = tangent_apply_f32(rdir, spherical_f32(u, v));
        float w = 1.0f;
        if (vdot_f32(od, nrm) < 0.0f) {
            od = vneg_f32(od);
            w = 0.0f;
        }
        *odir = od;
        *wei = vs_f32(w);
    }
}

static vec_f32 screen_px(const float *img, int W, int x, int y)
{
    const float *p = img + ((long)y * W + x) * 3;
    return v3_f32(p[0], p[1], p[2]);
}

/* bilerp (tina/common.py:140-149) over the [H][W][3] image, texels clamped */
static vec_f32 screen_bilerp(const float *img, int W, int H, float px, float py)
{
    int ix = f2i_sat(floorf(px)), iy = f2i_sat(floorf(py));
    float x0 = px - (float)ix, x1 = py - (float)iy;
    float y0 = 1.0f - x0, y1 = 1.0f - x1;
    int xa = ix < 0 ? 0 : (ix > W - 1 ? W - 1 : ix), xb = ix + 1 < 0 ? 0 : (ix + 1 > W - 1 ? W - 1 : ix + 1);
    int ya = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy), yb = iy + 1 < 0 ? 0 : (iy + 1 > H - 1 ? H - 1 : iy + 1);
    vec_f32 r = vsmul_f32(x1, vsmul_f32(x0, screen_px(img, W, xb, yb)));
    r = vadd_f32(r, vsmul_f32(y1, vsmul_f32(x0, screen_px(img, W, xb, ya))));
    r = vadd_f32(r, vsmul_f32(y1, vsmul_f32(y0, screen_px(img, W, xa, ya))));
    r = vadd_f32(r, vsmul_f32(x1, vsmul_f32(y0, screen_px(img, W, xa, yb))));
    return r;
}

void orc_ssr(int W, int H, const tina_camera *cam, const tina_mesh *mesh, const tina_material *mat,
             const int32_t *depth, const int32_t *occup, const float *norm, const float *coor, const float *image,
             int nsamples, int nsteps, float stepsize, float tolerance, int blur, float *out)
{
    int size[TINA_MAX_OPS];
    for (int i = 0; i < mat->nops; i++) {
        int op = mat->ops[i].op;
        if (op <= TINA_OP_EMISSION) size[i] = 1;
        else if (op == TINA_OP_SCALE) size[i] = 1 + size[i - 1];
        else size[i] = 1 + size[i - 1] + size[i - 1 - size[i - 1]];
    }
    const float resx = (float)W, resy = (float)H;
    for (long P = 0; P < (long)W * H; P++) {
        if (occup[P] < 0) continue;
        uint32_t fu = (uint32_t)occup[P] - mesh->id_base;
        if (fu >= (uint32_t)mesh->nfaces) continue;
        int x = (int)(P % W), y = (int)(P / W);
        vec_f32 normal = screen_px(norm, W, x, y);
        float *o = out + 4 * P;
        if (vdot_f32(normal, normal) < 1e-6f) {
            o[0] = o[1] = o[2] = o[3] = 0.0f;
            continue;
        }
        vec_f32 tc = coor ? v3_f32(coor[2 * P], coor[2 * P + 1], 0.0f) : vs_f32(0.0f);
        float p_x = (float)x + cam->bias[0], p_y = (float)y + cam->bias[1];
        vec_f32 vpos = v3_f32(((p_x / resx) * 2.0f) - 1.0f, ((p_y / resy) * 2.0f) - 1.0f,
                              (float)depth[P] / 1073741824.0f);
        vec_f32 pos = mapply_pos_f32(cam->V2W, vpos, NULL);
        /* calc_viewdir (shader.py:81-93) */
        vec_f32 ro0 = mapply_pos_f32(cam->V2W, v3_f32(vpos.x, vpos.y, -1.0f), NULL);
        vec_f32 ro1 = mapply_pos_f32(cam->V2W, v3_f32(vpos.x, vpos.y, 1.0f), NULL);
        vec_f32 viewdir = vneg_f32(vnormalized_f32(vsub_f32(ro1, ro0)));
        vec_f32 pv[TINA_MAX_PARAMS];
        eval_params_f32(mat, (long)fu, pos, normal, tc, vs_f32(1.0f), pv);
        uint32_t rng = wang_noise((uint32_t)(y % blur) ^ wang_noise((uint32_t)(x % blur)));
        float nst = (float)nsteps;
        vec_f32 vpz = mapply_pos_f32(cam->W2V, pos, NULL);
        float vtol = tolerance * (mapply_pos_f32(cam->W2V, vsub_f32(pos, vdivs_f32(viewdir, nst)), NULL).z - vpz.z);
        float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int i = 0; i < nsamples; i++) {
            vec_f32 odir, wei;
            sample_node(mat, pv, mat->nops - 1, size, viewdir, normal, &rng, &odir, &wei);
            float c = vdot_f32(odir, viewdir);
            float step = stepsize / (sqrtf(1.0f - c * c) * nst);
            float jit = wang_random(&rng);
            vec_f32 ro = vadd_f32(pos, vsmul_f32(step, vsmul_f32(jit, odir)));
            for (int j = 0; j < nsteps; j++) {
                ro = vadd_f32(ro, vsmul_f32(step, odir));
                vec_f32 vro = mapply_pos_f32(cam->W2V, ro, NULL);
                if (!((-1.0f <= vro.x) && (vro.x <= 1.0f) && (-1.0f <= vro.y) && (vro.y <= 1.0f) &&
                      (-1.0f <= vro.z) && (vro.z <= 1.0f)))
                    break;
                float Dx = (vro.x * 0.5f + 0.5f) * resx, Dy = (vro.y * 0.5f + 0.5f) * resy;
                int ix = (int)Dx, iy = (int)Dy;
                if (ix > W - 1) ix = W - 1;
                if (iy > H - 1) iy = H - 1;
                float d = (float)depth[(long)iy * W + ix] / 1073741824.0f;
                if ((vro.z - vtol < d) && (d < vro.z)) {
                    vec_f32 clr = vmul_f32(screen_bilerp(image, W, H, Dx, Dy), wei);
                    r[0] = r[0] + clr.x;
                    r[1] = r[1] + clr.y;
                    r[2] = r[2] + clr.z;
                    r[3] = r[3] + 1.0f;
                    break;
                }
            }
        }
        for (int k = 0; k < 4; k++) o[k] = r[k] / (float)nsamples;
    }
}

void orc_ssr_apply(int W, int H, float *img, const float *buf, int blur)
{
    const int offs = blur / 2;
    for (long P = 0; P < (long)W * H; P++) {
        int x = (int)(P % W), y = (int)(P / W);
        float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < blur; k++)
            for (int l = 0; l < blur; l++) {
                int xx = x + k - offs, yy = y + l - offs;
                xx = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
                yy = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
                const float *b = buf + 4 * ((long)yy * W + xx);
                for (int c = 0; c < 4; c++) r[c] = r[c] + b[c];
            }
        for (int c = 0; c < 4; c++) r[c] = r[c] / (float)(blur * blur);
        float m = 1.0f - r[3];
        for (int c = 0; c < 3; c++) img[3 * P + c] = img[3 * P + c] * m + r[c];
    }
}

/* ---------------------------------------------------------------- mesh nodes */

/* mapply_pos (tina/common.py:169-183): res = mat[i][3] * 1, then res += mat[i][j] *
 * pos[j] for j = 0..2; w likewise from row 3; res / w.  Normals: trans_normal @ n,
 * each entry ((M[i][0] n0 + M[i][1] n1) + M[i][2] n2) (Matrix @ Vector). */
void orc_mesh_transform(const float trans[16], const float trans_normal[9], const float *verts, float *out_verts,
                        long nverts, const float *norms, float *out_norms, long nnorms)
{
    for (long i = 0; i < nverts; i++) {
        const float *p = verts + 3 * i;
        float r[3], w;
        mapply1(trans, p, r, &w);
        for (int k = 0; k < 3; k++) out_verts[3 * i + k] = r[k] / w;
    }
    for (long i = 0; i < nnorms; i++) {
        const float *n = norms + 3 * i;
        for (int k = 0; k < 3; k++) {
            const float *m = trans_normal + 3 * k;
            out_norms[3 * i + k] = ((m[0] * n[0]) + (m[1] * n[1])) + (m[2] * n[2]);
        }
    }
}

/* (b - a).cross(c - a).normalized(): Taichi's cross product, then v * (1 / |v|)
 * with |v|^2 summed x, y, z in order. */
static void face_normal_f32(const float *v, float n[3])
{
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float inv = 1.0f / sqrtf(((c[0] * c[0]) + (c[1] * c[1])) + (c[2] * c[2]));
    for (int k = 0; k < 3; k++) n[k] = inv * c[k];
}

void orc_face_normals(const float *soup, long nfaces, float *out)
{
    for (long f = 0; f < nfaces; f++) face_normal_f32(soup + 9 * f, out + 3 * f);
}

/* norm[v] = 0; for n in range(nfaces): norm[i] += nrm; norm[j] += nrm;
 * norm[k] += nrm; then norm[v] = norm[v].normalized()  (norm.py:37-47). */
void orc_smooth_normals(const float *soup, const int32_t *faces, long nfaces, int nverts, float *out)
{
    for (long v = 0; v < 3L * nverts; v++) out[v] = 0.0f;
    for (long f = 0; f < nfaces; f++) {
        float n[3];
        face_normal_f32(soup + 9 * f, n);
        for (int c = 0; c < 3; c++) {
            float *o = out + 3L * faces[3 * f + c];
            for (int k = 0; k < 3; k++) o[k] = o[k] + n[k];
        }
    }
    for (long v = 0; v < nverts; v++) {
        float *o = out + 3 * v;
        const float inv = 1.0f / sqrtf(((o[0] * o[0]) + (o[1] * o[1])) + (o[2] * o[2]));
        for (int k = 0; k < 3; k++) o[k] = inv * o[k];
    }
}
