/*
 * shade_impl.h -- TEST INFRASTRUCTURE (oracle).  Deferred shading of Tina,
 * instantiated twice by tina_oracle.c: REAL = float (op-exact binary32) and
 * REAL = double (the FP64 oracle).  Every expression keeps the grouping of the
 * Python source it restates; the comment on each helper names that source.
 */

typedef struct { REAL x, y, z; } RSUF(vec);
#define VEC RSUF(vec)
#define FN(n) RSUF(n)

static inline REAL FN(r_sqrt)(REAL x) { return sizeof(REAL) == 4 ? (REAL)sqrtf((float)x) : (REAL)sqrt((double)x); }
static inline REAL FN(r_pow)(REAL x, REAL y) { return sizeof(REAL) == 4 ? (REAL)powf((float)x, (float)y) : (REAL)pow((double)x, (double)y); }
static inline REAL FN(r_floor)(REAL x) { return sizeof(REAL) == 4 ? (REAL)floorf((float)x) : (REAL)floor((double)x); }
static inline REAL FN(r_atan2)(REAL y, REAL x) { return sizeof(REAL) == 4 ? (REAL)atan2f((float)y, (float)x) : (REAL)atan2((double)y, (double)x); }
static inline REAL FN(r_max)(REAL a, REAL b) { return a > b ? a : b; }
static inline REAL FN(r_min)(REAL a, REAL b) { return a < b ? a : b; }

static inline VEC FN(v3)(REAL x, REAL y, REAL z) { VEC r; r.x = x; r.y = y; r.z = z; return r; }
static inline VEC FN(vs)(REAL s) { return FN(v3)(s, s, s); }
static inline VEC FN(vadd)(VEC a, VEC b) { return FN(v3)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline VEC FN(vsub)(VEC a, VEC b) { return FN(v3)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline VEC FN(vmul)(VEC a, VEC b) { return FN(v3)(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline VEC FN(vsmul)(REAL s, VEC a) { return FN(v3)(s * a.x, s * a.y, s * a.z); }
static inline VEC FN(vdivs)(VEC a, REAL s) { return FN(v3)(a.x / s, a.y / s, a.z / s); }
static inline VEC FN(vneg)(VEC a) { return FN(v3)(-a.x, -a.y, -a.z); }
/* Matrix.dot: (a * b).sum(), summed left to right */
static inline REAL FN(vdot)(VEC a, VEC b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
/* Matrix.normalized(): v * (1 / sqrt(sum v^2)) */
static inline VEC FN(vnormalized)(VEC a) { REAL inv = RC(1) / FN(r_sqrt)(FN(vdot)(a, a)); return FN(vsmul)(inv, a); }
static inline VEC FN(vcross)(VEC a, VEC b)
{
    return FN(v3)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* reflect(I, N) = I - 2 * N.dot(I) * N   (tina/common.py:197-199) */
static inline VEC FN(reflect)(VEC I, VEC N) { REAL d = RC(2) * FN(vdot)(N, I); return FN(vsub)(I, FN(vsmul)(d, N)); }

/* mapply_pos (tina/common.py:169-183) in REAL precision on f32 matrix entries */
static inline VEC FN(mapply_pos)(const float *M, VEC p, REAL *w_out)
{
    REAL r[3];
    REAL pp[3] = {p.x, p.y, p.z};
    for (int i = 0; i < 3; i++) {
        REAL acc = (REAL)M[4 * i + 3] * RC(1);
        acc = acc + (REAL)M[4 * i + 0] * pp[0];
        acc = acc + (REAL)M[4 * i + 1] * pp[1];
        acc = acc + (REAL)M[4 * i + 2] * pp[2];
        r[i] = acc;
    }
    REAL w = (REAL)M[15] * RC(1);
    w = w + (REAL)M[12] * pp[0];
    w = w + (REAL)M[13] * pp[1];
    w = w + (REAL)M[14] * pp[2];
    if (w_out) *w_out = w;
    return FN(v3)(r[0] / w, r[1] / w, r[2] / w);
}

/* bilerp (tina/common.py:140-149) over a [w][h][c] array; indices clamped to the
 * array (the reference reads past the edge with weight 0, or out of bounds). */
static inline VEC FN(bilerp)(const float *img, int w, int h, int c, REAL px, REAL py)
{
    REAL fx = FN(r_floor)(px), fy = FN(r_floor)(py);
    int ix = f2i_sat((float)fx), iy = f2i_sat((float)fy);
    REAL x0 = px - (REAL)ix, x1 = py - (REAL)iy;
    REAL y0 = RC(1) - x0, y1 = RC(1) - x1;
    int xs[2] = {ix, ix + 1}, ys[2] = {iy, iy + 1};
    for (int k = 0; k < 2; k++) {
        xs[k] = xs[k] < 0 ? 0 : (xs[k] > w - 1 ? w - 1 : xs[k]);
        ys[k] = ys[k] < 0 ? 0 : (ys[k] > h - 1 ? h - 1 : ys[k]);
    }
    VEC f[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const float *t = img + ((long)xs[a] * h + ys[b]) * c;
            f[a][b] = c >= 3 ? FN(v3)(t[0], t[1], t[2]) : FN(vs)(t[0]);
        }
    VEC r = FN(vsmul)(x1, FN(vsmul)(x0, f[1][1]));
    r = FN(vadd)(r, FN(vsmul)(y1, FN(vsmul)(x0, f[1][0])));
    r = FN(vadd)(r, FN(vsmul)(y1, FN(vsmul)(y0, f[0][0])));
    r = FN(vadd)(r, FN(vsmul)(x1, FN(vsmul)(y0, f[0][1])));
    return r;
}

/* Skybox.sample for an equirect map (tina/skybox.py:23-28, 138-145):
 * spheremap swaps to (x, -z, y), coor = (atan2(-z, x)/tau % 1, y*0.5 + 0.5),
 * texel = (shape - 1) * coor. */
static inline VEC FN(skybox_sample)(const float *img, int w, int h, VEC dir)
{
    REAL tau = sizeof(REAL) == 4 ? (REAL)6.2831855f : (REAL)6.283185307179586;
    REAL p = FN(r_atan2)(-dir.z, dir.x) / tau;
    REAL cx = p - FN(r_floor)(p);
    REAL cy = dir.y * RC(0.5) + RC(0.5);
    return FN(bilerp)(img, w, h, 3, (REAL)(w - 1) * cx, (REAL)(h - 1) * cy);
}

/* ---- lobes ---------------------------------------------------------------- */

/* Phong.brdf (material.py:449-454): max(0, V.reflect(-L, N))**m * (m + 2) / 2 */
static inline VEC FN(phong)(REAL m, VEC n, VEC L, VEC V)
{
    VEC rdir = FN(reflect)(FN(vneg)(L), n);
    REAL VoR = FN(r_max)(RC(0), FN(vdot)(V, rdir));
    return FN(vs)((FN(r_pow)(VoR, m) * (m + RC(2))) / RC(2));
}

/* BlinnPhong (extension): max(0, N.normalize(L + V))**m * (m + 8) / 8 */
static inline VEC FN(blinnphong)(REAL m, VEC n, VEC L, VEC V)
{
    VEC h = FN(vnormalized)(FN(vadd)(L, V));
    REAL NoH = FN(r_max)(RC(0), FN(vdot)(n, h));
    return FN(vs)((FN(r_pow)(NoH, m) * (m + RC(8))) / RC(8));
}

/* CookTorrance.sub_brdf / brdf (material.py:322-362) with Tina's own terms. */
static inline VEC FN(cooktorrance)(REAL roughness, VEC f0, VEC n, VEC L, VEC V)
{
    const REAL EPS = RC(1e-10);
    const REAL ONE_M_EPS = sizeof(REAL) == 4 ? (REAL)(float)(1.0 - 1e-10) : (REAL)(1.0 - 1e-10);
    const REAL eps6 = RC(1e-6);                        /* tina/advans.py:5 */
    const REAL four_pi = RC(4.0 * 3.141592653589793);
    VEC half = FN(vnormalized)(FN(vadd)(L, V));
    REAL NoH = FN(r_max)(EPS, FN(vdot)(half, n));
    REAL NoL = FN(r_max)(EPS, FN(vdot)(L, n));
    REAL NoV = FN(r_max)(EPS, FN(vdot)(V, n));
    REAL VoH = FN(r_min)(ONE_M_EPS, FN(r_max)(EPS, FN(vdot)(half, V)));
    REAL alpha2 = FN(r_max)(eps6, roughness * roughness);
    REAL denom = RC(1) - (NoH * NoH) * (RC(1) - alpha2);
    REAL ndf = alpha2 / (denom * denom);
    REAL k = alpha2 / RC(2);
    REAL vdf = RC(1) / (((NoV * k) + RC(1)) - k);
    vdf = vdf * (RC(1) / (((NoL * k) + RC(1)) - k));
    vdf = vdf / ((RC(1) * (RC(1) - alpha2)) + (four_pi * alpha2));
    /* (1 - VoH)**5 by squaring: ((x*x)*(x*x))*x (how Taichi lowers an integer
     * power is not pinned here; squaring keeps both restatements identical) */
    REAL x1 = RC(1) - VoH, x2 = x1 * x1;
    REAL p5 = (x2 * x2) * x1;
    VEC fdf = FN(vadd)(f0, FN(vsmul)(p5, FN(vsub)(FN(vs)(RC(1)), f0)));
    return FN(vsmul)(ndf, FN(vsmul)(vdf, fdf));
}

/* CookTorrance.sample_ibl (material.py:255-277) */
static inline VEC FN(cooktorrance_ibl)(const tina_ibl *ibl, REAL roughness, VEC f0, VEC idir, VEC nrm)
{
    const REAL EPS = RC(1e-10);
    const REAL ONE_M_EPS = sizeof(REAL) == 4 ? (REAL)(float)(1.0 - 1e-10) : (REAL)(1.0 - 1e-10);
    VEC rdir = FN(reflect)(FN(vneg)(idir), nrm);
    VEC wei = FN(vs)(RC(1));
    for (int id = 0; id < 6; id++) {          /* static loop: a later match overwrites */
        REAL lo = (REAL)ibl->rough_lo[id], hi = (REAL)ibl->rough_hi[id], st = (REAL)ibl->rough_step[id];
        if (lo <= roughness && roughness <= hi) {
            VEC w1 = FN(skybox_sample)((const float *)(uintptr_t)ibl->spec[id], ibl->spec_w[id], ibl->spec_h[id], rdir);
            VEC w2 = FN(skybox_sample)((const float *)(uintptr_t)ibl->spec[id + 1], ibl->spec_w[id + 1], ibl->spec_h[id + 1], rdir);
            REAL fac2 = (hi - roughness) / st;
            /* lerp(fac2, wei2, wei1) = wei2 * (1 - fac2) + wei1 * fac2 */
            wei = FN(vadd)(FN(vsmul)(RC(1) - fac2, w2), FN(vsmul)(fac2, w1));
        }
    }
    VEC half = FN(vnormalized)(FN(vadd)(idir, rdir));
    REAL VoH = FN(r_min)(ONE_M_EPS, FN(r_max)(EPS, FN(vdot)(half, rdir)));
    VEC AB = FN(bilerp)(&ibl->lut[0][0][0], 3, 3, 2, VoH, roughness);
    /* bilerp over a 2-channel table: the helper returns (A, A, A) for c < 3, so
     * evaluate B separately on the second channel. */
    REAL A = AB.x;
    VEC ABb = FN(bilerp)(&ibl->lut[0][0][1], 3, 3, 2, VoH, roughness);
    REAL B = ABb.x;
    VEC s = FN(vadd)(FN(vsmul)(A, f0), FN(vs)(B));    /* f0 * AB.x + AB.y */
    return FN(vmul)(wei, s);
}

enum { FN(M_BRDF) = 0, FN(M_AMB) = 1, FN(M_EMI) = 2, FN(M_IBL) = 3 };

/* Evaluates the material program in one of four modes: brdf(nrm, L, V), ambient(),
 * emission(), sample_ibl(tab, V, nrm).  Post-order: leaves push, combinators pop. */
static VEC FN(run_material)(const tina_material *m, const VEC *pv, int mode, VEC n,
                            VEC L, VEC V, const tina_ibl *ibl)
{
    VEC st[TINA_MAX_STACK];
    int sp = 0;
    const REAL inv_pi = RC(1.0 / 3.141592653589793);
    for (int i = 0; i < m->nops; i++) {
        const tina_op *o = &m->ops[i];
        VEC v = FN(vs)(RC(0));
        switch (o->op) {
        case TINA_OP_LAMBERT:
            if (mode == FN(M_BRDF)) v = FN(vs)(inv_pi);
            else if (mode == FN(M_AMB)) v = FN(vs)(RC(1));
            else if (mode == FN(M_IBL))
                v = FN(skybox_sample)((const float *)(uintptr_t)ibl->diff, ibl->diff_w, ibl->diff_h, n);
            st[sp++] = v;
            break;
        case TINA_OP_PHONG:
        case TINA_OP_BLINNPHONG:
            if (mode == FN(M_BRDF))
                v = o->op == TINA_OP_PHONG ? FN(phong)(pv[o->a].x, n, L, V) : FN(blinnphong)(pv[o->a].x, n, L, V);
            else if (mode == FN(M_AMB)) v = FN(vs)(RC(1));
            st[sp++] = v;
            break;
        case TINA_OP_COOKTORRANCE:
            if (mode == FN(M_BRDF)) v = FN(cooktorrance)(pv[o->a].x, pv[o->b], n, L, V);
            else if (mode == FN(M_AMB)) v = FN(vs)(RC(1));
            else if (mode == FN(M_IBL)) v = FN(cooktorrance_ibl)(ibl, pv[o->a].x, pv[o->b], V, n);
            st[sp++] = v;
            break;
        case TINA_OP_EMISSION:
            if (mode == FN(M_EMI)) v = FN(vs)(RC(1));
            st[sp++] = v;
            break;
        case TINA_OP_SCALE:      /* fac * wei */
            st[sp - 1] = FN(vmul)(pv[o->a], st[sp - 1]);
            break;
        case TINA_OP_MIX: {      /* (1 - fac) * wei1 + fac * wei2 */
            VEC w2 = st[--sp], w1 = st[--sp];
            VEC fac = pv[o->a];
            VEC omf = FN(vsub)(FN(vs)(RC(1)), fac);
            st[sp++] = FN(vadd)(FN(vmul)(omf, w1), FN(vmul)(fac, w2));
            break;
        }
        case TINA_OP_ADD: {
            VEC w2 = st[--sp], w1 = st[--sp];
            st[sp++] = FN(vadd)(w1, w2);
            break;
        }
        default:
            break;
        }
    }
    return sp > 0 ? st[0] : FN(vs)(RC(0));
}

/* Lighting.shade_color (tina/core/lighting.py:84-98); with TINA_IBL the
 * SkyboxLighting form (lighting.py:20-21) replaces emission + ambient. */
static VEC FN(lighting)(const tina_material *m, const VEC *pv, const tina_lights *lt,
                        const tina_ibl *ibl, uint32_t flags, VEC pos, VEC n, VEC V)
{
    VEC res = FN(vs)(RC(0));
    if (flags & TINA_IBL) {
        res = FN(run_material)(m, pv, FN(M_IBL), n, n, V, ibl);
        res = FN(vadd)(res, FN(run_material)(m, pv, FN(M_EMI), n, n, V, ibl));
    } else {
        res = FN(vadd)(res, FN(run_material)(m, pv, FN(M_EMI), n, n, V, ibl));
        VEC amb = FN(v3)(lt->ambient[0], lt->ambient[1], lt->ambient[2]);
        res = FN(vadd)(res, FN(vmul)(amb, FN(run_material)(m, pv, FN(M_AMB), n, n, V, ibl)));
    }
    if (flags & TINA_NOLIGHT) return res;
    for (int l = 0; l < lt->n; l++) {
        VEC lxyz = FN(v3)(lt->dirs[l][0], lt->dirs[l][1], lt->dirs[l][2]);
        REAL lw = lt->dirs[l][3];
        VEC ldir = FN(vsub)(lxyz, FN(v3)(pos.x * lw, pos.y * lw, pos.z * lw));
        REAL dist = FN(r_sqrt)(FN(vdot)(ldir, ldir));
        ldir = FN(vdivs)(ldir, dist);
        REAL cos_i = FN(vdot)(n, ldir);
        if (cos_i > RC(0)) {
            VEC lcolor = FN(v3)(lt->colors[l][0], lt->colors[l][1], lt->colors[l][2]);
            lcolor = FN(vdivs)(lcolor, dist * dist);
            VEC mcolor = FN(run_material)(m, pv, FN(M_BRDF), n, ldir, V, ibl);
            res = FN(vadd)(res, FN(vmul)(FN(vsmul)(cos_i, lcolor), mcolor));
        }
    }
    return res;
}

/* ACES (tina/advans.py:31-34): c * (2.51c + 0.03) / (c * (2.43c + 0.59) + 0.14) */
static inline REAL FN(aces1)(REAL c)
{
    return (c * ((RC(2.51) * c) + RC(0.03))) / ((c * ((RC(2.43) * c) + RC(0.59))) + RC(0.14));
}

static inline VEC FN(load3)(uint64_t base, long idx)
{
    const float *p = (const float *)(uintptr_t)base + idx * 3;
    return FN(v3)(p[0], p[1], p[2]);
}

/* Parameter node values for one pixel (tina/matr/nodes.py:42-111,
 * FresnelFactor material.py:68-83).  Slots are in dependency order. */
static void FN(eval_params)(const tina_material *m, long f, VEC pos, VEC nrm, VEC tc,
                            VEC color, VEC *pv)
{
    for (int i = 0; i < m->nparams; i++) {
        const tina_param *p = &m->params[i];
        VEC v = FN(vs)(RC(0));
        switch (p->kind) {
        case TINA_P_CONST:
            v = p->dim == 1 ? FN(vs)(p->value[0]) : FN(v3)(p->value[0], p->value[1], p->value[2]);
            break;
        case TINA_P_PARAM: {
            const float *q = (const float *)(uintptr_t)p->ptr;
            v = p->dim == 1 ? FN(vs)(q[0]) : FN(v3)(q[0], q[1], q[2]);
            break;
        }
        case TINA_P_COLOR: v = color; break;
        case TINA_P_POS: v = pos; break;
        case TINA_P_NORMAL: v = nrm; break;
        case TINA_P_TEXCOORD: v = tc; break;
        case TINA_P_TEXTURE: {
            VEC t = pv[p->a];
            REAL cx = t.x * (REAL)(p->tex_w - 1), cy = t.y * (REAL)(p->tex_h - 1);
            v = FN(bilerp)((const float *)(uintptr_t)p->ptr, p->tex_w, p->tex_h, p->tex_c, cx, cy);
            break;
        }
        case TINA_P_FACE: {
            const float *q = (const float *)(uintptr_t)p->ptr;
            v = p->dim == 1 ? FN(vs)(q[f]) : FN(v3)(q[3 * f], q[3 * f + 1], q[3 * f + 2]);
            break;
        }
        case TINA_P_FRESNEL: {
            /* metallic * albedo + (1 - metallic) * 0.16 * specular**2 */
            VEC mt = pv[p->a], al = pv[p->b], sp = pv[p->c];
            VEC t1 = FN(vmul)(mt, al);
            VEC t2 = FN(vmul)(FN(vsmul)(RC(0.16), FN(vsub)(FN(vs)(RC(1)), mt)), FN(vmul)(sp, sp));
            v = FN(vadd)(t1, t2);
            break;
        }
        default: break;
        }
        pv[i] = v;
    }
}

void FN(orc_shade)(int W, int H, const tina_camera *cam, const tina_mesh *mesh,
                   const tina_material *mat, const tina_lights *lights,
                   const tina_ibl *ibl, const int32_t *occup, REAL *img,
                   const REAL bg[3], uint32_t flags, int nthreads)
{
    const long npix = (long)W * H;
    const REAL resx = (REAL)W, resy = (REAL)H;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1024)
    for (long P = 0; P < npix; P++) {
        int32_t id = occup[P];
        int px = (int)(P % W), py = (int)(P / W);
        REAL p_x = (REAL)px + (REAL)cam->bias[0];
        REAL p_y = (REAL)py + (REAL)cam->bias[1];
        if (id < 0) {
            if (flags & TINA_BG) {
                VEC c = FN(v3)(bg[0], bg[1], bg[2]);
                if (flags & TINA_IBL) {       /* Engine.render_background (engine.py:41-49) */
                    REAL ux = ((p_x / resx) * RC(2)) - RC(1), uy = ((p_y / resy) * RC(2)) - RC(1);
                    VEC ro = FN(mapply_pos)(cam->V2W, FN(v3)(ux, uy, RC(-1)), NULL);
                    VEC ro1 = FN(mapply_pos)(cam->V2W, FN(v3)(ux, uy, RC(1)), NULL);
                    VEC rd = FN(vnormalized)(FN(vsub)(ro1, ro));
                    c = FN(skybox_sample)((const float *)(uintptr_t)ibl->spec[0], ibl->spec_w[0], ibl->spec_h[0], rd);
                }
                if (flags & TINA_TONEMAP) c = FN(v3)(FN(aces1)(c.x), FN(aces1)(c.y), FN(aces1)(c.z));
                img[3 * P] = c.x; img[3 * P + 1] = c.y; img[3 * P + 2] = c.z;
            }
            continue;
        }
        uint32_t fu = (uint32_t)id - mesh->id_base;
        if (fu >= (uint32_t)mesh->nfaces) continue;
        long f = (long)fu;

        VEC pos, nrm, tc, color;
        if (flags & TINA_PARTICLES) {
            /* ParticleRaster.render_color (particle.py:125-146): the front-hemisphere
             * point behind the pixel on the sphere of centre Al, radius Rl */
            VEC Al = FN(load3)(mesh->verts, f);
            REAL Rl = (REAL)((const float *)(uintptr_t)mesh->norms)[f];
            VEC Av = FN(mapply_pos)(cam->W2V, Al, NULL);
            REAL vx = ((p_x / resx) * RC(2)) - RC(1), vy = ((p_y / resy) * RC(2)) - RC(1);
            VEC Pl = FN(mapply_pos)(cam->V2W, FN(v3)(vx, vy, Av.z), NULL);
            VEC Dl = FN(vdivs)(FN(vsub)(Pl, Al), Rl);
            VEC Zl = FN(v3)((REAL)cam->V2W[3] * RC(0), (REAL)cam->V2W[7] * RC(0), (REAL)cam->V2W[11] * RC(0));
            Zl = FN(v3)(Zl.x + (REAL)cam->V2W[0] * RC(0), Zl.y + (REAL)cam->V2W[4] * RC(0), Zl.z + (REAL)cam->V2W[8] * RC(0));
            Zl = FN(v3)(Zl.x + (REAL)cam->V2W[1] * RC(0), Zl.y + (REAL)cam->V2W[5] * RC(0), Zl.z + (REAL)cam->V2W[9] * RC(0));
            Zl = FN(v3)(Zl.x + (REAL)cam->V2W[2] * RC(1), Zl.y + (REAL)cam->V2W[6] * RC(1), Zl.z + (REAL)cam->V2W[10] * RC(1));
            Zl = FN(vnormalized)(Zl);                       /* mapply_dir(V2W, (0, 0, 1)).normalized() */
            Dl = FN(vsub)(Dl, FN(vsmul)(FN(r_sqrt)(RC(1) - FN(vdot)(Dl, Dl)), Zl));
            Dl = FN(vnormalized)(Dl);
            nrm = Dl;
            pos = FN(vadd)(Al, FN(vsmul)(Rl, Dl));
            tc = FN(vs)(RC(0));
            color = (flags & TINA_VCOLOR) && mesh->colors ? FN(load3)(mesh->colors, f) : FN(vs)(RC(1));
        } else {
            /* render_color (triangle.py:135-151): the setup of :93-113 recomputed from
             * the vertices, then the same weights as the coverage pass. */
            VEC A = FN(load3)(mesh->verts, 3 * f), B = FN(load3)(mesh->verts, 3 * f + 1),
                C = FN(load3)(mesh->verts, 3 * f + 2);
            REAL wA, wB, wC;
            VEC Bv = FN(mapply_pos)(cam->W2V, B, &wB);
            VEC Cv = FN(mapply_pos)(cam->W2V, C, &wC);
            VEC Av = FN(mapply_pos)(cam->W2V, A, &wA);
            REAL ax = (Av.x * RC(0.5) + RC(0.5)) * resx, ay = (Av.y * RC(0.5) + RC(0.5)) * resy;
            REAL bx = (Bv.x * RC(0.5) + RC(0.5)) * resx, by = (Bv.y * RC(0.5) + RC(0.5)) * resy;
            REAL cx = (Cv.x * RC(0.5) + RC(0.5)) * resx, cy = (Cv.y * RC(0.5) + RC(0.5)) * resy;
            REAL nn = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
            REAL bcnx = (bx - cx) / nn, bcny = (by - cy) / nn;
            REAL canx = (cx - ax) / nn, cany = (cy - ay) / nn;
            REAL ws0 = RC(1) / wA, ws1 = RC(1) / wB, ws2 = RC(1) / wC;
            REAL w_bc = (p_x - bx) * bcny - (p_y - by) * bcnx;
            REAL w_ca = (p_x - cx) * cany - (p_y - cy) * canx;
            REAL t = (RC(1) - w_bc) - w_ca;
            REAL wx = w_bc * ws0, wy = w_ca * ws1, wz = t * ws2;
            REAL s = (wx + wy) + wz;
            wx = wx / s; wy = wy / s; wz = wz / s;

            /* interpolate (triangle.py:31-49) */
            pos = FN(vadd)(FN(vadd)(FN(vsmul)(wx, A), FN(vsmul)(wy, B)), FN(vsmul)(wz, C));
            if ((flags & TINA_SMOOTH) && mesh->norms) {
                VEC An = FN(load3)(mesh->norms, 3 * f), Bn = FN(load3)(mesh->norms, 3 * f + 1),
                    Cn = FN(load3)(mesh->norms, 3 * f + 2);
                nrm = FN(vadd)(FN(vadd)(FN(vsmul)(wx, An), FN(vsmul)(wy, Bn)), FN(vsmul)(wz, Cn));
            } else {
                nrm = FN(vcross)(FN(vsub)(B, A), FN(vsub)(C, A));
            }
            nrm = FN(vnormalized)(nrm);
            tc = FN(vs)(RC(0));
            if ((flags & TINA_TEX) && mesh->coors) {
                const float *q = (const float *)(uintptr_t)mesh->coors + 6 * f;
                REAL tx = ((wx * (REAL)q[0]) + (wy * (REAL)q[2])) + (wz * (REAL)q[4]);
                REAL ty = ((wx * (REAL)q[1]) + (wy * (REAL)q[3])) + (wz * (REAL)q[5]);
                tc = FN(v3)(tx, ty, RC(0));
            }
            color = FN(vs)(RC(1));
            if ((flags & TINA_VCOLOR) && mesh->colors) {
                VEC Ac = FN(load3)(mesh->colors, 3 * f), Bc = FN(load3)(mesh->colors, 3 * f + 1),
                    Cc = FN(load3)(mesh->colors, 3 * f + 2);
                color = FN(vadd)(FN(vadd)(FN(vsmul)(wx, Ac), FN(vsmul)(wy, Bc)), FN(vsmul)(wz, Cc));
            }
        }

        /* calc_viewdir (shader.py:81-93) */
        REAL ux = ((p_x / resx) * RC(2)) - RC(1), uy = ((p_y / resy) * RC(2)) - RC(1);
        VEC ro = FN(mapply_pos)(cam->V2W, FN(v3)(ux, uy, RC(-1)), NULL);
        VEC ro1 = FN(mapply_pos)(cam->V2W, FN(v3)(ux, uy, RC(1)), NULL);
        VEC viewdir = FN(vneg)(FN(vnormalized)(FN(vsub)(ro1, ro)));

        VEC pv[TINA_MAX_PARAMS];
        FN(eval_params)(mat, f, pos, nrm, tc, color, pv);
        VEC res = FN(lighting)(mat, pv, lights, ibl, flags, pos, nrm, viewdir);
        if (flags & TINA_TONEMAP) res = FN(v3)(FN(aces1)(res.x), FN(aces1)(res.y), FN(aces1)(res.z));
        img[3 * P] = res.x; img[3 * P + 1] = res.y; img[3 * P + 2] = res.z;
    }
}

void FN(orc_aces)(REAL *img, int npixels)
{
    for (long i = 0; i < 3L * npixels; i++) img[i] = FN(aces1)(img[i]);
}

void FN(orc_shade_point)(const tina_material *mat, const tina_lights *lights,
                         const tina_ibl *ibl, uint32_t flags, const REAL pos[3],
                         const REAL nrm[3], const REAL viewdir[3],
                         const REAL *slot_values, REAL out[3])
{
    VEC pv[TINA_MAX_PARAMS];
    VEC P = FN(v3)(pos[0], pos[1], pos[2]);
    VEC N = FN(v3)(nrm[0], nrm[1], nrm[2]);
    VEC V = FN(v3)(viewdir[0], viewdir[1], viewdir[2]);
    if (slot_values) {
        for (int i = 0; i < mat->nparams; i++)
            pv[i] = FN(v3)(slot_values[3 * i], slot_values[3 * i + 1], slot_values[3 * i + 2]);
    } else {
        FN(eval_params)(mat, 0, P, N, FN(vs)(RC(0)), FN(vs)(RC(1)), pv);
    }
    VEC r = FN(lighting)(mat, pv, lights, ibl, flags, P, N, V);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

#undef VEC
#undef FN
