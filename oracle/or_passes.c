/*
 * or_passes.c -- CPU restatement of the voxel cone tracing passes.
 * TEST INFRASTRUCTURE ONLY (see vct_oracle.h).
 *
 * Rasterization (shared by the shadow map, the voxelizer and the G-buffer)
 * follows the GL 4.5 core rules, restated:
 *   - viewport transform x_w = (w/2) x_d + w/2, y_w likewise (y up);
 *   - window coordinates snapped to 8 sub-pixel bits (1/256 px, round to
 *     nearest even), the precision Mesa's rasterizer uses;
 *   - a sample is inside when all three integer edge functions are > 0, or
 *     == 0 on a top or left edge (D3D/GL top-left rule);
 *   - zero-area triangles produce no fragments;
 *   - 4x multisample positions (0.375,0.125) (0.875,0.375) (0.125,0.625)
 *     (0.625,0.875) within the pixel (the standard 4x pattern); single-sample
 *     rasterization samples the pixel centre;
 *   - a fragment exists when any sample is covered; its varyings are evaluated
 *     at the pixel centre (no centroid qualifier in voxelize.frag), with
 *     barycentrics lambda_i = E_i(centre) / area in float.
 * Depth clipping is per sample (z_ndc in [-1, 1]); the camera pass clips
 * triangles against the near plane geometrically first.
 */
#include "vct_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small float helpers, fixed evaluation order, no contraction               */
static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static void nrm3(const float v[3], float o[3]) {
    float l = sqrtf(dot3(v, v));
    o[0] = v[0] / l;
    o[1] = v[1] / l;
    o[2] = v[2] / l;
}
static void cross3(const float a[3], const float b[3], float o[3]) {
    float t0 = a[1] * b[2] - a[2] * b[1];
    float t1 = a[2] * b[0] - a[0] * b[2];
    float t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
}
static void mat3_vec3(const float m[9], const float v[3], float o[3]) {
    float t[3];
    for (int r = 0; r < 3; ++r) t[r] = (m[0 * 3 + r] * v[0] + m[1 * 3 + r] * v[1]) + m[2 * 3 + r] * v[2];
    o[0] = t[0];
    o[1] = t[1];
    o[2] = t[2];
}
static float interp(const float l[3], float a0, float a1, float a2) { return (l[0] * a0 + l[1] * a1) + l[2] * a2; }
static float clampf(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }

/* ------------------------------------------------------------------------ */
/* raster setup                                                               */
typedef struct tri_setup {
    int64_t A[3], B[3], C[3];
    int64_t bias[3];
    int64_t area;
    int x0, y0, x1, y1; /* inclusive pixel bbox clipped to the viewport */
} tri_setup;

static const int MSAA_X[4] = {96, 224, 32, 160};
static const int MSAA_Y[4] = {32, 96, 160, 224};

static float clamp_win(float v) { return v < -524288.0f ? -524288.0f : (v > 524288.0f ? 524288.0f : v); }

/* win: window coordinates (x, y) of the three vertices.  cull: 1 culls
 * clockwise (back-facing with CCW front faces).  Returns 0 when the triangle
 * produces no fragments. */
static int tri_setup_fn(const float win[3][2], int W, int H, int cull, tri_setup *s) {
    int64_t X[3], Y[3];
    for (int i = 0; i < 3; ++i) {
        X[i] = (int64_t)rintf(clamp_win(win[i][0]) * 256.0f);
        Y[i] = (int64_t)rintf(clamp_win(win[i][1]) * 256.0f);
    }
    s->A[0] = Y[1] - Y[2];
    s->B[0] = X[2] - X[1];
    s->C[0] = X[1] * Y[2] - X[2] * Y[1];
    s->A[1] = Y[2] - Y[0];
    s->B[1] = X[0] - X[2];
    s->C[1] = X[2] * Y[0] - X[0] * Y[2];
    s->A[2] = Y[0] - Y[1];
    s->B[2] = X[1] - X[0];
    s->C[2] = X[0] * Y[1] - X[1] * Y[0];
    int64_t area = s->C[0] + s->C[1] + s->C[2];
    if (area == 0) return 0;
    if (area < 0) {
        if (cull) return 0;
        for (int i = 0; i < 3; ++i) {
            s->A[i] = -s->A[i];
            s->B[i] = -s->B[i];
            s->C[i] = -s->C[i];
        }
        area = -area;
    }
    s->area = area;
    for (int i = 0; i < 3; ++i) {
        int tl = (s->A[i] > 0) || (s->A[i] == 0 && s->B[i] < 0);
        s->bias[i] = tl ? 0 : -1;
    }
    int64_t mnx = X[0], mxx = X[0], mny = Y[0], mxy = Y[0];
    for (int i = 1; i < 3; ++i) {
        if (X[i] < mnx) mnx = X[i];
        if (X[i] > mxx) mxx = X[i];
        if (Y[i] < mny) mny = Y[i];
        if (Y[i] > mxy) mxy = Y[i];
    }
    int64_t x0 = mnx >> 8, x1 = mxx >> 8, y0 = mny >> 8, y1 = mxy >> 8;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > W - 1) x1 = W - 1;
    if (y1 > H - 1) y1 = H - 1;
    if (x0 > x1 || y0 > y1) return 0;
    s->x0 = (int)x0;
    s->x1 = (int)x1;
    s->y0 = (int)y0;
    s->y1 = (int)y1;
    return 1;
}

static int inside(const tri_setup *s, int64_t sx, int64_t sy) {
    for (int i = 0; i < 3; ++i)
        if (s->A[i] * sx + s->B[i] * sy + s->C[i] + s->bias[i] < 0) return 0;
    return 1;
}

static void bary_at(const tri_setup *s, int64_t sx, int64_t sy, float l[3]) {
    float fa = (float)s->area;
    for (int i = 0; i < 3; ++i) l[i] = (float)(s->A[i] * sx + s->B[i] * sy + s->C[i]) / fa;
}

/* ------------------------------------------------------------------------ */
/* texture helpers                                                            */

/* GL LINEAR sample of an fp32 depth texture with CLAMP_TO_BORDER (border 1.0)
 * and an integer texel offset (textureOffset).  Shadow maps are stored
 * row-major with row 0 at t = 0. */
static float shadow_bilinear(const float *sm, int N, float s, float t, int ox, int oy) {
    float u = s * (float)N;
    u = u + (float)ox;
    u = u - 0.5f;
    float v = t * (float)N;
    v = v + (float)oy;
    v = v - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    int i0 = (int)fu, j0 = (int)fv;
    float a = u - fu, b = v - fv;
    float t00, t10, t01, t11;
#define TEX(i, j) (((i) >= 0 && (i) < N && (j) >= 0 && (j) < N) ? sm[(size_t)(j) * N + (i)] : 1.0f)
    t00 = TEX(i0, j0);
    t10 = TEX(i0 + 1, j0);
    t01 = TEX(i0, j0 + 1);
    t11 = TEX(i0 + 1, j0 + 1);
#undef TEX
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return ((w00 * t00 + w10 * t10) + w01 * t01) + w11 * t11;
}

/* calcShadowFactor (shaders/voxelize.frag:160-184; the same function is
 * phong.frag:183-207) */
float or_calc_shadow_factor(const float lsp[4], const float *sm, int N) {
    float sh[3];
    for (int k = 0; k < 3; ++k) sh[k] = (lsp[k] / lsp[3] + 1.0f) * 0.5f;
    float frag = sh[2] - 0.01f;
    if (frag > 1.0f) return 0.0f;
    static const int OX[5] = {0, 1, 0, -1, 0}, OY[5] = {0, 0, 1, 0, -1};
    float f = 0.0f;
    for (int i = 0; i < 5; ++i)
        if (frag > shadow_bilinear(sm, N, sh[0], sh[1], OX[i], OY[i])) f += 1.0f;
    f /= 5.0f;
    return f;
}

/* ------------------------------------------------------------------------ */
/* diffuse maps: GL 4.5 core 8.14 with LINEAR_MIPMAP_NEAREST / NEAREST / REPEAT
 * (ResourceLoader.h:79-82, GLHelper.cpp:180-183) */
static int tex_wrap(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}
static int tex_floor(float x) { /* clamped to +-2^30: whole numbers there */
    float f = floorf(x);
    f = f < -1073741824.0f ? -1073741824.0f : (f > 1073741824.0f ? 1073741824.0f : f);
    return (int)f;
}
static void tex_unorm(uint32_t v, float o[4]) {
    for (int k = 0; k < 4; ++k) o[k] = (float)((v >> (8 * k)) & 0xffu) / 255.0f;
}

void or_tex_sample(const or_texture *tex, int ntex, int id, float s, float t, float rho2, float out[4]) {
    out[0] = out[1] = out[2] = 0.0f;
    out[3] = 1.0f;
    if (id < 0 || id >= ntex) return; /* an unbound unit */
    const or_texture *T = &tex[id];
    int q = 0;
    int mx = T->width > T->height ? T->width : T->height;
    while ((mx >> (q + 1)) > 0) ++q;
    if (T->levels != q + 1) return; /* mipmap-incomplete (8.17) */
    if (!(rho2 > 1.0f)) {           /* lambda <= 0: magnification, NEAREST on level 0 */
        int i = tex_wrap(tex_floor(s * (float)T->width), T->width);
        int j = tex_wrap(tex_floor(t * (float)T->height), T->height);
        tex_unorm(T->texels[(size_t)j * T->width + i], out);
        return;
    }
    /* level d = ceil(lambda + 1/2) - 1 (0 for lambda <= 1/2), clamped to q:
     * lambda > k - 1/2 is rho^2 > 2^(2k-1) */
    int d = 0;
    for (int k = 1; k <= q; ++k)
        if (rho2 > ldexpf(1.0f, 2 * k - 1)) d = k;
    size_t off = 0;
    for (int l = 0; l < d; ++l) {
        int w = T->width >> l, h = T->height >> l;
        off += (size_t)(w < 1 ? 1 : w) * (size_t)(h < 1 ? 1 : h);
    }
    int W = T->width >> d, H = T->height >> d;
    W = W < 1 ? 1 : W;
    H = H < 1 ? 1 : H;
    const uint32_t *L = T->texels + off;
    float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    float a = u - floorf(u), b = v - floorf(v);
    int i0 = tex_floor(u), j0 = tex_floor(v);
    int ia = tex_wrap(i0, W), ib = tex_wrap(i0 + 1, W), ja = tex_wrap(j0, H), jb = tex_wrap(j0 + 1, H);
    float t00[4], t10[4], t01[4], t11[4];
    tex_unorm(L[(size_t)ja * W + ia], t00);
    tex_unorm(L[(size_t)ja * W + ib], t10);
    tex_unorm(L[(size_t)jb * W + ia], t01);
    tex_unorm(L[(size_t)jb * W + ib], t11);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    for (int k = 0; k < 4; ++k) out[k] = ((w00 * t00[k] + w10 * t10[k]) + w01 * t01[k]) + w11 * t11[k];
}

float or_tex_rho2_affine(const float win[3][2], const float st[6], float W, float H) {
    float e1x = win[1][0] - win[0][0], e1y = win[1][1] - win[0][1];
    float e2x = win[2][0] - win[0][0], e2y = win[2][1] - win[0][1];
    float den = e1x * e2y - e2x * e1y;
    float du1 = (st[2] - st[0]) * W, du2 = (st[4] - st[0]) * W;
    float dv1 = (st[3] - st[1]) * H, dv2 = (st[5] - st[1]) * H;
    float dudx = (du1 * e2y - du2 * e1y) / den, dvdx = (dv1 * e2y - dv2 * e1y) / den;
    float dudy = (du2 * e1x - du1 * e2x) / den, dvdy = (dv2 * e1x - dv1 * e2x) / den;
    float x = dudx * dudx + dvdx * dvdx, y = dudy * dudy + dvdy * dvdy;
    return fmaxf(x, y);
}

/* the triangle's diffuse map index, -1 for none */
static int tri_tex(const or_scene *sc, int t) { return sc->texid ? sc->texid[t] : -1; }

static void world_vertex(const or_scene *sc, int t, int v, float w[3]) {
    const float *M = sc->model + (size_t)sc->obj[t] * 16;
    const float *p = sc->pos + (size_t)t * 9 + v * 3;
    float p4[4] = {p[0], p[1], p[2], 1.0f}, o[4];
    or_mat4_vec4(M, p4, o);
    w[0] = o[0];
    w[1] = o[1];
    w[2] = o[2];
}

/* ------------------------------------------------------------------------ */
/* shadow map: simple.vert (gl_Position = ls * (model * p)), depth only        */
void or_shadowmap(const or_scene *sc, const or_light *L, float *depth) {
    int N = (int)L->shadow_size;
    for (size_t i = 0; i < (size_t)N * N; ++i) depth[i] = 1.0f;
    float half = (float)N * 0.5f;
    /* triangles on all host threads: depth z in [0, 1] (+0 for -0), so the
     * order of the bit patterns is the order of the depths and the depth test
     * is an atomic minimum on them -- independent of the thread schedule */
    uint32_t *dbits = (uint32_t *)depth;
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < sc->ntri; ++t) {
        float clip[3][4];
        int behind = 0;
        for (int v = 0; v < 3; ++v) {
            float w[3];
            world_vertex(sc, t, v, w);
            float w4[4] = {w[0], w[1], w[2], 1.0f};
            or_mat4_vec4(L->ls, w4, clip[v]);
            if (clip[v][2] < -clip[v][3] || clip[v][3] <= 0.0f) behind = 1;
        }
        if (behind) continue; /* triangles crossing the light's near plane are dropped */
        float win[3][2], zw[3];
        for (int v = 0; v < 3; ++v) {
            float xd = clip[v][0] / clip[v][3], yd = clip[v][1] / clip[v][3], zd = clip[v][2] / clip[v][3];
            win[v][0] = xd * half + half;
            win[v][1] = yd * half + half;
            zw[v] = zd * 0.5f + 0.5f;
        }
        tri_setup s;
        /* GL_CULL_FACE (GL_BACK, CCW front) is still enabled from the previous
         * frame's passes when the shadow map is drawn (Application.cpp:737,942,979) */
        if (!tri_setup_fn(win, N, N, 1, &s)) continue;
        for (int py = s.y0; py <= s.y1; ++py)
            for (int px = s.x0; px <= s.x1; ++px) {
                int64_t cx = (int64_t)px * 256 + 128, cy = (int64_t)py * 256 + 128;
                if (!inside(&s, cx, cy)) continue;
                float l[3];
                bary_at(&s, cx, cy, l);
                float z = interp(l, zw[0], zw[1], zw[2]);
                z = z + 0.0f; /* -0 -> +0 */
                if (!(z >= 0.0f && z <= 1.0f)) continue;
                uint32_t zb;
                memcpy(&zb, &z, 4);
                uint32_t *d = &dbits[(size_t)py * N + px];
                uint32_t cur = __atomic_load_n(d, __ATOMIC_RELAXED);
                while (zb < cur && !__atomic_compare_exchange_n(d, &cur, zb, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                }
            }
    }
}

/* ------------------------------------------------------------------------ */
/* voxelize (GS path)                                                         */
/* The voxelize loop runs over triangles on all host threads (OpenMP); the
 * stores are the order-independent imageAtomicMax / imageAtomicOr of the
 * shader, done here with compare-and-swap, so the result does not depend on
 * the thread count or schedule. */
static void atomic_max_u32(uint32_t *p, uint32_t v) {
    uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

/* imageAtomicRGBA8Avg (voxelize.frag:111-139, the live `#if 1` branch): the
 * word's rgb bytes hold the running average of 255 c, its alpha byte the
 * count; convVec4ToRGBA8 truncates each component to an integer and keeps its
 * low byte.  Applied in fragment order (the oracle runs this mode on one
 * thread). */
static uint32_t rgba8_from_floats(float x, float y, float z, float w) {
    return (((uint32_t)w & 0xffu) << 24) | (((uint32_t)z & 0xffu) << 16) | (((uint32_t)y & 0xffu) << 8) |
           ((uint32_t)x & 0xffu);
}
static void rgba8_avg(uint32_t *p, const float c[3]) {
    const float v[4] = {c[0] * 255.0f, c[1] * 255.0f, c[2] * 255.0f, 1.0f};
    const uint32_t cur = *p;
    if (cur == 0u) { /* the first compare-and-swap against 0 succeeds */
        *p = rgba8_from_floats(v[0], v[1], v[2], v[3]);
        return;
    }
    float r[4] = {(float)(cur & 0xffu), (float)((cur >> 8) & 0xffu), (float)((cur >> 16) & 0xffu),
                  (float)(cur >> 24)};
    for (int k = 0; k < 3; ++k) r[k] = r[k] * r[3];
    float sm[4];
    for (int k = 0; k < 4; ++k) sm[k] = r[k] + v[k];
    for (int k = 0; k < 3; ++k) sm[k] = sm[k] / sm[3];
    *p = rgba8_from_floats(sm[0], sm[1], sm[2], sm[3]);
}

/* the sum-accumulate mode's fixed point: round(65535 clamp(c, 0, 1)) */
static uint32_t fix16(float c) { return (uint32_t)rintf(clampf(c, 0.0f, 1.0f) * 65535.0f); }

static void light_fragment(const or_light *lights, int nlights, const float *const *shadowmaps,
                           const float color[3], const float nrm_interp[3], const float wpos[3],
                           float out[3]) {
    float fin[3] = {0, 0, 0};
    float nl[3];
    nrm3(nrm_interp, nl);
    for (int i = 0; i < nlights; ++i) {
        const or_light *L = &lights[i];
        if (!L->enabled) continue;
        float lit[3] = {0, 0, 0};
        if (L->type == 0u) { /* POINT_LIGHT */
            float d[3] = {L->position[0] - wpos[0], L->position[1] - wpos[1], L->position[2] - wpos[2]};
            float dn[3];
            nrm3(d, dn);
            float ndl = fmaxf(0.0f, dot3(nl, dn));
            for (int k = 0; k < 3; ++k) lit[k] = ((color[k] * L->intensity) * L->color[k]) * ndl;
            float dist = sqrtf(dot3(d, d));
            float e0 = 0.75f * L->range, e1 = L->range;
            float tt = clampf((dist - e0) / (e1 - e0), 0.0f, 1.0f);
            float ss = (tt * tt) * (3.0f - 2.0f * tt);
            float att = 1.0f - ss;
            for (int k = 0; k < 3; ++k) lit[k] = lit[k] * att;
        } else { /* DIRECTIONAL_LIGHT */
            float nd[3] = {-L->direction[0], -L->direction[1], -L->direction[2]}, ndn[3];
            nrm3(nd, ndn);
            float ndl = fmaxf(0.0f, dot3(nl, ndn));
            for (int k = 0; k < 3; ++k) lit[k] = ((color[k] * L->color[k]) * L->intensity) * ndl;
        }
        if (L->shadow_caster && shadowmaps && shadowmaps[i]) {
            float w4[4] = {wpos[0], wpos[1], wpos[2], 1.0f}, lsp[4];
            or_mat4_vec4(L->ls, w4, lsp);
            float sf = 1.0f - or_calc_shadow_factor(lsp, shadowmaps[i], (int)L->shadow_size);
            for (int k = 0; k < 3; ++k) lit[k] = lit[k] * sf;
        }
        for (int k = 0; k < 3; ++k) fin[k] = fin[k] + lit[k];
    }
    for (int k = 0; k < 3; ++k) out[k] = clampf(fin[k], 0.0f, 1.0f);
}

/* occ != NULL: the occupancy mode of voxelize.frag:187-193 (imageAtomicOr 1 at
 * the unwarped position, nothing else). */
static void voxelize_impl(const or_scene *sc, const or_light *lights, int nlights, const float *const *shadowmaps,
                          const or_voxelize_params *p, uint32_t *albedo, uint32_t *normal, or_stats *stats,
                          uint32_t *occ) {
    int dim = p->dim;
    float half = (float)dim * 0.5f;
    uint32_t nfrag = 0;
    const int serial = !occ && p->accum == 0; /* the running average depends on fragment order */
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nfrag) if (!serial)
    for (int t = 0; t < sc->ntri; ++t) {
        const float *Nm = sc->nmat + (size_t)sc->obj[t] * 9;
        float W[3][3], Nw[3][3];
        for (int v = 0; v < 3; ++v) {
            world_vertex(sc, t, v, W[v]);
            mat3_vec3(Nm, sc->nrm + (size_t)t * 9 + v * 3, Nw[v]);
        }
        /* voxelize.geom:51-81 */
        float fs[3], fn[3];
        for (int k = 0; k < 3; ++k) fs[k] = (Nw[0][k] + Nw[1][k]) + Nw[2][k];
        nrm3(fs, fn);
        float an[3] = {fabsf(fn[0]), fabsf(fn[1]), fabsf(fn[2])};
        int axis;
        if (an[0] > an[1] && an[0] > an[2]) axis = 0;
        else if (an[1] > an[0] && an[1] > an[2]) axis = 1;
        else axis = 2;
        if (p->axis_override >= 0 && p->axis_override <= 2) axis = p->axis_override;
        const float *mvp = axis == 0 ? p->mvp_x : (axis == 1 ? p->mvp_y : p->mvp_z);
        float ndc[3][3], win[3][2];
        int needs_zclip = 0;
        for (int v = 0; v < 3; ++v) {
            float w4[4] = {W[v][0], W[v][1], W[v][2], 1.0f}, c[4];
            or_mat4_vec4(mvp, w4, c);
            for (int k = 0; k < 3; ++k) ndc[v][k] = c[k] / c[3];
            if (ndc[v][2] < -1.0f || ndc[v][2] > 1.0f) needs_zclip = 1;
            win[v][0] = ndc[v][0] * half + half;
            win[v][1] = ndc[v][1] * half + half;
        }
        tri_setup s;
        if (!tri_setup_fn(win, dim, dim, 0, &s)) continue;
        float col[4];
        or_unpack_unorm4x8(sc->albedo[t], col);
        const int tid = tri_tex(sc, t);
        float rho2 = 0.0f; /* the orthographic views: one footprint per triangle */
        if (tid >= 0 && tid < sc->ntex)
            rho2 = or_tex_rho2_affine(win, sc->uv + (size_t)t * 6, (float)sc->tex[tid].width,
                                      (float)sc->tex[tid].height);
        int nsamp = p->msaa ? 4 : 1;
        for (int py = s.y0; py <= s.y1; ++py)
            for (int px = s.x0; px <= s.x1; ++px) {
                int cov = 0;
                for (int k = 0; k < nsamp && !cov; ++k) {
                    int64_t sx = (int64_t)px * 256 + (p->msaa ? MSAA_X[k] : 128);
                    int64_t sy = (int64_t)py * 256 + (p->msaa ? MSAA_Y[k] : 128);
                    if (!inside(&s, sx, sy)) continue;
                    if (needs_zclip) {
                        float ls[3];
                        bary_at(&s, sx, sy, ls);
                        float z = interp(ls, ndc[0][2], ndc[1][2], ndc[2][2]);
                        if (!(z >= -1.0f && z <= 1.0f)) continue;
                    }
                    cov = 1;
                }
                if (!cov) continue;
                int64_t cx = (int64_t)px * 256 + 128, cy = (int64_t)py * 256 + 128;
                float l[3];
                bary_at(&s, cx, cy, l);
                float pos[3], wp[3], nr[3];
                for (int k = 0; k < 3; ++k) {
                    pos[k] = interp(l, ndc[0][k], ndc[1][k], ndc[2][k]);
                    wp[k] = interp(l, W[0][k], W[1][k], W[2][k]);
                    nr[k] = interp(l, Nw[0][k], Nw[1][k], Nw[2][k]);
                }
                float color[3] = {col[0], col[1], col[2]};
                if (!occ && tid >= 0) { /* texture(diffuseMap, texcoord).rgb (voxelize.frag:197) */
                    const float *st = sc->uv + (size_t)t * 6;
                    float tc[4];
                    or_tex_sample(sc->tex, sc->ntex, tid, interp(l, st[0], st[2], st[4]), interp(l, st[1], st[3], st[5]),
                                  rho2, tc);
                    color[0] = tc[0];
                    color[1] = tc[1];
                    color[2] = tc[2];
                }
                float nn[3], nenc[4];
                if (!occ) {
                    nfrag += 1; /* voxelize.frag:195 */
                    nrm3(nr, nn);
                    for (int k = 0; k < 3; ++k) nenc[k] = (nn[k] + 1.0f) * 0.5f;
                    nenc[3] = 1.0f;
                    if (p->lighting) light_fragment(lights, nlights, shadowmaps, color, nr, wp, color);
                }
                /* getVoxelPosition (voxelize.frag:79-108) */
                float u[3];
                for (int k = 0; k < 3; ++k) u[k] = (pos[k] + 1.0f) * 0.5f;
                float un[3];
                if (axis == 0) {
                    un[0] = 1.0f - u[2];
                    un[1] = u[1];
                    un[2] = u[0];
                } else if (axis == 1) {
                    un[0] = u[0];
                    un[1] = 1.0f - u[2];
                    un[2] = u[1];
                } else {
                    un[0] = u[0];
                    un[1] = u[1];
                    un[2] = u[2];
                }
                un[2] = 1.0f - un[2];
                if (!occ && p->warpmap) or_warp_sample(p->warpmap, un, un); /* warpTexture (:103-105) */
                int idx[3];
                for (int k = 0; k < 3; ++k) idx[k] = (int)((float)dim * un[k]);
                if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0 || idx[0] >= dim || idx[1] >= dim || idx[2] >= dim)
                    continue; /* out-of-range image atomics are no-ops */
                size_t vi = ((size_t)idx[2] * dim + idx[1]) * dim + idx[0];
                if (occ) {
                    __atomic_fetch_or(&occ[vi], 1u, __ATOMIC_RELAXED);
                    continue;
                }
                if (p->accum == 0) {
                    rgba8_avg(&albedo[vi], color);
                    rgba8_avg(&normal[vi], nenc);
                    continue;
                }
                if (p->accum == 2) {
                    uint32_t *w = p->sums + vi * 8;
                    const uint32_t add[8] = {fix16(color[0]), fix16(color[1]), fix16(color[2]), 1u,
                                             fix16(nenc[0]), fix16(nenc[1]), fix16(nenc[2]), 1u};
                    for (int k = 0; k < 8; ++k) __atomic_fetch_add(&w[k], add[k], __ATOMIC_RELAXED);
                    continue;
                }
                float c4[4] = {color[0], color[1], color[2], 1.0f};
                uint32_t pc = or_pack_unorm4x8(c4), pn = or_pack_unorm4x8(nenc);
                atomic_max_u32(&albedo[vi], pc);
                atomic_max_u32(&normal[vi], pn);
            }
    }
    if (stats) stats->total_voxel_fragments += nfrag;
}

void or_voxelize(const or_scene *sc, const or_light *lights, int nlights, const float *const *shadowmaps,
                 const or_voxelize_params *p, uint32_t *albedo, uint32_t *normal, or_stats *stats) {
    voxelize_impl(sc, lights, nlights, shadowmaps, p, albedo, normal, stats, NULL);
}

/* ------------------------------------------------------------------------ */
/* warp texture (src/Application.cpp:235-575)                                */
void or_warp_occupancy(const or_scene *sc, const or_voxelize_params *p, uint32_t *occ) {
    memset(occ, 0, sizeof(uint32_t) * (size_t)p->dim * p->dim * p->dim); /* glClearTexImage (:257) */
    voxelize_impl(sc, NULL, 0, NULL, p, NULL, NULL, NULL, occ);
}

/* generateWarpmap.frag:49-99 at every texel centre, with the inclusive
 * per-axis prefix sums (Application.cpp:312-344) and the low/high weights per
 * occupied-cell count (:346-370), stored as RGBA16 unorm. */
void or_warpmap(const uint32_t *occ, const or_warp_params *p, float *out) {
    enum { D = OR_WARP_DIM };
    static int part[D][D][D][3];
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < D; ++y) {
            int s = 0;
            for (int x = 0; x < D; ++x) part[z][y][x][0] = (s += occ[(z * D + y) * D + x] > 0 ? 1 : 0);
        }
    for (int z = 0; z < D; ++z)
        for (int x = 0; x < D; ++x) {
            int s = 0;
            for (int y = 0; y < D; ++y) part[z][y][x][1] = (s += occ[(z * D + y) * D + x] > 0 ? 1 : 0);
        }
    for (int y = 0; y < D; ++y)
        for (int x = 0; x < D; ++x) {
            int s = 0;
            for (int z = 0; z < D; ++z) part[z][y][x][2] = (s += occ[(z * D + y) * D + x] > 0 ? 1 : 0);
        }
    float wt[2][D + 1];
    for (int n = 0; n <= D; ++n) {
        if (n == 0 || n == D) {
            wt[0][n] = wt[1][n] = 1.0f;
            continue;
        }
        int empty = D - n;
        float h = p->high_resolution;
        float l = ((float)D - h * (float)n) / (float)empty;
        if (l < p->low_resolution) {
            l = p->low_resolution;
            h = ((float)D - l * (float)empty) / (float)n;
        }
        wt[0][n] = l;
        wt[1][n] = h;
    }
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < D; ++y)
            for (int x = 0; x < D; ++x) {
                float tc[3] = {((float)x + 0.5f) / (float)D, ((float)y + 0.5f) / (float)D, ((float)z + 0.5f) / (float)D};
                int o = occ[(z * D + y) * D + x] > 0;
                const int total[3] = {part[z][y][D - 1][0], part[z][D - 1][x][1], part[D - 1][y][x][2]};
                float *t = out + (((size_t)z * D + y) * D + x) * 4;
                for (int k = 0; k < 3; ++k) {
                    float lin = tc[k] * (float)D;
                    float cell = floorf(lin), inner = lin - cell; /* modf */
                    float lo = wt[0][total[k]], hi = wt[1][total[k]];
                    if (p->use_weights_texture) { /* through the rgba16f weights texture (:485-488) */
                        lo = or_half_to_float(or_float_to_half(lo));
                        hi = or_half_to_float(or_float_to_half(hi));
                    }
                    float res = o ? hi : lo;
                    float prev = (float)(o ? part[z][y][x][k] - 1 : part[z][y][x][k]);
                    float off = lo * (cell - prev) + hi * prev;
                    float inoff = inner * res;
                    float w = (off + inoff) / (float)D;
                    float v = (!p->linear && p->axes[k]) ? w : tc[k];
                    t[k] = (float)(int)(fminf(fmaxf(v, 0.0f), 1.0f) * 65535.0f + 0.5f) / 65535.0f;
                }
                /* w: packCellInfo (occupied counts and flag), as stored */
                uint32_t bits = ((uint32_t)total[0] & 0x1Fu) | (((uint32_t)total[1] & 0x1Fu) << 5) |
                                (((uint32_t)total[2] & 0x1Fu) << 10) | ((uint32_t)o << 15);
                t[3] = (float)bits / 65535.0f;
            }
}

/* texture(warpmap, tc) with LINEAR filtering and CLAMP_TO_EDGE (xyz) */
void or_warp_sample(const float *warpmap, const float tc[3], float out[3]) {
    enum { D = OR_WARP_DIM };
    int i0[3], i1[3];
    float f[3];
    for (int k = 0; k < 3; ++k) {
        float u = tc[k] * (float)D - 0.5f;
        float fl = floorf(u);
        f[k] = u - fl;
        int a = (int)fl, b = a + 1;
        i0[k] = a < 0 ? 0 : (a > D - 1 ? D - 1 : a);
        i1[k] = b < 0 ? 0 : (b > D - 1 ? D - 1 : b);
    }
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int c = 0; c < 8; ++c) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        float w = ((dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1])) * (dz ? f[2] : 1.0f - f[2]);
        int x = dx ? i1[0] : i0[0], y = dy ? i1[1] : i0[1], z = dz ? i1[2] : i0[2];
        const float *t = warpmap + (((size_t)z * D + y) * D + x) * 4;
        for (int k = 0; k < 3; ++k) acc[k] = acc[k] + w * t[k];
    }
    for (int k = 0; k < 3; ++k) out[k] = acc[k];
}

/* ------------------------------------------------------------------------ */
/* normalize (transferVoxels) + inject                                       */
static float q16(float x) { return or_half_to_float(or_float_to_half(x)); }

static void inject_litmask(int dim, const float center[3], const float vmin[3], const float vmax[3],
                           const or_light *lights, int nlights, const float *const *shadowmaps,
                           uint32_t light_mask, const float *warpmap, uint8_t *litmask) {
    size_t nv = (size_t)dim * dim * dim;
    memset(litmask, 0, nv);
    /* injectRadiance.comp:152-168 per shadow texel, as a lit-mask scatter */
    for (int li = 0; li < nlights; ++li) {
        const or_light *L = &lights[li];
        if (!((light_mask >> li) & 1u) || !L->shadow_caster || !shadowmaps || !shadowmaps[li]) continue;
        int N = (int)L->shadow_size;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                float tc[2] = {(float)i / (float)N, (float)j / (float)N};
                float d = shadow_bilinear(shadowmaps[li], N, tc[0], tc[1], 0, 0);
                float ndc4[4] = {tc[0] * 2.0f - 1.0f, tc[1] * 2.0f - 1.0f, d * 2.0f - 1.0f, 1.0f}, w[4];
                or_mat4_vec4(L->ls_inverse, ndc4, w);
                if (L->perspective) {
                    w[0] = w[0] / w[3];
                    w[1] = w[1] / w[3];
                    w[2] = w[2] / w[3];
                }
                int idx[3];
                float lin[3];
                for (int k = 0; k < 3; ++k) lin[k] = ((w[k] - center[k]) - vmin[k]) / (vmax[k] - vmin[k]);
                if (warpmap) or_warp_sample(warpmap, lin, lin); /* getVoxelPosition, common.glsl:48-50 */
                for (int k = 0; k < 3; ++k) idx[k] = (int)((float)dim * lin[k]);
                if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0 || idx[0] >= dim || idx[1] >= dim || idx[2] >= dim)
                    continue;
                size_t vi = ((size_t)idx[2] * dim + idx[1]) * dim + idx[0];
                __atomic_store_n(&litmask[vi], (uint8_t)1, __ATOMIC_RELAXED); /* every writer stores 1 */
            }
    }
}

void or_normalize_inject(uint32_t *albedo, int dim, const float center[3], const float vmin[3],
                         const float vmax[3], const or_light *lights, int nlights,
                         const float *const *shadowmaps, uint32_t light_mask, float opacity,
                         int storage_fp32, const float *warpmap, float *rad0, uint8_t *litmask, or_stats *stats) {
    size_t nv = (size_t)dim * dim * dim;
    inject_litmask(dim, center, vmin, vmax, lights, nlights, shadowmaps, light_mask, warpmap, litmask);
    /* transferVoxels.comp:37-62 (+ normalizeVoxels.comp:101-105 for rgba16f
     * radiance) and injectRadiance.comp:186,212 per voxel */
    uint32_t uniq = 0, maxf = 0, nlit = 0;
#pragma omp parallel for schedule(static) reduction(+ : uniq, nlit) reduction(max : maxf)
    for (size_t vi = 0; vi < nv; ++vi) {
        uint32_t a32 = albedo[vi];
        float c[4];
        or_unpack_unorm4x8(a32, c);
        float *r = rad0 + vi * 4;
        if (c[3] > 0.0f) {
            uniq += 1;
            uint32_t mf = (uint32_t)(255.0f * c[3]);
            if (mf > maxf) maxf = mf;
            if (opacity > 0.0f) c[3] = opacity;
            albedo[vi] = or_pack_unorm4x8(c);
            int lit = litmask[vi] != 0;
            if (lit) nlit += 1;
            r[0] = lit ? c[0] : 0.0f;
            r[1] = lit ? c[1] : 0.0f;
            r[2] = lit ? c[2] : 0.0f;
            r[3] = c[3];
        } else {
            r[0] = r[1] = r[2] = r[3] = 0.0f;
        }
        if (!storage_fp32)
            for (int k = 0; k < 4; ++k) r[k] = q16(r[k]);
    }
    stats->unique_voxels += uniq;
    if (maxf > stats->max_fragments_per_voxel) stats->max_fragments_per_voxel = maxf;
    stats->lit_voxels += nlit;
}

/* The temporalFilterRadiance path in the reference's own two steps: level 0
 * is not cleared (Application.cpp:762-764); transferVoxels.comp:37-62 with
 * temporalFilter stores mix(previous, (0,0,0,a'), 1 - decay) into every voxel;
 * then injectRadiance.comp:99-101 overwrites every lit voxel with
 * packUnorm4x8(voxelColor[v]), i.e. (albedo, a') or 0 for an empty voxel.
 * prev: last frame's stored level 0, float4 per voxel. */
void or_normalize_inject_temporal(uint32_t *albedo, int dim, const float center[3], const float vmin[3],
                                  const float vmax[3], const or_light *lights, int nlights,
                                  const float *const *shadowmaps, uint32_t light_mask, float opacity,
                                  float decay, int storage_fp32, const float *warpmap, const float *prev,
                                  float *rad0, uint8_t *litmask, or_stats *stats) {
    size_t nv = (size_t)dim * dim * dim;
    inject_litmask(dim, center, vmin, vmax, lights, nlights, shadowmaps, light_mask, warpmap, litmask);
    const float a = 1.0f - decay; /* GLSL mix(x, y, a) = x * (1 - a) + y * a */
    for (size_t vi = 0; vi < nv; ++vi) { /* transfer */
        float c[4];
        or_unpack_unorm4x8(albedo[vi], c);
        if (c[3] > 0.0f) {
            stats->unique_voxels += 1;
            uint32_t mf = (uint32_t)(255.0f * c[3]);
            if (mf > stats->max_fragments_per_voxel) stats->max_fragments_per_voxel = mf;
            if (opacity > 0.0f) c[3] = opacity;
            albedo[vi] = or_pack_unorm4x8(c);
        }
        c[0] = c[1] = c[2] = 0.0f;
        for (int k = 0; k < 4; ++k) rad0[vi * 4 + k] = prev[vi * 4 + k] * (1.0f - a) + c[k] * a;
    }
    for (size_t vi = 0; vi < nv; ++vi) { /* inject: lit voxels take their (rewritten) albedo */
        if (!litmask[vi]) continue;
        float c[4];
        or_unpack_unorm4x8(albedo[vi], c);
        if (c[3] > 0.0f) {
            stats->lit_voxels += 1;
            /* the alpha is a' itself, as in or_normalize_inject: the fp16/fp32
             * radiance keeps it, where the reference's rgba8 radiance holds
             * round(255 a') / 255 */
            if (opacity > 0.0f) c[3] = opacity;
        }
        for (int k = 0; k < 4; ++k) rad0[vi * 4 + k] = c[k];
    }
    if (!storage_fp32)
        for (size_t i = 0; i < nv * 4; ++i) rad0[i] = q16(rad0[i]);
}

/* ------------------------------------------------------------------------ */
/* fill holes (voxelFillHoles.comp:8-36)                                      */
void or_fill_holes(const float *src, float *dst, int dim, int storage_fp32) {
    for (int z = 0; z < dim; ++z)
        for (int y = 0; y < dim; ++y)
            for (int x = 0; x < dim; ++x) {
                size_t vi = ((size_t)z * dim + y) * dim + x;
                float cur[4] = {src[vi * 4 + 0], src[vi * 4 + 1], src[vi * 4 + 2], src[vi * 4 + 3]};
                if (cur[3] == 0.0f) {
                    float count = 0.0f;
                    for (int i = -1; i <= 1; ++i)
                        for (int j = -1; j <= 1; ++j)
                            for (int k = -1; k <= 1; ++k) {
                                int xx = x + i, yy = y + j, zz = z + k;
                                if (xx < 0 || yy < 0 || zz < 0 || xx >= dim || yy >= dim || zz >= dim) continue;
                                const float *v = src + (((size_t)zz * dim + yy) * dim + xx) * 4;
                                if (v[3] > 0.0f) {
                                    for (int q = 0; q < 4; ++q) cur[q] = cur[q] + v[q];
                                    count += 1.0f;
                                }
                            }
                    if (count > 0.0f)
                        for (int q = 0; q < 4; ++q) cur[q] = cur[q] / count;
                }
                for (int q = 0; q < 4; ++q) dst[vi * 4 + q] = storage_fp32 ? cur[q] : q16(cur[q]);
            }
}

/* ------------------------------------------------------------------------ */
/* mip chain (filterRadiance.comp KERNEL_BOX2)                                */
static const int BOX2[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

void or_mip(float **levels, int dim, int nlevels, int storage_fp32) {
    for (int l = 0; l + 1 < nlevels; ++l) {
        int s = dim >> l;
        if (s < 1) s = 1;
        int d = s >> 1;
        if (d < 1) break;
        const float *src = levels[l];
        float *dst = levels[l + 1];
#pragma omp parallel for schedule(static)
        for (int z = 0; z < d; ++z)
            for (int y = 0; y < d; ++y)
                for (int x = 0; x < d; ++x) {
                    float v[4] = {0, 0, 0, 0};
                    for (int o = 0; o < 8; ++o) {
                        int xx = 2 * x + BOX2[o][0], yy = 2 * y + BOX2[o][1], zz = 2 * z + BOX2[o][2];
                        const float *p = src + (((size_t)zz * s + yy) * s + xx) * 4;
                        for (int q = 0; q < 4; ++q) v[q] = v[q] + p[q];
                    }
                    float *o = dst + (((size_t)z * d + y) * d + x) * 4;
                    for (int q = 0; q < 4; ++q) {
                        float r = v[q] * 0.125f;
                        o[q] = storage_fp32 ? r : q16(r);
                    }
                }
    }
}

void or_normalize_voxels(const uint32_t *sums, int dim, float opacity, uint32_t *albedo, uint32_t *normal,
                         float *rad0, or_stats *stats) {
    const size_t nv = (size_t)dim * dim * dim;
    uint32_t uniq = 0, maxc = 0;
    for (size_t v = 0; v < nv; ++v) {
        const uint32_t *w = sums + v * 8;
        float *r = rad0 + v * 4;
        r[0] = r[1] = r[2] = r[3] = 0.0f;
        if (w[3] > 0u) { /* normalizeVoxels.comp:26-35 */
            uniq += 1;
            if (w[3] > maxc) maxc = w[3];
            const float den = 65535.0f * (float)w[3];
            const float c[4] = {(float)w[0] / den, (float)w[1] / den, (float)w[2] / den, opacity > 0.0f ? opacity : 1.0f};
            albedo[v] = or_pack_unorm4x8(c);
            r[3] = c[3];
        }
        if (w[7] > 0u) { /* :38-41 */
            const float den = 65535.0f * (float)w[7];
            const float n[4] = {(float)w[4] / den, (float)w[5] / den, (float)w[6] / den, 1.0f};
            normal[v] = or_pack_unorm4x8(n);
        }
    }
    stats->unique_voxels += uniq;
    if (maxc > stats->max_fragments_per_voxel) stats->max_fragments_per_voxel = maxc;
}

/* The voxelColor chain of the same mipmap pass (src/Application.cpp:904-917):
 * filterRadiance.comp KERNEL_BOX2 with rgba8 image bindings, so each level is
 * packUnorm4x8 of 0.125 * the sum of its children read back as k/255. */
void or_mip_albedo(uint32_t **levels, int dim, int nlevels) {
    for (int l = 0; l + 1 < nlevels; ++l) {
        int s = dim >> l;
        if (s < 1) s = 1;
        int d = s >> 1;
        if (d < 1) break;
        const uint32_t *src = levels[l];
        uint32_t *dst = levels[l + 1];
#pragma omp parallel for schedule(static)
        for (int z = 0; z < d; ++z)
            for (int y = 0; y < d; ++y)
                for (int x = 0; x < d; ++x) {
                    float v[4] = {0, 0, 0, 0};
                    for (int o = 0; o < 8; ++o) {
                        int xx = 2 * x + BOX2[o][0], yy = 2 * y + BOX2[o][1], zz = 2 * z + BOX2[o][2];
                        float t[4];
                        or_unpack_unorm4x8(src[((size_t)zz * s + yy) * s + xx], t);
                        for (int q = 0; q < 4; ++q) v[q] = v[q] + t[q];
                    }
                    for (int q = 0; q < 4; ++q) v[q] = v[q] * 0.125f;
                    dst[((size_t)z * d + y) * d + x] = or_pack_unorm4x8(v);
                }
    }
}

/* 6-direction anisotropic chain.  Face f: axis f/2, sign + (f even) means the
 * light/ray travels toward +axis, so the near child is the lower coordinate.
 * Each output is the mean of the four front-to-back composites
 * near + (1 - near.a) * far along that axis. */
void or_mip_aniso(const float *level0, float ***aniso, int dim, int nlevels, int storage_fp32) {
    for (int l = 0; l + 1 < nlevels; ++l) {
        int s = dim >> l;
        if (s < 1) s = 1;
        int d = s >> 1;
        if (d < 1) break;
        for (int f = 0; f < 6; ++f) {
            const float *src = (l == 0) ? level0 : aniso[l][f];
            float *dst = aniso[l + 1][f];
            int ax = f >> 1, neg = f & 1;
#pragma omp parallel for schedule(static)
            for (int z = 0; z < d; ++z)
                for (int y = 0; y < d; ++y)
                    for (int x = 0; x < d; ++x) {
                        float acc[4] = {0, 0, 0, 0};
                        for (int a = 0; a < 2; ++a)
                            for (int b = 0; b < 2; ++b) {
                                int cn[3], cf[3];
                                int oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;
                                int base[3] = {2 * x, 2 * y, 2 * z};
                                for (int k = 0; k < 3; ++k) cn[k] = cf[k] = base[k];
                                cn[oa] += a;
                                cf[oa] += a;
                                cn[ob] += b;
                                cf[ob] += b;
                                cn[ax] += neg ? 1 : 0;
                                cf[ax] += neg ? 0 : 1;
                                const float *pn = src + (((size_t)cn[2] * s + cn[1]) * s + cn[0]) * 4;
                                const float *pf = src + (((size_t)cf[2] * s + cf[1]) * s + cf[0]) * 4;
                                float t = 1.0f - pn[3];
                                for (int q = 0; q < 4; ++q) acc[q] = acc[q] + (pn[q] + t * pf[q]);
                            }
                        float *o = dst + (((size_t)z * d + y) * d + x) * 4;
                        for (int q = 0; q < 4; ++q) {
                            float r = acc[q] * 0.25f;
                            o[q] = storage_fp32 ? r : q16(r);
                        }
                    }
        }
    }
}

void or_fp8_quantize_level(float *lv, int size) {
    int nb = (size + 3) / 4;
    for (int bz = 0; bz < nb; ++bz)
        for (int by = 0; by < nb; ++by)
            for (int bx = 0; bx < nb; ++bx) {
                float mx = 0.0f;
                for (int z = bz * 4; z < bz * 4 + 4 && z < size; ++z)
                    for (int y = by * 4; y < by * 4 + 4 && y < size; ++y)
                        for (int x = bx * 4; x < bx * 4 + 4 && x < size; ++x)
                            for (int q = 0; q < 4; ++q) {
                                float a = fabsf(lv[(((size_t)z * size + y) * size + x) * 4 + q]);
                                if (a > mx) mx = a;
                            }
                float scale = mx > 0.0f ? mx / 448.0f : 1.0f;
                for (int z = bz * 4; z < bz * 4 + 4 && z < size; ++z)
                    for (int y = by * 4; y < by * 4 + 4 && y < size; ++y)
                        for (int x = bx * 4; x < bx * 4 + 4 && x < size; ++x)
                            for (int q = 0; q < 4; ++q) {
                                float *v = &lv[(((size_t)z * size + y) * size + x) * 4 + q];
                                *v = or_fp8_to_float(or_float_to_fp8(*v / scale)) * scale;
                            }
            }
}

/* ------------------------------------------------------------------------ */
/* G-buffer: visibility (depth pre-pass) + attribute resolve                 */
typedef struct { float c[4]; } cv4;

static int clip_near(const float in[3][4], cv4 out[4]) {
    float d[3];
    for (int i = 0; i < 3; ++i) d[i] = in[i][2] + in[i][3];
    int n = 0;
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        int ii = d[i] >= 0.0f, jj = d[j] >= 0.0f;
        if (ii) {
            memcpy(out[n].c, in[i], 16);
            n++;
        }
        if (ii != jj) {
            float t = d[i] / (d[i] - d[j]);
            for (int k = 0; k < 4; ++k) out[n].c[k] = in[i][k] + t * (in[j][k] - in[i][k]);
            n++;
        }
    }
    return n;
}

void or_gbuffer(const or_scene *sc, const or_camera *cam, or_gbuf_texel *out) {
    int W = cam->width, H = cam->height;
    size_t np = (size_t)W * H;
    uint64_t *vis = (uint64_t *)malloc(np * sizeof(uint64_t));
    for (size_t i = 0; i < np; ++i) vis[i] = ~(uint64_t)0;
    float PV[16];
    or_mul4(cam->projection, cam->view, PV);
    float hw = (float)W * 0.5f, hh = (float)H * 0.5f;
    for (int t = 0; t < sc->ntri; ++t) {
        float clip[3][4];
        for (int v = 0; v < 3; ++v) {
            float w[3];
            world_vertex(sc, t, v, w);
            float w4[4] = {w[0], w[1], w[2], 1.0f};
            or_mat4_vec4(PV, w4, clip[v]);
        }
        cv4 poly[4];
        int n = clip_near(clip, poly);
        for (int k = 1; k + 1 < n; ++k) {
            const cv4 *vv[3] = {&poly[0], &poly[k], &poly[k + 1]};
            float win[3][2], zw[3];
            int bad = 0;
            for (int v = 0; v < 3; ++v) {
                float w = vv[v]->c[3];
                if (!(w > 0.0f)) bad = 1;
                float xd = vv[v]->c[0] / w, yd = vv[v]->c[1] / w, zd = vv[v]->c[2] / w;
                win[v][0] = xd * hw + hw;
                win[v][1] = yd * hh + hh;
                zw[v] = zd * 0.5f + 0.5f;
            }
            if (bad) continue;
            tri_setup s;
            if (!tri_setup_fn(win, W, H, cam->cull_back, &s)) continue;
            for (int py = s.y0; py <= s.y1; ++py)
                for (int px = s.x0; px <= s.x1; ++px) {
                    int64_t cx = (int64_t)px * 256 + 128, cy = (int64_t)py * 256 + 128;
                    if (!inside(&s, cx, cy)) continue;
                    float l[3];
                    bary_at(&s, cx, cy, l);
                    float z = interp(l, zw[0], zw[1], zw[2]);
                    z = z + 0.0f;
                    if (!(z >= 0.0f && z <= 1.0f)) continue;
                    uint32_t zb;
                    memcpy(&zb, &z, 4);
                    uint64_t key = ((uint64_t)zb << 32) | (uint64_t)(0xffffffffu - (uint32_t)t);
                    size_t pi = (size_t)py * W + px;
                    if (key < vis[pi]) vis[pi] = key;
                }
        }
    }
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            size_t pi = (size_t)py * W + px;
            or_gbuf_texel *g = &out[pi];
            memset(g, 0, sizeof(*g));
            if (vis[pi] == ~(uint64_t)0) continue;
            int t = (int)(0xffffffffu - (uint32_t)(vis[pi] & 0xffffffffu));
            float xn = (2.0f * ((float)px + 0.5f)) / (float)W - 1.0f;
            float yn = (2.0f * ((float)py + 0.5f)) / (float)H - 1.0f;
            float Wp[3][3], clip[3][4];
            for (int v = 0; v < 3; ++v) {
                world_vertex(sc, t, v, Wp[v]);
                float w4[4] = {Wp[v][0], Wp[v][1], Wp[v][2], 1.0f};
                or_mat4_vec4(PV, w4, clip[v]);
            }
            float ax[3], ay[3];
            for (int v = 0; v < 3; ++v) {
                ax[v] = clip[v][0] - xn * clip[v][3];
                ay[v] = clip[v][1] - yn * clip[v][3];
            }
            float b[3];
            b[0] = ax[1] * ay[2] - ay[1] * ax[2];
            b[1] = ax[2] * ay[0] - ay[2] * ax[0];
            b[2] = ax[0] * ay[1] - ay[0] * ax[1];
            float sum = (b[0] + b[1]) + b[2];
            for (int v = 0; v < 3; ++v) b[v] = b[v] / sum;
            const float *Nm = sc->nmat + (size_t)sc->obj[t] * 9;
            float Nv[3][3], Tv[3][3], Bv[3][3];
            float Tw[3];
            mat3_vec3(Nm, sc->tan + (size_t)t * 3, Tw);
            for (int v = 0; v < 3; ++v) {
                mat3_vec3(Nm, sc->nrm + (size_t)t * 9 + v * 3, Nv[v]);
                /* phong.vert:48-52 per vertex, with the vertex's tangent when there is one */
                if (sc->vtan) mat3_vec3(Nm, sc->vtan + (size_t)t * 9 + v * 3, Tw);
                float dtn = dot3(Tw, Nv[v]);
                float tt[3] = {Tw[0] - dtn * Nv[v][0], Tw[1] - dtn * Nv[v][1], Tw[2] - dtn * Nv[v][2]};
                nrm3(tt, Tv[v]);
                cross3(Nv[v], Tv[v], Bv[v]);
            }
            for (int k = 0; k < 3; ++k) {
                g->position[k] = interp(b, Wp[0][k], Wp[1][k], Wp[2][k]);
                g->normal[k] = interp(b, Nv[0][k], Nv[1][k], Nv[2][k]);
                g->tangent[k] = interp(b, Tv[0][k], Tv[1][k], Tv[2][k]);
                g->bitangent[k] = interp(b, Bv[0][k], Bv[1][k], Bv[2][k]);
            }
            g->albedo = sc->albedo[t];
            const int tid = tri_tex(sc, t);
            if (tid >= 0) { /* texture(diffuseMap, fragTexcoord) (phong.frag:449), kept as RGBA8 */
                const float *st = sc->uv + (size_t)t * 6;
                float ts = interp(b, st[0], st[2], st[4]), tt = interp(b, st[1], st[3], st[5]);
                float rho2 = 0.0f;
                if (tid < sc->ntex) {
                    /* the unnormalized barycentrics are affine in NDC: their
                     * derivatives give ds/dxn = sum_i db_i/dxn (s_i - s) / sum */
                    float gx[3], gy[3];
                    for (int i = 0; i < 3; ++i) {
                        int j = (i + 1) % 3, k = (i + 2) % 3;
                        gx[i] = clip[j][1] * clip[k][3] - clip[j][3] * clip[k][1];
                        gy[i] = clip[j][3] * clip[k][0] - clip[j][0] * clip[k][3];
                    }
                    float TW = (float)sc->tex[tid].width, TH = (float)sc->tex[tid].height;
                    float kx = 2.0f / (float)W, ky = 2.0f / (float)H;
                    float sx = (gx[0] * (st[0] - ts) + gx[1] * (st[2] - ts)) + gx[2] * (st[4] - ts);
                    float tx = (gx[0] * (st[1] - tt) + gx[1] * (st[3] - tt)) + gx[2] * (st[5] - tt);
                    float sy = (gy[0] * (st[0] - ts) + gy[1] * (st[2] - ts)) + gy[2] * (st[4] - ts);
                    float ty = (gy[0] * (st[1] - tt) + gy[1] * (st[3] - tt)) + gy[2] * (st[5] - tt);
                    float dudx = ((sx / sum) * kx) * TW, dvdx = ((tx / sum) * kx) * TH;
                    float dudy = ((sy / sum) * ky) * TW, dvdy = ((ty / sum) * ky) * TH;
                    float x = dudx * dudx + dvdx * dvdx, y = dudy * dudy + dvdy * dvdy;
                    rho2 = fmaxf(x, y);
                }
                float c[4];
                or_tex_sample(sc->tex, sc->ntex, tid, ts, tt, rho2, c);
                g->albedo = or_pack_unorm4x8(c);
            }
            g->roughness = sc->rough ? sc->rough[sc->obj[t]] : 0.5f;
            g->covered = 1;
            g->triangle = (uint32_t)t;
        }
    free(vis);
}

/* ------------------------------------------------------------------------ */
/* tessellation voxelizer: simpleTesselated.vert + testTesselation.tesc/.tese
 * (Application.cpp:585-654), the alternative the reference runs when
 * voxelizeTesselation is set.  Per triangle (a GL_PATCHES patch of 3):
 *   - world positions and normals as simpleTesselated.vert:16-21;
 *   - .tesc:40-49: the patch is discarded (levels 0) when all three vertices
 *     fall in one voxel, ivec3(dim * (p - center - min) / (max - min));
 *   - .tesc:51-75: IL = max(1, max altitude / voxelSize.x), outer level of an
 *     edge = max(1, length / length(normalize(edge) * voxelSize)), assigned
 *     OL0 = edge b - a, OL1 = edge c - b, OL2 = edge c - a (as written there);
 *   - point-mode triangle tessellation with equal_spacing (GL 4.5 core,
 *     11.2.2.1): each level clamped to [1, 64] and rounded up; an inner level
 *     of 1 with any outer level above 1 becomes 2; levels all 1 give the three
 *     corners; otherwise the corners, the OL_e - 1 interior points of outer
 *     edge e (edge 0: u = 0, 1: v = 0, 2: w = 0), and the concentric inner
 *     triangles k = 1 .. n/2 whose corners are the intersections of the
 *     perpendiculars through the k-th subdivision points of the domain's edges
 *     (in barycentric coordinates (1 - 4k/3n, 2k/3n, 2k/3n) and rotations),
 *     each edge cut into n - 2k segments (n - 2k = 0: the centre point);
 *   - .tese:121-140: position, normal (and texcoord) interpolated with the
 *     tessellation coordinate, voxelStore at ivec3(dim * linear position)
 *     with the triangle's albedo (the diffuse map stands for a constant) and
 *     normalize(n) * 0.5 + 0.5; atomicMax (voxelizeAtomicMax) or the RGBA8
 *     running average; no lighting, no fragment counter.
 * Fixed evaluation order, no contraction. */
int or_tess_levels(const float P[3][3], int dim, const float center[3], const float vmin[3], const float vmax[3],
                   int lv[4]) {
    int vox[3][3];
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) {
            const float tc = ((P[v][k] - center[k]) - vmin[k]) / (vmax[k] - vmin[k]);
            vox[v][k] = (int)((float)dim * tc);
        }
    int same = 1;
    for (int k = 0; k < 3; ++k) same = same && vox[0][k] == vox[1][k] && vox[0][k] == vox[2][k];
    if (same) return 0;
    const float *a = P[0], *b = P[1], *c = P[2];
    float E[3][3]; /* A = c - b, B = c - a, C = b - a */
    for (int k = 0; k < 3; ++k) {
        E[0][k] = c[k] - b[k];
        E[1][k] = c[k] - a[k];
        E[2][k] = b[k] - a[k];
    }
    float len[3];
    for (int i = 0; i < 3; ++i) len[i] = sqrtf(dot3(E[i], E[i]));
    const float s = ((len[0] + len[1]) + len[2]) * 0.5f;
    const float area = sqrtf(((s * (s - len[0])) * (s - len[1])) * (s - len[2]));
    float maxalt = 0.0f;
    for (int i = 0; i < 3; ++i) {
        const float alt = (2.0f * area) / len[i];
        maxalt = i == 0 ? alt : fmaxf(maxalt, alt);
    }
    float vs[3];
    for (int k = 0; k < 3; ++k) vs[k] = (vmax[k] - vmin[k]) / (float)dim;
    float outer[3];
    for (int i = 0; i < 3; ++i) {
        float n[3];
        nrm3(E[i], n);
        const float d[3] = {n[0] * vs[0], n[1] * vs[1], n[2] * vs[2]};
        const float div = fabsf(sqrtf(dot3(d, d)));
        outer[i] = fmaxf(1.0f, len[i] / div);
    }
    const float lev[4] = {fmaxf(1.0f, maxalt / vs[0]), outer[2], outer[0], outer[1]};
    for (int i = 0; i < 4; ++i) {
        const float x = lev[i] > 64.0f ? 64.0f : lev[i]; /* >= 1 already (NaN became 1) */
        lv[i] = (int)ceilf(x);
    }
    const int any_outer = lv[1] > 1 || lv[2] > 1 || lv[3] > 1;
    if (lv[0] == 1 && any_outer) lv[0] = 2;
    return or_tess_count(lv);
}

int or_tess_count(const int lv[4]) {
    const int n = lv[0];
    if (n == 1 && lv[1] == 1 && lv[2] == 1 && lv[3] == 1) return 3;
    int cnt = 3 + (lv[1] - 1) + (lv[2] - 1) + (lv[3] - 1);
    for (int k = 1; 2 * k <= n; ++k) {
        const int m = n - 2 * k;
        cnt += m == 0 ? 1 : 3 * m;
    }
    return cnt;
}

void or_tess_point(const int lv[4], int j, float uvw[3]) {
    static const float corner[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    if (j < 3) {
        for (int k = 0; k < 3; ++k) uvw[k] = corner[j][k];
        return;
    }
    j -= 3;
    for (int e = 0; e < 3; ++e) {
        const int m = lv[1 + e];
        if (j < m - 1) {
            const float f = (float)(j + 1) / (float)m;
            const float g = 1.0f - f;
            if (e == 0) { uvw[0] = 0.0f; uvw[1] = g; uvw[2] = f; }
            else if (e == 1) { uvw[0] = f; uvw[1] = 0.0f; uvw[2] = g; }
            else { uvw[0] = g; uvw[1] = f; uvw[2] = 0.0f; }
            return;
        }
        j -= m - 1;
    }
    const int n = lv[0];
    for (int k = 1; 2 * k <= n; ++k) {
        const int m = n - 2 * k;
        const float cc = (2.0f * (float)k) / (3.0f * (float)n);
        const float oo = 1.0f - 2.0f * cc;
        const float R[3][3] = {{oo, cc, cc}, {cc, oo, cc}, {cc, cc, oo}};
        const int cnt = m == 0 ? 1 : 3 * m;
        if (j < cnt) {
            if (m == 0) {
                for (int q = 0; q < 3; ++q) uvw[q] = R[0][q];
                return;
            }
            const int e = j / m, s = j % m;
            const float f = (float)s / (float)m;
            const float *ra = R[e], *rb = R[(e + 1) % 3];
            for (int q = 0; q < 3; ++q) uvw[q] = ra[q] + f * (rb[q] - ra[q]);
            return;
        }
        j -= cnt;
    }
    uvw[0] = uvw[1] = uvw[2] = 0.0f; /* beyond the count: not reached */
}

void or_voxelize_tess(const or_scene *sc, int dim, const float center[3], const float vmin[3], const float vmax[3],
                      int accum, uint32_t *albedo, uint32_t *normal) {
    /* accum 1 (atomicMax) on all host threads; accum 0 (the running average)
     * in patch order, points in the order of or_tess_point */
#pragma omp parallel for schedule(dynamic, 16) if (accum == 1)
    for (int t = 0; t < sc->ntri; ++t) {
        float P[3][3], N[3][3];
        const float *Nm = sc->nmat + (size_t)sc->obj[t] * 9;
        for (int v = 0; v < 3; ++v) {
            world_vertex(sc, t, v, P[v]);
            mat3_vec3(Nm, sc->nrm + (size_t)t * 9 + v * 3, N[v]);
        }
        int lv[4];
        const int cnt = or_tess_levels(P, dim, center, vmin, vmax, lv);
        float col[4];
        or_unpack_unorm4x8(sc->albedo[t], col);
        const int tid = tri_tex(sc, t);
        for (int j = 0; j < cnt; ++j) {
            float w[3];
            or_tess_point(lv, j, w);
            float p[3], nr[3];
            for (int k = 0; k < 3; ++k) {
                p[k] = (w[0] * P[0][k] + w[1] * P[1][k]) + w[2] * P[2][k];
                nr[k] = (w[0] * N[0][k] + w[1] * N[1][k]) + w[2] * N[2][k];
            }
            if (tid >= 0) { /* texture() in the evaluation shader: zero footprint, level 0 NEAREST */
                const float *st = sc->uv + (size_t)t * 6;
                or_tex_sample(sc->tex, sc->ntex, tid, (w[0] * st[0] + w[1] * st[2]) + w[2] * st[4],
                              (w[0] * st[1] + w[1] * st[3]) + w[2] * st[5], 0.0f, col);
            }
            int idx[3];
            for (int k = 0; k < 3; ++k)
                idx[k] = (int)((float)dim * (((p[k] - center[k]) - vmin[k]) / (vmax[k] - vmin[k])));
            if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0 || idx[0] >= dim || idx[1] >= dim || idx[2] >= dim) continue;
            float nn[3], nenc[4];
            nrm3(nr, nn);
            for (int k = 0; k < 3; ++k) nenc[k] = nn[k] * 0.5f + 0.5f;
            nenc[3] = 1.0f;
            const size_t vi = ((size_t)idx[2] * dim + idx[1]) * dim + idx[0];
            if (accum == 1) {
                atomic_max_u32(&albedo[vi], or_pack_unorm4x8(col));
                atomic_max_u32(&normal[vi], or_pack_unorm4x8(nenc));
            } else {
                const float c3[3] = {col[0], col[1], col[2]};
                rgba8_avg(&albedo[vi], c3);
                rgba8_avg(&normal[vi], nenc);
            }
        }
    }
}
