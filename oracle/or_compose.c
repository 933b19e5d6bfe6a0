/*
 * or_compose.c -- CPU restatement of the final composition of phong.frag
 * (shaders/phong.frag:209-344 for the direct lighting and post-process,
 * :486-518 for how indirect light, reflection and occlusion are combined,
 * :521-549 for the Cook-Torrance terms).  TEST INFRASTRUCTURE ONLY: the HIP
 * compose pass is checked against this file, nothing in the product uses it.
 *
 * Operation order is fixed (no FP contraction) and is the one the HIP kernel
 * uses.  powf and sqrtf are the C library's; they may differ from the GPU's
 * by an ulp, so parity of this pass is stated as a relative tolerance.
 */
#include "vct_oracle.h"

#include <math.h>

static const float PI_SHADER = 3.1415982f; /* PI of shaders/common.glsl:1, as written there */

static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static void nrm3(const float v[3], float o[3]) {
    float l = sqrtf(dot3(v, v));
    o[0] = v[0] / l;
    o[1] = v[1] / l;
    o[2] = v[2] / l;
}
static float sq(float x) { return x * x; }
static float clampf(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }
static float unorm8(uint32_t v, int i) { return (float)((v >> (8 * i)) & 0xffu) / 255.0f; }

/* GLSL smoothstep */
static float smoothstep_gl(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}

/* cook_torrance with metallic = 0 (phong.frag:229-253; D_ggxtr :524-528,
 * G_schlickggx :531-545 with k_direct = (r+1)^2/8, schlicks :547-549). */
static void cook_torrance(const float alb[3], const float col[3], const float N[3], const float V[3],
                          const float L[3], const float H[3], float rough, float dif[3], float spe[3]) {
    float ndl = fmaxf(0.0f, dot3(N, L));
    float ndv = fmaxf(0.0f, dot3(N, V));
    float ndh = fmaxf(0.0f, dot3(N, H));
    float hdv = fmaxf(0.0f, dot3(H, V));
    float a2 = sq(rough);
    float D = a2 / (PI_SHADER * sq(sq(ndh) * (a2 - 1.0f) + 1.0f));
    float k = sq(rough + 1.0f) / 8.0f;
    float G = (ndv / (ndv * (1.0f - k) + k)) * (ndl / (ndl * (1.0f - k) + k));
    float F = 0.04f + (1.0f - 0.04f) * powf(1.0f - hdv, 5.0f);
    float ct = ((D * G) * F) / fmaxf((4.0f * ndl) * ndv, 0.001f);
    float kd = 1.0f - F;
    for (int c = 0; c < 3; ++c) {
        float lambert = alb[c] / PI_SHADER;
        dif[c] = ((col[c] * ndl) * kd) * lambert;
        spe[c] = (col[c] * ndl) * ct;
    }
}

void or_compose(const or_gbuf_texel *g, int npix, const float *indirect, const float *specular,
                const or_light *lights, int nlights, const float *const *shadowmaps, int has_roughness,
                int occluded, const or_compose_params *p, float *out) {
    for (int i = 0; i < npix; ++i) {
        const or_gbuf_texel *t = &g[i];
        float *o = out + (size_t)i * 4;
        if (!t->covered) {
            o[0] = o[1] = o[2] = o[3] = 0.0f;
            continue;
        }
        const float *P = t->position;
        float alb[3] = {unorm8(t->albedo, 0), unorm8(t->albedo, 1), unorm8(t->albedo, 2)};
        float N[3], V[3];
        nrm3(t->normal, N);
        /* roughness .5 unless the material has a roughness map (phong.frag:306-309) */
        float rough = has_roughness ? t->roughness : 0.5f;
        float ve[3] = {p->eye[0] - P[0], p->eye[1] - P[1], p->eye[2] - P[2]};
        nrm3(ve, V);
        float dif[3] = {0, 0, 0}, spe[3] = {0, 0, 0};
        /* calculateDirectLighting (phong.frag:304-344) */
        for (int l = 0; l < nlights; ++l) {
            const or_light *L = &lights[l];
            if (!L->enabled) continue;
            float Ld[3], att = 1.0f;
            if (L->type == 0u) { /* calculatePointLight (phong.frag:269-288) */
                float d[3] = {L->position[0] - P[0], L->position[1] - P[1], L->position[2] - P[2]};
                float dist = sqrtf(dot3(d, d));
                if (dist > L->range) continue;
                att = 1.0f - smoothstep_gl(0.75f * L->range, L->range, dist); /* :264-267 */
                nrm3(d, Ld);
            } else { /* calculateDirectionalLight (phong.frag:290-302) */
                float nd[3] = {-L->direction[0], -L->direction[1], -L->direction[2]};
                nrm3(nd, Ld);
            }
            float vl[3] = {V[0] + Ld[0], V[1] + Ld[1], V[2] + Ld[2]}, H[3];
            nrm3(vl, H);
            float ld[3], ls[3];
            if (p->cook_torrance) {
                cook_torrance(alb, L->color, N, V, Ld, H, rough, ld, ls);
                if (L->type == 0u)
                    for (int c = 0; c < 3; ++c) {
                        ld[c] = ld[c] * att;
                        ls[c] = ls[c] * att;
                    }
            } else { /* calculateDiffuse / calculateSpecular (phong.frag:255-262) */
                float kdf = fmaxf(dot3(N, Ld), 0.0f);
                float ksp = powf(fmaxf(dot3(N, H), 0.0f), p->shininess);
                for (int c = 0; c < 3; ++c) {
                    ld[c] = (((kdf * att) * L->color[c]) * L->intensity) * alb[c];
                    ls[c] = (((ksp * att) * L->color[c]) * L->intensity) * alb[c];
                }
            }
            /* every shadow caster uses its own map (the shader has one, :328-333) */
            if (L->shadow_caster && shadowmaps && shadowmaps[l]) {
                float w4[4] = {P[0], P[1], P[2], 1.0f}, lsp[4];
                or_mat4_vec4(L->ls, w4, lsp);
                float sf = 1.0f - or_calc_shadow_factor(lsp, shadowmaps[l], (int)L->shadow_size);
                for (int c = 0; c < 3; ++c) {
                    ld[c] = ld[c] * sf;
                    ls[c] = ls[c] * sf;
                }
            }
            for (int c = 0; c < 3; ++c) {
                dif[c] = dif[c] + ld[c];
                spe[c] = spe[c] + ls[c];
            }
        }
        if (!p->enable_diffuse) dif[0] = dif[1] = dif[2] = 0.0f;
        if (!p->enable_specular) spe[0] = spe[1] = spe[2] = 0.0f;
        float col[3];
        if (p->enable_indirect) {
            /* indirect holds (indirect.rgb * m, occlusion) with m the occlusion
             * when the trace drew it: ((ind + refl*rs) * as * albedo + direct) * m
             * (phong.frag:486-515) with m distributed over the terms. */
            const float *ind = indirect + (size_t)i * 4;
            float m = occluded ? ind[3] : 1.0f;
            float refl[3] = {0, 0, 0};
            if (p->enable_reflections) {
                const float *sp = specular + (size_t)i * 4;
                for (int c = 0; c < 3; ++c) refl[c] = (sp[c] * p->reflect_scale) * m;
            }
            for (int c = 0; c < 3; ++c) col[c] = (ind[c] + refl[c]) * (p->ambient_scale * alb[c]) + (dif[c] + spe[c]) * m;
        } else {
            for (int c = 0; c < 3; ++c) col[c] = (p->ambient_scale * alb[c] + dif[c]) + spe[c];
        }
        if (p->postprocess) /* postprocess (phong.frag:209-218) */
            for (int c = 0; c < 3; ++c) col[c] = powf(col[c] / (col[c] + 1.0f), 1.0f / 2.2f);
        o[0] = col[0];
        o[1] = col[1];
        o[2] = col[2];
        o[3] = 1.0f;
    }
}
