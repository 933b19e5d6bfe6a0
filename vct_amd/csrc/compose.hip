// compose.hip -- the final image of the phong pass (shaders/phong.frag:183-344,
// 455-520): direct light per light (Cook-Torrance or Blinn-Phong) scaled by
// its PCF shadow factor, the traced indirect light and reflection, occlusion,
// and the Reinhard + gamma post-process.  One thread per pixel; the operation
// order is fixed (no FP contraction) and matches the oracle's or_compose, whose
// powf(x, 5) in the Fresnel term is x^4 * x here.
#pragma clang fp contract(off)

#include "vct_device.h"

namespace vctk {

namespace {

constexpr float kPi = 3.1415982f;  // PI of shaders/common.glsl:1, as the shader has it

__device__ __forceinline__ float pow2(float x) { return x * x; }

// albedo / PI: the IEEE quotient by the constant's FP64 reciprocal
__device__ __forceinline__ float div_pi(float x) { return quot_f32(x, 1.0 / (double)kPi); }

// cook_torrance (phong.frag:229-253) with metallic 0: diffuse and specular
__device__ __forceinline__ void cook_torrance(const float lambert[3], const float col[3], const float N[3],
                                              const float V[3], const float L[3], const float H[3], float rough,
                                              float dif[3], float spe[3]) {
    const float n_dot_l = fmaxf(0.0f, dot3(N, L));
    const float n_dot_v = fmaxf(0.0f, dot3(N, V));
    const float n_dot_h = fmaxf(0.0f, dot3(N, H));
    const float h_dot_v = fmaxf(0.0f, dot3(H, V));
    // D_ggxtr (phong.frag:524-528)
    const float a2 = pow2(rough);
    const float D = a2 / (kPi * pow2(pow2(n_dot_h) * (a2 - 1.0f) + 1.0f));
    // G_schlickggx (phong.frag:531-545)
    const float k = pow2(rough + 1.0f) / 8.0f;
    const float G = (n_dot_v / (n_dot_v * (1.0f - k) + k)) * (n_dot_l / (n_dot_l * (1.0f - k) + k));
    // F0 = mix(0.04, albedo, metallic = 0); schlicks (phong.frag:547-549)
    const float F0 = 0.04f;
    // pow(x, 5) as x^4 * x: within 2 ulp of the power (the C library's powf
    // and a GPU's pow are themselves approximations)
    const float u = 1.0f - h_dot_v;
    const float F = F0 + (1.0f - F0) * (pow2(pow2(u)) * u);
    const float ct = ((D * G) * F) / fmaxf((4.0f * n_dot_l) * n_dot_v, 0.001f);
    const float kd = 1.0f - F;
    for (int c = 0; c < 3; ++c) {
        dif[c] = ((col[c] * n_dot_l) * kd) * lambert[c];
        spe[c] = (col[c] * n_dot_l) * ct;
    }
}

__device__ __forceinline__ float smoothstep_gl(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}

}  // namespace

struct ComposeArgs {
    const vct_gbuffer_texel *gbuf;
    const float4 *ind, *spec;
    float4 *out;
    int npix;
    const DevLight *lights;
    int nlights;
    int has_rough;  // per-object roughness present (the material's roughness map)
    int occluded;   // the trace multiplied the indirect rgb by the occlusion
    vct_compose_params p;
};

__global__ __launch_bounds__(256) void k_compose(ComposeArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npix) return;
    const float4 *g = (const float4 *)&a.gbuf[i];
    const float4 g0 = g[0], g1 = g[1], g2 = g[2];
    if (__float_as_uint(g2.w) == 0u) {
        a.out[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float P[3] = {g0.x, g0.y, g0.z};
    const uint32_t ab = __float_as_uint(g0.w);
    const float alb[3] = {unorm8(ab, 0), unorm8(ab, 1), unorm8(ab, 2)};
    const float nraw[3] = {g1.x, g1.y, g1.z};
    float N[3];
    nrm3(nraw, N);
    const float rough = a.has_rough ? g1.w : 0.5f;  // phong.frag:306-309
    const float ve[3] = {a.p.eye[0] - P[0], a.p.eye[1] - P[1], a.p.eye[2] - P[2]};
    float V[3];
    nrm3(ve, V);
    float dif[3] = {0.f, 0.f, 0.f}, spe[3] = {0.f, 0.f, 0.f};
    const float lambert[3] = {div_pi(alb[0]), div_pi(alb[1]), div_pi(alb[2])};  // the same in every light
    for (int l = 0; l < a.nlights; ++l) {
        const DevLight &Lt = a.lights[l];
        if (!Lt.enabled) continue;
        float Ld[3], att = 1.0f;
        if (Lt.type == 0u) {  // point light (phong.frag:268-288)
            const float d[3] = {Lt.position[0] - P[0], Lt.position[1] - P[1], Lt.position[2] - P[2]};
            const float dist = sqrtf(dot3(d, d));
            if (dist > Lt.range) continue;
            att = 1.0f - smoothstep_gl(0.75f * Lt.range, Lt.range, dist);
            nrm3(d, Ld);
        } else {  // directional (phong.frag:290-302)
            const float nd[3] = {-Lt.direction[0], -Lt.direction[1], -Lt.direction[2]};
            nrm3(nd, Ld);
        }
        const float vl[3] = {V[0] + Ld[0], V[1] + Ld[1], V[2] + Ld[2]};
        float H[3];
        nrm3(vl, H);
        float ld[3], ls[3];
        if (a.p.cook_torrance) {
            cook_torrance(lambert, Lt.color, N, V, Ld, H, rough, ld, ls);
            if (Lt.type == 0u)
                for (int c = 0; c < 3; ++c) {
                    ld[c] = ld[c] * att;
                    ls[c] = ls[c] * att;
                }
        } else {  // calculateDiffuse / calculateSpecular (phong.frag:255-262)
            const float kdf = fmaxf(dot3(N, Ld), 0.0f);
            const float ksp = powf(fmaxf(dot3(N, H), 0.0f), a.p.shininess);
            for (int c = 0; c < 3; ++c) {
                ld[c] = (((kdf * att) * Lt.color[c]) * Lt.intensity) * alb[c];
                ls[c] = (((ksp * att) * Lt.color[c]) * Lt.intensity) * alb[c];
            }
        }
        if (Lt.shadow_caster && Lt.shadowquad) {
            float w4[4] = {P[0], P[1], P[2], 1.0f}, lsp[4];
            mat4_vec4(Lt.ls, w4, lsp);
            const float sf = 1.0f - calc_shadow_factor(lsp, Lt.shadowquad, (int)Lt.shadow_size);
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
    if (!a.p.enable_diffuse) dif[0] = dif[1] = dif[2] = 0.0f;
    if (!a.p.enable_specular) spe[0] = spe[1] = spe[2] = 0.0f;
    float col[3];
    if (a.p.enable_indirect) {
        // the trace stores (indirect.rgb * m, occlusion) with m = occlusion when
        // it drew the occlusion: ((indirect + reflect * rs) * as * albedo + direct) * m
        const float4 ind = a.ind[i];
        const float m = a.occluded ? ind.w : 1.0f;
        float refl[3] = {0.f, 0.f, 0.f};
        if (a.p.enable_reflections) {
            const float4 sp = a.spec[i];
            refl[0] = (sp.x * a.p.reflect_scale) * m;
            refl[1] = (sp.y * a.p.reflect_scale) * m;
            refl[2] = (sp.z * a.p.reflect_scale) * m;
        }
        const float indc[3] = {ind.x, ind.y, ind.z};
        for (int c = 0; c < 3; ++c)
            col[c] = (indc[c] + refl[c]) * (a.p.ambient_scale * alb[c]) + (dif[c] + spe[c]) * m;
    } else {
        for (int c = 0; c < 3; ++c) col[c] = (a.p.ambient_scale * alb[c] + dif[c]) + spe[c];
    }
    if (a.p.postprocess) {  // postprocess (phong.frag:209-218)
        for (int c = 0; c < 3; ++c) {
            const float t = col[c] / (col[c] + 1.0f);
            col[c] = powf(t, 1.0f / 2.2f);
        }
    }
    a.out[i] = make_float4(col[0], col[1], col[2], 1.0f);
}

hipError_t launch_compose(const vct_gbuffer_texel *gbuf, const float4 *ind, const float4 *spec, float4 *out, int npix,
                          const DevLight *lights, int nlights, int has_rough, int occluded,
                          const vct_compose_params &p, hipStream_t s) {
    ComposeArgs a;
    a.gbuf = gbuf;
    a.ind = ind;
    a.spec = spec;
    a.out = out;
    a.npix = npix;
    a.lights = lights;
    a.nlights = nlights;
    a.has_rough = has_rough;
    a.occluded = occluded;
    a.p = p;
    hipLaunchKernelGGL(k_compose, dim3((npix + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace vctk
