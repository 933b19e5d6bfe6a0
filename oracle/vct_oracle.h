/*
 * vct_oracle.h -- CPU restatement of the reference's voxel cone tracing passes.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the parity tests,
 * the smoke check and the benchmark's cpu_baseline leg compare the HIP path
 * against.  Nothing in the product (vct_amd/, include/) may link, load or call
 * it; the product fails loudly when its HIP library is missing instead.
 *
 * Every function restates a reference shader or host routine; the citation is
 * given beside each declaration (paths relative to the reference root).  The
 * float arithmetic is written in a fixed operation order without fused
 * multiply-adds so that the HIP kernels, which use the same order, can be
 * compared bit for bit where the north-star asks for it (occupancy, the
 * packed RGBA8 accumulators, shadow depth, radiance level 0, the mip chain).
 *
 * Pinning status (what checks each function independently of this file):
 *   or_ortho/or_lookat/or_perspective/or_inverse4/or_mul4   closed forms
 *   or_voxel_axis_projections                               brute-force mapping to
 *                                                           (p-c-min)/(max-min)
 *   or_shadowmap, or_voxelize                               closed forms on
 *       axis-aligned boxes, shuffle invariance, popcount == unique voxels,
 *       exact rational raster on tiny inputs (tests/test_oracle_raster.py)
 *   or_normalize_inject                                     per-voxel closed form,
 *       brute-force per-texel unprojection, idempotence
 *   or_normalize_inject_temporal                            decay 0 = the plain pass,
 *       decay 1 keeps unlit voxels, geometric decay on a static scene,
 *       lit voxels = the plain pass
 *   or_fill_holes                                           brute force 3x3x3
 *   or_mip                                                  constant/ramp closed
 *       forms, mean preservation, numpy float16 rounding
 *   or_mip_albedo                                           round(255 * mean of the
 *       eight children's bytes / 255) per level in float64 (exact ties listed)
 *   or_trace / or_trace64                                   empty volume -> 0,
 *       uniform volume closed form alpha_n = 1-(1-a)^n; LINEAR_MIPMAP_LINEAR
 *       level blend and the cone's lambda schedule on pyramids whose levels
 *       (and anisotropic faces) are uniform with distinct values
 *   or_voxelize accum 0 / 2, or_normalize_voxels            a quad whose voxels get
 *       one fragment each (average = that fragment), two coincident quads of
 *       known colours (running average and sum closed forms), the alpha byte
 *       = fragment count
 *   voxelize lighting (inside or_voxelize)                  closed forms on a quad:
 *       directional N.(-d), point-light smoothstep band, shadow * (1 - PCF)
 *   or_fp8_quantize_level                                   scale = brick max / 448,
 *       nearest E4M3 code by table search, per-brick independence
 *   or_half / or_fp8                                        numpy float16; OCP
 *       E4M3 code table round trip
 *   or_gbuffer                                              closed form plane hits
 *   or_tex_sample (diffuse maps, inside or_voxelize, or_gbuffer and
 *   or_voxelize_tess)                                       NEAREST / LINEAR texel
 *       and weight closed forms, REPEAT wrap, the LINEAR_MIPMAP_NEAREST level at
 *       footprints on both sides of each threshold rho = 2^(k-1/2), and the
 *       footprint of affine and perspective uv maps against finite differences;
 *       incomplete textures (0,0,0,1)
 *   or_compose                                              closed forms: N=L=V
 *       Cook-Torrance and Blinn-Phong values, point-light range cut and
 *       smoothstep, shadow on/off, Reinhard+gamma, indirect/occlusion algebra
 *   The reference ships no tests, golden vectors or fixtures for any of these,
 *   so none is pinned by reference data; each is pinned by the checks above.
 */
#ifndef VCT_ORACLE_H
#define VCT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- GLM-convention matrices (column-major) ----------------- */
void or_identity(float m[16]);
void or_mul4(const float a[16], const float b[16], float out[16]);        /* a*b */
void or_ortho(float l, float r, float b, float t, float n, float f, float out[16]);
void or_lookat(const float eye[3], const float center[3], const float up[3], float out[16]);
void or_perspective(float fovy, float aspect, float n, float f, float out[16]);
int  or_inverse4(const float m[16], float out[16]);
void or_normal_matrix(const float model[16], float out9[9]);
void or_mat4_vec4(const float m[16], const float v[4], float out[4]);
/* mvp_x/y/z of src/Application.cpp:689-697 */
void or_voxel_axis_projections(const float center[3], const float vmin[3], const float vmax[3],
                               float mvp_x[16], float mvp_y[16], float mvp_z[16]);

/* ---------------- storage formats --------------------------------------- */
uint16_t or_float_to_half(float f);          /* IEEE binary16, round to nearest even */
float    or_half_to_float(uint16_t h);
uint8_t  or_float_to_fp8(float f);           /* OCP FP8 E4M3, RNE, saturating to 448 */
float    or_fp8_to_float(uint8_t v);
uint32_t or_pack_unorm4x8(const float c[4]); /* GLSL packUnorm4x8, RNE               */
void     or_unpack_unorm4x8(uint32_t v, float out[4]);

/* ---------------- scene input ------------------------------------------ */
typedef struct or_scene {
    int ntri, nobj;
    const float *pos;       /* [ntri][3][3] object space          */
    const float *nrm;       /* [ntri][3][3] object space          */
    const float *tan;       /* [ntri][3]                          */
    const uint32_t *albedo; /* [ntri] RGBA8                       */
    const uint32_t *obj;    /* [ntri]                             */
    const float *rough;     /* [nobj] or NULL                     */
    const float *model;     /* [nobj][16]                         */
    const float *nmat;      /* [nobj][9] normal matrices          */
    const float *vtan;      /* [ntri][3][3] per-corner (vertex) tangents
                               (Mesh.cpp:190-197), or NULL: `tan` */
    const float *uv;        /* [ntri][3][2] texture coordinates, or NULL     */
    const int32_t *texid;   /* [ntri] diffuse map, -1 none; or NULL          */
    const struct or_texture *tex; /* the diffuse maps                        */
    int ntex;
} or_scene;

/* A diffuse map: RGBA8 levels back to back, level l max(1, w >> l) x
 * max(1, h >> l), row-major from row t = 0. */
typedef struct or_texture {
    int width, height, levels;
    const uint32_t *texels;
} or_texture;

/* texture(diffuseMap, (s, t)) with the reference's sampler state
 * (ResourceLoader.h:79-82, GLHelper.cpp:180-183) by GL 4.5 core 8.14:
 * rho2 = rho^2 of the footprint in level-0 texels; lambda <= 0: NEAREST on
 * level 0; else LINEAR_MIPMAP_NEAREST (level ceil(lambda + 1/2) - 1, 0 for
 * lambda <= 1/2, clamped); REPEAT; byte / 255; an incomplete texture or an
 * index outside the table gives (0, 0, 0, 1). */
void or_tex_sample(const or_texture *tex, int ntex, int id, float s, float t, float rho2, float out[4]);
/* rho^2 of an affine (s, t) over a window-space triangle (orthographic views):
 * st = (s0, t0, s1, t1, s2, t2), W x H level 0's size. */
float or_tex_rho2_affine(const float win[3][2], const float st[6], float W, float H);

typedef struct or_light {
    float position[3]; float range;
    float direction[3]; float intensity;
    float color[3]; uint32_t type;
    uint32_t enabled, shadow_caster, shadow_size, perspective;
    float ls[16];
    float ls_inverse[16];
} or_light;

/* Shadow map of one light (src/Application.cpp:212-233; simple.vert).
 * Depth cleared to 1.0, LESS test, single sample at pixel centres. */
void or_shadowmap(const or_scene *sc, const or_light *L, float *depth /* [size^2] */);

typedef struct or_voxelize_params {
    const float *mvp_x, *mvp_y, *mvp_z;
    int axis_override, lighting, msaa;
    int dim;
    const float *warpmap; /* warpTexture: the 32^3 warp texture (or_warpmap), or NULL */
    int accum;            /* VCT_ACCUM_*: 1 atomicMax, 0 the RGBA8 running average
                             (fragments in triangle order, pixel rows in order),
                             2 fixed-point sums into `sums`                       */
    uint32_t *sums;       /* accum 2: 8 words per voxel (linear x-fastest), zeroed
                             by the caller                                        */
} or_voxelize_params;

typedef struct or_stats {
    uint32_t total_voxel_fragments, unique_voxels, max_fragments_per_voxel, lit_voxels;
} or_stats;

/* voxelize.vert/geom/frag, GS path, atomicMax mode (shaders/voxelize.*;
 * src/Application.cpp:666-754).  albedo/normal must be zeroed by the caller
 * (glClearTexImage, :686-687).  shadowmaps[i] belongs to lights[i]. */
void or_voxelize(const or_scene *sc, const or_light *lights, int nlights,
                 const float *const *shadowmaps, const or_voxelize_params *p,
                 uint32_t *albedo, uint32_t *normal, or_stats *stats);

/* transferVoxels.comp + injectRadiance.comp, fp16 radiance
 * (shaders/transferVoxels.comp:29-70, normalizeVoxels.comp:89-106,
 * injectRadiance.comp:40-103).  albedo is updated in place (alpha rewrite),
 * rad0 receives float4 radiance already rounded to the storage precision
 * (storage_fp32 = 1: no rounding; else fp16). litmask [dim^3] (1 = lit) is written. */
void or_normalize_inject(uint32_t *albedo, int dim, const float center[3],
                         const float vmin[3], const float vmax[3],
                         const or_light *lights, int nlights, const float *const *shadowmaps,
                         uint32_t light_mask, float opacity, int storage_fp32,
                         const float *warpmap, float *rad0, uint8_t *litmask, or_stats *stats);

/* The same with temporalFilterRadiance (Application.h:75-76): mix with the
 * previous level 0 `prev` (float4 per voxel, as stored), then lit voxels
 * overwritten (transferVoxels.comp:55-62, injectRadiance.comp:99-101). */
void or_normalize_inject_temporal(uint32_t *albedo, int dim, const float center[3],
                                  const float vmin[3], const float vmax[3],
                                  const or_light *lights, int nlights, const float *const *shadowmaps,
                                  uint32_t light_mask, float opacity, float decay, int storage_fp32,
                                  const float *warpmap, const float *prev, float *rad0, uint8_t *litmask,
                                  or_stats *stats);

/* Warp texture (src/Application.cpp:235-575): occupancy voxelization at
 * 32^3 (voxelize.frag:187-193: the voxelizer's fragments OR 1 into the cell of
 * their unwarped position), then per texel centre the warped coordinate of
 * generateWarpmap.frag:49-99 from the inclusive per-axis prefix sums
 * (Application.cpp:312-344) and the low/high weights per occupied-cell count
 * (:346-370; through the rgba16f weights texture when use_weights_texture),
 * stored as RGBA16 unorm (k / 65535).  w holds packCellInfo. */
#define OR_WARP_DIM 32
typedef struct or_warp_params {
    float high_resolution, low_resolution; /* warpTextureHigh/LowResolution 2.0, 0.5 */
    int axes[3];                           /* warpTextureAxes                       */
    int linear;                            /* warpTextureLinear                     */
    int use_weights_texture;               /* useWarpmapWeightsTexture              */
} or_warp_params;
void or_warp_occupancy(const or_scene *sc, const or_voxelize_params *p /* dim 32 */, uint32_t *occ);

/* tessellation voxelizer (simpleTesselated.vert, testTesselation.tesc/.tese):
 * or_tess_levels -> the patch's point count (0: discarded) and lv = (inner,
 * outer 0, 1, 2) after clamping and rounding; or_tess_point -> barycentric
 * (u, v, w) of point j < count; or_voxelize_tess -> albedo/normal grids
 * (linear x-fastest, zeroed by the caller), accum 1 atomicMax, 0 running avg */
int or_tess_levels(const float P[3][3], int dim, const float center[3], const float vmin[3], const float vmax[3],
                   int lv[4]);
int or_tess_count(const int lv[4]);
void or_tess_point(const int lv[4], int j, float uvw[3]);
void or_voxelize_tess(const or_scene *sc, int dim, const float center[3], const float vmin[3], const float vmax[3],
                      int accum, uint32_t *albedo, uint32_t *normal);
void or_warpmap(const uint32_t *occ, const or_warp_params *p, float *out /* [32^3][4] */);
/* texture(warpmap, tc).xyz: LINEAR, CLAMP_TO_EDGE */
void or_warp_sample(const float *warpmap, const float tc[3], float out[3]);

/* voxelFillHoles.comp (shaders/voxelFillHoles.comp:8-36) on a float4 grid whose
 * values are quantised by `quant` (0 = fp16 rounding, 1 = none). */
void or_fill_holes(const float *src, float *dst, int dim, int storage_fp32);

/* filterRadiance.comp KERNEL_BOX2 chain (shaders/filterRadiance.comp:14-65;
 * src/Application.cpp:885-919).  levels[0] is the input; levels[l] for
 * l = 1..nlevels-1 are written (size max(1, dim>>l)^3 float4), each rounded to
 * fp16 unless storage_fp32.  For anisotropic chains, aniso[l][f] (l >= 1,
 * f = 0..5 for +x,-x,+y,-y,+z,-z) receive the six directional volumes. */
void or_mip(float **levels, int dim, int nlevels, int storage_fp32);
void or_mip_aniso(const float *level0, float ***aniso, int dim, int nlevels, int storage_fp32);
/* The albedo (voxelColor) chain of the same pass (src/Application.cpp:904-917):
 * levels[0] is the RGBA8 grid (linear, x fastest), levels[l] (l >= 1) receive
 * packUnorm4x8 of the BOX2 mean of the unorm8 children. */
void or_mip_albedo(uint32_t **levels, int dim, int nlevels);
/* normalizeVoxels.comp:19-42 on the fixed-point sums of an accum-2
 * voxelization: albedo/normal (linear, zeroed by the caller) receive
 * packUnorm4x8 of the averages (colour alpha 1, or opacity when > 0), rad0
 * (float4 per voxel) (0, 0, 0, alpha) where there are fragments, 0 elsewhere;
 * stats->unique_voxels and max_fragments_per_voxel (the largest count). */
void or_normalize_voxels(const uint32_t *sums, int dim, float opacity, uint32_t *albedo, uint32_t *normal,
                         float *rad0, or_stats *stats);
/* FP8 E4M3 + per-4^3-brick fp32 scale round trip of one level (values
 * replaced in place by their decoded storage). */
void or_fp8_quantize_level(float *level, int size);

/* G-buffer (depth pre-pass, src/Application.cpp:936-965; phong.vert:36-59). */
typedef struct or_gbuf_texel {
    float position[3]; uint32_t albedo;
    float normal[3]; float roughness;
    float tangent[3]; uint32_t covered;
    float bitangent[3]; uint32_t triangle;
} or_gbuf_texel;

typedef struct or_camera {
    int width, height;
    float eye[3];
    float view[16], projection[16];
    int cull_back;
} or_camera;

void or_gbuffer(const or_scene *sc, const or_camera *cam, or_gbuf_texel *out);

typedef struct or_cone {
    int steps; float angle, bias, h0, lod_offset;
} or_cone;

typedef struct or_trace_params {
    or_cone diffuse, specular;
    int enable_reflections, angle_from_roughness, draw_occlusion;
    int has_roughness;
    float ambient_scale, reflect_scale;
    float eye[3];
    int dim, nlevels;
    float center[3], vmin[3], vmax[3];
    int anisotropic;
    const float *warpmap; /* warpTexture: the 32^3 warp texture (or_warpmap), or NULL */
} or_trace_params;

/* Cone gather of phong.frag (shaders/phong.frag:135-180, 455-512).
 * levels[l]: float4 volume of size max(1,dim>>l)^3 (already storage-rounded).
 * aniso: NULL for isotropic; else aniso[l][f] for l >= 1.
 * out_indirect: float4 (indirect.rgb*occlusion, occlusion) per pixel;
 * out_specular: float4 specular cone (may be NULL). */
void or_trace(const or_gbuf_texel *g, int npix, float *const *levels, float *const *const *aniso,
              const or_trace_params *p, float *out_indirect, float *out_specular);
/* Same integral with every operation in double precision (the FP64 oracle). */
void or_trace64(const or_gbuf_texel *g, int npix, float *const *levels, float *const *const *aniso,
                const or_trace_params *p, double *out_indirect, double *out_specular);

/* One cone, exposed for closed-form tests. pos/normal/dir in texture space. */
void or_trace_cone(float *const *levels, float *const *const *aniso, int dim, int nlevels,
                   const float pos[3], const float nrm[3], const float dir[3],
                   const or_cone *c, float out[4]);

/* Trilinear LINEAR_MIPMAP_LINEAR sample with border 0 and lambda clamp. */
void or_sample(float *const *levels, int dim, int nlevels, const float p[3], float lambda,
               float out[4]);

/* FP64 64-cone reference of the diffuse gather (labels of the learned gather;
 * see or_label.c for the exact definition).  out: [npix][4] = (rgb *
 * occlusion, occlusion).  Isotropic pyramids only. */
void or_label_dirs(double dirs[64][3], double w[64]);
void or_label_cone64(float *const *levels, int dim, int nlevels, const double pos[3], const double nrm[3],
                     const double dir[3], double tan_half, double out[4]);
void or_label64(const or_gbuf_texel *g, int npix, float *const *levels, int dim, int nlevels,
                const float center[3], const float vmin[3], const float vmax[3], double *out);

/* calcShadowFactor (shaders/voxelize.frag:160-184 = phong.frag:183-207):
 * fraction of the 5 PCF taps in shadow. lsp: light clip-space position. */
float or_calc_shadow_factor(const float lsp[4], const float *sm, int N);

/* Final composition of phong.frag (shaders/phong.frag:209-344, 486-518):
 * per-light Cook-Torrance (metallic 0) or Blinn-Phong direct light with
 * per-light PCF shadows, combined with the traced indirect (as or_trace
 * writes it) and reflection, then Reinhard + gamma 2.2.  out: float4 per
 * pixel (rgb, 1); uncovered pixels are 0.  shadowmaps[l] belongs to
 * lights[l] (NULL for non-casters).  Layout equals vct_compose_params. */
typedef struct or_compose_params {
    int cook_torrance, enable_diffuse, enable_specular, enable_indirect;
    int enable_reflections, postprocess;
    float ambient_scale, reflect_scale, shininess;
    float eye[3];
} or_compose_params;
void or_compose(const or_gbuf_texel *g, int npix, const float *indirect, const float *specular,
                const or_light *lights, int nlights, const float *const *shadowmaps, int has_roughness,
                int occluded, const or_compose_params *p, float *out);

#ifdef __cplusplus
}
#endif
#endif
