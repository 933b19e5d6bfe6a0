/*
 * vct.h -- C ABI of the MI355X voxel cone tracing passes.
 *
 * The reference (sfreed141/vct) has no plugin or FFI layer: each pass is a GL
 * program bound and dispatched from Application::render
 * (src/Application.cpp:196-1085).  Every entry point below replaces exactly one
 * of those pass invocations (program bind + uniforms + image/texture bindings +
 * glDispatchCompute / scene->draw + glMemoryBarrier).  The replaced call site is
 * cited beside each declaration.
 *
 * Conventions
 *  - All matrices are column-major float[16] (GLM / GLSL layout: m[col*4+row]).
 *  - A context owns every device buffer.  Host pointers passed in are borrowed
 *    for the duration of the call only.
 *  - Every call returns an int status: VCT_OK (0) or a negative VCT_ERR_*.
 *    Nothing aborts or throws across the ABI; vct_last_error() gives the text.
 *  - One host thread per context.  All passes are enqueued in order on the
 *    context's single HIP stream (stream order replaces glMemoryBarrier).
 *    Only vct_download, vct_get_stats, vct_get_pass_times and vct_sync block.
 *  - Parameter validation happens before any HIP call, so invalid arguments
 *    are reported without touching the GPU.
 */
#ifndef VCT_H
#define VCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCT_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define VCT_OK 0
#define VCT_ERR_INVALID (-1)   /* bad argument (null, out of range, size)   */
#define VCT_ERR_HIP (-2)       /* a HIP runtime call failed                 */
#define VCT_ERR_STATE (-3)     /* pass called before its inputs exist       */
#define VCT_ERR_NOMEM (-4)     /* device allocation failed                  */
#define VCT_ERR_NODEVICE (-5)  /* no usable gfx950 device                   */

/* ---- storage formats of the radiance mip chain ---------------------------- */
#define VCT_STORE_FP16 0  /* rgba16f every level (the reference's use_rgba16f)   */
#define VCT_STORE_FP32 1  /* rgba32f every level                                 */
#define VCT_STORE_FP8 2   /* rgba16f levels 0-2, FP8 E4M3 + per-4^3-brick fp32
                              scale on levels >= 3                               */

/* ---- light types (src/Scene.h:13-34, shaders/voxelize.frag:29-30) -------- */
#define VCT_LIGHT_POINT 0u
#define VCT_LIGHT_DIRECTIONAL 1u

/* ---- what vct_download / vct_upload address ------------------------------ */
#define VCT_BUF_ALBEDO 0     /* u32 packed RGBA8 [s^3], x fastest (voxelColor);
                                level 0 the grid, levels >= 1 its mip chain      */
#define VCT_BUF_NORMAL 1     /* u32 packed RGBA8 [dim^3] (voxelNormal)            */
#define VCT_BUF_RADIANCE 2   /* float4 [s^3] of mip `level` (decoded to f32); for
                                 anisotropic levels >= 1 pass l | (face << 4),
                                 face 0..5 = +x,-x,+y,-y,+z,-z                  */
#define VCT_BUF_SHADOWMAP 3  /* float [size^2] depth of light `level`            */
#define VCT_BUF_GBUFFER 4    /* vct_gbuffer_texel [width*height]                 */
#define VCT_BUF_INDIRECT 5   /* float4 [width*height]: (indirect.rgb*occlusion,
                                 occlusion) -- phong.frag:489 debugIndirect view  */
#define VCT_BUF_SPECULAR 6   /* float4 [width*height]: specular cone rgba          */
#define VCT_BUF_LITMASK 7    /* u8 [dim^3]: 1 where a voxel is lit by any light   */
#define VCT_BUF_RADIANCE_RAW 8 /* raw stored bytes of mip `level` (layout-native:   */
                               /* 4^3 bricks with a zero border brick each side)    */
#define VCT_BUF_LABELS 10    /* float4 [npoints]: 64-cone reference (rgb*occ, occ)  */
#define VCT_BUF_FEATURES 11  /* u16 (bf16) [npoints][160]: learned-gather inputs   */
#define VCT_BUF_REFERENCE 12 /* float4 [width*height]: 64-cone reference gather    */
#define VCT_BUF_WEIGHTS 13   /* float [VCT_MLP_PARAMS + 1]: FP32 weights, out scale */
#define VCT_BUF_GRADIENT 14  /* float [VCT_MLP_PARAMS + 1]: last step's minibatch
                                 gradient (sum of the workgroup partials) + loss   */
#define VCT_BUF_WARPMAP 16   /* float4 [32^3] z-major: the warp texture (xyz warped
                                texture coordinate, w packCellInfo), unorm16 values */
#define VCT_BUF_WARP_OCCUPANCY 17 /* u32 [32^3] z-major: the occupancy grid behind it */
#define VCT_BUF_COLOR 15     /* float4 [height][width]: the composed image (rgb, 1)
                                before framebuffer quantization; 0 where uncovered */

/* Grid description: the VCT resource owner (src/Application.h:107-156). */
typedef struct vct_grid_desc {
    int dim;           /* voxelDim: power of two, 4..512                        */
    int levels;        /* voxelLevels: clamped to log2(dim)+1 like VCT::remake  */
    float center[3];   /* VCT::center                                           */
    float vmin[3];     /* VCT::min                                              */
    float vmax[3];     /* VCT::max                                              */
    int storage;       /* VCT_STORE_*                                           */
    int anisotropic;   /* 0: isotropic mips (reference); 1: 6-direction mips    */
    int device;        /* HIP device ordinal                                    */
} vct_grid_desc;

/* Triangle soup in object space.  Replaces the VBO/IBO drawn by Mesh::draw
 * (src/Graphics/Mesh.cpp:268-374) with one actor per object. */
typedef struct vct_mesh {
    int num_tris;
    int num_objects;
    const float *positions;    /* [num_tris][3][3] object-space vertices          */
    const float *normals;      /* [num_tris][3][3] object-space vertex normals     */
    const float *tangents;     /* [num_tris][3]   object-space tangent per triangle
                                  (Mesh.cpp:174-197 rule with uv (0,0),(1,0),(0,1)) */
    const uint32_t *albedo;    /* [num_tris] RGBA8 (R in low byte) = diffuseMap texel */
    const uint32_t *object;    /* [num_tris] actor index in [0, num_objects)       */
    const float *roughness;    /* [num_objects] roughness map value, or NULL (0.5)  */
    const float *vertex_tangents; /* [num_tris][3][3] object-space tangent of each
                                  corner's vertex (Mesh.cpp:190-197: the normalized
                                  sum over the faces sharing it), or NULL: every
                                  corner takes its triangle's `tangents` entry    */
    const float *texcoords;    /* [num_tris][3][2] texture coordinates (s, t) of
                                  each corner (Mesh.cpp:160-162), or NULL         */
    const int32_t *texture;    /* [num_tris] the triangle's diffuse map, an index
                                  into the vct_set_textures table, or -1: its
                                  `albedo` is its colour (material.diffuse, or a
                                  1x1 map); NULL: -1 for every triangle         */
} vct_mesh;

/* One diffuse map (Mesh.cpp:60-81 loads them; ResourceLoader.h:26-108 for DDS),
 * sampled as the reference's sampler state sets it (ResourceLoader.h:79-82,
 * GLHelper.cpp:180-183): LINEAR_MIPMAP_NEAREST minification, NEAREST
 * magnification, REPEAT in s and t; isotropic (the reference asks for the
 * driver's maximum anisotropy, whose footprint is implementation-defined).
 * A texture with fewer than floor(log2(max(width, height))) + 1 levels is
 * mipmap-incomplete and samples as (0, 0, 0, 1), as in GL. */
typedef struct vct_texture {
    int width, height;        /* level 0, each in [1, 32768]                       */
    int levels;               /* levels in `texels`; level l is max(1, width >> l)
                                 x max(1, height >> l)                             */
    const uint32_t *texels;   /* RGBA8 (R in the low byte), the levels one after
                                 another, each row-major from row t = 0 (the first
                                 row glCompressedTexImage2D / glTexSubImage2D take) */
} vct_texture;

/* One light, replacing the std140 SSBO record of Light::writeSSBO
 * (src/Scene.h:13-34, src/Scene.cpp:64-76) plus its light-space matrix
 * (src/Application.cpp:207-210) and shadow-map size. */
typedef struct vct_light {
    float position[3];
    float range;
    float direction[3];
    float intensity;
    float color[3];
    uint32_t type;          /* VCT_LIGHT_POINT | VCT_LIGHT_DIRECTIONAL          */
    uint32_t enabled;
    uint32_t shadow_caster; /* renders a shadow map, shadows voxelize, injects    */
    uint32_t shadow_size;   /* shadow map is shadow_size^2 fp32                   */
    uint32_t perspective;   /* 0: ortho light (w == 1), 1: perspective (w-divide)  */
    float ls[16];           /* light projection * light view                       */
    float ls_inverse[16];   /* inverse(ls), the lsInverse uniform of
                               injectRadiance.comp (src/Application.cpp:804-805)   */
} vct_light;

/* How fragments accumulate into a voxel (vct_voxelize_params.atomic_max). */
#define VCT_ACCUM_AVG 0 /* voxelizeAtomicMax off, the reference's RGBA8 build:
                           imageAtomicRGBA8Avg, a compare-and-swap running
                           average with the fragment count in the alpha byte
                           (voxelize.frag:111-139); the colour depends on the
                           order in which fragments land, the count does not */
#define VCT_ACCUM_MAX 1 /* voxelizeAtomicMax (default): atomicMax of the packed
                           RGBA8 word, order independent (voxelize.frag:271-273) */
#define VCT_ACCUM_SUM 2 /* the fp16 imageAtomicAdd of (c, 1) (voxelize.frag:
                           257-260, an NV extension), here exact 32-bit fixed
                           point sums, order independent; vct_normalize_voxels
                           turns them into the averages                        */

/* Voxelize pass uniforms (src/Application.cpp:689-733). */
typedef struct vct_voxelize_params {
    float mvp_x[16], mvp_y[16], mvp_z[16]; /* dominant-axis projections            */
    int axis_override;      /* -1 = dominant axis, 0/1/2 forces x/y/z             */
    int lighting;           /* voxelizeLighting                                    */
    int atomic_max;         /* voxelizeAtomicMax: VCT_ACCUM_MAX (1, the default),
                               VCT_ACCUM_AVG (0) or VCT_ACCUM_SUM (2)             */
    int msaa;               /* 1: 4x MSAA coverage (ConservativeRasterizeMode MSAA)
                               0: single sample at the pixel centre (OFF)          */
    float eye[3];
    int warp_texture;       /* warpTexture (Application.h:89, off): the voxel of a
                               fragment is dim * texture(warpmap, position)
                               (voxelize.frag:103-105); needs vct_warpmap        */
} vct_voxelize_params;

/* transferVoxels + injectRadiance uniforms (src/Application.cpp:757-837). */
typedef struct vct_inject_params {
    float voxel_set_opacity; /* voxelSetOpacity (Application.h:77), 0.5          */
    uint32_t light_mask;     /* bit i: light i injects (shadow casters only)       */
    /* temporalFilterRadiance / temporalDecay (Application.h:75-76; off, 0.8):
     * radiance level 0 is not cleared and every voxel not lit this frame
     * becomes mix(previous, (0,0,0,a), 1 - decay) (transferVoxels.comp:55-62,
     * Application.cpp:762-764); lit voxels are overwritten with their albedo as
     * without the filter (injectRadiance.comp:99-101: the filtered branch
     * belongs to radianceLighting, which is off).  The previous value is the
     * stored level 0 of the last frame (fp16 or fp32, per the storage format). */
    int temporal_filter;
    float temporal_decay;
    int warp_texture;        /* warpTexture: each shadow texel lands in the voxel of
                                its warped position (getVoxelPosition,
                                common.glsl:48-50); needs vct_warpmap             */
} vct_inject_params;

/* Mip chain (src/Application.cpp:877-921, shaders/filterRadiance.comp). */
typedef struct vct_mip_params {
    int kernel_mode;   /* 0 = KERNEL_BOX2 (the only live mode)                     */
    int albedo_chain;  /* 1: also filter the albedo (voxelColor) chain, as the
                          reference's mipmap pass does (Application.cpp:904-917):
                          RGBA8 levels, each packUnorm4x8 of the BOX2 mean      */
} vct_mip_params;

/* Camera for the G-buffer (depth pre-pass, src/Application.cpp:936-965). */
typedef struct vct_camera {
    int width, height;
    float eye[3];
    float view[16];
    float projection[16];
    int cull_back;   /* GL_CULL_FACE with CCW front faces                          */
} vct_camera;

/* Cone settings: VCTSettings (src/Application.h:28-34). */
typedef struct vct_cone_settings {
    int steps;
    float cone_angle;          /* radians                                          */
    float bias;
    float cone_initial_height;
    float lod_offset;
} vct_cone_settings;

/* Trace pass uniforms (src/Application.cpp:982-1048, shaders/phong.frag:455-512). */
typedef struct vct_trace_params {
    vct_cone_settings diffuse;    /* {16, 60deg, 1.0, 1.0, 0.5}                    */
    vct_cone_settings specular;   /* {32, 30deg, 1.7, 0.5, 0.1}                    */
    int enable_reflections;       /* the specular cone                             */
    int cone_angle_from_roughness;/* vctSpecularConeAngleFromRoughness             */
    int draw_occlusion;           /* drawOcclusion                                 */
    float ambient_scale;          /* ambientScale                                  */
    float reflect_scale;          /* reflectScale                                  */
    float eye[3];
    int warp_texture;             /* warpTexture: every cone sample position goes
                                     through the warp texture (phong.frag:150-152);
                                     needs vct_warpmap                             */
} vct_trace_params;

/* One G-buffer texel, 64 bytes (four float4). */
typedef struct vct_gbuffer_texel {
    float position[3];  uint32_t albedo;     /* world position, RGBA8 albedo       */
    float normal[3];    float roughness;     /* interpolated normal (unnormalized) */
    float tangent[3];   uint32_t covered;    /* TBN column 0; 1 if a surface hit   */
    float bitangent[3]; uint32_t triangle;   /* TBN column 1; visible triangle id  */
} vct_gbuffer_texel;

/* VoxelizeInfo (src/Application.h:194-198). */
typedef struct vct_stats {
    uint32_t total_voxel_fragments;
    uint32_t unique_voxels;
    uint32_t max_fragments_per_voxel;
    uint32_t lit_voxels;
} vct_stats;

/* Per-pass GPU milliseconds of the last frame (GLBufferedTimer brackets,
 * src/Graphics/GLTimer.h:49-87 and Application.cpp:213-1077). */
typedef struct vct_pass_times {
    double shadowmap_ms;
    double voxelize_ms;
    double radiance_ms;   /* normalize + inject                                    */
    double fill_ms;
    double mipmap_ms;
    double gbuffer_ms;
    double render_ms;     /* cone trace                                            */
    double learned_ms;    /* learned gather (when used instead of render)          */
    double compose_ms;    /* direct light + composition                            */
    double warpmap_ms;    /* occupancy voxelization + warp texture                 */
} vct_pass_times;

typedef struct vct_ctx vct_ctx;

/* ---- library ------------------------------------------------------------ */
int vct_abi_version(void);
/* Text of the last error on this context (ctx may be NULL for errors raised
 * by vct_create).  No reference counterpart: GL reports through LOG_ERROR and
 * the debug callback (src/Graphics/GLHelper.cpp:45-59). */
const char *vct_last_error(const vct_ctx *ctx);

/* ---- context lifetime: VCT::VCT + VCT::make (src/Application.h:109-148) --- */
int vct_create(vct_ctx **out, const vct_grid_desc *desc);
/* VCT::remake (src/Application.h:118-129): levels clamped to [1, log2(dim)+1]
 * with a warning in vct_last_error (status stays VCT_OK). */
int vct_remake(vct_ctx *ctx, int dim, int levels);
void vct_destroy(vct_ctx *ctx);
int vct_get_grid(const vct_ctx *ctx, vct_grid_desc *out);
/* VCT::center/min/max assignment, e.g. the voxelTrackCamera snap
 * (src/Application.cpp:187-191). */
int vct_set_extent(vct_ctx *ctx, const float center[3], const float vmin[3],
                   const float vmax[3]);

/* ---- scene: Scene::addActor / Mesh upload / Scene::addLight --------------- */
/* Uploads geometry (src/Scene.cpp:31-36 draws every actor; Mesh.cpp:268-374). */
int vct_set_scene(vct_ctx *ctx, const vct_mesh *mesh);
/* The diffuse maps the scene's triangles index (Mesh::draw binds each
 * drawable's map to unit 0, Mesh.cpp:340): replaces the table.  A triangle
 * whose index is not in the table samples (0, 0, 0, 1), as an unbound unit. */
int vct_set_textures(vct_ctx *ctx, const vct_texture *textures, int count);
/* Per-actor model matrices, the `model` uniform set per actor by Scene::draw
 * (src/Scene.cpp:33; Transform::getMatrix, src/Transform.h:98-104).  The
 * normal matrix mat3(transpose(inverse(model))) of voxelize.vert:18 and
 * phong.vert:37 is derived here once per actor. */
int vct_set_transforms(vct_ctx *ctx, const float *model, int num_objects);
/* Light SSBO at binding 3 (src/Scene.cpp:38-76). */
int vct_set_lights(vct_ctx *ctx, const vct_light *lights, int num_lights);

/* Warp texture parameters (Settings, src/Application.h:80-92). */
typedef struct vct_warpmap_params {
    float high_resolution;   /* warpTextureHighResolution, 2.0: occupied cells     */
    float low_resolution;    /* warpTextureLowResolution, 0.5: lower bound, empty   */
    int axes[3];             /* warpTextureAxes, {1,1,1}: the axes that are warped */
    int linear;              /* warpTextureLinear, 0: 1 = identity map             */
    int use_weights_texture; /* useWarpmapWeightsTexture, 1: the weights pass
                                through an rgba16f texture (fp16-rounded)          */
} vct_warpmap_params;

/* ---- passes (order of Application::render) -------------------------------- */
/* Shadow map of one light: simple.vert + depth only (src/Application.cpp:212-233). */
int vct_shadowmap(vct_ctx *ctx, int light_index);
/* Shadow maps of every shadow-casting light.  The reference draws one map,
 * the main light's (src/Application.cpp:212-233); the multi-light maps repeat
 * that pass per caster, here up to 8 lights per rasterizer walk.  Each map is
 * bit-identical to what vct_shadowmap gives that light. */
int vct_shadowmaps(vct_ctx *ctx);
/* Voxelize, GS path: voxelize.vert/geom/frag drawn by scene->draw
 * (src/Application.cpp:666-754). Clears albedo/normal first (:686-687). */
int vct_voxelize(vct_ctx *ctx, const vct_voxelize_params *p);
/* Voxelize, tessellation path: simpleTesselated.vert + testTesselation.tesc /
 * .tese in point mode with rasterizer discard (src/Application.cpp:585-654,
 * the reference's choice when voxelizeTesselation is set).  Clears albedo and
 * normal (:627-628); patches whose three vertices share a voxel are dropped
 * (.tesc:40-49); every tessellation point stores the triangle's albedo and its
 * normal normalize(n) * 0.5 + 0.5 in the voxel under it.  Of p only
 * atomic_max is read: VCT_ACCUM_MAX (voxelizeAtomicMax) or VCT_ACCUM_AVG (the
 * RGBA8 running average).  No lighting and no fragment count (the stats'
 * total_voxel_fragments is 0); the voxelizeTesselationWarp projection is not
 * provided. */
int vct_voxelize_tessellated(vct_ctx *ctx, const vct_voxelize_params *p);
/* transferVoxels.comp then injectRadiance.comp (src/Application.cpp:757-837),
 * fused: lit-mask scatter over shadow texels, then one per-voxel pass that
 * normalizes, rewrites opacity and writes radiance level 0. */
int vct_normalize_inject(vct_ctx *ctx, const vct_inject_params *p);
/* normalizeVoxels.comp (shaders/normalizeVoxels.comp:19-42), which the
 * reference links but never dispatches: after a VCT_ACCUM_SUM voxelization,
 * every voxel with fragments gets albedo = packUnorm4x8(sum / count, alpha 1
 * or voxel_set_opacity when > 0), normal = packUnorm4x8(normal sum / count),
 * radiance level 0 = (0, 0, 0, alpha); uniqueVoxels and maxFragmentsPerVoxel
 * (the largest count) are updated.  VCT_ERR_STATE unless the last
 * vct_voxelize accumulated sums. */
int vct_normalize_voxels(vct_ctx *ctx, float voxel_set_opacity);
/* vct_normalize_inject without any device allocation of its own.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The same passes
 * and results as vct_normalize_inject; its scratch (the per-voxel lit mask and
 * the statistics partials) is a static device array of the library, one per
 * GPU, 16 MiB + 32 KiB, so the grid's dim must be <= 256 (VCT_ERR_INVALID
 * otherwise).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (contexts on different streams) race.  The lit mask of such a call
 * stays in the scratch (VCT_BUF_LITMASK keeps the last vct_normalize_inject's). */
int vct_normalize_inject_noalloc(vct_ctx *ctx, const vct_inject_params *p);
/* The warp texture (src/Application.cpp:235-575): the scene voxelized into a
 * 32^3 occupancy grid with the projections and coverage mode of `vox`
 * (voxelize.frag:187-193), per-axis prefix sums, the weight table and the
 * 32^3 RGBA16 warpmap of generateWarpmap.frag, all on the GPU (the reference
 * reads the occupancy back to the host).  Run it before the passes that set
 * warp_texture; the reference builds it every frame whether or not it is used. */
int vct_warpmap(vct_ctx *ctx, const vct_voxelize_params *vox, const vct_warpmap_params *p);
/* voxelFillHoles.comp + glCopyImageSubData (src/Application.cpp:839-875). */
int vct_fill_holes(vct_ctx *ctx);
/* filterRadiance.comp loop over levels (src/Application.cpp:877-921). */
int vct_mip(vct_ctx *ctx, const vct_mip_params *p);
/* Depth pre-pass + per-pixel surface attributes consumed by phong.frag
 * (src/Application.cpp:936-965, shaders/phong.vert:36-59). */
int vct_gbuffer(vct_ctx *ctx, const vct_camera *cam);
/* Cone-traced indirect light: traceCone + the 6-cone gather + specular cone
 * (src/Application.cpp:967-1067, shaders/phong.frag:135-180,455-512). */
int vct_trace(vct_ctx *ctx, const vct_trace_params *p);

/* Final image of the phong pass (shaders/phong.frag:183-344,455-520):
 * direct light of every enabled light (Cook-Torrance with GGX, Schlick-GGX and
 * Schlick Fresnel, metallic 0, roughness from the object or 0.5; or
 * Blinn-Phong), each shadow caster's light scaled by 1 - its 5-tap PCF
 * shadow factor, plus (indirect + reflection * reflect_scale) * ambient_scale
 * * albedo from the last vct_trace, times the occlusion when that trace drew
 * it, then Reinhard tone mapping and gamma 2.2.  Needs a vct_trace of the
 * current G-buffer (with reflections when enable_reflections is set).  The
 * reference shades every shadow caster with the main light's map
 * (phong.frag:331); here each light uses its own. */
typedef struct vct_compose_params {
    int cook_torrance;        /* settings.cooktorrance (1); 0: Blinn-Phong      */
    int enable_diffuse;       /* enableDiffuse (1)                              */
    int enable_specular;      /* enableSpecular (1)                             */
    int enable_indirect;      /* enableIndirect (1)                             */
    int enable_reflections;   /* enableReflections (1)                          */
    int postprocess;          /* enablePostprocess (1)                          */
    float ambient_scale;      /* ambientScale (1)                               */
    float reflect_scale;      /* reflectScale (1)                               */
    float shininess;          /* material.shininess (Blinn-Phong only)          */
    float eye[3];             /* camera position                                */
} vct_compose_params;
int vct_compose(vct_ctx *ctx, const vct_compose_params *p);

/* ---- readback, timing ----------------------------------------------------- */
int vct_sync(vct_ctx *ctx);
/* VoxelizeInfo readback (src/Overlay.cpp:106-110). Blocks. */
int vct_get_stats(vct_ctx *ctx, vct_stats *out);
/* Per-pass times of the most recent pass calls (GLBufferedTimer::getTime,
 * src/Graphics/GLTimer.h:80-87). Blocks. Zero when timing is disabled. */
int vct_get_pass_times(vct_ctx *ctx, vct_pass_times *out);
/* 1: bracket every pass with HIP events (default); 0: no events. */
int vct_set_timing(vct_ctx *ctx, int enable);
/* Byte size vct_download/vct_upload expect for (what, level); 0 if invalid. */
size_t vct_buffer_bytes(const vct_ctx *ctx, int what, int level);
/* Copies a context buffer to host memory in the documented linear layout.
 * Blocks. No reference counterpart beyond glGetTextureImage
 * (src/Application.cpp:306-310). */
int vct_download(vct_ctx *ctx, int what, int level, void *host, size_t bytes);
/* Overwrites a context buffer from host memory (test and replay hook:
 * VCT_BUF_ALBEDO, VCT_BUF_NORMAL, VCT_BUF_RADIANCE level 0,
 * VCT_BUF_SHADOWMAP, VCT_BUF_GBUFFER, VCT_BUF_LITMASK, VCT_BUF_WEIGHTS). */
int vct_upload(vct_ctx *ctx, int what, int level, const void *host, size_t bytes);

/* ---- learned diffuse gather ------------------------------------------------
 * No reference counterpart: a small MLP stands in for the 6-cone diffuse loop
 * of phong.frag:480-486, and the plain 6-cone tracer remains the per-8x8-tile
 * fallback.  Network 129 -> 64 -> 64 -> 64 -> 3 (ReLU), BF16 MFMA operands,
 * FP32 accumulation.  Inputs: 27 trilinear samples of mip level 2 on a 3x3x3
 * stencil one level-2 voxel apart, the gather axis TBN*(0,1,0), and
 * sin/cos(pi f p) of the texture-space position for f = 1, 2, 4.  Output: the
 * gather's rgb * occlusion.  Labels: the FP64-accumulated 64-cone reference
 * (64 cosine-weighted 20-degree cones around the same axis, quarter-diameter
 * steps).  Needs an isotropic pyramid with at least 3 levels. */
#define VCT_MLP_PARAMS 16835

typedef struct vct_learned_params {
    int batch;        /* rows per Adam step, a multiple of 64 (4096)              */
    int steps;        /* Adam steps (20000)                                       */
    float lr, beta1, beta2, eps; /* 1e-3, 0.9, 0.999, 1e-8                        */
    uint32_t seed;    /* minibatch sampling and weight initialisation             */
    int reset;        /* 1: He-uniform weights and zero Adam state before training */
    float lr_final;   /* learning rate of the call's last step: the rate follows a
                         cosine from lr to lr_final over the call's steps (0 or
                         lr_final == lr: constant lr)                             */
} vct_learned_params;

/* Training surface points (G-buffer texels, host memory). */
int vct_learned_set_points(vct_ctx *ctx, const vct_gbuffer_texel *points, int npoints);
/* 64-cone labels and network inputs of the points on the current mip chain. */
int vct_learned_prepare(vct_ctx *ctx);
/* Trains on the prepared points; loss (host, [steps]) may be NULL. Blocks.
 * Adam's step count t runs on across calls until reset or a VCT_BUF_WEIGHTS
 * upload (which also zeroes the moments); row j of step t's minibatch is
 * point h(seed ^ h(t * 0x9e3779b9 + j)) % npoints, h the 32-bit mixer
 * x ^= x>>16; x *= 0x7feb352d; x ^= x>>15; x *= 0x846ca68b; x ^= x>>16. */
int vct_learned_train(vct_ctx *ctx, const vct_learned_params *p, float *loss);
/* Learned gather over the current G-buffer into VCT_BUF_INDIRECT: (rgb, 1)
 * from the network, or the 6-cone tracer's (rgb*occ, occ) on tiles whose
 * predicted-luminance variance exceeds variance_threshold (an infinite or NaN
 * threshold sends no tile to the tracer, which is then not launched).
 * fallback_tiles (may be NULL; non-NULL blocks) receives the number of such
 * tiles. */
int vct_trace_learned(vct_ctx *ctx, const vct_trace_params *p, float variance_threshold, int *fallback_tiles);
/* The 64-cone reference of the current G-buffer into VCT_BUF_REFERENCE. */
int vct_reference_gather(vct_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* VCT_H */
