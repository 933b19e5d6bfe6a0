/*
 * vct_app.h -- headless application layer above the pass ABI (vct.h).
 *
 * Replaces the part of the reference's Application (src/Application.{h,cpp})
 * that feeds the voxel cone tracing passes: the Settings/VCTSettings defaults
 * (Application.h:28-103), the light and voxel matrices (Application.cpp:
 * 200-210, 689-692), the scene (here: seeded synthetic scenes instead of
 * Sponza, Application.cpp:93-136) and the per-frame pass order of
 * Application::render (Application.cpp:196-1085).  It calls the GPU only
 * through vct.h, as the reference's Application calls GL.
 */
#ifndef VCT_APP_H
#define VCT_APP_H

#include "vct.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vct_app vct_app;

/* Workload description.  Fields left at 0 (or -1 for the flags) take the
 * defaults of the chosen scene:
 *  scene 1: room [-1,1]^3 + 2 seeded boxes (36 tris), 64^3 x 7 levels,
 *           1 directional light with a 512^2 shadow map, 256x256 camera,
 *           6 diffuse cones.
 *  scene 2: 2000 boxes + 64 icospheres (subdivision 4), 351,680 tris, moving
 *           rigidly every frame, 256^3 x 8 levels with 6-direction mips,
 *           1 directional + 4 point lights with 1024^2 shadow maps,
 *           1920x1080 camera, 6 diffuse + 1 specular cone.
 *  scene 5: scene 2's distributions at 4x the extent: 8000 boxes + 512
 *           icospheres, 2,717,440 tris in [-4,4]^3 (positions x4, sizes x2.5:
 *           the same volume fraction as scene 2), moving rigidly, 256^3 x 8
 *           levels with 6-direction mips, FP8 E4M3 storage on levels >= 3,
 *           the warp texture at high resolution 4.0 (voxelize, inject and
 *           trace through it), 1 directional + 7 point lights with 1024^2
 *           shadow maps, 3840x2160 camera, 6 diffuse + 1 specular cone. */
typedef struct vct_app_config {
    int scene;          /* 1, 2 or 5                                            */
    uint64_t seed;      /* 0 = scene index                                      */
    int dim, levels;    /* voxel grid                                           */
    int storage;        /* VCT_STORE_*, -1 = default (FP16)                     */
    int anisotropic;    /* -1 = default                                         */
    int width, height;  /* camera                                               */
    int shadow_size;    /* per light                                            */
    int lighting;       /* voxelizeLighting, -1 = default (1)                   */
    int msaa;           /* 4x MSAA voxelization, -1 = default (1)               */
    int reflections;    /* specular cone, -1 = default                          */
    int fill_holes;     /* voxelFillHoles, -1 = default (0, Application.h:95)   */
    int device;         /* HIP device; -1 builds the scene and uniforms only
                           (no GPU context: render calls fail)                */
    int temporal_filter; /* temporalFilterRadiance (Application.h:75), 0       */
    float temporal_decay;/* temporalDecay (Application.h:76), 0.8              */
    int track_camera;    /* voxelTrackCamera (Application.h:86), 0: the grid
                            centre follows the eye in steps of
                            2^levels * (max - min) / dim (Application.cpp:187-191) */
    int warp_texture;    /* warpTexture (Application.h:89), 0: every frame builds the
                            warp texture (vct_warpmap) and voxelize, inject and
                            trace go through it                                */
    float warp_high_resolution; /* warpTextureHighResolution (Application.h:87):
                            2.0 (0 = default); 4.0, the overlay's maximum
                            (Overlay.cpp:289), for scene 5                     */
    int tessellation;    /* voxelizeTesselation (Application.h:100), 0: 1 draws
                            the voxels with the tessellation voxelizer
                            (vct_voxelize_tessellated, Application.cpp:585-654)
                            instead of the GS path                             */
} vct_app_config;

/* ---- DDS textures: ResourceLoader::loadDDS (src/ResourceLoader.h:26-108) ----------
 * "DDS " + the 124-byte header, then DXT1 / DXT3 / DXT5 (S3TC) blocks, level
 * after level: level l is (width >> l) x (height >> l), ((w + 3) / 4) *
 * ((h + 3) / 4) blocks, for l < dwMipMapCount while a dimension is non-zero.
 * Decoded by EXT_texture_compression_s3tc into RGBA8 (each channel the exact
 * value rounded to the nearest byte; DXT1 opaque, as COMPRESSED_RGB_S3TC_DXT1;
 * DXT3/DXT5 colour blocks always in four-colour mode).  Other four-character
 * codes are rejected, as the reference returns 0 for them.  The reference's
 * loop uploads a level with a zero dimension as an empty image, which leaves
 * the texture incomplete; here the level list stops there and the shorter
 * chain is incomplete in the same way. */
typedef struct vct_image vct_image;
int vct_dds_load(const char *path, vct_image **out, char *err, int errlen);
/* a view of the decoded levels (valid until vct_dds_free) */
int vct_dds_texture(const vct_image *img, vct_texture *out);
void vct_dds_free(vct_image *img);

/* ---- other images: GLHelper::createTextureFromImage (GLHelper.cpp:165-210) ---------
 * A PNG file (every PNG colour type and depth, interlaced or not) as stb_image
 * returns it with its own channel count, uploaded at GL's default unpack
 * alignment of 4 as GL_R8 (1 channel: samples (r, 0, 0, 1)), GL_RGB8 or
 * GL_RGBA8 (2 channels get no storage: an incomplete texture, no levels), and
 * floor(log2(max(w, h))) + 1 levels by a 2x2 box (glGenerateTextureMipmap's
 * filter is the implementation's).  The image is viewed and freed with
 * vct_dds_texture and vct_dds_free.  JPEG, TGA and the other stb_image
 * formats are not decoded. */
int vct_png_load(const char *path, vct_image **out, char *err, int errlen);

/* ---- OBJ/MTL assets: Mesh::loadMesh (src/Graphics/Mesh.cpp:42-266) ------------
 * Faces triangulated (fans), texture v flipped, per-face tangents from the uv
 * deltas (Mesh.cpp:174-189), albedo = the material's Kd as RGBA8 (white
 * without a material), roughness = the material's Pr standing in for a
 * constant roughness map.  A material's map_Kd that names a .dds or .png file
 * is loaded (vct_dds_load / vct_png_load, path relative to the .obj) and its
 * faces carry their corner uvs and that texture; other image formats are not
 * decoded and those faces keep Kd.  A face without normals gets its face
 * normal, one without usable uvs a tangent orthogonal to it.  The file is one
 * object (one actor). */
typedef struct vct_obj vct_obj;
int vct_obj_load(const char *path, vct_obj **out, char *err, int errlen);
/* views into the loaded arrays (valid until vct_obj_free) */
int vct_obj_mesh(const vct_obj *obj, vct_mesh *out);
int vct_obj_bounds(const vct_obj *obj, float bmin[3], float bmax[3], int *num_materials);
/* the diffuse maps vct_obj_mesh's texture indices refer to */
int vct_obj_textures(const vct_obj *obj, const vct_texture **textures, int *count);
void vct_obj_free(vct_obj *obj);

int vct_app_default_config(int scene, vct_app_config *out);
/* An application whose scene is an OBJ file (`cfg->scene` is ignored): the
 * mesh as one static actor, the grid extent a cube around its bounds (5%
 * margin), one shadow-casting directional light from above and a camera
 * outside the bounds looking at their centre. */
int vct_app_create_from_obj(vct_app **out, const vct_app_config *cfg, const char *path);
int vct_app_create(vct_app **out, const vct_app_config *cfg);
void vct_app_destroy(vct_app *app);
const char *vct_app_last_error(const vct_app *app);
vct_ctx *vct_app_ctx(vct_app *app);
int vct_app_get_config(const vct_app *app, vct_app_config *out);

/* Scene sizes and the object-space geometry the app uploaded. */
int vct_app_scene_sizes(const vct_app *app, int *num_tris, int *num_objects, int *num_lights);
/* Any pointer may be NULL.  Returns 1 if the scene has per-object roughness. */
int vct_app_get_mesh(const vct_app *app, float *pos, float *nrm, float *tan, uint32_t *albedo, uint32_t *obj,
                     float *rough);
/* The scene's per-corner (vertex) tangents, [num_tris][3][3], when it has
 * them (OBJ scenes): returns 1 and fills vtan (may be NULL), else 0. */
int vct_app_get_vertex_tangents(const vct_app *app, float *vtan);

/* Everything the passes of frame `frame` receive (uniforms and matrices);
 * any pointer may be NULL. model: [num_objects][16]; lights: [num_lights]. */
int vct_app_frame_state(vct_app *app, int frame, float *model, vct_light *lights, vct_voxelize_params *vox,
                        vct_inject_params *inj, vct_mip_params *mip, vct_camera *cam, vct_trace_params *trace);

/* `n` seeded surface points of frame `frame` as G-buffer texels (training
 * points of the learned gather): triangles picked with probability
 * proportional to their world-space area, uniform barycentrics, attributes
 * interpolated exactly as the G-buffer resolve interpolates them. */
int vct_app_surface_points(vct_app *app, int frame, int n, uint64_t seed, vct_gbuffer_texel *out);

/* Another seeded camera of the scene's camera distribution (the rule that
 * placed the scene's own camera: an eye outside every object, a random yaw
 * and a small pitch) at width x height.  The learned gather trains on the
 * views of some seeds and is scored on the views of others. */
int vct_app_camera(const vct_app *app, uint64_t seed, int width, int height, vct_camera *out);

/* The composition uniforms of frame `frame` (Settings defaults of
 * src/Application.h:60-71: Cook-Torrance, diffuse, specular, indirect,
 * reflections as configured, post-processing, ambient and reflect scale 1). */
int vct_app_compose_params(vct_app *app, int frame, vct_compose_params *out);
/* The warp texture settings (Settings, src/Application.h:80-92). */
int vct_app_warpmap_params(const vct_app *app, vct_warpmap_params *out);

/* One frame in the order of Application::render: transforms, shadow maps,
 * voxelize, normalize+inject, [fill holes], mip, G-buffer, trace, compose. */
int vct_app_render(vct_app *app, int frame);
/* The subset of a frame up to and including the mip chain. */
int vct_app_build_radiance(vct_app *app, int frame);

/* GLM-convention host math (column-major), exported for tests. */
void vct_math_ortho(float l, float r, float b, float t, float n, float f, float out[16]);
void vct_math_lookat(const float eye[3], const float center[3], const float up[3], float out[16]);
void vct_math_perspective(float fovy, float aspect, float n, float f, float out[16]);
int vct_math_inverse4(const float m[16], float out[16]);
void vct_math_mul4(const float a[16], const float b[16], float out[16]);
void vct_math_voxel_projections(const float center[3], const float vmin[3], const float vmax[3], float mvp_x[16],
                                float mvp_y[16], float mvp_z[16]);

#ifdef __cplusplus
}
#endif

#endif /* VCT_APP_H */
