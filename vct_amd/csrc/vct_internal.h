// vct_internal.h -- structures shared between the host context (context.cpp)
// and the gfx950 kernels, plus the launch wrappers each .hip file exports.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "vct.h"

#define VCT_MAX_LEVELS 10

namespace vctk {
// edge of the warp texture and its occupancy grid (warpDim, src/Application.h:178)
constexpr int kWarpDim = 32;
}  // namespace vctk

namespace vctk {

// Light as the kernels see it: the ABI record plus its shadow map pointer.
constexpr int kMaxLights = 32;  // vct_set_lights

struct DevLight {
    float position[3];
    float range;
    float direction[3];
    float intensity;
    float color[3];
    uint32_t type;
    uint32_t enabled, shadow_caster, shadow_size, perspective;
    float ls[16];
    float ls_inverse[16];
    const float *shadowmap;  // device, shadow_size^2 fp32, row 0 at t = 0
    // the same map as bilinear footprints: element (i, j), i, j in [-1, N-1],
    // at (j + 1) * (N + 1) + (i + 1) holds texels (i, j), (i+1, j), (i, j+1),
    // (i+1, j+1), border texels 1.0 -- one aligned 16-byte load per PCF tap
    const float4 *shadowquad;
};

// Radiance pyramid: one device allocation, addressed by 32-bit byte offsets
// (so a sampler needs a single buffer descriptor for every level and face).
// Level l has edge `size`, nb = ceil(size/4) bricks per axis and `nfaces`
// volumes (1, or 6 for anisotropic levels >= 1: +x,-x,+y,-y,+z,-z).  A brick
// holds 4x4x4 elements x-fastest and the faces of one brick are adjacent.
// Every level carries a border one brick thick on each side (nb + 2 bricks
// per axis) that holds zeros, so that the fp16-pair sampler can fetch texel
// -1 and texel size of any axis without clamping or masking.  With `base` at
// the first interior brick, the element of texel (x, y, z) in face f is
//   (x>>2)*sx + (x&3) + (y>>2)*sy + (y&3)*4 + (z>>2)*sz + (z&3)*16 + 64*f
// with sx = 64*nfaces, sy = (nb+2)*sx, sz = (nb+2)*sy (vct_device.h
// elem_index); the level's storage starts `pad` = sx + sy + sz elements
// before `base`.
// Element formats (fmt):
//   0  fp16 pair, 16 B: texel (x,y,z) and texel (x+1,y,z) as fp16x4 each, so
//      one aligned 16-byte load feeds both x-taps of a trilinear sample.  The
//      writers of texel 0 also write the second half of element -1 (in the
//      border); the second half of element size-1 is texel size, the border
//      (0), and is never written.
//   1  fp32x4, 16 B, one texel.
//   2  fp8 E4M3 x4, 4 B, one texel, times the fp32 scale of its (brick, face):
//      scales[scale + element / 64].
// `stage` is where the mip chain reads and writes the level: `base` for fmt 0
// and 1, an fp16-pair copy for fmt 2 (filtered in fp16, encoded after).
struct LevelInfo {
    uint32_t base, stage;  // byte offsets into the pyramid allocation (first interior brick)
    uint32_t scale;        // float offset into the scale array (fmt 2)
    uint32_t sx, sy, sz;   // brick strides in elements along x, y, z
    uint32_t pad;          // sx + sy + sz: elements from the storage start to `base`
    uint32_t nelem;        // elements of the stored level, border included
    int size, nb, fmt, nfaces;
};

struct PyrDesc {
    int dim, nlevels, storage, aniso;
    char *pyr;       // the allocation (all levels, faces and fp8 stage copies)
    uint32_t bytes;  // its size (< 4 GiB)
    float *scales;   // per-(brick, face) fp8 scales, or null
    LevelInfo lv[VCT_MAX_LEVELS];
    const LevelInfo *d_lv;  // device copy of lv[0..nlevels)
    // One byte per level-0 brick (x fastest, nb^3 rounded up to whole words):
    // bit 0 = every element the injection writes for the brick (its 64
    // elements, and for bricks at x = 0 the second halves of the border
    // elements before them) holds +0 in memory; bit 1 = every word of the
    // same brick of the albedo grid was 0 after the injection; bit 2 = it was
    // 0 before the injection's opacity rewrite, as the voxelizer left it
    // (where the standard voxelizer wrote nothing to either grid).  The
    // injection keeps the bytes; any other writer of level 0 resets them to 0
    // (unknown).
    uint8_t *zero0;
};
// Resets PyrDesc::zero0 to "unknown" (after a write of level 0 other than the injection).
hipError_t launch_zero0_reset(const PyrDesc &pyr, hipStream_t s);

// Grid extent as the kernels use it (VCT::center/min/max).
struct Extent {
    float center[3], vmin[3], vmax[3];
};

// The scene's diffuse maps (vct_set_textures): info[i] = (first level entry,
// width, height, complete | q << 8) with q = floor(log2(max(width, height)));
// levels[e] = (first texel, width, height, 0).
struct DevTex {
    const uint32_t *texels;
    const int4 *info;
    const int4 *levels;
    int count;
};

// Scene geometry on the device (object space, triangle soup).
struct DevScene {
    int ntri, nobj;
    const float *pos;       // [ntri][9]
    const float *nrm;       // [ntri][9]
    const float *tan;       // [ntri][3]
    const float *vtan;      // [ntri][3][3] per-corner (vertex) tangents, or null
    const uint32_t *albedo; // [ntri]
    const uint32_t *obj;    // [ntri]
    const float *rough;     // [nobj] or null
    const float *model;     // [nobj][16]
    const float *nmat;      // [nobj][9]
    const float *uv;        // [ntri][3][2] texture coordinates, or null
    const int32_t *texid;   // [ntri] diffuse map (-1: none), or null
    DevTex tex;
};

// Scratch for the binned rasterizer: per-slot setup records (4 x int4), an
// attribute block of `attr_f4` float4 per slot, per-slot fragment-candidate
// counts (below 2^32: a bounding box is at most 65535 x 65535 pixels), their
// exclusive scan (nslots + 1 entries, last = total) and the first slot of
// every 64-item chunk.
struct RasterScratch {
    int4 *rec;
    float4 *attr;
    uint32_t *counts;
    uint64_t *offs;
    uint32_t *chunk;
    size_t chunk_cap;
    uint64_t *tsum;        // the scan's tile sums and tile prefixes, 2 x 2 x scan_tiles_cap
    size_t scan_tiles_cap;
    size_t cap_slots;
    // the 8x8-block walk of the large triangles (shadow maps, G-buffer): per
    // slot block counts, their scan and its chunk starts (same capacities as above)
    uint32_t *bcounts;
    uint64_t *boffs;
    uint32_t *bchunk;
    uint32_t *bslots;   // the slot of each block, for the first bslots_cap blocks
    size_t bslots_cap;
};

// The diffuse cones' step schedule.  Every diffuse cone of every pixel starts
// at the same initial height with the same aperture, so the sample distance
// h_i of step i and its level class (D_i, fr_i: levels D_i and D_i + 1 blended
// with weights 1 - fr_i, fr_i; D_i = -1 NEAREST on level 0, D_i = q the last
// level alone) depend on i only.  The host computes them once per launch in
// traceCone's own order (phong.frag:145-168), so the kernel reads them from
// its arguments instead of deriving them for every cone of every pixel.
constexpr int kMaxSchedule = 32;
struct ConeSchedule {
    int n;                    // steps covered: the cone's step count, or 0 (compute in the kernel)
    float h[kMaxSchedule];    // h_i, in voxels
    int cls[kMaxSchedule];    // D_i
    float fr[kMaxSchedule];   // fr_i
};

struct TraceLaunch {
    const vct_gbuffer_texel *gbuf;
    int width, height;
    float4 *out_ind, *out_spec;
    int enable_reflections, angle_from_roughness, draw_occlusion, has_roughness;
    float ambient_scale, reflect_scale;
    float eye[3];
    // diffuse / specular cone settings; tan(angle/2) precomputed on the host
    int d_steps;
    float d_tan_half, d_bias, d_h0, d_lod_offset;
    int s_steps;
    float s_angle, s_bias, s_h0, s_lod_offset;
    Extent ext;
    const float4 *warp;  // warpTexture: the 32^3 warp texture, or null
    // 1 when the pyramid was built by injection and filtering (not uploaded):
    // every texel is non-negative, and a filtered texel is zero exactly when
    // all the texels it filters are
    int nested;
    ConeSchedule d_sched;  // the diffuse cones' schedule (n = 0: none)
};

// Statistics are reduced per workgroup into `partials` (kStatPartials slots of
// 4 words) and summed by one small kernel: no contended same-address atomics.
constexpr int kStatPartials = 2048;

// radiance.hip: the static scratch of vct_normalize_inject_noalloc (lit mask of
// grids up to kNoAllocMaxDim^3 + statistics partials), one copy per device.
constexpr int kNoAllocMaxDim = 256;
struct NoAllocScratch {
    uint8_t *litmask;    // kNoAllocMaxDim^3 bytes
    uint32_t *partials;  // kStatPartials * 4 words
};
hipError_t noalloc_scratch_setup(int device);  // with `device` current
NoAllocScratch noalloc_scratch(int device);    // nulls before the setup

// ---- launch wrappers (return hipError_t of the last launch) ----------------
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t n, hipStream_t s);
// sums/maxes partials[0..n) (4 words each: sum, max, sum, sum) into the named stats fields
hipError_t launch_reduce_stats(const uint32_t *partials, int n, uint32_t *sum0, uint32_t *max1, uint32_t *sum2,
                               hipStream_t s);

// scan.hip: counts[0..n) -> offs[0..n] exclusive (offs[n] = total) and chunk
// starts; with `blocks` the same for the block counts (bcounts -> boffs, bchunk)
hipError_t raster_scan(RasterScratch &rs, size_t nslots, bool blocks, hipStream_t s);
// the scan's tiles for nslots slots (rs.tsum holds 4 words per tile)
size_t raster_scan_tiles(size_t nslots);

// raster.hip
// Shadow maps of the lights li[0 .. n): hlights / dlights are the same light
// array on the host / device (map pointers and sizes from there).  The lights
// are drawn shadow_batch(ntri) per walk, so the raster scratch must hold
// min(n, shadow_batch(ntri)) * ntri slots.
constexpr int kShadowBatch = 8;
inline int shadow_batch(int ntri) {  // slot indices stay below 2^31
    const long m = ntri > 0 ? 0x7ffffffeL / ntri : kShadowBatch;
    return m < kShadowBatch ? (int)m : kShadowBatch;
}
hipError_t launch_shadowmaps(const DevScene &sc, const DevLight *hlights, const DevLight *dlights, const int *li,
                             int n, RasterScratch &rs, hipStream_t s);
// rebuilds DevLight::shadowquad from the depth map
hipError_t launch_shadow_quads(const float *depth, int N, float4 *quads, hipStream_t s);
hipError_t launch_gbuffer(const DevScene &sc, const vct_camera &cam, uint64_t *vis, vct_gbuffer_texel *gbuf,
                          RasterScratch &rs, hipStream_t s);

// voxelize.hip
// The clear of both voxel grids (glClearTexImage); zero: PyrDesc::zero0 when
// its bit 2 describes both grids as they are (those bricks are then skipped),
// else null.
hipError_t launch_clear_grids(uint32_t *albedo, uint32_t *normal, int dim, const uint8_t *zero, hipStream_t s);
// The voxelizer; the grids are cleared by the caller, and grids_ready is
// recorded after that clear (the fragment walk waits for it).
hipError_t launch_voxelize(const DevScene &sc, const DevLight *dlights, int nlights, const vct_voxelize_params &p,
                           int dim, uint32_t *albedo, uint32_t *normal, hipEvent_t grids_ready, uint32_t *sums,
                           vct_stats *stats, uint32_t *partials, const float4 *warp, RasterScratch &rs,
                           hipStream_t s);
// zero_grids: as launch_clear_grids' zero
// the tessellation voxelizer (accum: VCT_ACCUM_MAX or VCT_ACCUM_AVG) into the
// brick-ordered albedo / normal grids of a dim^3 volume with the given extent
hipError_t launch_voxelize_tess(const DevScene &sc, const float center[3], const float vmin[3], const float vmax[3],
                                int dim, int accum, uint32_t *albedo, uint32_t *normal, const uint8_t *zero_grids,
                                vct_stats *stats, RasterScratch &rs, hipStream_t s);
// normalizeVoxels.comp for the sum-accumulate mode (sums: 8 words per voxel)
hipError_t launch_normalize_voxels(const uint32_t *sums, uint32_t *albedo, uint32_t *normal, int dim, float opacity,
                                   const PyrDesc &pyr, vct_stats *stats, uint32_t *partials, hipStream_t s);
// the 32^3 occupancy grid of the warp texture (the voxelizer in occupancy mode)
hipError_t launch_warp_occupancy(const DevScene &sc, const vct_voxelize_params &p, uint32_t *occ, RasterScratch &rs,
                                 hipStream_t s);
// warp.hip: prefix sums + the 32^3 warp texture from the occupancy grid
hipError_t launch_warpmap(const uint32_t *occ, int4 *part, const vct_warpmap_params &p, float4 *out, hipStream_t s);

// radiance.hip
hipError_t launch_inject_scatter(const DevLight *dlights_host, const DevLight *dlights_dev, int nlights,
                                 uint32_t light_mask, int dim, Extent ext, const float4 *warp, uint8_t *litmask,
                                 hipStream_t s);
hipError_t launch_normalize_inject(uint32_t *albedo, const uint8_t *litmask, int dim, float opacity, bool temporal,
                                   float decay, const PyrDesc &pyr, vct_stats *stats, uint32_t *partials,
                                   hipStream_t s);
hipError_t launch_fill_holes(const PyrDesc &pyr, void *tmp, hipStream_t s);

// mip.hip
hipError_t launch_mip(const PyrDesc &pyr, hipStream_t s);
// the albedo chain: levels[1 .. nlevels) (linear RGBA8) from the bricked grid
// zero0: PyrDesc::zero0 when the albedo grid has not been written since the last
// injection (its bit 1 then marks the grid's all-zero bricks), else null
hipError_t launch_albedo_chain(const uint32_t *grid, const uint8_t *zero0, uint32_t *const *levels, int dim, int nlevels,
                               hipStream_t s);

// trace.hip
hipError_t launch_trace(const PyrDesc &pyr, const TraceLaunch &t, hipStream_t s);

// compose.hip
hipError_t launch_compose(const vct_gbuffer_texel *gbuf, const float4 *ind, const float4 *spec, float4 *out, int npix,
                          const DevLight *lights, int nlights, int has_rough, int occluded,
                          const vct_compose_params &p, hipStream_t s);

// bricked albedo/normal grid <-> linear x-fastest
hipError_t launch_grid_convert(const uint32_t *src, uint32_t *dst, int dim, bool to_linear, hipStream_t s);

// conversions for download (radiance level -> linear float4)
hipError_t launch_pyr_to_linear(const PyrDesc &pyr, int level, int face, float4 *out, hipStream_t s);
hipError_t launch_linear_to_pyr0(const PyrDesc &pyr, const float4 *in, hipStream_t s);

}  // namespace vctk
